"""CPU oracle for the thermochemistry and flamelet hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in ``pyrometheus_amd`` imports this
package; only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` may use it, and only as the checker (or the timed CPU
baseline), never as a fallback for the HIP path.

Contents:

* :mod:`oracle.thermochem_np` -- numpy restatement of the ``Thermochemistry``
  class that ``pyrometheus/codegen/python.py:106-642`` generates from the
  expression builders of ``pyrometheus/chem_expr.py``, one whole-array numpy
  expression per generated term (the reference's op structure).  Every
  function is dtype-generic so complex-step derivatives run through it.
* :mod:`oracle.flamelet_np` -- numpy restatement of
  ``pyrometheus/flamelets`` (Laplacian, residual, block Jacobian,
  ``block_thomas``, the Newton / Crank-Nicolson driver).

How each part is pinned (what a plausible mistake would fail):

* NASA-7 thermo, molecular weights and the T-from-e Newton: the reference's
  stored autoignition trajectory (``examples/free_flame/output/
  autoignition.h5``), whose internal energy is constant along every row.
* Rate coefficients, equilibrium constants and net production rates (the
  sandiego mechanism): re-integrating the stored RK4 windows of that same
  trajectory reproduces the next stored state.
* Closed forms and invariants: single-reaction hand values, mass and element
  conservation, creation - destruction = net, detailed balance.
* Derivatives: complex-step against central finite differences.
* block_thomas: dense ``numpy.linalg.solve`` of the assembled matrix, and the
  scalar Thomas algorithm for nv = 1.
"""
