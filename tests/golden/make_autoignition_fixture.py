"""Extract the reference's stored autoignition trajectory into a fixture.

Source: ``examples/free_flame/output/autoignition.h5`` in the reference tree
(written by ``examples/python_driver.py:16-67``: constant-volume 9-species
H2/air ignition from 1200 K, stoichiometric, 1 atm; RK4 with 10 steps per
stored window for the ``explicit`` group, Crank-Nicolson for ``implicit``).
The HDF5 file stores its datasets contiguously as little-endian float64, so
they are read at their byte offsets without an HDF5 library:

=================  =========  ==========
dataset            shape      offset (B)
=================  =========  ==========
explicit/time      (1000,)    2432
explicit/state     (1000,10)  12480
implicit/time      (100,)     94528
implicit/state     (100,10)   95328
=================  =========  ==========

State columns: mass fractions of H2, H, O2, O, OH, HO2, H2O2, H2O, N2, then T.
Run from the repository root: ``python tests/golden/make_autoignition_fixture.py``.
"""
import os
import numpy as np

SRC = "/root/reference/examples/free_flame/output/autoignition.h5"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                   "autoignition.npz")


def main():
    raw = open(SRC, "rb").read()
    assert len(raw) == 103328, len(raw)

    def arr(off, shape):
        n = int(np.prod(shape))
        return np.frombuffer(raw[off:off + 8 * n], "<f8").reshape(shape).copy()

    np.savez(OUT,
             explicit_time=arr(2432, (1000,)),
             explicit_state=arr(12480, (1000, 10)),
             implicit_time=arr(94528, (100,)),
             implicit_state=arr(95328, (100, 10)))
    print("wrote", OUT)


if __name__ == "__main__":
    main()
