"""Mixture-fraction grids (``pyrometheus/flamelets/domain.py:5-86``).

``Domain`` keeps the reference's uniform grid (``x``, ``dx``, ``jac=[dx]``)
and adds an optional interior clustering about ``z_c`` (Vinokur/Roberts
transform; Anderson, *Computational Fluid Dynamics*, Sec. 5.6).  The 3-point
Laplacian on either grid is exposed as per-row coefficients ``(a, b, c)``
with identity Dirichlet rows scaled by ``1/h_end^2`` -- on a uniform grid
exactly the reference's ``1/dx^2`` stencil (``onedim_laplacian.py:115-212``).
"""

from dataclasses import dataclass
from typing import Optional

import numpy as np


@dataclass
class DomainConfig:
    num_x: int
    x_l: float
    x_r: float
    dtype: np.dtype = np.float64
    num_dim: int = 1
    cluster_center: Optional[float] = None
    cluster_beta: float = 3.0


def clustered_points(num_x, z_c, beta=3.0, x_l=0.0, x_r=1.0):
    s = np.linspace(0.0, 1.0, num_x)
    zc = (z_c - x_l) / (x_r - x_l)
    A = np.log((1 + (np.exp(beta) - 1) * zc)
               / (1 + (np.exp(-beta) - 1) * zc)) / (2 * beta)
    u = zc * (1 + np.sinh(beta * (s - A)) / np.sinh(beta * A))
    u[0], u[-1] = 0.0, 1.0
    return x_l + (x_r - x_l) * u


def laplacian_coefficients(x):
    n = len(x)
    h = np.diff(x)
    a, b, c = np.zeros(n), np.zeros(n), np.zeros(n)
    hm, hp = h[:-1], h[1:]
    a[1:-1] = 2.0 / (hm * (hm + hp))
    c[1:-1] = 2.0 / (hp * (hm + hp))
    b[1:-1] = -(a[1:-1] + c[1:-1])
    b[0] = 1.0 / h[0] ** 2
    b[-1] = 1.0 / h[-1] ** 2
    return a, b, c


class Domain:
    def __init__(self, domain_config: DomainConfig):
        cfg = domain_config
        self.dtype = cfg.dtype
        self.num_dim = cfg.num_dim
        self.num_x = cfg.num_x
        self.dx = (cfg.x_r - cfg.x_l) / (self.num_x - 1)
        if cfg.cluster_center is None:
            self.x = np.linspace(cfg.x_l, cfg.x_r, self.num_x, endpoint=True)
        else:
            self.x = clustered_points(self.num_x, cfg.cluster_center,
                                      cfg.cluster_beta, cfg.x_l, cfg.x_r)
        self.jac = np.array([self.dx])
        self.laplacian = laplacian_coefficients(self.x)

    @property
    def boundary_dissipation(self):
        """chi values at the two ends that turn the Dirichlet rows into
        identity rows (2 dx^2 on a uniform grid)."""
        b = self.laplacian[1]
        return 2.0 / b[0], 2.0 / b[-1]
