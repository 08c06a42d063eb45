"""Boundary diagnostics used as error metrics of the stepper (host, NumPy).

* :func:`order_parameter` -- O = (func_Lz(Z1))(:,:,end), the boundary condensate
  (``tevol_driven_cluster.m:136,148``; ``scripts/analysis/movie_show.m:91``)
* :func:`countN`          -- periodic plaquette winding count with the reference's 1e-4
  tolerance (``src/diagnostics/countN.m:7-21``)
"""

from __future__ import annotations

import math

import numpy as np


def order_parameter(grid, Z1):
    """Boundary order parameter of ``Z1`` [nx, ny, nz] on ``grid`` (an ``ic.Grid``)."""
    return np.tensordot(np.asarray(Z1), grid.Lz[-1, :], axes=([2], [0]))


def countN(x1):
    """``[n, np, nm] = countN(x1)``: number of plaquettes with winding +1 or -1, and each sign."""
    x1 = np.asarray(x1)
    x2 = np.roll(x1, -1, axis=0)
    x3 = np.roll(x2, -1, axis=1)
    x4 = np.roll(x1, -1, axis=1)
    with np.errstate(divide="ignore", invalid="ignore"):
        w = (np.angle(x1 / x2) + np.angle(x2 / x3) + np.angle(x3 / x4) + np.angle(x4 / x1)) / (2 * math.pi)
    lp = (w >= 1 - 1e-4) & (w <= 1 + 1e-4)
    lm = (w <= -1 + 1e-4) & (w >= -1 - 1e-4)
    return int(np.count_nonzero(lp | lm)), int(np.count_nonzero(lp)), int(np.count_nonzero(lm))


def vortex_positions(x1):
    """Plaquette indices (i, j, charge) of the windings counted by :func:`countN`."""
    x1 = np.asarray(x1)
    x2 = np.roll(x1, -1, axis=0)
    x3 = np.roll(x2, -1, axis=1)
    x4 = np.roll(x1, -1, axis=1)
    with np.errstate(divide="ignore", invalid="ignore"):
        w = (np.angle(x1 / x2) + np.angle(x2 / x3) + np.angle(x3 / x4) + np.angle(x4 / x1)) / (2 * math.pi)
    out = []
    for q in (1, -1):
        ii, jj = np.where(np.abs(w - q) <= 1e-4)
        out += [(int(i), int(j), q) for i, j in zip(ii, jj)]
    return sorted(out)
