"""Initial fields for the stepper (host set-up, untimed).

Restates ``tevol_driven_cluster.m:81-102`` (stationary profile -> not-a-knot spline to the
run's radial grid -> gauge rotation -> perturbation) and ``tevol_lattice.m:101-135`` (vortex
imprint), with the perturbations of the benchmark workloads: seeded Gaussian phase noise and
seeded +-1 windings.  The radial seed is the reference's own stationary profile
``data/reference/mu=6,nz=30,vy=7_20.mat``, shipped as ``data/mu6_nz30_kk7.json``.
"""

from __future__ import annotations

import json
import math
import os

import numpy as np

from .api import NewtonRaphson_vy, chebdif, fourdif

_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")


def load_profile(name: str = "mu6_nz30_kk7") -> dict:
    """The stationary radial profile ``Z1, Z2, Vy1, rhoi`` (nz0 = 30, chebdif nodes)."""
    with open(os.path.join(_DATA, name + ".json")) as fh:
        d = json.load(fh)
    return {k: np.asarray(d[k], dtype=np.float64) for k in ("Z1", "Z2", "Vy1", "rhoi")} | {"mu": float(d["mu"])}


def continue_profile(prof: dict, vy_from: float, vy_to: float, nsteps: int = 10, mu: float | None = None) -> dict:
    """Newton continuation in the boundary flow ``vy`` (the loop form of
    ``scripts/legacy/Statinary_state_with_v.m:10-52``).  A single jump from the kk=7 seed to
    vy=0 (``tests/hqt_smoke_test.m:182``) lands on the normal branch; stepping vy in
    ``nsteps`` increments stays on the superfluid branch."""
    nz0 = len(prof["Z1"])
    Z1, Z2, V = prof["Z1"].copy(), prof["Z2"].copy(), prof["Vy1"].copy()
    if mu is not None and mu != prof["mu"]:
        Z2 = Z2 / prof["mu"] * mu                       # Statinary_state_with_v.m:50
    for s in range(1, nsteps + 1):
        vy = vy_from + (vy_to - vy_from) * s / nsteps
        Z1, Z2, V, it, err = NewtonRaphson_vy(nz0, 1, 0, 0, Z1, Z2, V, vy)
    _, Dz = chebdif(nz0, 1)
    rhoi = -(2 * Dz[:, :, 0]) @ Z2                      # rhoi = -Lz*Z2 (Statinary_state_with_v.m:53)
    return {"Z1": Z1, "Z2": Z2, "Vy1": V, "rhoi": rhoi, "mu": mu if mu is not None else prof["mu"]}


def continue_mu(prof: dict, mu_to: float, vy: float = 0.0, dmu: float = 0.1) -> dict:
    """Newton continuation in ``mu`` at fixed boundary flow ``vy``.  A_t(z=0) = mu is the
    Dirichlet value the initial guess carries (``NewtonRaphson_vy.m:50-52``), so each step
    rescales Z2 by mu_new/mu_old (``scripts/legacy/Statinary_state_with_v.m:50``) and solves
    again; steps of at most ``dmu`` keep the iteration on the superfluid branch down to the
    critical mu (about 4.07 at vy = 0)."""
    nz0 = len(prof["Z1"])
    mu0 = float(prof["mu"])
    n = max(1, int(math.ceil(abs(mu_to - mu0) / dmu - 1e-12)))
    Z1, Z2, V = prof["Z1"].copy(), prof["Z2"].copy(), prof["Vy1"].copy()
    prev = mu0
    for s in range(1, n + 1):
        m = mu0 + (mu_to - mu0) * s / n
        Z2 = Z2 * (m / prev)
        prev = m
        Z1, Z2, V, it, err = NewtonRaphson_vy(nz0, 1, 0, 0, Z1, Z2, V, vy)
        if it >= 1000 or not np.all(np.isfinite(Z1)):
            raise RuntimeError(f"continue_mu: Newton failed at mu={m:.4f}")
    _, Dz = chebdif(nz0, 1)
    rhoi = -(2 * Dz[:, :, 0]) @ Z2
    return {"Z1": Z1, "Z2": Z2, "Vy1": V, "rhoi": rhoi, "mu": float(mu_to)}


_ZERO_FLOW_MU6 = None


def zero_flow_profile(mu: float = 6.0) -> dict:
    """Superfluid stationary profile at zero flow: the kk=7 fixture (mu = 6) continued to
    vy = 0 in 10 steps, then, for another mu, continued in mu (:func:`continue_mu`)."""
    global _ZERO_FLOW_MU6
    if _ZERO_FLOW_MU6 is None:
        _ZERO_FLOW_MU6 = continue_profile(load_profile(), -2 * math.pi * 7 / 20, 0.0, 10, mu=6.0)
    base = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in _ZERO_FLOW_MU6.items()}
    if float(mu) == 6.0:
        return base
    return continue_mu(base, float(mu))


def spline_interp(X, Y, Xi):
    """MATLAB ``interp1(X, Y, Xi, 'spline')`` (not-a-knot cubic spline)."""
    from scipy.interpolate import CubicSpline
    X = np.asarray(X, dtype=np.float64).ravel()
    Y = np.asarray(Y, dtype=np.float64).ravel()
    o = np.argsort(X)
    return CubicSpline(X[o], Y[o], bc_type="not-a-knot")(np.asarray(Xi, dtype=np.float64))


class Grid:
    """Coordinates and the few operators the set-up needs (``tevol_driven_cluster.m:13-52``)."""

    def __init__(self, nx, ny, nz, Lx, Ly):
        self.nx, self.ny, self.nz, self.Lx, self.Ly = nx, ny, nz, float(Lx), float(Ly)
        x, Dx = fourdif(nx, 1)
        y, _ = fourdif(ny, 1)
        z, Dz = chebdif(nz, 1)
        self.x = x * self.Lx / 2 / math.pi                 # xx (:42)
        self.y = (y * self.Ly) / 2 / math.pi               # yy (:43)
        self.z = (z - 1) * 1.0 / 2 + 1.0                   # zz (:44)
        self.Lx_mat = (2.0 * math.pi / self.Lx) * Dx
        self.Lz = 2 * Dz[:, :, 0]
        self.Lz1 = np.linalg.solve(self.Lz[:nz - 1, :nz - 1], np.eye(nz - 1))

    def Lzr(self, v):
        out = np.zeros_like(v)
        out[..., :-1] = v[..., :-1] @ self.Lz1.T
        return out


def stationary_fields(g: Grid, prof: dict):
    """Broadcast and gauge-rotate a stationary profile (``tevol_driven_cluster.m:82-98``).
    Returns ``(Z1, Z2, Z3, Z4, rhoi)`` of shape (nx, ny, nz), unperturbed."""
    nz0 = len(prof["Z1"])
    z0, _ = chebdif(nz0, 1)
    Zg = (z0 - 1) * 1.0 / 2 + 1.0
    Z1 = spline_interp(Zg, prof["Z1"], g.z)
    Z2 = spline_interp(Zg, prof["Z2"], g.z)
    Vy1 = spline_interp(Zg, prof["Vy1"], g.z)
    rh = spline_interp(Zg, prof["rhoi"], g.z)
    pZ2 = g.Lz @ Z2
    Z2_f = np.zeros(g.nz)
    Z2_f[1:] = Z2[1:] / (1 - (g.z[1:] / 1.0) ** 3)
    Z2_f[0] = -pZ2[0] / 3
    theta = -g.Lzr(Z2_f)                                   # :93 (x/y independent)
    shp = (g.nx, g.ny, g.nz)
    z1 = np.broadcast_to(Z1 * np.exp(1j * theta), shp).copy()
    z2 = np.broadcast_to(Z2, shp).copy()
    z3 = np.zeros(shp)                                     # func_Lx(theta) of an x-independent theta
    z4 = np.broadcast_to(-Vy1, shp).copy()                 # z4 = 0*vy - vy1 (:94)
    rhoi = np.broadcast_to(rh[-1], shp).copy()             # repmat(rho(:,:,nz)) (:98)
    return z1, z2, z3, z4, rhoi


def winding_phase(g: Grid, positions, charges):
    """Sum of ``q*atan2(y - Y, x - X)`` over the vortices (``tevol_lattice.m:125-133``)."""
    X, Y = np.meshgrid(g.x, g.y, indexing="ij")
    th = np.zeros((g.nx, g.ny))
    for (x0, y0), q in zip(positions, charges):
        th += q * np.arctan2(Y - y0, X - x0)
    return th


def random_vortices(g: Grid, n: int, rng: np.random.Generator):
    """n uniformly placed windings with alternating charge +1, -1, ..."""
    pos = [(rng.uniform(0, g.Lx), rng.uniform(0, g.Ly)) for _ in range(n)]
    q = [1 if i % 2 == 0 else -1 for i in range(n)]
    return pos, q


def lattice_vortices(g: Grid, nn: int, jitter: float, rng: np.random.Generator):
    """nn x nn jittered square lattice with checkerboard charges (``tevol_lattice.m:104-133``)."""
    cx = (2 * np.arange(1, nn + 1) - 1) / (2 * nn) * g.Lx
    cy = (2 * np.arange(1, nn + 1) - 1) / (2 * nn) * g.Ly
    pos, q = [], []
    for i in range(nn):
        for j in range(nn):
            pos.append((cx[i] + jitter * rng.standard_normal(), cy[j] + jitter * rng.standard_normal()))
            q.append(1 if (i % 2) == (j % 2) else -1)
    return pos, q


def initial_state(nx, ny, nz, Lx, Ly, mu=6.0, seed=0, n_vortices=4, phase_noise=1e-3, lattice=None, prof=None):
    """Seeded initial fields of the benchmark workloads: zero-flow stationary profile times
    ``exp(i*theta)``, theta = phase_noise * N(0,1) + windings (uniform or lattice)."""
    g = Grid(nx, ny, nz, Lx, Ly)
    if prof is None:
        prof = zero_flow_profile(mu)
    z1, z2, z3, z4, rhoi = stationary_fields(g, prof)
    rng = np.random.default_rng(seed)
    theta = phase_noise * rng.standard_normal((nx, ny))
    if lattice is not None:
        pos, q = lattice_vortices(g, lattice[0], lattice[1], rng)
    else:
        pos, q = random_vortices(g, n_vortices, rng)
    theta = theta + winding_phase(g, pos, q)
    z1 = z1 * np.exp(1j * theta)[:, :, None]
    return z1, z2, z3, z4, rhoi


def sinusoidal_drive(nx, ny, Ly, kk=7, amp=0.1):
    """A_x(z=0) = amp*sin(2*pi*kk*y/L): the superflow forcing of the driven-flow workload."""
    y = (2 * math.pi * np.arange(ny) / ny * Ly) / 2 / math.pi
    return np.broadcast_to(amp * np.sin(2 * math.pi * kk * y / Ly)[None, :], (nx, ny)).copy()
