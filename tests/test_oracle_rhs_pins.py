"""Pins of the oracle's transverse (x/y) right-hand-side terms (CPU, no GPU).

``F1``, ``F3`` and ``F4`` (``scripts/simulation/tevol_driven_cluster.m:72-75``) carry x/y terms
that vanish on every x/y-uniform state, so the stationary-profile test in ``test_oracle.py``
cannot see them.  Each term is pinned here by something other than its own transcription:

* **U(1) gauge covariance.**  Psi -> Psi e^{i lam(x,y)}, A_i -> A_i + d_i lam leaves the
  equations covariant: F1 -> F1 e^{i lam}, F2, F3, F4 unchanged.  This fixes the x/y terms of
  F1 to the covariant combination 1/2 (d - iA)^2 Psi and the F3/F4 currents to
  Im(Psi^* (d - iA) Psi), relative to the A^2 Psi and A |Psi|^2 terms.
* **Stationary superflow along x and along y.**  The Newton-solved profile of
  ``src/solvers/NewtonRaphson_vy.m:32-34`` (an independent restatement of the radial
  equations) is a zero of F1, F3, F4 for a flow in either direction; that fixes the A^2 Psi
  and A |Psi|^2 coefficients relative to the radial terms.
* **Normal-state charge diffusion.**  A small charge-density wave of wavenumber k in the
  normal phase decays as exp(-D k^2 t) with D = 3 / (4 pi T) = z_h = 1 for the planar AdS4
  black brane, f = 1 - z^3 (Herzog, Kovtun, Sachdev, Son, Phys. Rev. D 75, 085020 (2007)).
  The rate involves F2, the A_t constraint solve and the ratio of F3's 1/2 d_z d_x A_t and
  1/2 d_z (f d_z A_x) terms (and F4's in y).
* **Transverse-mode closed form.**  Psi = 0 and A = (a(z), b(z)) cos(p x + q y): every term
  reduces to polynomials in z; in particular for a divergence-free mode the x/y part of F3's
  integrand is 1/2 Laplacian(A_x), the same elliptic operator, with the same sign, as the
  radial part 1/2 d_z (f d_z A_x) at f = 1.  That fixes the sign of the curl terms.
* **Plane-wave closed form** for F1 (Psi = psi(z) e^{i(px+qy)}, smooth A): every x/y
  derivative of a single Fourier mode is exact, so F1 equals the continuum expression.

``test_pins_catch_planted_mistakes`` plants plausible transcription mistakes (a flipped
sign, a halved coefficient, a dropped term) in a copy of the oracle and checks that each one
fails at least one of these pins.
"""

import math

import numpy as np
import pytest
from numpy.polynomial import Polynomial as P

from conftest import rel_err
from oracle import hqt_oracle as O


# ---- a parametrised copy of the oracle's F1/F3/F4, used only to plant mistakes -------------------

class _Coeffs:
    adv = -1j       # -i (A_x d_x Psi + A_y d_y Psi)
    lap = 0.5       # 1/2 (d_xx + d_yy) Psi
    div = -0.5j     # -1/2 i Psi (d_x A_x + d_y A_y)
    cur = 1.0       # Im(Psi^* d Psi) current in F3/F4
    curl = 0.5      # 1/2 d_y (d_y A_x - d_x A_y) in F3 (and the x analogue in F4)
    at = 1.0        # d_x A_t inside 1/2 d_z (d_x A_t + f d_z A_x)


def _mutant_ops(nx, ny, nz, Lx, Ly, **change):
    """Operators whose F1/F3/F4 take their x/y coefficients from ``_Coeffs`` (with ``change``
    applied).  With no change they equal the oracle's closures (checked below)."""
    c = _Coeffs()
    for k, v in change.items():
        setattr(c, k, v)

    class M(O.Operators):
        def F1(self, y1, y2, y3, y4):
            Lx, Ly, Lz, Lzr = self.Lx_, self.Ly_, self.Lz_, self.Lzr_
            f = self.f1[None, None, :]
            dpsi = Lz(y1)
            return Lzr(1j * y2 * dpsi + 0.5j * y1 * Lz(y2) + 0.5 * f * Lz(dpsi)
                       + self.c_dpsi[None, None, :] * dpsi + self.c_psi[None, None, :] * y1
                       - 0.5 * (y3 * y3 + y4 * y4) * y1 + c.lap * (Lx(Lx(y1)) + Ly(Ly(y1)))
                       + c.div * y1 * (Lx(y3) + Ly(y4)) + c.adv * (y3 * Lx(y1) + y4 * Ly(y1)))

        def F3(self, y1, y2, y3, y4):
            Lx, Ly, Lz, Lzr = self.Lx_, self.Ly_, self.Lz_, self.Lzr_
            f = self.f1[None, None, :]
            cy = np.conj(y1)
            return np.real(Lzr(0.5 * Lz(c.at * Lx(y2) + f * Lz(y3)) + c.curl * Ly(Ly(y3) - Lx(y4))
                               + c.cur * np.imag(cy * Lx(y1)) - y3 * y1 * cy))

        def F4(self, y1, y2, y3, y4):
            Lx, Ly, Lz, Lzr = self.Lx_, self.Ly_, self.Lz_, self.Lzr_
            f = self.f1[None, None, :]
            cy = np.conj(y1)
            return np.real(Lzr(0.5 * Lz(c.at * Ly(y2) + f * Lz(y4)) + c.curl * Lx(Lx(y4) - Ly(y3))
                               + c.cur * np.imag(cy * Ly(y1)) - y4 * y1 * cy))

    return M(nx, ny, nz, Lx, Ly)


def _real_ops(nx, ny, nz, Lx, Ly, **_):
    return O.Operators(nx, ny, nz, Lx, Ly)


# ---- the pins, each returning an error measure (0 = exact) ------------------------------------------

def _lowmode(rng, shape, L, kmax, cplx):
    nx, ny, nz = shape
    x = np.arange(nx) * L[0] / nx
    y = np.arange(ny) * L[1] / ny
    X, Y = x[:, None, None], y[None, :, None]
    f = np.zeros(shape, complex)
    for mx in range(-kmax, kmax + 1):
        for my in range(-kmax, kmax + 1):
            prof = rng.standard_normal(nz) + 1j * rng.standard_normal(nz)
            f = f + prof[None, None, :] * np.exp(1j * (2 * np.pi * (mx * X / L[0] + my * Y / L[1])
                                                       + rng.uniform(0, 2 * np.pi)))
    return f if cplx else np.real(f)


def gauge_covariance_error(make=_real_ops):
    """max over F1..F4 of the relative violation of U(1) gauge covariance."""
    nx, ny, nz, Lx, Ly = 32, 32, 10, 40.0, 30.0
    ops = make(nx, ny, nz, Lx, Ly)
    rng = np.random.default_rng(1)
    X, Y, Z = ops.xx1[:, None, None], ops.yy1[None, :, None], ops.zz1[None, None, :]
    shp = (nx, ny, nz)
    psi = 0.3 * _lowmode(rng, shp, (Lx, Ly), 2, True) * (Z * (1 - Z) + 0.1 * Z)
    At = 6 * (1 - Z) + 0.2 * Z * _lowmode(rng, shp, (Lx, Ly), 1, False)
    Ax = 0.2 * _lowmode(rng, shp, (Lx, Ly), 2, False)
    Ay = 0.2 * _lowmode(rng, shp, (Lx, Ly), 2, False) + 0.5
    a, b, p, q = 0.3, 0.25, 2 * np.pi / Lx, 2 * np.pi / Ly
    lam = a * np.cos(p * X + 0.4) + b * np.sin(q * Y + 0.1)
    F = O.rhs(ops, psi, At, Ax, Ay)
    G = O.rhs(ops, psi * np.exp(1j * lam), At, Ax - a * p * np.sin(p * X + 0.4), Ay + b * q * np.cos(q * Y + 0.1))
    errs = [rel_err(G[0], np.exp(1j * lam) * F[0])] + [rel_err(G[i], F[i]) for i in (1, 2, 3)]
    return max(errs)


def stationary_flow_error(make=_real_ops, golden=None):
    """max |F1|, |F3|, |F4| on the stationary profile with the flow along y and along x."""
    ops = make(6, 6, 30, 50.0, 50.0)
    z1, z2, z3, z4, _ = O.stationary_fields(ops, golden)
    e_y = max(float(np.max(np.abs(F))) for F in (ops.F1(z1, z2, z3, z4), ops.F3(z1, z2, z3, z4),
                                                 ops.F4(z1, z2, z3, z4)))
    e_x = max(float(np.max(np.abs(F))) for F in (ops.F1(z1, z2, z4, z3), ops.F3(z1, z2, z4, z3),
                                                 ops.F4(z1, z2, z4, z3)))
    return max(e_x, e_y)


def charge_diffusion_constant(make=_real_ops, L=200.0, axis=0, T=40.0, h=0.02):
    """Decay rate / k^2 of a normal-state charge-density wave of wavenumber 2 pi / L along
    ``axis``, fitted over the second half of [0, T] (after the non-hydrodynamic modes died)."""
    n = 8
    nx, ny = (n, 2) if axis == 0 else (2, n)
    nz, mu, eps = 16, 3.0, 1e-3
    ops = make(nx, ny, nz, L, L)
    z = ops.zz1
    k = 2 * math.pi / L
    s = (ops.xx1[:, None] if axis == 0 else ops.yy1[None, :]) * np.ones((nx, ny))
    shp = (nx, ny, nz)
    rho = (mu + eps * np.cos(k * s))[:, :, None] * np.ones(shp)
    state = [np.zeros(shp, complex), mu - rho * z[None, None, :], np.zeros(shp), np.zeros(shp), rho]
    prm = O.Params(mu=mu, h=h, ax_b=0.0, ay_b=0.0, winding=False, filter=False)
    ts, amps = [], []
    nsteps = int(round(T / h))
    for i in range(nsteps):
        *state, _w = O.step(ops, prm, *state)
        if (i + 1) % 25 == 0:
            r = state[4][:, 0, -1] if axis == 0 else state[4][0, :, -1]
            ts.append((i + 1) * h)
            amps.append(abs(np.fft.fft(r)[1]) * 2 / n)
    ts, amps = np.array(ts), np.array(amps)
    if not np.all(np.isfinite(amps)) or np.any(amps <= 0):
        return float("nan")
    sel = ts >= T / 2
    rate = -np.polyfit(ts[sel], np.log(amps[sel]), 1)[0]
    return rate / k ** 2


def _transverse_closed_form(ops, p, q, r, s, eps, mu, a, b):
    """(F1, F2, F3, F4) of Psi = 0, A_t = mu (1 - z) + eps z^2 cos(r x + s y),
    A_x = a(z) cos(p x + q y), A_y = b(z) cos(p x + q y) from the continuum equations."""
    X, Y, z = ops.xx1[:, None, None], ops.yy1[None, :, None], ops.zz1[None, None, :]
    th, th2 = p * X + q * Y, r * X + s * Y
    f = P([1, 0, 0, -1])
    anti = lambda poly: poly.integ(lbnd=0)  # noqa: E731 -- radial integral from the boundary z = 0
    g3 = anti(0.5 * (f * a.deriv()).deriv() + 0.5 * q * (p * b - q * a))
    g4 = anti(0.5 * (f * b.deriv()).deriv() + 0.5 * p * (q * a - p * b))
    # 1/2 d_z d_x A_t = -eps r z sin(th2); its integral from z = 0 is -eps r z^2 / 2 sin(th2)
    F3 = np.cos(th) * g3(z) - 0.5 * eps * r * z ** 2 * np.sin(th2)
    F4 = np.cos(th) * g4(z) - 0.5 * eps * s * z ** 2 * np.sin(th2)
    F2 = (1 - z ** 3) * (-p * a.deriv()(z) - q * b.deriv()(z)) * np.sin(th)
    shp = (ops.nx, ops.ny, ops.nz)
    state = (np.zeros(shp, complex), mu * (1 - z) + eps * z ** 2 * np.cos(th2) + 0 * X * Y,
             a(z) * np.cos(th), b(z) * np.cos(th))
    return state, (np.zeros(shp, complex), np.broadcast_to(F2, shp), np.broadcast_to(F3, shp),
                   np.broadcast_to(F4, shp))


def transverse_closed_form_error(make=_real_ops):
    nx, ny, nz, Lx, Ly = 12, 10, 10, 40.0, 30.0
    ops = make(nx, ny, nz, Lx, Ly)
    p, q, r, s = 2 * np.pi * 2 / Lx, 2 * np.pi / Ly, 2 * np.pi / Lx, -2 * np.pi * 2 / Ly
    err = 0.0
    for a, b in ((P([0, 1]), P([0, 0, 1])),        # a general mode
                 (P([0, q]), P([0, -p]))):          # divergence-free: p a + q b = 0
        state, want = _transverse_closed_form(ops, p, q, r, s, 0.3, 6.0, a, b)
        got = O.rhs(ops, *state)
        err = max(err, max(float(np.max(np.abs(g - w))) / max(float(np.max(np.abs(w))), 1.0)
                           for g, w in zip(got, want)))
    return err


def plane_wave_closed_form_error(make=_real_ops):
    """F1 of Psi = psi(z) e^{i(px+qy)}, A_t = alpha(z), A_x = ax0 + ax1 sin(p1 x + c1),
    A_y = ay0 + ay1 cos(q1 y + c2), from the continuum expression."""
    nx, ny, nz, Lx, Ly = 16, 12, 10, 40.0, 30.0
    ops = make(nx, ny, nz, Lx, Ly)
    X, Y, z = ops.xx1[:, None, None], ops.yy1[None, :, None], ops.zz1[None, None, :]
    p, q, p1, q1 = 2 * np.pi * 3 / Lx, -2 * np.pi * 2 / Ly, 2 * np.pi / Lx, 2 * np.pi * 2 / Ly
    psi = P([0, 1, -1])                     # z (1 - z)
    alpha = P([6, -6, 0.5])                 # A_t(z)
    f = P([1, 0, 0, -1])
    ax = 0.4 + 0.3 * np.sin(p1 * X + 0.2)
    ay = -0.7 + 0.25 * np.cos(q1 * Y + 0.5)
    dxax = 0.3 * p1 * np.cos(p1 * X + 0.2)
    dyay = -0.25 * q1 * np.sin(q1 * Y + 0.5)
    # z-only part: i A_t psi' + 1/2 i psi A_t' + 1/2 f psi'' - 3/2 z^2 psi' - 1/2 z psi
    pz = P([0, 1])
    radial_re = 0.5 * f * psi.deriv(2) - 1.5 * pz ** 2 * psi.deriv() - 0.5 * pz * psi
    radial_im = alpha * psi.deriv() + 0.5 * psi * alpha.deriv()
    g = (-0.5 * (ax ** 2 + ay ** 2) - 0.5 * (p * p + q * q) - 0.5j * (dxax + dyay) + (ax * p + ay * q))
    want = np.exp(1j * (p * X + q * Y)) * (radial_re.integ(lbnd=0)(z) + 1j * radial_im.integ(lbnd=0)(z)
                                          + g * psi.integ(lbnd=0)(z))
    shp = (nx, ny, nz)
    y1 = psi(z) * np.exp(1j * (p * X + q * Y)) * np.ones(shp)
    got = ops.F1(y1, np.broadcast_to(alpha(z), shp), np.broadcast_to(ax, shp), np.broadcast_to(ay, shp))
    return rel_err(got, want)


# ---- tests on the oracle ---------------------------------------------------------------------

def test_mutant_copy_equals_oracle():
    """The parametrised copy used for the planted mistakes is the oracle when unchanged."""
    rng = np.random.default_rng(5)
    nx, ny, nz = 10, 8, 7
    a, b = O.Operators(nx, ny, nz, 40.0, 30.0), _mutant_ops(nx, ny, nz, 40.0, 30.0)
    s = (nx, ny, nz)
    Y = [rng.standard_normal(s) + 1j * rng.standard_normal(s), 6 + rng.standard_normal(s), rng.standard_normal(s),
         rng.standard_normal(s)]
    for u, v in zip(O.rhs(a, *Y), O.rhs(b, *Y)):
        assert rel_err(v, u) < 1e-13


def test_gauge_covariance():
    assert gauge_covariance_error() < 1e-13


def test_stationary_flow_along_x_and_y(golden_profile):
    assert stationary_flow_error(golden=golden_profile) < 1e-6


@pytest.mark.parametrize("axis", [0, 1])
def test_normal_state_charge_diffusion(axis):
    """D = 3 / (4 pi T) = 1; at k = 2 pi / 200 the O(k^2) correction is below 1e-3."""
    D = charge_diffusion_constant(axis=axis)
    assert abs(D - 1.0) < 3e-3


def test_transverse_mode_closed_form():
    assert transverse_closed_form_error() < 1e-12


def test_plane_wave_closed_form():
    assert plane_wave_closed_form_error() < 1e-12


_PINS = [
    ("gauge covariance", lambda mk, g: gauge_covariance_error(mk) < 1e-13),
    ("transverse closed form", lambda mk, g: transverse_closed_form_error(mk) < 1e-12),
    ("plane-wave closed form", lambda mk, g: plane_wave_closed_form_error(mk) < 1e-12),
    ("stationary flows", lambda mk, g: stationary_flow_error(mk, g) < 1e-6),
    ("charge diffusion", lambda mk, g: abs(charge_diffusion_constant(mk, T=20.0) - 1.0) < 1e-2),
]


@pytest.mark.parametrize("change", [
    {"adv": 1j},                 # sign of -i (A . grad Psi) flipped
    {"lap": 0.25},               # 1/2 d_xx Psi halved
    {"cur": -1.0},               # sign of F3/F4's current term flipped
    {"div": 0.5j},               # sign of -1/2 i Psi div A flipped
    {"curl": -0.5},              # sign of the curl term flipped
    {"at": 0.0},                 # 1/2 d_z d_x A_t dropped
    {"at": -1.0},                # ... or its sign flipped
    {"lap": 1.0, "adv": -2j, "div": -1j},   # the covariant x/y operator doubled as a whole
], ids=lambda c: ",".join(f"{k}={v}" for k, v in c.items()))
def test_pins_catch_planted_mistakes(change, golden_profile):
    """Each plausible transcription mistake, planted in a copy of the oracle, fails a pin."""
    def mk(nx, ny, nz, Lx, Ly):
        return _mutant_ops(nx, ny, nz, Lx, Ly, **change)
    failed = [name for name, ok in _PINS if not ok(mk, golden_profile)]
    assert failed, f"mutant {change} passed every pin"
