"""Pins of the oracle (CPU, no GPU).

Each function of ``oracle/hqt_oracle.py`` is checked against something other than its own
formula: the reference's known-answer tests (``tests/hqt_smoke_test.m``), the reference's
data fixture, closed forms, invariants of the equations, and an independent second
restatement (the FFT form of the x/y operators, the stationary equations of the Newton
solver).  A plausible transcription mistake -- a dropped term, a wrong sign, a transposed
operand -- fails at least one of these.
"""

import math

import numpy as np
import pytest

from conftest import rel_err
from oracle import hqt_oracle as O

L = 50.0


# ---- dmsuite ------------------------------------------------------------------------------

def test_fourdif_known_answer():
    """tests/hqt_smoke_test.m:129-130."""
    x, D = O.fourdif(33, 1)
    assert np.max(np.abs(D @ np.sin(3 * x) - 3 * np.cos(3 * x))) < 1e-10


@pytest.mark.parametrize("N", [8, 9, 16, 33, 64, 301])
def test_fourdif_is_spectral_derivative(N):
    """D e^{ikx} = ik e^{ikx} for |k| < N/2; the Nyquist mode of even N has eigenvalue 0."""
    x, D = O.fourdif(N, 1)
    for k in range(0, (N - 1) // 2 + 1):
        v = np.exp(1j * k * x)
        assert np.max(np.abs(D @ v - 1j * k * v)) < 1e-9 * max(1, k)
    if N % 2 == 0:
        v = np.cos(N // 2 * x)
        assert np.max(np.abs(D @ v)) < 1e-9


@pytest.mark.parametrize("N", [9, 33, 301])
def test_fourdif_odd_is_circulant(N):
    _, D = O.fourdif(N, 1)
    for s in (1, 5, N - 1):
        assert np.array_equal(np.roll(np.roll(D, s, 0), s, 1), D)


@pytest.mark.parametrize("N", [9, 16, 33])
def test_fourdif_second_derivative(N):
    """fourdif(N,2) is the spectral second derivative (-k^2, Nyquist -N^2/4 for even N)."""
    x, D2 = O.fourdif(N, 2)
    for k in range(0, N // 2 + 1):
        v = np.cos(k * x)
        assert np.max(np.abs(D2 @ v + k * k * v)) < 1e-8 * max(1, k * k)


def test_chebdif_known_answer():
    """tests/hqt_smoke_test.m:131-133."""
    z, DM = O.chebdif(20, 2)
    assert np.max(np.abs(DM[0] @ z ** 4 - 4 * z ** 3)) < 1e-10
    assert np.max(np.abs(DM[1] @ z ** 4 - 12 * z ** 2)) < 1e-8


@pytest.mark.parametrize("N", [5, 12, 20, 30])
def test_chebdif_polynomial_exactness(N):
    z, DM = O.chebdif(N, 2)
    assert np.allclose(z, np.cos(np.pi * np.arange(N) / (N - 1)), atol=1e-15)
    for p in range(0, N):
        d1 = p * z ** max(p - 1, 0) if p else 0 * z
        assert np.max(np.abs(DM[0] @ z ** p - d1)) < 1e-9 * N ** 2


# ---- operators -------------------------------------------------------------------------------

def test_lzr_inverts_lz():
    """func_Lzr: d_z Lzr(v) = v away from the boundary row, Lzr(v)(z=0) = 0; Lzr(3z^2) = z^3."""
    ops = O.Operators(4, 4, 12, L, L)
    rng = np.random.default_rng(0)
    v = rng.standard_normal((4, 4, 12))
    w = ops.Lzr_(v)
    assert np.all(w[:, :, -1] == 0)
    assert rel_err(ops.Lz_(w)[:, :, :-1], v[:, :, :-1]) < 1e-12
    z = ops.zz1
    w = ops.Lzr_(np.broadcast_to(3 * z ** 2, (4, 4, 12)).copy())
    assert np.max(np.abs(w - z ** 3)) < 1e-13


def test_z_grid_convention():
    ops = O.Operators(4, 4, 20, L, L)
    assert ops.zz1[0] == 1.0 and ops.zz1[-1] == 0.0      # horizon at k=0, boundary at k=nz-1


@pytest.mark.parametrize("nx,ny", [(16, 16), (17, 12)])
def test_dense_and_fft_forms_agree(nx, ny):
    """Two independent restatements of the x/y operators: dense fourdif products vs FFT."""
    rng = np.random.default_rng(nx)
    d = O.Operators(nx, ny, 8, L, 40.0)
    f = O.Operators(nx, ny, 8, L, 40.0, form="fft")
    X = rng.standard_normal((nx, ny, 8)) + 1j * rng.standard_normal((nx, ny, 8))
    assert rel_err(d.Lx_(X), f.Lx_(X)) < 1e-12
    assert rel_err(d.Ly_(X), f.Ly_(X)) < 1e-12
    Y = [X, 6 + rng.standard_normal((nx, ny, 8)), rng.standard_normal((nx, ny, 8)), rng.standard_normal((nx, ny, 8))]
    for a, b in zip(O.rhs(d, *Y), O.rhs(f, *Y)):
        assert rel_err(a, b) < 1e-11


# ---- filter ---------------------------------------------------------------------------------

def test_ff_empty_mask():
    """tests/hqt_smoke_test.m:139-141: ff(field, empty mask) = real(field)."""
    rng = np.random.default_rng(11)
    field = rng.standard_normal((17, 19)) + 1j * rng.standard_normal((17, 19))
    assert np.max(np.abs(O.ff(field, np.zeros((17, 19), bool)) - np.real(field))) < 1e-12


def _dft2(x):
    n0, n1 = x.shape
    F0 = np.exp(-2j * np.pi * np.outer(np.arange(n0), np.arange(n0)) / n0)
    F1 = np.exp(-2j * np.pi * np.outer(np.arange(n1), np.arange(n1)) / n1)
    return F0 @ x @ F1.T


@pytest.mark.parametrize("nx,ny", [(7, 6), (9, 10), (12, 12)])
def test_ff_against_direct_dft(nx, ny):
    """ff(x) = real(IDFT(masked DFT(x))) with the DFT written out as O(N^2) sums."""
    rng = np.random.default_rng(nx * ny)
    x = rng.standard_normal((nx, ny))
    fx, fy = O.filter_masks(nx, ny)
    m = fx[:, None] | fy[None, :]
    b = _dft2(x)
    b[m] = 0
    direct = np.real(np.conj(_dft2(np.conj(b)))) / (nx * ny)
    assert rel_err(O.ff(x, m), direct) < 1e-12


@pytest.mark.parametrize("N", [64, 301])
def test_mask_asymmetry_halves_modes(N):
    """The reference mask is asymmetric: for a real field a masked mode whose conjugate is
    kept is halved, not removed (M_eff = (K(k) + K(-k))/2)."""
    fx, fy = O.filter_masks(N, N)
    K = ~fx
    k = np.arange(N)
    Meff = (K.astype(float) + K[(-k) % N]) / 2
    assert set(np.unique(Meff)) == {0.0, 0.5, 1.0}
    half = np.where(Meff == 0.5)[0]
    assert len(half) == 4
    x = np.cos(2 * np.pi * half[0] * np.arange(N) / N)[:, None] * np.ones((1, N))
    # y is constant (mode 0 kept), x-mode is half-masked -> amplitude halves
    assert np.max(np.abs(O.ff(x, fx[:, None] | fy[None, :]) - 0.5 * x)) < 1e-12


def test_filter_mask_indices():
    """fx(round(round(nx/3)):round(nx/3*2)) with MATLAB rounding; 0-based 99..200 at nx=301."""
    fx, _ = O.filter_masks(301, 301)
    assert np.where(fx)[0][0] == 99 and np.where(fx)[0][-1] == 200
    fx, _ = O.filter_masks(64, 64)
    assert np.where(fx)[0][0] == 20 and np.where(fx)[0][-1] == 42


# ---- countN ----------------------------------------------------------------------------------

def test_countN_known_answer():
    """tests/hqt_smoke_test.m:147-151."""
    x, y = np.meshgrid(np.arange(33) * 2 * np.pi / 33, np.arange(33) * 2 * np.pi / 33, indexing="ij")
    phi = np.sin(x + 0.271) + 1j * np.sin(y + 0.613)
    assert O.countN(phi) == (4, 2, 2)


def test_countN_single_vortex_pair():
    n = 40
    x, y = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    phi = np.exp(1j * (np.arctan2(y - 10.3, x - 10.6) - np.arctan2(y - 25.2, x - 27.7)))
    assert O.countN(phi) == (2, 1, 1)


# ---- Newton and the fixture -------------------------------------------------------------------

def test_fixture_identities(golden_profile):
    """Internal identities of data/reference/mu=6,nz=30,vy=7_20.mat."""
    p = golden_profile
    _, Dz = O.chebdif(30, 1)
    assert np.max(np.abs(p["rhoi"] + 2 * Dz[0] @ p["Z2"])) < 1e-12
    assert p["Z2"][-1] == 6.0
    assert p["Vy1"][-1] == -2 * math.pi * 7 / 20
    assert abs(p["Z1"][-1]) < 1e-12 and abs(p["Z2"][0]) < 1e-12


def test_newton_reference_cases(golden_profile):
    """tests/hqt_smoke_test.m:180-183: converges from the seed for vy=-2*pi*7/20 and vy=0."""
    p = golden_profile
    z1, z2, v, it, err = O.NewtonRaphson_vy(30, 1, 0, 0, p["Z1"], p["Z2"], p["Vy1"], -2 * np.pi * 7 / 20)
    assert it == 0 and np.isfinite(err)             # the fixture is already converged
    assert np.max(np.abs(z1 - p["Z1"])) < 1e-6
    z1, z2, v, it, err = O.NewtonRaphson_vy(30, 1, 0, 0, p["Z1"], p["Z2"], p["Vy1"], 0.0)
    assert it < 1000 and np.all(np.isfinite(np.r_[z1, z2, v]))


def test_newton_normal_phase_exact():
    """Psi=0, A_t = mu(1-z), V = vy solves the stationary equations exactly."""
    nz = 20
    zc, _ = O.chebdif(nz, 1)
    z = (zc - 1) / 2 + 1
    for mu, vy in ((3.5, 0.0), (8.0, -1.3)):
        z1, z2, v, it, err = O.NewtonRaphson_vy(nz, 1, 0, 0, 0 * z, mu * (1 - z), vy + 0 * z, vy)
        assert it == 0
        assert np.max(np.abs(z1)) == 0 and np.max(np.abs(z2 - mu * (1 - z))) < 1e-13
        assert np.max(np.abs(v - vy)) < 1e-13


def test_stationary_state_is_rhs_null(golden_profile):
    """Independent pin of F1/F3/F4: the Newton-converged stationary profile (equations of
    src/solvers/NewtonRaphson_vy.m:32-34), gauge-rotated as tevol_driven_cluster.m:90-96, is a
    zero of the time-evolution right-hand sides up to Newton tolerance and truncation."""
    ops = O.Operators(6, 6, 30, L, L)
    z1, z2, z3, z4, rh = O.stationary_fields(ops, golden_profile)
    F1, F2, F3, F4 = O.rhs(ops, z1, z2, z3, z4)
    assert np.max(np.abs(F1)) < 1e-6 and np.max(np.abs(F3)) < 1e-12 and np.max(np.abs(F4)) < 1e-6
    assert np.max(np.abs(F2)) < 1e-12
    # F1 is linear in Psi at fixed gauge fields; perturbing A_t or Psi's modulus breaks stationarity
    assert np.max(np.abs(O.rhs(ops, z1, z2 * 1.01, z3, z4)[0])) > 1e-3
    assert np.max(np.abs(O.rhs(ops, z1 * 1.01, z2, z3, z4)[3])) > 1e-3


# ---- invariants of the step -------------------------------------------------------------------

def _ic(nx, ny, nz, seed=3, nv=2):
    from hqt import ic
    return ic.initial_state(nx, ny, nz, L, L, mu=6.0, seed=seed, n_vortices=nv)


def test_normal_state_fixed_point():
    """Psi=0, A_t=mu(1-z), rho=mu, A_x=0, A_y=const is a fixed point (winding removal off)."""
    nx, ny, nz, mu = 12, 12, 10, 5.0
    ops = O.Operators(nx, ny, nz, L, L)
    z = ops.zz1
    shp = (nx, ny, nz)
    S = [np.zeros(shp, complex), np.broadcast_to(mu * (1 - z), shp).copy(), np.zeros(shp),
         np.full(shp, 0.7), np.full(shp, mu)]
    prm = O.Params(mu=mu, h=0.02, ax_b=0.0, ay_b=0.7, winding=False)
    out = O.step(ops, prm, *S)
    for a, b in zip(out[:5], S):
        assert np.max(np.abs(a - b)) < 1e-12


def test_normal_state_with_winding_gives_nan():
    """With winding removal on, O == 0 makes angle(0/0) = NaN (tevol_driven_cluster.m:137)."""
    nx, ny, nz, mu = 8, 8, 6, 5.0
    ops = O.Operators(nx, ny, nz, L, L)
    z = ops.zz1
    shp = (nx, ny, nz)
    S = [np.zeros(shp, complex), np.broadcast_to(mu * (1 - z), shp).copy(), np.zeros(shp), np.zeros(shp),
         np.full(shp, mu)]
    out = O.step(ops, O.Params(mu=mu, winding=True), *S)
    assert np.isnan(out[5]) and np.isnan(out[0]).all()


def test_translation_covariance():
    nx, ny, nz = 12, 10, 8
    ops = O.Operators(nx, ny, nz, L, L)
    S = _ic(nx, ny, nz)
    prm = O.Params(mu=6.0, h=0.02, ax_b=0.0, ay_b=0.2, winding=False)
    a = O.step(ops, prm, *S)
    Ss = [np.roll(np.roll(s, 3, 0), 2, 1) for s in S]
    b = O.step(ops, prm, *Ss)
    for u, v in zip(a[:5], b[:5]):
        assert rel_err(np.roll(np.roll(u, 3, 0), 2, 1), v) < 1e-12


def test_charge_conservation():
    """sum_xy rho is conserved: the boundary flux is a divergence (d_x A_x + d_y A_y)."""
    nx, ny, nz = 16, 16, 8
    ops = O.Operators(nx, ny, nz, L, L)
    S = list(_ic(nx, ny, nz))
    S[2] = S[2] + 0.05 * np.random.default_rng(1).standard_normal(S[2].shape)
    q0 = S[4][:, :, -1].sum()
    prm = O.Params(mu=6.0, h=0.02, ax_b=0.0, ay_b=0.0, winding=False)
    for _ in range(5):
        *S, _w = O.step(ops, prm, *S)
    assert abs(S[4][:, :, -1].sum() - q0) < 1e-10 * abs(q0)


def test_time_order_of_the_reference_scheme():
    """The reference freezes A_t over the four RK stages and re-solves it from the constraint
    after the step (tevol_driven_cluster.m:120-131), so the scheme converges at first order
    in h although each stage combination is classic RK4: halving h halves the error.  An
    oracle that advanced A_t inside the stages would converge faster and fail here."""
    nx, ny, nz = 8, 8, 8
    ops = O.Operators(nx, ny, nz, L, L)
    S0 = list(_ic(nx, ny, nz, nv=0))
    rng = np.random.default_rng(2)
    S0[0] = S0[0] * (1 + 0.05 * np.cos(2 * np.pi * np.arange(nx) / nx))[:, None, None]
    S0[3] = S0[3] + 0.02 * rng.standard_normal(S0[3].shape) * 0

    def run(h, n):
        prm = O.Params(mu=6.0, h=h, ax_b=0.0, ay_b=0.0, winding=False, filter=False)
        S = list(S0)
        for _ in range(n):
            *S, _w = O.step(ops, prm, *S)
        return S[0]

    T = 0.08
    ref = run(T / 64, 64)
    e1 = np.max(np.abs(run(T / 4, 4) - ref))
    e2 = np.max(np.abs(run(T / 8, 8) - ref))
    assert 1.6 < e1 / e2 < 2.6


def test_philox_known_answers():
    """Random123's published Philox4x32-10 known-answer vectors pin the generator restatement."""
    cases = [((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
             ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
             ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
              (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]
    for ctr, key, want in cases:
        got = O.philox4x32_10(*[np.array([c], dtype=np.uint64) for c in ctr], *key)
        assert tuple(int(g[0]) for g in got) == want


def test_column_normals_are_standard_normal():
    """Box-Muller of the Philox uniforms: mean 0, variance 1, and independent of grid shape
    (column i + nx*j gets the same deviate for any ny)."""
    g1, g2 = O.column_normals(256, 256, seed=11)
    for g in (g1, g2):
        assert abs(g.mean()) < 0.02 and abs(g.var() - 1) < 0.02
    h1, _ = O.column_normals(256, 3, seed=11)
    assert np.array_equal(h1, g1[:, :3])
