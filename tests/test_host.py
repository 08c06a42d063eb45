"""CPU-side tests (no GPU): the C ABI library loads and exports every declared symbol, its
host-only entry points (operator construction, Newton) agree with the oracle, argument
errors are reported through the status codes, and the host-side set-up is pinned."""

import math
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, rel_err

HEADER = os.path.join(ROOT, "include", "hqt.h")


def _declared_symbols():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(hqt_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    import hqt
    from hqt import _lib
    L = hqt.lib()
    declared = _declared_symbols()
    assert len(declared) >= 17
    for name in declared:
        assert hasattr(L, name), name
    assert set(declared) == set(_lib.EXPORTS)
    out = subprocess.run(["nm", "-D", "--defined-only", hqt.LIB_PATH], capture_output=True, text=True).stdout
    for name in declared:
        assert re.search(rf"\bT {name}\b", out), name


def test_version_and_error_string():
    import hqt
    assert b"hqt" in hqt.lib().hqt_version()
    assert isinstance(hqt.lib().hqt_last_error(), bytes)


@pytest.mark.parametrize("N", [9, 16, 33, 64, 301])
def test_fourdif_matches_oracle(N):
    import hqt
    from oracle import hqt_oracle as O
    for m in (0, 1, 2, 3):
        x, D = hqt.fourdif(N, m)
        xo, Do = O.fourdif(N, m)
        assert np.array_equal(x, xo)
        # m <= 1 bit-exact; m = 2 differs by one rounding of csc^2; m = 3 is built by a DFT
        tol = {0: 0.0, 1: 0.0, 2: 1e-15 * np.abs(Do).max()}.get(m, 1e-9 * max(1.0, np.abs(Do).max()))
        assert np.max(np.abs(D - Do)) <= tol, (N, m)


def test_fourdif_known_answer():
    """tests/hqt_smoke_test.m:129-130 through the C ABI."""
    import hqt
    x, D = hqt.fourdif(33, 1)
    assert np.max(np.abs(D @ np.sin(3 * x) - 3 * np.cos(3 * x))) < 1e-10


@pytest.mark.parametrize("N", [5, 12, 20, 30])
def test_chebdif_matches_oracle(N):
    import hqt
    from oracle import hqt_oracle as O
    x, D = hqt.chebdif(N, 2)
    xo, Do = O.chebdif(N, 2)
    assert np.array_equal(x, xo)
    for ell in range(2):
        assert rel_err(D[:, :, ell], Do[ell]) < 1e-14


def test_chebdif_known_answer():
    """tests/hqt_smoke_test.m:131-133 through the C ABI."""
    import hqt
    z, D = hqt.chebdif(20, 2)
    assert np.max(np.abs(D[:, :, 0] @ z ** 4 - 4 * z ** 3)) < 1e-10
    assert np.max(np.abs(D[:, :, 1] @ z ** 4 - 12 * z ** 2)) < 1e-8


def test_newton_matches_oracle(golden_profile):
    """src/solvers/NewtonRaphson_vy.m through the C ABI vs the oracle restatement, on the
    reference's own solver test cases (tests/hqt_smoke_test.m:180-183)."""
    import hqt
    from oracle import hqt_oracle as O
    p = golden_profile
    for vy in (-2 * np.pi * 7 / 20, 0.0, -1.0):
        a = hqt.NewtonRaphson_vy(30, 1, 0, 0, p["Z1"], p["Z2"], p["Vy1"], vy)
        b = O.NewtonRaphson_vy(30, 1, 0, 0, p["Z1"], p["Z2"], p["Vy1"], vy)
        assert a[3] == b[3] and a[3] < 1000
        for u, v in zip(a[:3], b[:3]):
            assert np.max(np.abs(u - v)) < 1e-9
        assert np.isfinite(a[4])


def test_newton_normal_phase_exact():
    import hqt
    nz = 20
    zc, _ = hqt.chebdif(nz, 1)
    z = (zc - 1) / 2 + 1
    z1, z2, v, it, err = hqt.NewtonRaphson_vy(nz, 1, 0, 0, 0 * z, 3.5 * (1 - z), 0 * z, 0.0)
    assert it == 0 and np.max(np.abs(z1)) == 0 and np.max(np.abs(z2 - 3.5 * (1 - z))) < 1e-13


def test_newton_nan_guard():
    """A NaN initial guess stops the iteration (NewtonRaphson_vy.m:88-90)."""
    import hqt
    nz = 12
    g = np.full(nz, np.nan)
    z1, z2, v, it, err = hqt.NewtonRaphson_vy(nz, 1, 0, 0, g, g, g, 0.0)
    assert it == 0 and math.isnan(err)
    with pytest.raises(hqt.HQTError) as ei:
        hqt.NewtonRaphson_vy(nz, 1, 0, 0, g, g, g, 0.0, raise_on_failure=True)
    assert ei.value.code == hqt.HQT_ERR_NOCONV


@pytest.mark.parametrize("nz,nb", [(2, 1), (65, 1), (20, 0)])
def test_newton_batch_rejects_bad_arguments(nz, nb):
    """hqt_newton_vy_batch validates nz in [3, 64] and nb >= 1 before touching the GPU."""
    import ctypes as C
    import hqt
    L = hqt.lib()
    buf = np.zeros(max(nz, 1) * max(nb, 1))
    d = buf.ctypes.data_as(C.POINTER(C.c_double))
    it = np.zeros(max(nb, 1), dtype=np.int32)
    conv = np.zeros(max(nb, 1), dtype=np.uint8)
    rc = L.hqt_newton_vy_batch(nz, 1.0, 0.0, 0.0, nb, d, d, d, d, 0, 0.0, d, d, d,
                               it.ctypes.data_as(C.POINTER(C.c_int32)), d, conv.ctypes.data_as(C.POINTER(C.c_uint8)))
    assert rc == hqt.HQT_ERR_ARG
    assert "hqt_newton_vy_batch" in L.hqt_last_error().decode()


@pytest.mark.parametrize("seed,target,dmu", [(0.0, 6.0, 0.25), (6.0, -6.0, 0.25), (6.0, 8.0, 0.0),
                                              (math.inf, 8.0, 0.25), (6.0, 8.0, math.nan), (6.0, 0.0, 0.25)])
def test_newton_continuation_rejects_bad_arguments(seed, target, dmu):
    """hqt_newton_vy_continue_batch checks each continued problem (seed and target A_t(z=0) finite,
    nonzero and of one sign; step finite and positive) before touching the GPU; the server's
    continued submission makes the same checks."""
    import ctypes as C
    import hqt
    L = hqt.lib()
    nz, nb = 8, 2
    Z2 = np.ones((nz, nb))
    Z2[-1, :] = 6.0
    Z2[-1, 1] = target
    other = np.zeros(nz * nb)
    ms = np.array([math.nan, seed])
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    it = np.zeros(nb, dtype=np.int32)
    steps = np.zeros(nb, dtype=np.int32)
    conv = np.zeros(nb, dtype=np.uint8)
    Z2f = np.asfortranarray(Z2)
    rc = L.hqt_newton_vy_continue_batch(nz, 1.0, 0.0, 0.0, nb, dp(other), dp(Z2f), dp(other), dp(np.zeros(nb)),
                                        dp(ms), dmu, 0, 0.0, dp(other), dp(other), dp(other),
                                        it.ctypes.data_as(C.POINTER(C.c_int32)), dp(np.zeros(nb)),
                                        conv.ctypes.data_as(C.POINTER(C.c_uint8)),
                                        steps.ctypes.data_as(C.POINTER(C.c_int32)))
    assert rc == hqt.HQT_ERR_ARG
    assert "problem 1" in L.hqt_last_error().decode()


@pytest.mark.parametrize("nx,ny,ns,ms", [(8, 8, 3, 2), (0, 8, 1, 1), (8, 8, 0, 1), (8, 8, 2, 0)])
def test_ff_mask_rejects_bad_arguments(nx, ny, ns, ms):
    """hqt_ff_mask checks its shape and that the mask covers one slice or every slice before touching the GPU."""
    import ctypes as C
    import hqt
    L = hqt.lib()
    x = np.zeros(max(nx * ny * ns, 1))
    m = np.zeros(max(nx * ny * max(ms, 1), 1), dtype=np.uint8)
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    rc = L.hqt_ff_mask(nx, ny, ns, dp(x), None, m.ctypes.data_as(C.POINTER(C.c_uint8)), ms, dp(x))
    assert rc == hqt.HQT_ERR_ARG


def test_ff_mask_shape_checked_in_python():
    import hqt
    with pytest.raises(hqt.HQTError):
        hqt.ff(np.zeros((8, 6, 3)), np.zeros((8, 6, 2), dtype=bool))
    with pytest.raises(hqt.HQTError):
        hqt.ff(np.zeros((8, 6)), np.zeros((6, 8), dtype=bool))


@pytest.mark.parametrize("kw", [dict(nx=1), dict(nz=3), dict(nz=33), dict(nx=600), dict(Lx=0.0), dict(Ly=-1.0)])
def test_ctx_create_rejects_bad_arguments(kw):
    """Argument checks happen before any device call, so they run without a GPU."""
    import hqt
    a = dict(nx=16, ny=16, nz=8, Lx=50.0, Ly=50.0)
    a.update(kw)
    with pytest.raises(hqt.HQTError) as ei:
        hqt.Stepper(a["nx"], a["ny"], a["nz"], a["Lx"], a["Ly"])
    assert ei.value.code == hqt.HQT_ERR_ARG


def test_ff_mask_factorisation():
    import hqt
    from oracle import hqt_oracle as O
    fx, fy = O.filter_masks(31, 20)
    m = fx[:, None] | fy[None, :]
    a, b = hqt.mask_factors(m)
    assert np.array_equal(a, fx) and np.array_equal(b, fy)
    a, b = hqt.mask_factors(np.repeat(m[:, :, None], 3, axis=2))
    assert np.array_equal(a, fx)
    bad = np.zeros((5, 5), bool)
    bad[1, 2] = True
    with pytest.raises(hqt.HQTError):
        hqt.mask_factors(bad)


def test_spline_is_not_a_knot():
    """interp1(...,'spline') (tevol_driven_cluster.m:84) reproduces cubics exactly."""
    from hqt import ic
    z0, _ = __import__("hqt").chebdif(30, 1)
    Z = (z0 - 1) / 2 + 1
    zi = np.linspace(0, 1, 17)
    f = lambda t: 2 - t + 3 * t ** 2 - 0.7 * t ** 3
    assert np.max(np.abs(ic.spline_interp(Z, f(Z), zi) - f(zi))) < 1e-12


def test_initial_fields_match_oracle_restatement():
    """The product's set-up of tevol_driven_cluster.m:82-98 vs the oracle's restatement."""
    from hqt import ic
    from oracle import hqt_oracle as O
    prof = ic.zero_flow_profile(6.0)
    ops = O.Operators(8, 6, 12, 50.0, 50.0)
    a = O.stationary_fields(ops, prof)
    b = ic.stationary_fields(ic.Grid(8, 6, 12, 50.0, 50.0), prof)
    for u, v in zip(a, b):
        assert np.max(np.abs(u - v)) < 1e-13


def test_zero_flow_profile_is_superfluid_and_stationary():
    from hqt import ic
    from oracle import hqt_oracle as O
    prof = ic.zero_flow_profile(6.0)
    assert prof["Z1"].max() > 1.5                       # condensed, not the normal branch
    assert np.max(np.abs(prof["Vy1"])) < 1e-8            # V = vy = 0 up to the Newton tolerance
    ops = O.Operators(4, 4, 30, 50.0, 50.0)
    F = O.rhs(ops, *O.stationary_fields(ops, prof)[:4])
    assert max(float(np.max(np.abs(f))) for f in F) < 1e-5


def test_config1_fixture_is_sampled_oracle_state():
    """The 200-step fixture's samples are consistent with its own maxima and stride (a quick check that
    the file holds what make_config1.py writes; the oracle run itself takes about a minute)."""
    g = np.load(os.path.join(GOLDEN, "config1_200steps_64x64x12.npz"))
    n = 64 * 64 * 12
    k = int(g["stride"])
    for name in ("ReZ1", "ImZ1", "Z2", "Z3", "Z4", "rhoi"):
        assert g[name].shape == ((n + k - 1) // k,)
        assert np.all(np.isfinite(g[name])) and np.max(np.abs(g[name])) <= float(g[name + "_max"])
    # A_t = mu at the boundary row (k = nz) of every column after the constraint solve
    idx = np.arange(0, n, k)
    top = idx >= 64 * 64 * 11
    assert np.max(np.abs(g["Z2"][top] - 6.0)) < 1e-12


def test_golden_fixture_reproduced_by_oracle():
    """The committed golden trajectory is what the oracle computes (pins oracle drift)."""
    from oracle import hqt_oracle as O
    g = np.load(os.path.join(GOLDEN, "golden_step_12x10x6.npz"))
    ops = O.Operators(12, 10, 6, 50.0, 50.0)
    S = [g["in_Z1"], g["in_Z2"], g["in_Z3"], g["in_Z4"], g["in_rhoi"]]
    F = O.rhs(ops, *S[:4])
    for name, f in zip(("F1", "F2", "F3", "F4"), F):
        assert rel_err(f, g[name]) < 1e-13
    prm = O.Params(mu=6.0, h=0.02, ax_b=g["ax_b"], ay_b=0.0, winding=True)
    *S, W = O.step(ops, prm, *S)
    for name, v in zip(("Z1", "Z2", "Z3", "Z4", "rhoi"), S):
        assert rel_err(v, g["step1_" + name]) < 1e-13


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_diagnostics_countN_matches_reference_case():
    """countN known answer (tests/hqt_smoke_test.m:147-151) for the product's diagnostics."""
    from hqt import diagnostics
    x, y = np.meshgrid(np.arange(33) * 2 * np.pi / 33, np.arange(33) * 2 * np.pi / 33, indexing="ij")
    phi = np.sin(x + 0.271) + 1j * np.sin(y + 0.613)
    assert diagnostics.countN(phi) == (4, 2, 2)
    assert len(diagnostics.vortex_positions(phi)) == 4


def test_continue_mu_stays_on_superfluid_branch():
    """Zero-flow profiles continued in mu: each is a Newton fixed point (re-solving takes zero
    iterations), carries A_t(z=0) = mu, and the condensate grows with mu above the critical
    value (about 4.07)."""
    import hqt
    from hqt import ic
    prev = 0.0
    for mu in (4.2, 5.0, 6.0, 8.0):
        p = ic.zero_flow_profile(mu)
        assert p["Z2"][-1] == pytest.approx(mu, abs=1e-12)
        amp = float(np.max(np.abs(p["Z1"])))
        assert amp > prev + 0.1
        prev = amp
        z1, z2, v, it, err = hqt.NewtonRaphson_vy(30, 1, 0, 0, p["Z1"], p["Z2"], p["Vy1"], 0.0)
        assert it == 0


def test_footprint_formula_scales_with_points():
    """hqt_ctx_footprint (host only): 42 state-precision arrays plus 3 FP64 arrays per grid
    point (each z-plane padded to a multiple of 32 points; 64 x 64 needs none), plus per-plane,
    per-line and constant terms, plus the filter's cross-term partials (5 fields x one 16-line
    block of x-lines x 32 doubles); mixed precision halves the first part except the filter's ten
    intermediate fields, which stay FP64 (its filter computes in FP64)."""
    import hqt
    a = hqt.footprint(64, 64, 12, "fp64")
    b = hqt.footprint(64, 64, 24, "fp64")
    per_z = b - a
    blocks = (64 * 24 + 15) // 16 - (64 * 12 + 15) // 16
    assert per_z == (12 * 64 * 64 * (42 * 8 + 24) + ((24 * 24 + 23 * 23 + 120) - (12 * 12 + 11 * 11 + 60)) * 8
                     + 5 * blocks * 32 * 8)
    m = hqt.footprint(64, 64, 12, "mixed")
    assert a - m > 64 * 64 * 12 * 32 * 4


def test_server_rejects_bad_arguments():
    import ctypes as C
    import hqt
    h = C.c_void_p()
    assert hqt.lib().hqt_server_create(0, 10, C.byref(h)) == hqt.HQT_ERR_ARG
    assert hqt.lib().hqt_server_create(1 << 30, 0, C.byref(h)) == hqt.HQT_ERR_ARG



@pytest.mark.parametrize("t,name", [(0.0, "0"), (1.0, "1"), (0.02, "0.02"), (1.5, "1.5"), (math.pi, "3.1416"),
                                    (123.456, "123.456"), (12345.678, "12345.678"), (2000.0, "2000")])
def test_checkpoint_name_is_num2str(t, name):
    """num2str of a scalar double (MATLAB's documented %.<d>g rule) as used in the
    checkpoint file name (tevol_driven_cluster.m:179)."""
    import hqt
    assert hqt.checkpoint_name(t) == f"check_point_t={name}.mat"


def test_checkpoint_name_of_accumulated_time():
    """t = t + h with h = 0.02, 50 steps, as the reference's loop accumulates it (:110,126)."""
    import hqt
    t = 0.0
    for _ in range(50):
        t = t + 0.02
    assert t != 1.0 and hqt.checkpoint_name(t) == "check_point_t=1.mat"


def test_checkpoint_file_reads_back_with_scipy(tmp_path):
    """The MAT level 5 writer: scipy.io.loadmat returns the five variables with the reference's
    names, shapes, complex Z1, and the exact values."""
    import hqt
    import scipy.io
    rng = np.random.default_rng(3)
    nx, ny, nz = 5, 7, 3
    Z1 = rng.standard_normal((nx, ny, nz)) + 1j * rng.standard_normal((nx, ny, nz))
    Z2, Z3, Z4, rh = (rng.standard_normal((nx, ny, nz)) for _ in range(4))
    path = tmp_path / hqt.checkpoint_name(1.0)
    hqt.write_checkpoint(path, Z1, Z2, Z3, Z4, rh)
    d = scipy.io.loadmat(str(path))
    assert {"Z1", "Z2", "Z3", "Z4", "rhoi"} <= set(d)
    assert d["Z1"].dtype == np.complex128 and d["Z1"].shape == (nx, ny, nz)
    for k, v in (("Z1", Z1), ("Z2", Z2), ("Z3", Z3), ("Z4", Z4), ("rhoi", rh)):
        assert np.array_equal(d[k], v), k


def test_spline_matches_scipy_not_a_knot():
    """hqt_spline (interp1 'spline') against scipy's not-a-knot CubicSpline on the reference's
    radial use (chebdif(30) nodes to chebdif(20) nodes), and exactness on a cubic."""
    import hqt
    from scipy.interpolate import CubicSpline
    from oracle import hqt_oracle as O
    z0 = (O.chebdif(30, 1)[0] - 1) / 2 + 1
    z = (O.chebdif(20, 1)[0] - 1) / 2 + 1
    y = np.sin(3 * z0) + z0 ** 5
    o = np.argsort(z0)
    ref = CubicSpline(z0[o], y[o], bc_type="not-a-knot")(z)
    assert np.max(np.abs(hqt.spline(z0, y, z) - ref)) < 1e-13
    c = 1 - 2 * z0 + 0.5 * z0 ** 2 - 3 * z0 ** 3
    assert np.max(np.abs(hqt.spline(z0, c, z) - (1 - 2 * z + 0.5 * z ** 2 - 3 * z ** 3))) < 1e-13
