"""Batched NewtonRaphson_vy on the GPU (hqt_newton_vy_batch) against the oracle's restatement of
src/solvers/NewtonRaphson_vy.m.

Each problem of a batch must reproduce the oracle's iteration: the same iteration count and
the converged profiles within 1e-9 (both stop at max|Fun_next| < 1e-6 after the same number of
steps; the remaining difference is LU rounding).  Covered: the reference's own solver cases
(tests/hqt_smoke_test.m:180-183), a vy sweep, the LDS-resident (nz <= 41) and the
global-workspace (nz > 41) Jacobian, the exact normal-phase solution, a NaN problem inside a
batch, and batch-order independence.  The continuation in mu (hqt_newton_vy_continue_batch)
against the oracle's newton_mu_continuation on the superfluid branch: the same number of solves and
retries, the same last iteration count, the profiles within 1e-9; below the branch's end, a
solution at the target in the normal phase."""

import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

VY_REF = -2 * math.pi * 7 / 20


def _oracle(nz, Z1, Z2, V, vy):
    from oracle import hqt_oracle as O
    return O.NewtonRaphson_vy(nz, 1, 0, 0, Z1, Z2, V, vy)


def _check_against_oracle(nz, Z1, Z2, V, vys, tol=1e-9):
    import hqt
    z1, z2, VY1, it, err, conv = hqt.NewtonRaphson_vy_batch(nz, 1, 0, 0, Z1, Z2, V, vys)
    for b, vy in enumerate(vys):
        a1, a2, av = (np.asarray(x)[:, b] if np.ndim(x) == 2 else x for x in (Z1, Z2, V))
        r1, r2, rv, rit, rerr = _oracle(nz, a1, a2, av, vy)
        assert it[b] == rit, f"vy={vy}: it {it[b]} vs oracle {rit}"
        assert conv[b] == (rit < 1000)
        scale = max(1.0, float(np.max(np.abs(r2))))
        for name, u, v in (("z1", z1[:, b], r1), ("z2", z2[:, b], r2), ("VY1", VY1[:, b], rv)):
            d = float(np.max(np.abs(u - v)))
            assert d < tol * scale, f"vy={vy} {name}: |diff| {d:.3e}"
        # err is the last update's norm: tiny at convergence, same order as the oracle's
        assert np.isfinite(err[b]) and err[b] < max(10 * rerr, 1e-9)
    return z1, z2, VY1, it, err, conv


def test_reference_solver_cases(golden_profile):
    """tests/hqt_smoke_test.m:180-183 on the fixture seed, plus vy=-1, in one launch."""
    p = golden_profile
    _check_against_oracle(30, p["Z1"], p["Z2"], p["Vy1"], np.array([VY_REF, 0.0, -1.0]))


def test_vy_sweep(golden_profile):
    p = golden_profile
    vys = np.linspace(-0.2, -1.6, 8)
    z1, z2, VY1, it, err, conv = _check_against_oracle(30, p["Z1"], p["Z2"], p["Vy1"], vys)
    assert conv.all()
    assert np.allclose(VY1[-1, :], vys, atol=1e-12)          # boundary value V(z=0) = vy (:28-29)


def test_global_workspace_path(golden_profile):
    """nz = 48: the 144 x 144 Jacobian does not fit in LDS and lives in the global workspace."""
    from hqt import ic
    p = golden_profile
    from oracle import hqt_oracle as O
    z30 = (O.chebdif(30, 1)[0] - 1) / 2 + 1
    nz = 48
    z48 = (O.chebdif(nz, 1)[0] - 1) / 2 + 1
    seed = [ic.spline_interp(z30, p[k], z48) for k in ("Z1", "Z2", "Vy1")]
    _check_against_oracle(nz, *seed, np.array([VY_REF, -0.5]))


def test_normal_phase_exact():
    """Psi = 0, A_t = mu(1-z), V = vy is an exact solution: zero iterations."""
    import hqt
    from oracle import hqt_oracle as O
    nz = 20
    z = (O.chebdif(nz, 1)[0] - 1) / 2 + 1
    mus = np.array([3.5, 8.0])
    vys = np.array([0.0, -1.3])
    Z1 = np.zeros((nz, 2))
    Z2 = np.stack([m * (1 - z) for m in mus], axis=1)
    V = np.stack([v + 0 * z for v in vys], axis=1)
    z1, z2, VY1, it, err, conv = hqt.NewtonRaphson_vy_batch(nz, 1, 0, 0, Z1, Z2, V, vys)
    assert (it == 0).all() and conv.all()
    assert np.max(np.abs(z1)) == 0
    assert np.max(np.abs(z2 - Z2)) < 1e-13 and np.max(np.abs(VY1 - V)) < 1e-13


def test_nan_problem_in_batch(golden_profile):
    """A NaN seed stops its own problem at K=0 with err=NaN (:88-90); the others are untouched."""
    import hqt
    p = golden_profile
    Z1 = np.repeat(p["Z1"][:, None], 3, axis=1)
    Z1[:, 1] = np.nan
    vys = np.array([VY_REF, VY_REF, -1.0])
    z1, z2, VY1, it, err, conv = hqt.NewtonRaphson_vy_batch(30, 1, 0, 0, Z1, p["Z2"], p["Vy1"], vys)
    assert it[1] == 0 and math.isnan(err[1]) and not conv[1]
    assert conv[0] and conv[2]
    with pytest.raises(hqt.HQTError) as ei:
        hqt.NewtonRaphson_vy_batch(30, 1, 0, 0, Z1, p["Z2"], p["Vy1"], vys, raise_on_failure=True)
    assert ei.value.code == hqt.HQT_ERR_NOCONV
    a = hqt.NewtonRaphson_vy_batch(30, 1, 0, 0, p["Z1"], p["Z2"], p["Vy1"], np.array([-1.0]))
    assert np.array_equal(a[0][:, 0], z1[:, 2]) and a[3][0] == it[2]


def test_batch_order_independent(golden_profile):
    """A problem's result does not depend on its position or neighbours (bitwise)."""
    import hqt
    p = golden_profile
    vys = np.linspace(-0.3, -1.2, 5)
    a = hqt.NewtonRaphson_vy_batch(30, 1, 0, 0, p["Z1"], p["Z2"], p["Vy1"], vys)
    b = hqt.NewtonRaphson_vy_batch(30, 1, 0, 0, p["Z1"], p["Z2"], p["Vy1"], vys[::-1].copy())
    for k in range(3):
        assert np.array_equal(a[k], b[k][:, ::-1])
    assert np.array_equal(a[3], b[3][::-1])


MU_FOLD = 5.8     # the superfluid branch at vy = -2*pi*7/20 ends at a fold near mu = 5.79


def _is_solution_at(nz, z1, z2, v, target, vy=VY_REF):
    """The oracle's NewtonRaphson_vy started from (z1, z2, v) stops at iteration 0 (its residual is
    already below tolerance) and A_t(z=0) is the target."""
    from oracle import hqt_oracle as O
    r = O.NewtonRaphson_vy(nz, 1, 0, 0, z1, z2, v, vy)
    return r[3] == 0 and abs(z2[-1] - target) <= 1e-12 * abs(target)


def _check_continuation(nz, mus, seeds, dmu=0.25, tol=1e-9):
    """hqt_newton_vy_continue_batch on serving-workload guesses against the oracle's
    newton_mu_continuation.  On the superfluid branch (target above the fold) each solve converges
    in a few iterations from the last one: the same convergence, number of solves and last iteration
    count, and the profiles within tol.  Below the fold the walk leaves the branch through solves
    that wander for tens of iterations, where rounding decides the path; there the result must be a
    solution at the target (checked by the oracle's residual) in the normal phase."""
    import hqt
    from hqt import loadgen
    from oracle import hqt_oracle as O
    G = [loadgen.newton_guess(nz, mu) for mu in mus]
    Z1, Z2, V = (np.stack([g[k] for g in G], axis=1) for k in range(3))
    vys = np.full(len(mus), VY_REF)
    z1, z2, VY1, it, err, conv, steps = hqt.NewtonRaphson_vy_batch(nz, 1, 0, 0, Z1, Z2, V, vys,
                                                                   mu_seed=np.asarray(seeds, dtype=float), dmu=dmu)
    for b, (mu, sd) in enumerate(zip(mus, seeds)):
        if not math.isnan(sd) and mu < MU_FOLD:
            assert conv[b], (nz, mu)
            assert _is_solution_at(nz, z1[:, b], z2[:, b], VY1[:, b], Z2[-1, b]), (nz, mu)
            assert np.max(np.abs(z1[:, b])) < 1e-6, (nz, mu)
            continue
        if math.isnan(sd):
            r1, r2, rv, rit, rerr = O.NewtonRaphson_vy(nz, 1, 0, 0, Z1[:, b], Z2[:, b], V[:, b], VY_REF)
            rconv, rsteps = (rit < 1000 and np.isfinite(rerr)), 0
        else:
            r1, r2, rv, rit, rerr, rconv, rsteps = O.newton_mu_continuation(nz, 1, 0, 0, Z1[:, b], Z2[:, b], V[:, b],
                                                                            VY_REF, sd, dmu)
        assert (conv[b], steps[b], it[b]) == (rconv, rsteps, rit), (nz, mu, sd, conv[b], steps[b], it[b], rsteps, rit)
        scale = max(1.0, float(np.max(np.abs(r2))))
        for name, u, v in (("z1", z1[:, b], r1), ("z2", z2[:, b], r2), ("VY1", VY1[:, b], rv)):
            d = float(np.max(np.abs(u - v)))
            assert d < tol * scale, f"nz={nz} mu={mu} {name}: |diff| {d:.3e}"
    return conv, steps, z1


def test_mu_continuation_matches_oracle():
    """Continued problems (above and below the end of the superfluid branch near mu = 5.79) and a
    plain one (NaN seed) in one launch, on the LDS-resident Jacobian (nz = 32)."""
    mus = [8.25, 9.25, 7.5, 10.0, 5.5, 3.5, 6.5]
    seeds = [6.0] * 6 + [math.nan]
    print("continuation nz=32 mus", mus, flush=True)
    conv, steps, z1 = _check_continuation(32, mus, seeds)
    assert conv.all()
    assert (steps[:6] >= 2).all() and steps[6] == 0
    assert (np.max(np.abs(z1[:, :4]), axis=0) > 1.0).all()       # superfluid above the fold
    assert (np.max(np.abs(z1[:, 4:6]), axis=0) < 1e-6).all()     # normal phase below it


def test_mu_continuation_global_workspace():
    """nz = 48 (Jacobian in the global workspace), where the plain solve from these guesses wanders
    for its 1000 iterations at mu = 9.25."""
    mus = [9.25, 8.5, 5.75, 6.25]
    print("continuation nz=48 mus", mus, flush=True)
    conv, steps, _ = _check_continuation(48, mus, [6.0] * 4)
    assert conv.all()


def test_mu_continuation_halves_failed_steps():
    """Steps of up to 3.0 from mu = 6 to 8.5: the first step fails within its 100 iterations and is
    retried at half the length, the oracle and the GPU taking the same retries."""
    for nz in (24, 32):
        print("continuation halving nz", nz, flush=True)
        conv, steps, z1 = _check_continuation(nz, [8.5], [6.0], dmu=3.0)
        assert conv.all() and steps[0] > 2
        assert np.max(np.abs(z1)) > 1.0
