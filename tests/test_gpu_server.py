"""Serving loop (hqt_server_*): admission by HBM budget, continuous batching of evolutions
through the ensemble, Newton batches beside them.  Every result must equal the request run
alone (bitwise): hqt_step_n for evolutions, hqt_newton_vy_batch for Newton solves."""

import math
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

L = 50.0
VY_REF = -2 * math.pi * 7 / 20


def _state(nx, nz, mu, seed):
    from hqt import ic
    return ic.initial_state(nx, nx, nz, L, L, mu=mu, seed=seed, n_vortices=4, phase_noise=1e-3)


def _alone(nx, nz, mu, prec, seed, nsteps, **kw):
    import hqt
    with hqt.Stepper(nx, nx, nz, L, L, mu=mu, dt=0.02, precision=prec, **kw) as st:
        st.set_state(*_state(nx, nz, mu, seed))
        st.step(nsteps)
        return st.get_state()


@pytest.mark.parametrize("nx,ny,nz,prec", [(64, 64, 12, "fp64"), (96, 80, 16, "mixed"), (301, 301, 20, "fp64")])
def test_footprint_matches_allocation(nx, ny, nz, prec):
    import hqt
    with hqt.Stepper(nx, ny, nz, L, L, precision=prec) as st:
        assert st.device_bytes == hqt.footprint(nx, ny, nz, prec)


def test_mixed_stream_results_equal_alone(golden_profile):
    """Three evolutions (budget for two at a time) and four Newton solves, submitted at
    staggered times; each result is bitwise the stand-alone one."""
    import hqt
    from hqt import ic
    evo = [(64, 12, 6.0, "fp64", 11, 30), (96, 16, 5.0, "mixed", 12, 45), (48, 8, 7.0, "fp64", 13, 20)]
    budget = hqt.footprint(96, 96, 16, "mixed") + hqt.footprint(64, 64, 12, "fp64") + 1
    p = golden_profile
    vys = [VY_REF, 0.0, -0.5, -1.0]
    ref_evo = [_alone(nx, nz, mu, prec, seed, ns, ax_b=ic.sinusoidal_drive(nx, nx, L, 5, 0.05))
               for nx, nz, mu, prec, seed, ns in evo]
    ref_new = [hqt.NewtonRaphson_vy_batch(30, 1, 0, 0, p["Z1"], p["Z2"], p["Vy1"], np.array([v])) for v in vys]
    with hqt.Server(budget, quantum=7) as srv:
        ids = []
        for k, (nx, nz, mu, prec, seed, ns) in enumerate(evo):
            ids.append(("evo", k, srv.submit_evolution(_state(nx, nz, mu, seed), ns, nx, nx, nz, L, L, mu=mu, dt=0.02,
                                                       precision=prec, ax_b=ic.sinusoidal_drive(nx, nx, L, 5, 0.05))))
            ids.append(("new", k, srv.submit_newton(30, 1, 0, 0, p["Z1"], p["Z2"], p["Vy1"], vys[k])))
            time.sleep(0.01)
        ids.append(("new", 3, srv.submit_newton(30, 1, 0, 0, p["Z1"], p["Z2"], p["Vy1"], vys[3])))
        for kind, k, rid in ids:
            st, lat, out = srv.result(rid, timeout=60.0)
            assert st == 0 and lat > 0
            if kind == "evo":
                for a, b in zip(out, ref_evo[k]):
                    assert np.array_equal(a, b), f"evolution {k}"
            else:
                z1, z2, v, it, err = out
                r = ref_new[k]
                assert np.array_equal(z1, r[0][:, 0]) and np.array_equal(z2, r[1][:, 0]) and np.array_equal(v, r[2][:, 0])
                assert it == r[3][0]
        c = srv.counters()
        p50, p99, n = srv.stats()
    assert n == 7 and 0 < p50 <= p99
    assert c["admitted"] == 3 and c["evolutions_done"] == 3 and c["newton_done"] == 4
    assert c["peak_bytes"] <= budget
    assert c["peak_resident"] <= 2 and c["ensemble_builds"] >= 2


def test_late_arrival_joins_running_batch():
    """A request submitted while another evolution runs is admitted into the running ensemble
    (two resident at once) and still matches its stand-alone result."""
    import hqt
    with hqt.Server(1 << 40, quantum=5) as srv:
        a = srv.submit_evolution(_state(64, 12, 6.0, 21), 1500, 64, 64, 12, L, L, mu=6.0, dt=0.02)
        time.sleep(0.05)
        b = srv.submit_evolution(_state(48, 10, 6.0, 22), 15, 48, 48, 10, L, L, mu=6.0, dt=0.02)
        st_b, lat_b, out_b = srv.result(b, timeout=60.0)
        st_a, lat_a, out_a = srv.result(a, timeout=60.0)
        c = srv.counters()
    assert st_a == 0 and st_b == 0
    assert c["peak_resident"] == 2
    for x, y in zip(out_b, _alone(48, 10, 6.0, "fp64", 22, 15)):
        assert np.array_equal(x, y)
    for x, y in zip(out_a, _alone(64, 12, 6.0, "fp64", 21, 1500)):
        assert np.array_equal(x, y)


def test_oversize_and_nan_requests():
    import hqt
    budget = hqt.footprint(64, 64, 12) + 1
    with hqt.Server(budget, quantum=4) as srv:
        big = srv.submit_evolution(_state(96, 12, 6.0, 1), 5, 96, 96, 12, L, L)
        bad_state = list(_state(64, 12, 6.0, 2))
        bad_state[3] = bad_state[3].copy()
        bad_state[3][1, 2, 3] = np.nan
        bad = srv.submit_evolution(tuple(bad_state), 10, 64, 64, 12, L, L)
        good = srv.submit_evolution(_state(64, 12, 6.0, 3), 10, 64, 64, 12, L, L)
        st_big, _, _ = srv.result(big, timeout=60.0)
        st_bad, _, _ = srv.result(bad, timeout=60.0)
        st_good, _, out = srv.result(good, timeout=60.0)
    assert st_big == hqt.HQT_ERR_OOM
    assert st_bad == hqt.HQT_ERR_NAN
    assert st_good == 0
    for x, y in zip(out, _alone(64, 12, 6.0, "fp64", 3, 10)):
        assert np.array_equal(x, y)


def test_mixed_size_newton_batch_equals_alone():
    """Continued Newton requests of different radial sizes, submitted together, run in one problem
    list (nz <= 41 with the Jacobian in LDS and nz > 41 in global memory): each result is bitwise the
    one hqt_newton_vy_continue_batch gives for that problem alone."""
    import hqt
    from hqt import loadgen
    cases = [(16, 7.3), (24, 9.1), (33, 6.4), (41, 8.8), (44, 7.7), (48, 9.6)]
    print("mixed-size Newton batch", cases, flush=True)
    alone = []
    for nz, mu in cases:
        Z1, Z2, V = loadgen.newton_guess(nz, mu)
        alone.append(hqt.NewtonRaphson_vy_batch(nz, 1, 0, 0, Z1, Z2, V, np.array([VY_REF]), mu_seed=6.0, dmu=0.25))
    with hqt.Server(1 << 30, quantum=5) as srv:
        ids = []
        for nz, mu in cases:
            Z1, Z2, V = loadgen.newton_guess(nz, mu)
            ids.append(srv.submit_newton(nz, 1, 0, 0, Z1, Z2, V, VY_REF, mu_seed=6.0, dmu=0.25))
        outs = [srv.result(i, timeout=60.0) for i in ids]
        c = srv.counters()
    for (nz, mu), (st, lat, (z1, z2, v, it, err)), r in zip(cases, outs, alone):
        assert st == 0 and r[5][0]
        assert np.array_equal(z1, r[0][:, 0]) and np.array_equal(z2, r[1][:, 0]) and np.array_equal(v, r[2][:, 0])
        assert it == r[3][0] and err == r[4][0]
    assert c["newton_continued"] == len(cases)
    assert c["newton_solves"] == sum(int(r[6][0]) for r in alone)
