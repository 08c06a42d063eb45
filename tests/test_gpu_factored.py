"""The RK stages' x/y derivative passes in their factored-DFT form (line lengths 32, 64, 96, 128, 192,
256 and 301, forced here; the FP64 default from 192 up: one DFT per source line, the operators'
eigenvalues, one inverse per output) against the even/odd folded matrix products of the same operators
(the fourdif matrices of tevol_driven_cluster.m:60-61, which the oracle parity tests pin).

FP64: F1..F4 on a generic state and the state after 5 steps agree to 1e-12 relative on every supported
line length, mixed with unsupported ones (non-square grids), so each direction takes its own form.
Mixed precision: after 20 steps each field's error against the FP64 run stays within 3x the folded
FP32 form's error (and below 1e-4).

The filter (ff.m:7-11) follows the stages' form: its factored-DFT passes (the even part of the mask as
a spectral multiplier, the rank-4 cross term through the slice matrices added to the y-spectrum at
the unpaired modes) against its folded passes, with the stages factored in both runs: 1 and 5 steps of
a generic (all-mode) state agree to 1e-12 relative, with winding removal off so that only the filter
differs between the runs."""

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

GRIDS = [(301, 301, 20), (64, 48, 8), (32, 33, 6), (96, 128, 7), (192, 256, 5), (256, 64, 4), (45, 301, 6)]


def _state(nx, ny, nz):
    from hqt import ic
    state = ic.initial_state(nx, ny, nz, 50.0, 50.0, mu=6.0, seed=8, n_vortices=2)
    rng = np.random.default_rng(nx + ny)
    shp = (nx, ny, nz)
    gen = (state[0] * (1 + 0.05 * rng.standard_normal(shp)) + 0.01j * rng.standard_normal(shp),
           state[1] + 0.05 * rng.standard_normal(shp), 0.05 * rng.standard_normal(shp),
           state[3] + 0.05 * rng.standard_normal(shp))
    return state, gen


@pytest.mark.parametrize("grid", GRIDS, ids=lambda g: "x".join(map(str, g)))
def test_factored_matches_folded(grid):
    import hqt
    from hqt import ic
    nx, ny, nz = grid
    print("factored vs folded", nx, ny, nz, flush=True)
    state, gen = _state(nx, ny, nz)
    axb = ic.sinusoidal_drive(nx, ny, 50.0, 3, 0.1)
    res, rhs = [], []
    for factored in (True, False):
        with hqt.Stepper(nx, ny, nz, 50.0, 40.0, mu=6.0, dt=0.02, ax_b=axb, factored=factored) as st:
            rhs.append(st.rhs(*gen))
            st.set_state(*state)
            st.step(5)
            res.append(st.get_state())
    for a, b in zip(*rhs):
        assert rel_err(a, b) < 1e-12
    for a, b in zip(*res):
        assert rel_err(a, b) < 1e-12


@pytest.mark.parametrize("grid", [(301, 301, 20), (64, 64, 12)], ids=lambda g: "x".join(map(str, g)))
def test_factored_mixed_accuracy(grid):
    import hqt
    from hqt import ic
    nx, ny, nz = grid
    state, _ = _state(nx, ny, nz)
    axb = ic.sinusoidal_drive(nx, ny, 50.0, 3, 0.1)
    out = {}
    for name, prec, factored in (("ref", "fp64", True), ("fac", "mixed", True), ("fold", "mixed", False)):
        with hqt.Stepper(nx, ny, nz, 50.0, 50.0, mu=6.0, dt=0.02, ax_b=axb, precision=prec, factored=factored,
                         winding=False) as st:
            st.set_state(*state)
            st.step(20)
            out[name] = st.get_state()
    for k in range(5):
        ef = rel_err(out["fac"][k], out["ref"][k])
        eo = rel_err(out["fold"][k], out["ref"][k])
        assert ef < 1e-4 and ef < 3 * eo + 1e-7, (k, ef, eo)


@pytest.mark.parametrize("grid", GRIDS, ids=lambda g: "x".join(map(str, g)))
def test_factored_filter_matches_folded_filter(grid):
    import hqt
    from hqt import ic
    nx, ny, nz = grid
    print("factored vs folded filter", nx, ny, nz, flush=True)
    state, gen = _state(nx, ny, nz)
    rho = state[4]
    axb = ic.sinusoidal_drive(nx, ny, 50.0, 3, 0.1)
    res = {1: [], 5: []}
    for folded_filter in (False, True):
        with hqt.Stepper(nx, ny, nz, 50.0, 40.0, mu=6.0, dt=0.02, ax_b=axb, winding=False, factored=True,
                         folded_filter=folded_filter) as st:
            st.set_state(*gen, rho)
            st.step(1)
            res[1].append(st.get_state())
            st.step(4)
            res[5].append(st.get_state())
    for n in (1, 5):
        for k, (a, b) in enumerate(zip(*res[n])):
            assert rel_err(a, b) < 1e-12, (n, k, rel_err(a, b))
