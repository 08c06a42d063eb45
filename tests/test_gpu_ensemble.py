"""Batched ensemble (hqt_ensemble_*): members of uneven grids, nz, mu, drive, switches and
precision stepped together must give exactly (bitwise) what each member gives stepped alone
through hqt_step_n, which the parity tests tie to the oracle."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

L = 50.0

# (nx, ny, nz, mu, precision, winding, filter, drive amplitude)
MEMBERS = [
    (48, 48, 8, 6.0, "fp64", True, True, 0.0),
    (64, 64, 12, 6.0, "fp64", True, True, 0.1),
    (96, 80, 16, 5.0, "mixed", True, True, 0.05),
    (64, 64, 12, 7.0, "fp64", False, False, 0.0),
    (33, 40, 10, 4.5, "fp64", True, True, 0.0),
    (128, 128, 20, 6.0, "mixed", False, True, 0.1),
]


def _make(spec, seed):
    import hqt
    from hqt import ic
    nx, ny, nz, mu, prec, wind, filt, amp = spec
    axb = ic.sinusoidal_drive(nx, ny, L, kk=5, amp=amp) if amp else 0.0
    st = hqt.Stepper(nx, ny, nz, L, L, mu=mu, dt=0.02, winding=wind, filter=filt, precision=prec, ax_b=axb)
    st.set_state(*ic.initial_state(nx, ny, nz, L, L, mu=mu, seed=seed, n_vortices=4, phase_noise=1e-3))
    return st


def _alone(specs, nsteps):
    out = []
    for i, spec in enumerate(specs):
        with _make(spec, 100 + i) as st:
            st.step(nsteps)
            out.append(st.get_state())
    return out


@pytest.mark.parametrize("graph", [True, False])
def test_ensemble_equals_members_alone(graph):
    import hqt
    nsteps = 12
    ref = _alone(MEMBERS, nsteps)
    sts = [_make(spec, 100 + i) for i, spec in enumerate(MEMBERS)]
    with hqt.Ensemble(sts, graph=graph) as ens:
        assert ens.launches_per_step < 20 * len(MEMBERS)
        ens.step(nsteps)
        for i, st in enumerate(sts):
            got = st.get_state()
            for a, b in zip(got, ref[i]):
                assert np.array_equal(a, b), f"member {i} {MEMBERS[i]}"
    for st in sts:
        st.close()


def test_ensemble_many_members_per_class():
    """Twelve members, two or more per kernel class and lane: the four lanes run on their own
    streams between the step's fork and join.  Still bitwise equal to each member stepped alone,
    over repeated graph replays."""
    import hqt
    specs = MEMBERS + MEMBERS[::-1]
    nsteps = 7
    ref = _alone(specs, nsteps)
    sts = [_make(spec, 100 + i) for i, spec in enumerate(specs)]
    with hqt.Ensemble(sts, graph=True) as ens:
        for _ in range(nsteps):
            ens.step(1)
            ens.sync()
        for i, st in enumerate(sts):
            got = st.get_state()
            for a, b in zip(got, ref[i]):
                assert np.array_equal(a, b), f"member {i} {specs[i]}"
    for st in sts:
        st.close()


def test_member_calls_are_ordered_after_ensemble_steps():
    """Ensemble steps then the member's own steps: same as stepping the member alone."""
    import hqt
    spec = MEMBERS[1]
    ref = _alone([spec], 9)[0]
    a = _make(spec, 100)
    b = _make(MEMBERS[0], 101)
    ens = hqt.Ensemble([a, b])
    ens.step_async(5)
    a.step(4)                       # waits for the ensemble's queued steps first
    got = a.get_state()
    for x, y in zip(got, ref):
        assert np.array_equal(x, y)
    ens.close()
    a.close()
    b.close()


def test_nan_member_stops_alone():
    """A member with a NaN in its state trips its own guard at step 1; the others go on and
    match their stand-alone results."""
    import hqt
    specs = MEMBERS[:3]
    ref = _alone(specs, 6)
    sts = [_make(spec, 100 + i) for i, spec in enumerate(specs)]
    bad = _make(MEMBERS[0], 7)
    Z = list(bad.get_state())
    Z[2][3, 4, 2] = np.nan
    bad.set_state(*Z)
    ens = hqt.Ensemble(sts + [bad])
    ens.step_async(6)
    ns = ens.nan_steps()
    assert list(ns) == [0, 0, 0, 1]
    with pytest.raises(hqt.HQTError) as ei:
        ens.sync()
    assert ei.value.code == hqt.HQT_ERR_NAN
    for i, st in enumerate(sts):
        for x, y in zip(st.get_state(), ref[i]):
            assert np.array_equal(x, y)
    ens.close()
    for st in sts + [bad]:
        st.close()


def test_ensemble_rejects_bad_membership():
    import hqt
    a = _make(MEMBERS[0], 1)
    ens = hqt.Ensemble([a])
    with pytest.raises(hqt.HQTError) as ei:
        hqt.Ensemble([a])                   # already a member elsewhere
    assert ei.value.code == hqt.HQT_ERR_STATE
    ens.close()
    with pytest.raises(hqt.HQTError) as ei:
        hqt.Ensemble([a, a])
    assert ei.value.code == hqt.HQT_ERR_ARG
    a.close()


def test_member_diagnostics_wait_for_ensemble_steps():
    """A member's order parameter and census read right after queueing ensemble steps see the
    state after those steps (the same values as read after the ensemble's sync)."""
    import hqt
    a = _make(MEMBERS[1], 100)
    b = _make(MEMBERS[4], 101)
    ens = hqt.Ensemble([a, b])
    ens.step_async(15)
    O_early = a.order_parameter()
    c_early = b.vortex_census()
    ens.sync()
    assert np.array_equal(O_early, a.order_parameter())
    assert c_early == b.vortex_census()
    with _make(MEMBERS[1], 100) as alone:
        alone.step(15)
        assert np.array_equal(O_early, alone.order_parameter())
    ens.close()
    a.close()
    b.close()
