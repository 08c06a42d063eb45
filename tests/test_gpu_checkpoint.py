"""Checkpoints of a device-resident evolution (hqt_checkpoint_save / _save_async): the file
holds exactly the state hqt_get_state returns at the time of the call, also when the write
runs on a host thread while the evolution continues, and for an ensemble member."""

import numpy as np
import pytest
import scipy.io

pytestmark = pytest.mark.gpu


def _stepper(n=64, nz=12, precision="fp64", seed=5):
    import hqt
    from hqt import ic
    st = hqt.Stepper(n, n, nz, 50.0, 50.0, mu=6.0, dt=0.02, precision=precision,
                     ax_b=ic.sinusoidal_drive(n, n, 50.0, 7, 0.1))
    st.set_state(*ic.initial_state(n, n, nz, 50.0, 50.0, mu=6.0, seed=seed, n_vortices=4))
    return st


def _read(path):
    d = scipy.io.loadmat(str(path))
    return d["Z1"], d["Z2"], d["Z3"], d["Z4"], d["rhoi"]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_sync_and_async_checkpoints(tmp_path, precision):
    import hqt
    with _stepper(precision=precision) as st:
        st.step(10)
        ref10 = st.get_state()
        st.save_checkpoint(tmp_path / "sync.mat")
        st.save_checkpoint(tmp_path / hqt.checkpoint_name(0.2), wait=False)   # snapshot at step 10
        st.step_async(10)                                                     # state moves on meanwhile
        st.checkpoint_wait()
        st.sync()
        ref20 = st.get_state()
    for a, b in zip(_read(tmp_path / "sync.mat"), ref10):
        assert np.array_equal(a, b)
    for a, b in zip(_read(tmp_path / "check_point_t=0.2.mat"), ref10):
        assert np.array_equal(a, b)
    assert not np.array_equal(ref10[0], ref20[0])


def test_async_checkpoint_of_ensemble_member(tmp_path):
    import hqt
    a, b = _stepper(seed=1), _stepper(n=48, nz=8, seed=2)
    with hqt.Ensemble([a, b]) as ens:
        ens.step_async(6)
        a.save_checkpoint(tmp_path / "a6.mat", wait=False)     # ordered after the queued ensemble steps
        ens.step_async(6)                                      # ordered after the snapshot
        a.checkpoint_wait()
        ens.sync()
    with _stepper(seed=1) as alone:
        alone.step(6)
        ref6 = alone.get_state()
        alone.step(6)
        ref12 = alone.get_state()
    for x, y in zip(_read(tmp_path / "a6.mat"), ref6):
        assert np.array_equal(x, y)
    for x, y in zip(a.get_state(), ref12):
        assert np.array_equal(x, y)
    a.close()
    b.close()
