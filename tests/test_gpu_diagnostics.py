"""Device-side boundary diagnostics (hqt_order_parameter, hqt_vortex_census) against the
oracle: O = (func_Lz(Z1))(:,:,end) (tevol_driven_cluster.m:136,148) and countN
(src/diagnostics/countN.m:7-21) evaluated on the host from the downloaded state."""

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


def _charges_oracle(O):
    """Plaquette charge map (+1/-1/0) with countN's tolerance, from the oracle's w."""
    x1 = np.asarray(O)
    x2 = np.roll(x1, (-1, 0), axis=(0, 1))
    x3 = np.roll(x1, (-1, -1), axis=(0, 1))
    x4 = np.roll(x1, (0, -1), axis=(0, 1))
    with np.errstate(divide="ignore", invalid="ignore"):
        w = (np.angle(x1 / x2) + np.angle(x2 / x3) + np.angle(x3 / x4) + np.angle(x4 / x1)) / (2 * np.pi)
    return (np.abs(w - 1) <= 1e-4).astype(np.int8) - (np.abs(w + 1) <= 1e-4).astype(np.int8)


def test_census_reference_known_answer():
    """tests/hqt_smoke_test.m:147-151: phi = sin(x+0.271) + i sin(y+0.613) on 33x33 has 4
    windings (2 positive, 2 negative).  The state Psi = phi(x,y) * z has O = phi exactly
    (Lz is exact on the linear polynomial z)."""
    import hqt
    from oracle import hqt_oracle as O
    n, nz = 33, 8
    x, y = np.meshgrid(np.arange(n) * 2 * np.pi / n, np.arange(n) * 2 * np.pi / n, indexing="ij")
    phi = np.sin(x + 0.271) + 1j * np.sin(y + 0.613)
    z = (O.chebdif(nz, 1)[0] + 1) / 2
    Z1 = phi[:, :, None] * z[None, None, :]
    zero = np.zeros((n, n, nz))
    with hqt.Stepper(n, n, nz, 50.0, 50.0, mu=6.0, dt=0.02) as st:
        st.set_state(Z1, zero, zero, zero, zero)
        Od = st.order_parameter()
        assert rel_err(Od, phi) < 1e-13
        cn, cp, cm, omax, omean, cmap = st.vortex_census(charges=True)
    assert (cn, cp, cm) == (4, 2, 2) == O.countN(phi)
    assert np.array_equal(cmap, _charges_oracle(Od))
    assert abs(omax - np.max(np.abs(Od))) <= 1e-15 * omax
    assert abs(omean - np.mean(np.abs(Od))) <= 1e-13 * omean


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_census_matches_oracle_on_evolved_state(precision):
    """64x64x12 with 8 seeded windings, 40 steps: device O, counts and charge map equal the
    oracle's countN of O computed from the downloaded state."""
    import hqt
    from hqt import ic
    from oracle import hqt_oracle as O
    nx = ny = 64
    nz = 12
    L = 50.0
    state = ic.initial_state(nx, ny, nz, L, L, mu=6.0, seed=3, n_vortices=8, phase_noise=1e-3)
    with hqt.Stepper(nx, ny, nz, L, L, mu=6.0, dt=0.02, precision=precision) as st:
        st.set_state(*state)
        st.step(40)
        Z1 = st.get_state()[0]
        Od = st.order_parameter()
        cn, cp, cm, omax, omean, cmap = st.vortex_census(charges=True)
    g = ic.Grid(nx, ny, nz, L, L)
    Oh = np.tensordot(Z1, g.Lz[-1, :], axes=([2], [0]))
    assert rel_err(Od, Oh) < 1e-12
    assert (cn, cp, cm) == O.countN(Oh)
    assert cn >= 2
    assert np.array_equal(cmap, _charges_oracle(Oh))
    assert abs(omax - np.max(np.abs(Oh))) <= 1e-12 * omax


def test_census_of_normal_state_is_empty():
    """Psi = 0: every plaquette ratio is 0/0 (NaN), so nothing is counted (countN.m:4)."""
    import hqt
    n, nz = 16, 6
    zero = np.zeros((n, n, nz))
    with hqt.Stepper(n, n, nz, 50.0, 50.0, mu=6.0, dt=0.02) as st:
        st.set_state(zero.astype(complex), zero + 6.0, zero, zero, zero + 6.0)
        cn, cp, cm, omax, omean, cmap = st.vortex_census(charges=True)
    assert (cn, cp, cm) == (0, 0, 0) and omax == 0.0 and omean == 0.0
    assert not cmap.any()
