"""The time-step limit of the reference scheme (CPU, no GPU).

``tests/golden/instability_32x32x12.npz`` (written by ``tests/golden/make_instability.py``) holds
the oracle's long runs at two grid spacings with dt = 0.02: at spacing 0.166 (the spacing of the
301^2 grid on L = 50) the state turns non-finite after some 800 steps; at spacing 0.48 (the
reference's own boxes, L = (nx-1)/2) it stays finite for 1500 steps.  This test re-derives the
first 100 steps of both runs from the current oracle, so the fixture cannot drift from it.
"""

import os

import numpy as np
import pytest

from conftest import GOLDEN


@pytest.fixture(scope="module")
def ref():
    return np.load(os.path.join(GOLDEN, "instability_32x32x12.npz"))


def test_fixture_states_the_limit(ref):
    assert 0 < int(ref["fine_nan_step"]) <= 1000
    assert float(ref["fine_L"]) / 32 < 0.17
    assert int(ref["coarse_nan_step"]) == 0 and ref["coarse_rows"][-1][0] == 1500
    assert float(ref["coarse_L"]) / 32 > 0.48
    # the fine run grows: before the NaN, max |A_t| has left mu = 6 by more than tenfold and max |A_x|
    # has grown more than tenfold; the coarse run keeps max |A_t| = mu
    rows = ref["fine_rows"]
    assert rows[-1][2] > 60 and rows[-1][3] > 10 * rows[3][3]
    assert np.all(np.abs(ref["coarse_rows"][:, 2] - 6.0) < 1e-9)


@pytest.mark.parametrize("case", ["fine", "coarse"])
def test_fixture_prefix_is_the_oracle(ref, case):
    import sys
    sys.path.insert(0, GOLDEN)
    import make_instability as M
    nan_step, rows = M.run(float(ref[f"{case}_L"]), 1500, upto=100)
    assert nan_step == 0
    assert np.allclose(rows, ref[f"{case}_rows"][:len(rows)], rtol=1e-12, atol=1e-14)
