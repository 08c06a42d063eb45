"""Pins of the mixed-precision error-budget model (oracle/mixed_error_model.py).

* With every FP32 knob off the model is the FP64 step: two loop bodies on a 12 x 10 x 6 grid with a vortex
  pair (winding removal off) agree with oracle.hqt_oracle.step, whose pins are in test_oracle*.py, to 1e-12.
* Each knob rounds: one step with it on differs from the FP64 model by an FP32-sized amount (between 1e-9 and
  1e-5 relative), and the knobs off give exactly the FP64 model's state.
"""

import numpy as np
import pytest

from oracle import hqt_oracle as O
from oracle.mixed_error_model import VARIANTS, MixedModel


def _state():
    from hqt import ic
    return ic.initial_state(12, 10, 6, 20.0, 20.0, mu=6.0, seed=5, n_vortices=2)


def test_fp64_model_is_the_oracle_step():
    z1, z2, z3, z4, rh = _state()
    m = MixedModel(12, 10, 6, 20.0, 6.0, 0.02)
    s = [z1.real.copy(), z1.imag.copy(), z2.copy(), z3.copy(), z4.copy(), rh.copy()]
    ops = O.Operators(12, 10, 6, 20.0, 20.0)
    prm = O.Params(mu=6.0, h=0.02, ax_b=0.0, ay_b=0.0, winding=False)
    ref = (z1, z2, z3, z4, rh)
    for _ in range(2):
        s = m.step(s)
        ref = O.step(ops, prm, *ref)[:5]
    got = (s[0] + 1j * s[1], s[2], s[3], s[4], s[5])
    for a, b in zip(got, ref):
        assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(b))


@pytest.mark.parametrize("variant", ["mixed", "deriv32", "state32", "mixed_filter64", "mixed_filter64_start64", "deriv_io32"])
def test_knobs_round(variant):
    z1, z2, z3, z4, rh = _state()
    s0 = [z1.real.copy(), z1.imag.copy(), z2.copy(), z3.copy(), z4.copy(), rh.copy()]
    ref = MixedModel(12, 10, 6, 20.0, 6.0, 0.02).step([a.copy() for a in s0])
    got = MixedModel(12, 10, 6, 20.0, 6.0, 0.02, **VARIANTS[variant]).step([a.copy() for a in s0])
    err = max(np.max(np.abs(a - b)) / np.max(np.abs(b)) for a, b in zip(got[:5], ref[:5]))
    assert 1e-9 < err < 1e-5, err
