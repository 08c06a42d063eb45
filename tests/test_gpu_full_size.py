"""The full 301 x 301 x 20 grid on the GPU, through the C ABI.

* Step parity against the oracle: the sampled fixtures of ``tests/golden/make_full_size.py``
  (1 and 3 steps) for the driven-flow benchmark workload (L = 50, kk = 7 sinusoidal A_x drive)
  and for the reference's own run (L = (nx-1)/2 = 150, uniform A_y = -vy, 1e-2 amplitude noise;
  ``scripts/simulation/tevol_driven_cluster.m:23-24,81-101,133``).
* U(1) gauge covariance of the right-hand sides (``:72-76``): Psi -> Psi e^{i lam},
  A_i -> A_i + d_i lam gives F1 -> F1 e^{i lam} and leaves F2, F3, F4 unchanged, at full size.
"""

import math
import os

import numpy as np
import pytest

from conftest import GOLDEN, rel_err

pytestmark = pytest.mark.gpu

N, NZ = 301, 20
FIELDS = ("ReZ1", "ImZ1", "Z2", "Z3", "Z4", "rhoi")


def _fields(state):
    z1, z2, z3, z4, rh = state
    return dict(zip(FIELDS, (z1.real, z1.imag, z2, z3, z4, rh)))


def _stepper_for(case):
    """Stepper and initial state of a fixture case (the inputs of make_full_size.case_state)."""
    import hqt
    from hqt import ic
    if case == "driven_L50":
        L = 50.0
        st = hqt.Stepper(N, N, NZ, L, L, mu=6.0, dt=0.02, winding=True,
                         ax_b=ic.sinusoidal_drive(N, N, L, kk=7, amp=0.1), ay_b=0.0)
        st.set_state(*ic.initial_state(N, N, NZ, L, L, mu=6.0, seed=11, n_vortices=4, phase_noise=1e-3))
        return st
    L = (N - 1) / 2
    st = hqt.Stepper(N, N, NZ, L, L, mu=6.0, dt=0.02, winding=True, ax_b=0.0, ay_b=2 * math.pi * 7 / 20)
    st.init_state(ic.load_profile(), amp_noise=1e-2, phase_noise=0.0, seed=5)
    return st


@pytest.mark.parametrize("case", ["driven_L50", "reference_L150"])
def test_full_size_steps_match_oracle(case):
    path = os.path.join(GOLDEN, f"full_301x301x20_{case}.npz")
    ref = np.load(path)
    idx = ref["index"]
    print(f"full-size parity {case}", flush=True)
    with _stepper_for(case) as st:
        for step, todo in ((1, 1), (3, 2)):
            st.step(todo)
            got = _fields(st.get_state())
            for k, v in got.items():
                flat = np.asarray(v).reshape(-1, order="F")
                err = float(np.max(np.abs(flat[idx] - ref[f"s{step}_{k}"]))) / float(ref[f"s{step}_{k}_max"])
                assert err < 1e-11, (case, step, k, err)
                m = float(np.max(np.abs(flat)))
                assert abs(m - float(ref[f"s{step}_{k}_max"])) <= 1e-11 * m, (case, step, k, "max")
        assert st.last_winding == float(ref["W"][-1])


def test_gauge_covariance_full_size():
    """F(Psi e^{i lam}, A_t, A + grad lam) = (e^{i lam} F1, F2, F3, F4)(Psi, A_t, A) at 301^2 x 20.
    The state is the driven workload after one filtered step, so Psi's x/y content stops at
    |k| = 100 and the product with e^{i lam} (content |k| <= 15 above 1e-25) is resolved."""
    L = 50.0
    with _stepper_for("driven_L50") as st:
        st.step(1)
        z1, z2, z3, z4, _ = st.get_state()
        x = np.arange(N) * L / N
        X, Y = x[:, None, None], x[None, :, None]
        a, b, p = 0.3, 0.25, 2 * np.pi / L
        lam = a * np.cos(p * X + 0.4) + b * np.sin(p * Y + 0.1)
        F = st.rhs(z1, z2, z3, z4)
        G = st.rhs(z1 * np.exp(1j * lam), z2, z3 - a * p * np.sin(p * X + 0.4), z4 + b * p * np.cos(p * Y + 0.1))
    assert rel_err(G[0], np.exp(1j * lam) * F[0]) < 1e-11
    for i in (1, 2, 3):
        assert rel_err(G[i], F[i]) < 1e-11, i
