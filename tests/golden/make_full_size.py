"""Write ``full_301x301x20_<case>.npz``: the oracle's state after 1 and 3 RK4 steps at the full
301 x 301 x 20 grid, sampled, for two workloads.

* ``driven_L50``: the driven-flow benchmark workload: L = 50, mu = 6, dt = 0.02, the kk = 7
  sinusoidal drive A_x(z=0) = 0.1 sin(2 pi 7 y / L), winding removal on, filter on.  Initial
  fields ``hqt.ic.initial_state(301, 301, 20, 50, 50, mu=6, seed=11, n_vortices=4,
  phase_noise=1e-3)`` (zero-flow stationary profile, 4 windings, 1e-3 phase noise).
* ``reference_L150``: the reference's own run, ``scripts/simulation/tevol_driven_cluster.m``:
  box L = L0 = (nx - 1)/2 = 150 (:23-24,85), the kk = 7 stationary profile
  ``data/reference/mu=6,nz=30,vy=7_20.mat`` with its flow A_y = -Vy1 (:81-98), the uniform
  boundary value A_y(z=0) = -vy = 2 pi 7 / 20 (:86,133), winding removal and filter on, and the
  1e-2 amplitude noise of :101 drawn from the seeded Philox stream of the device initial-state
  builder (``oracle.hqt_oracle.device_initial_state(..., amp_noise=1e-2, seed=5)``).

Outputs per case: 4096 seeded sample positions (flat index over [nx, ny, nz], column-major),
Re Z1, Im Z1, Z2, Z3, Z4 and rhoi at those positions after step 1 and step 3, each field's max
|value| over the full grid (the normaliser of the tests' relative error), and the winding
number removed at each step.  About 4 minutes of CPU (37 s per oracle step).

Run from the repository root:  python tests/golden/make_full_size.py
"""

import math
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "ads4-holographic-quantum-turbulence_amd"))
sys.path.insert(0, ROOT)

from hqt import ic  # noqa: E402
from oracle import hqt_oracle as O  # noqa: E402

N, NZ, NSAMP = 301, 20, 4096
HERE = os.path.dirname(os.path.abspath(__file__))


def case_state(case):
    """(ops, params, initial fields) of a case; the GPU tests rebuild the same inputs."""
    if case == "driven_L50":
        L = 50.0
        ops = O.Operators(N, N, NZ, L, L)
        S = list(ic.initial_state(N, N, NZ, L, L, mu=6.0, seed=11, n_vortices=4, phase_noise=1e-3))
        prm = O.Params(mu=6.0, h=0.02, ax_b=ic.sinusoidal_drive(N, N, L, kk=7, amp=0.1), ay_b=0.0, winding=True)
        return ops, prm, S
    if case == "reference_L150":
        L = (N - 1) / 2
        ops = O.Operators(N, N, NZ, L, L)
        prof = ic.load_profile()
        S = list(O.device_initial_state(ops, prof, amp_noise=1e-2, phase_noise=0.0, seed=5))
        prm = O.Params(mu=6.0, h=0.02, ax_b=0.0, ay_b=2 * math.pi * 7 / 20, winding=True)
        return ops, prm, S
    raise ValueError(case)


def sample_index():
    rng = np.random.default_rng(2024)
    return np.sort(rng.choice(N * N * NZ, size=NSAMP, replace=False))


def main(cases=("driven_L50", "reference_L150")):
    idx = sample_index()
    for case in cases:
        ops, prm, S = case_state(case)
        out = {"index": idx}
        Ws = []
        for step in (1, 2, 3):
            *S, W = O.step(ops, prm, *S)
            Ws.append(W)
            print(case, "step", step, "W", W, flush=True)
            if step in (1, 3):
                z1, z2, z3, z4, rh = S
                fields = {"ReZ1": z1.real, "ImZ1": z1.imag, "Z2": z2, "Z3": z3, "Z4": z4, "rhoi": rh}
                for k, v in fields.items():
                    flat = np.asarray(v, dtype=np.float64).reshape(-1, order="F")
                    out[f"s{step}_{k}"] = flat[idx].copy()
                    out[f"s{step}_{k}_max"] = np.array(np.max(np.abs(flat)))
        out["W"] = np.array(Ws)
        path = os.path.join(HERE, f"full_301x301x20_{case}.npz")
        np.savez_compressed(path, **out)
        print("wrote", path, flush=True)


if __name__ == "__main__":
    main(tuple(sys.argv[1:]) or ("driven_L50", "reference_L150"))
