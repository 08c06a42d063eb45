"""Poisson load generator for the serving loop (:class:`hqt.Server`).

Two request types arrive as one Poisson stream of rate ``rate`` (requests/s), each type with
probability 1/2:

* stationary Newton solves (``src/solvers/NewtonRaphson_vy.m``): nz uniform in [16, 48],
  mu ~ U[3.5, 10]; the guess is the reference's mu=6 profile splined to nz with A_t scaled to
  mu (``scripts/legacy/Statinary_state_with_v.m:50``), at the profile's own flow
  vy = -2*pi*7/20.  Below mu of about 4.07 the iteration lands on the normal phase;
* evolution segments: nx = ny in {64, 96, 128, 192, 256, 301}, nz uniform in [12, 24],
  50-1000 steps, mu ~ U[4.1, 8], kk in 3..9 sinusoidal A_x drive, 4 seeded windings, on the
  reference's default box L = (nx-1)/2 (``tevol_driven_cluster.m:23-24``; grid spacing 0.5).
  On a finer grid (L = 50 at nx = 301, spacing 0.17) the dt = 0.02 scheme grows unstable
  after some 500 steps and the NaN guard stops the run.

Initial states are generated before the clock starts; the arrival process, the request
parameters and every seed come from one NumPy generator, so a run is reproducible.
Latency = completion - submission on the host clock.
"""

from __future__ import annotations

import math
import time

import numpy as np

VY_REF = -2 * math.pi * 7 / 20


def newton_guess(nz: int, mu: float):
    """Spline of the mu=6 kk=7 profile to nz Chebyshev nodes, A_t scaled to mu."""
    from . import ic
    from .api import chebdif
    p = ic.load_profile()
    z0 = (chebdif(30, 1)[0] - 1) / 2 + 1
    z = (chebdif(nz, 1)[0] - 1) / 2 + 1
    Z1 = ic.spline_interp(z0, p["Z1"], z)
    Z2 = ic.spline_interp(z0, p["Z2"], z) / p["mu"] * mu
    V = ic.spline_interp(z0, p["Vy1"], z)
    return Z1, Z2, V


def make_requests(rate: float, duration: float, seed: int = 0, max_evolutions: int | None = None):
    """The request list ``[(t_arrival, kind, payload), ...]`` of one run."""
    from . import ic
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    nevo = 0
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            break
        if rng.random() < 0.5:
            nz = int(rng.integers(16, 49))
            mu = float(rng.uniform(3.5, 10.0))
            out.append((t, "newton", dict(nz=nz, mu=mu, guess=newton_guess(nz, mu))))
        else:
            if max_evolutions is not None and nevo >= max_evolutions:
                continue
            nevo += 1
            n = int(rng.choice([64, 96, 128, 192, 256, 301]))
            nz = int(rng.integers(12, 25))
            steps = int(rng.integers(50, 1001))
            mu = float(rng.uniform(4.1, 8.0))
            kk = int(rng.integers(3, 10))
            s = int(rng.integers(1 << 30))
            box = (n - 1) / 2
            state = ic.initial_state(n, n, nz, box, box, mu=mu, seed=s, n_vortices=4, phase_noise=1e-3)
            out.append((t, "evo", dict(n=n, nz=nz, steps=steps, mu=mu, kk=kk, L=box, state=state)))
    return out


def run(rate: float = 2.0, duration: float = 15.0, seed: int = 0, budget: int = 64 << 30, quantum: int = 25,
        precision: str = "fp64", max_evolutions: int | None = None):
    """Drive a :class:`hqt.Server` with the Poisson stream; return latency and throughput figures."""
    from . import ic
    from .api import Server
    reqs = make_requests(rate, duration, seed, max_evolutions)
    srv = Server(budget, quantum=quantum)
    ids = []
    t0 = time.perf_counter()
    for t_arr, kind, d in reqs:
        delay = t0 + t_arr - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        if kind == "newton":
            Z1, Z2, V = d["guess"]
            ids.append((kind, d, srv.submit_newton(d["nz"], 1, 0, 0, Z1, Z2, V, VY_REF)))
        else:
            n = d["n"]
            axb = ic.sinusoidal_drive(n, n, d["L"], d["kk"], 0.1)
            ids.append((kind, d, srv.submit_evolution(d["state"], d["steps"], n, n, d["nz"], d["L"], d["L"], mu=d["mu"],
                                                      dt=0.02, precision=precision, ax_b=axb)))
    t_sub = time.perf_counter() - t0
    lat = {"newton": [], "evo": []}
    status = {"newton": {}, "evo": {}}
    steps = 0
    points = 0
    for kind, d, rid in ids:
        st, latency, _ = srv.result(rid)
        lat[kind].append(latency)
        status[kind][st] = status[kind].get(st, 0) + 1
        if kind == "evo" and st == 0:
            steps += d["steps"]
            points += d["steps"] * d["n"] * d["n"] * d["nz"]
    wall = time.perf_counter() - t0
    counters = srv.counters()
    srv.close()

    def pct(v, q):
        if not v:
            return None
        v = sorted(v)
        return v[min(len(v), max(1, math.ceil(q * len(v)))) - 1]

    allv = lat["newton"] + lat["evo"]
    return {
        "rate_per_s": rate, "duration_s": duration, "seed": seed, "precision": precision, "quantum": quantum,
        "hbm_budget_bytes": budget, "requests": len(ids), "submit_span_s": round(t_sub, 3), "wall_s": round(wall, 3),
        "p50_s": pct(allv, 0.5), "p99_s": pct(allv, 0.99),
        "newton": {"n": len(lat["newton"]), "p50_s": pct(lat["newton"], 0.5), "p99_s": pct(lat["newton"], 0.99),
                   "solves_per_s": round(len(lat["newton"]) / wall, 3),
                   "status": {str(k): v for k, v in status["newton"].items()}},
        "evolution": {"n": len(lat["evo"]), "p50_s": pct(lat["evo"], 0.5), "p99_s": pct(lat["evo"], 0.99),
                      "steps_per_s": round(steps / wall, 1), "point_updates_per_s": round(points / wall, 1),
                      "status": {str(k): v for k, v in status["evo"].items()}},
        "counters": counters,
    }
