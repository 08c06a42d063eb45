"""Poisson load generator for the serving loop (:class:`hqt.Server`).

Two request types arrive as one Poisson stream of rate ``rate`` (requests/s), each type with
probability 1/2:

* stationary Newton solves (``src/solvers/NewtonRaphson_vy.m``): nz uniform in [16, 48],
  mu ~ U[3.5, 10]; the guess is the reference's mu=6 profile splined to nz with A_t scaled to
  mu (``scripts/legacy/Statinary_state_with_v.m:50``), at the profile's own flow
  vy = -2*pi*7/20.  A single solve from that guess wanders for its 1000 iterations at many mu,
  so each request is solved by continuation in mu from the profile's own mu = 6 in steps of at
  most 0.25 (``Statinary_state_with_v.m:11-17``; :func:`hqt.NewtonRaphson_vy_batch` with
  ``mu_seed``).  Above the end of the superfluid branch at this flow (a fold near mu = 5.79)
  the result is the superfluid solution; below it, the normal phase;
* evolution segments: nx = ny in {64, 96, 128, 192, 256, 301}, nz uniform in [12, 24],
  50-1000 steps, mu ~ U[4.1, 8], kk in 3..9 sinusoidal A_x drive, 4 seeded windings, on the
  reference's default box L = (nx-1)/2 (``tevol_driven_cluster.m:23-24``; grid spacing 0.5).
  On a finer grid (L = 50 at nx = 301, spacing 0.17) the dt = 0.02 scheme grows unstable
  after some 500 steps and the NaN guard stops the run.

Initial states are generated before the clock starts; the arrival process, the request
parameters and every seed come from one NumPy generator, so a run is reproducible.
Latency = completion - submission on the host clock.

:func:`run` drives one open-loop stream; :func:`capacity` submits a burst of one kind at once and
reports the throughput the server sustains when never idle; :func:`sweep` runs the stream at a
list of rates up to and past that capacity.
"""

from __future__ import annotations

import math
import time

import numpy as np

VY_REF = -2 * math.pi * 7 / 20
MU_SEED = 6.0       # A_t(z=0) of the profile the Newton guesses are built from
DMU = 0.25          # largest mu step of the continuation


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


def make_requests(rate: float, duration: float, seed: int = 0, max_evolutions: int | None = None,
                  kinds=("newton", "evo")):
    """The request list ``[(t_arrival, kind, payload), ...]`` of one run; requests of a kind not in
    ``kinds`` are drawn (the stream stays the same) but left out."""
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
            if "newton" in kinds:
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
            if "evo" not in kinds:
                continue
            state = ic.initial_state(n, n, nz, box, box, mu=mu, seed=s, n_vortices=4, phase_noise=1e-3)
            out.append((t, "evo", dict(n=n, nz=nz, steps=steps, mu=mu, kk=kk, L=box, state=state)))
    return out


def _submit(srv, kind, d, precision):
    from . import ic
    if kind == "newton":
        Z1, Z2, V = d["guess"]
        return srv.submit_newton(d["nz"], 1, 0, 0, Z1, Z2, V, VY_REF, mu_seed=MU_SEED, dmu=DMU)
    n = d["n"]
    axb = ic.sinusoidal_drive(n, n, d["L"], d["kk"], 0.1)
    return srv.submit_evolution(d["state"], d["steps"], n, n, d["nz"], d["L"], d["L"], mu=d["mu"], dt=0.02,
                                precision=precision, ax_b=axb)


def _pct(v, q):
    """Nearest-rank percentile."""
    if not v:
        return None
    v = sorted(v)
    return v[min(len(v), max(1, math.ceil(q * len(v)))) - 1]


def _drive(reqs, budget, quantum, precision, paced=True):
    """Submit ``reqs`` (at their arrival times when ``paced``, else all at once), wait for every
    result and return the latency / throughput record."""
    from .api import Server
    srv = Server(budget, quantum=quantum)
    ids = []
    t0 = time.perf_counter()
    for t_arr, kind, d in reqs:
        if paced:
            delay = t0 + t_arr - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
        ids.append((kind, d, _submit(srv, kind, d, precision)))
    t_sub = time.perf_counter() - t0
    offered_points = sum(d["steps"] * d["n"] * d["n"] * d["nz"] for _, kind, d in reqs if kind == "evo")
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
    allv = lat["newton"] + lat["evo"]
    return {
        "precision": precision, "quantum": quantum, "hbm_budget_bytes": budget, "requests": len(ids),
        "submit_span_s": round(t_sub, 3), "wall_s": round(wall, 3), "drain_s": round(wall - t_sub, 3),
        "completed_per_s": round(len(ids) / wall, 3), "p50_s": _pct(allv, 0.5), "p99_s": _pct(allv, 0.99),
        "newton": {"n": len(lat["newton"]), "p50_s": _pct(lat["newton"], 0.5), "p99_s": _pct(lat["newton"], 0.99),
                   "solves_per_s": round(len(lat["newton"]) / wall, 3),
                   "status": {str(k): v for k, v in status["newton"].items()}},
        "evolution": {"n": len(lat["evo"]), "p50_s": _pct(lat["evo"], 0.5), "p99_s": _pct(lat["evo"], 0.99),
                      "steps_per_s": round(steps / wall, 1), "point_updates_per_s": round(points / wall, 1),
                      "status": {str(k): v for k, v in status["evo"].items()}},
        "offered_point_updates": offered_points, "counters": counters,
    }


def run(rate: float = 2.0, duration: float = 15.0, seed: int = 0, budget: int = 64 << 30, quantum: int = 25,
        precision: str = "fp64", max_evolutions: int | None = None):
    """Drive a :class:`hqt.Server` with the Poisson stream; return latency and throughput figures.
    ``completed_per_s`` below the offered rate and a ``drain_s`` that grows with the rate mark a
    stream past the server's capacity."""
    reqs = make_requests(rate, duration, seed, max_evolutions)
    rec = _drive(reqs, budget, quantum, precision)
    return {"rate_per_s": rate, "duration_s": duration, "seed": seed,
            "offered_per_s": round(len(reqs) / duration, 3)} | rec


def capacity(kind: str, count: int, seed: int = 0, budget: int = 64 << 30, quantum: int = 25,
             precision: str = "fp64"):
    """Saturated throughput of one request kind: ``count`` requests of the stream's distribution
    (``kind`` "newton" or "evo") submitted at once, so that the server never waits for work.
    Newton: solves/s; evolutions: steps/s and point updates/s over the whole drain."""
    rng_rate = 1000.0
    reqs = []
    s = seed
    while len(reqs) < count:
        for t, k, d in make_requests(rng_rate, 4 * count / rng_rate, s, kinds=(kind,)):
            if len(reqs) < count:
                reqs.append((0.0, k, d))
        s += 1
    rec = _drive(reqs, budget, quantum, precision, paced=False)
    return {"kind": kind, "count": count, "seed": seed} | rec


def sweep(rates=(1.0, 2.0, 4.0, 8.0), duration: float = 8.0, seed: int = 0, newton_burst: int = 128,
          evo_burst: int = 12, budget: int = 64 << 30, quantum: int = 25, precision: str = "fp64", log=None):
    """Capacity of each kind (bursts), then the mixed stream at each rate: offered vs completed
    requests/s, Newton solves/s, evolution steps/s and p50/p99 latency per rate.  ``load`` is the
    evolution work the stream offers per second (point updates over its arrival span) over the
    burst's capacity: the queue of a stream with load >= 1 grows without bound, and
    ``saturation_rate_per_s`` is the first rate where it does.  ``log``: called with a line of text
    after each part."""
    say = log if log is not None else (lambda msg: None)
    cap_n = capacity("newton", newton_burst, seed, budget, quantum, precision)
    say(f"serving: Newton burst of {newton_burst}: {cap_n['newton']['solves_per_s']} solves/s")
    cap_e = capacity("evo", evo_burst, seed, budget, quantum, precision)
    say(f"serving: evolution burst of {evo_burst}: {cap_e['evolution']['steps_per_s']} steps/s")
    per_rate = []
    sat = None
    cap_pu = cap_e["evolution"]["point_updates_per_s"]
    for r in rates:
        rec = run(rate=r, duration=duration, seed=seed, budget=budget, quantum=quantum, precision=precision)
        short = {k: rec[k] for k in ("rate_per_s", "offered_per_s", "requests", "completed_per_s", "wall_s",
                                      "drain_s", "p50_s", "p99_s")}
        load = rec["offered_point_updates"] / duration / cap_pu if cap_pu else None
        short["load"] = None if load is None else round(load, 3)
        short["newton"] = {k: rec["newton"][k] for k in ("n", "solves_per_s", "p50_s", "p99_s", "status")}
        short["evolution"] = {k: rec["evolution"][k] for k in ("n", "steps_per_s", "point_updates_per_s", "p50_s",
                                                               "p99_s", "status")}
        short["newton_solves_run"] = rec["counters"]["newton_solves"]
        per_rate.append(short)
        say(f"serving: rate {r}/s: load {short['load']}, p50 {short['p50_s']} s, p99 {short['p99_s']} s")
        if sat is None and load is not None and load >= 1.0:
            sat = r
    return {
        "capacity": {
            "newton_solves_per_s": cap_n["newton"]["solves_per_s"],
            "newton_burst": {k: cap_n[k] for k in ("count", "wall_s", "p50_s", "p99_s")} |
                            {"status": cap_n["newton"]["status"], "solves_run": cap_n["counters"]["newton_solves"],
                             "batches": cap_n["counters"]["newton_batches"]},
            "evolution_steps_per_s": cap_e["evolution"]["steps_per_s"],
            "evolution_point_updates_per_s": cap_e["evolution"]["point_updates_per_s"],
            "evolution_burst": {k: cap_e[k] for k in ("count", "wall_s", "p50_s", "p99_s")} |
                               {"status": cap_e["evolution"]["status"], "steps": cap_e["counters"]["steps"],
                                "peak_resident": cap_e["counters"]["peak_resident"]},
        },
        "rates": per_rate,
        "saturation_rate_per_s": sat,
        "duration_s": duration, "seed": seed, "precision": precision, "quantum": quantum,
    }
