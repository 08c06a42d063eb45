"""Benchmark: RK4 steps/s of the AdS4 stepper on one MI355X.

Workload (the reference's headline configuration): driven flow, nx=ny=301, nz=20, box side
L=50, mu=6, dt=0.02, boundary forcing A_x(z=0) = 0.1*sin(2*pi*7*y/L), seeded zero-flow
stationary profile with 1e-3 phase noise and 4 windings.  One "step" is one full RK4 step
(tevol_driven_cluster.m:120-142: 4 RHS stages, A_t constraint, boundary reset, winding
removal, filter, NaN guard) with the state resident in HBM.

Prints ONE JSON line.  ``value`` is FP64 steps/s; the mixed-precision form is reported beside
it with its per-field error against the FP64 run after the same steps.  Beside them:
``ensemble`` (many independent evolutions of uneven size stepped as one batch) and
``serving`` (the serving loop's saturated capacity for Newton solves and for evolution steps,
then a Poisson stream of both at a sweep of rates, with p50/p99 latency per rate).  ``roofline`` is the
dominant kernel's algorithmic FLOP rate from HIP events recorded around its launches on the
library's stream (a profiled repeat of the timed steps); ``cpu_baseline`` is the NumPy oracle
(the reference's dense-matrix form) timed on a bounded sample on this host.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "ads4-holographic-quantum-turbulence_amd"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)

PEAK_FP64_TFLOPS = 78.6      # MI355X dense FP64 (vector = matrix)
PEAK_FP32_TFLOPS = 157.3     # MI355X dense FP32 matrix
PEAK_HBM_GBS = 8000.0        # MI355X HBM3E


def line_pass_apps(name, stage_mix=(1, 3)):
    """Operator applications of one launch of a line pass (dense and rank-4 terms), averaged over
    the RK stages for the stage passes (stage 1 adds the derivative of A_t)."""
    if name == "line_x_stage":      # dx, dxx of Re/Im Psi; dx Ax; dx, dxx Ay; (+ dx At in stage 1)
        return (8 * stage_mix[0] + 7 * stage_mix[1]) / sum(stage_mix), 0
    if name == "line_y_stage":      # dy, dyy of Re/Im Psi; dy Ay; dyy Ax + dy(dxAy); dy(dxAx) (+ dy At)
        return (9 * stage_mix[0] + 8 * stage_mix[1]) / sum(stage_mix), 0
    if name in ("line_x_filter", "line_y_filter"):   # real part dense, imaginary part rank 4, on 5 fields
        return 5, 5
    if name in ("line_x_cons", "line_y_cons"):
        return 1, 0
    return None, None


def line_pass_flops(name, nx, ny, nz, folded=True):
    """Algorithmic FLOPs of one launch of a line pass.  Dense form (the reference's func_Lx: a
    dense N x N matrix times every line): 2 N flops per point per application.  Folded form (the
    kernels' even/odd split): each output pair (i, N-i) takes the even and the odd half-size dot
    products of N/2 + 1 terms, i.e. 2 (N/2 + 1) flops per point per application.  A rank-4 term
    (the filter's imaginary part) costs 2 x 2 x 4 flops per point (dots and expansion)."""
    S = nx * ny * nz
    N = nx if "_x_" in name else ny
    dense, lr = line_pass_apps(name)
    if dense is None:
        return None
    per = 2 * (N // 2 + 1) if folded else 2 * N
    lr_flops = 16 if folded else 2 * N          # the dense form applies the imaginary part densely
    return dense * per * S + lr * lr_flops * S


def line_pass_fields(name, stage_mix=(1, 3)):
    """Fields one launch of a line pass reads (sources and addends) and writes, averaged over the RK stages
    for the stage passes (stage 1 adds the derivatives of A_t)."""
    if name == "line_x_stage":      # reads Re/Im Psi, A_x, A_y (+ A_t); writes 7 (+ dx A_t)
        return (13 * stage_mix[0] + 11 * stage_mix[1]) / sum(stage_mix)
    if name == "line_y_stage":      # reads 5 sources + 4 addends (+ A_t); writes 7 (+ dy A_t)
        return (18 * stage_mix[0] + 16 * stage_mix[1]) / sum(stage_mix)
    if name in ("line_x_filter", "line_y_filter"):
        return 10
    if name == "zstage":            # 16 field reads + RK bookkeeping (4 evolved fields: 2 reads, 2 writes each)
        return 32
    if name == "line_cons":         # A_x, A_y in; d_x A_x, d_y A_y out
        return 4
    if name == "zpost":             # Psi (2), d_x A_x, d_y A_y, rho in; A_t out
        return 6
    return None


def factored_transforms(name, stage_mix=(1, 3)):
    """DFTs (forward + inverse) of one launch of an FP64 factored-DFT stage pass, per line.  The FP64 passes
    give every output its own work-group, which transforms the output's sources itself (one forward and one
    inverse per output, two forwards for the two-source output)."""
    if name == "line_x_stage":      # dx, dxx of R and I; dx A_x; dx, dxx A_y: 7 x (1 + 1) (+ dx A_t: 1 + 1)
        return (16 * stage_mix[0] + 14 * stage_mix[1]) / sum(stage_mix)
    if name == "line_y_stage":      # dy, dyy of R and I; dy A_y; dyy A_x - dy(dx A_y): 2 + 1; dy(dx A_x) (+ dy A_t)
        return (17 * stage_mix[0] + 15 * stage_mix[1]) / sum(stage_mix)
    if name in ("line_x_filter", "line_y_filter"):   # 5 fields: 1 + 1 each
        return 10
    return None


def factored_flops_per_point(N1, N2):
    """FLOPs per point of one factored real DFT of length N1 N2 (forward or inverse): stage A, the even and
    odd halves of N1 real DFTs of length N2, (H+1) H + HI^2 multiply-adds each; stage B, per kept column
    the twiddle (4 flops per point) and the folded complex DFT of length N1 (4 multiply-adds per term and
    output pair)."""
    H, HI, H1, HI1 = N2 // 2, (N2 - 1) // 2, N1 // 2, (N1 - 1) // 2
    stage_a = 2 * N1 * ((H + 1) * H + HI * HI)
    stage_b = (H + 1) * (4 * N1 + 2 * 4 * HI1 * (H1 + 1))
    return (stage_a + stage_b) / (N1 * N2)


def profiled_traffic(kernel):
    """HBM bytes per launch of ``kernel`` from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary
    (calibrated against a plain copy kernel), or None when no summary covers it."""
    import glob
    import re
    for f in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        for line in open(f):
            m = re.match(r"^traffic_bytes_per_launch\s+(\S+)\s+(\S+)", line)
            if m and m.group(1) == kernel:
                return float(m.group(2))
    return None


def step_flops(nx, ny, nz, folded):
    tot = 0.0
    for k, n in (("line_x_stage", 4), ("line_y_stage", 4), ("line_x_filter", 1), ("line_y_filter", 1),
                 ("line_x_cons", 1), ("line_y_cons", 1)):
        tot += n * line_pass_flops(k, nx, ny, nz, folded)
    return tot


def run_gpu(nx, ny, nz, L, precision, steps, warmup, state, axb, profile_steps, dense_lines=False, folded=False):
    import hqt
    st = hqt.Stepper(nx, ny, nz, L, L, mu=6.0, dt=0.02, winding=True, filter=True, precision=precision,
                     ax_b=axb, ay_b=0.0, dense_lines=dense_lines, factored=False if folded else None)
    st.set_state(*state)
    st.step(warmup)
    st.sync()
    t0 = time.perf_counter()
    st.step_async(steps)
    st.sync()
    t1 = time.perf_counter()
    final = st.get_state()
    stats = {}
    if profile_steps:
        st.profile(True)
        st.step(profile_steps)
        stats = st.kernel_stats()
        st.profile(False)
    st.close()
    return (t1 - t0), final, stats


def progress(msg):
    """One progress line on stderr (stdout carries only the JSON line)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def cpu_info():
    """CPU model, logical CPUs visible to this process and the BLAS/OpenMP thread settings."""
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = os.cpu_count()
    env = {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS")}
    return {"model": model, "logical_cpus": os.cpu_count(), "affinity_cpus": affinity, "env": env}


def cpu_baseline(nx, ny, nz, L, state, axb, warmup=2, timed=5):
    """The oracle (the reference's dense-matrix form, NumPy/OpenBLAS) on this host: ``warmup``
    untimed steps, then ``timed`` steps timed one by one; returns the median step time's rate and
    the per-step times."""
    from oracle import hqt_oracle as O
    ops = O.Operators(nx, ny, nz, L, L)
    prm = O.Params(mu=6.0, h=0.02, ax_b=axb, ay_b=0.0, winding=True)
    S = list(state)
    for i in range(warmup):
        t0 = time.perf_counter()
        *S, _w = O.step(ops, prm, *S)
        progress(f"cpu oracle warm-up step {i + 1}/{warmup}: {time.perf_counter() - t0:.1f} s")
    times = []
    for i in range(timed):
        t0 = time.perf_counter()
        *S, _w = O.step(ops, prm, *S)
        times.append(time.perf_counter() - t0)
        progress(f"cpu oracle timed step {i + 1}/{timed}: {times[-1]:.1f} s")
    return 1.0 / float(np.median(times)), times


def ensemble_leg(members, steps, seed=4):
    """Batched ensemble of independent evolutions of uneven size: nx = ny in {64, 96, 128, 192,
    256}, nz in {12, 16, 20, 24}, mu ~ U[4.1, 8] (below mu ~ 4.07 the zero-flow state is the
    normal phase), kk in 3..9 drive, 8-64 windings, each built on the device.  Timed: `steps`
    steps of the whole ensemble (replays of its 8-step graph) vs the same members stepped one by
    one (each its own graph and stream, the streams running concurrently) and one after another
    (each member's steps to completion before the next member starts, as separate runs go).
    The warm-up builds both of the ensemble's graphs."""
    import hqt
    from hqt import ic
    rng = np.random.default_rng(seed)
    L = 50.0
    specs = []
    for i in range(members):
        n = int(rng.choice([64, 96, 128, 192, 256]))
        nz = int(rng.choice([12, 16, 20, 24]))
        specs.append((n, nz, float(rng.uniform(4.1, 8.0)), int(rng.integers(3, 10)), int(rng.integers(4, 33)) * 2,
                      int(rng.integers(1 << 30))))
    pts = sum(n * n * nz for n, nz, *_ in specs)
    out = {"members": members, "steps": steps, "point_count": pts,
           "sizes": "nx=ny in {64,96,128,192,256}, nz in {12,16,20,24}, mu~U[4.1,8], kk 3..9, 8-64 windings"}
    profiles = [ic.zero_flow_profile(mu) for _, _, mu, *_ in specs]

    def fresh(sts, warm):
        # every timed run starts from the members' initial states: the finest grids (dx ~ 0.2 at L = 50) turn
        # unstable after ~900 steps, as the reference's loop does at that resolution
        for st, prof, (n, nz, mu, kk, nv, s) in zip(sts, profiles, specs):
            r = np.random.default_rng(s)
            vort = (r.uniform(0, L, nv), r.uniform(0, L, nv), np.where(np.arange(nv) % 2 == 0, 1.0, -1.0))
            st.init_state(prof, phase_noise=1e-3, seed=s, vortices=vort)
        for st in sts:
            if warm:
                st.step(warm)

    for prec in ("fp64", "mixed"):
        sts = [hqt.Stepper(n, n, nz, L, L, mu=mu, dt=0.02, precision=prec, ax_b=ic.sinusoidal_drive(n, n, L, kk, 0.1))
               for n, nz, mu, kk, nv, s in specs]
        fresh(sts, 2)
        t0 = time.perf_counter()
        for st in sts:
            st.step_async(steps)
        for st in sts:
            st.sync()
        tseq = (time.perf_counter() - t0) / steps
        fresh(sts, 2)
        t0 = time.perf_counter()
        for st in sts:          # one after another, as separate runs would go
            st.step(steps)
        tser = (time.perf_counter() - t0) / steps
        fresh(sts, 0)
        ens = hqt.Ensemble(sts)
        ens.step(8)
        t0 = time.perf_counter()
        ens.step(steps)
        tens = (time.perf_counter() - t0) / steps
        progress(f"ensemble {prec}: {1e3 * tens:.2f} ms/step, one by one {1e3 * tseq:.2f}, sequential {1e3 * tser:.2f}")
        out[prec] = {"ms_per_step": round(1e3 * tens, 3), "member_steps_per_s": round(members / tens, 1),
                     "point_updates_per_s": round(pts / tens, 1), "launches_per_step": ens.launches_per_step,
                     "members_one_by_one_ms_per_step": round(1e3 * tseq, 3), "speedup": round(tseq / tens, 3),
                     "members_sequential_ms_per_step": round(1e3 * tser, 3),
                     "speedup_vs_sequential": round(tser / tens, 3)}
        ens.close()
        for st in sts:
            st.close()
    # the batching case proper: many small evolutions (the 64x64x12 FP64 workload, 48 copies with
    # their own seeds), where a single evolution cannot fill the GPU
    sts = []
    for i in range(members):
        st = hqt.Stepper(64, 64, 12, L, L, mu=6.0, dt=0.02)
        r = np.random.default_rng(1000 + i)
        st.init_state(ic.zero_flow_profile(6.0), phase_noise=1e-3, seed=1000 + i,
                      vortices=(r.uniform(0, L, 4), r.uniform(0, L, 4), np.array([1.0, -1.0, 1.0, -1.0])))
        sts.append(st)
    for st in sts:
        st.step(2)
    t0 = time.perf_counter()
    for st in sts:
        st.step_async(steps)
    for st in sts:
        st.sync()
    tseq = (time.perf_counter() - t0) / steps
    with hqt.Ensemble(sts) as ens:
        ens.step(8)
        t0 = time.perf_counter()
        ens.step(steps)
        tens = (time.perf_counter() - t0) / steps
    out["small_64x64x12_fp64"] = {"members": members, "ms_per_step": round(1e3 * tens, 3),
                                  "member_steps_per_s": round(members / tens, 1),
                                  "members_one_by_one_ms_per_step": round(1e3 * tseq, 3),
                                  "speedup": round(tseq / tens, 3)}
    for st in sts:
        st.close()
    return out


def mixed_accuracy_leg(checkpoints=(500, 5000), seed=3):
    """The mixed-precision contract on the vortex-lattice workload: nx = ny = 64, nz = 12, L = 50,
    mu = 6, dt = 0.02, 64 +-1 windings on an 8 x 8 jittered square lattice (jitter 0.5), winding
    removal off (tevol_lattice.m:168-170).  The FP64 and mixed forms step side by side; at each
    checkpoint (t = 10 and t = 100) the per-field relative max-norm and L2 errors of the mixed state
    against the FP64 one, the condensate |O| error (max norm and L2 norm over the boundary plane, and of
    its spatial mean),
    and the vortex counts and plaquette positions.  Beside
    them, ``fp64_perturbed``: the same FP64 evolution from an initial Psi perturbed by 1e-7 relative
    (seeded, the size of an FP32 rounding), which measures how far the dynamics carry a perturbation
    of that size by each checkpoint."""
    import hqt
    from hqt import diagnostics, ic
    nx = ny = 64
    nz, L = 12, 50.0
    S = ic.initial_state(nx, ny, nz, L, L, mu=6.0, seed=seed, lattice=(8, 0.5))
    g = ic.Grid(nx, ny, nz, L, L)
    out = {"workload": "64x64x12 L=50 mu=6 dt=0.02, 8x8 lattice of +-1 windings (jitter 0.5), winding removal off",
           "checkpoints": {}}
    sts = {p: hqt.Stepper(nx, ny, nz, L, L, mu=6.0, dt=0.02, precision=p, winding=False) for p in ("fp64", "mixed")}
    sts["fp64_perturbed"] = hqt.Stepper(nx, ny, nz, L, L, mu=6.0, dt=0.02, precision="fp64", winding=False)
    Sp = list(S)
    Sp[0] = S[0] * (1 + 1e-7 * np.random.default_rng(seed + 1).standard_normal(S[0].shape))
    try:
        for p, st in sts.items():
            st.set_state(*(Sp if p == "fp64_perturbed" else S))
        done = 0
        for n in checkpoints:
            for st in sts.values():
                st.step_async(n - done)
            states = {}
            for p, st in sts.items():
                st.sync()
                states[p] = st.get_state()
            done = n
            O64 = diagnostics.order_parameter(g, states["fp64"][0])
            Om = diagnostics.order_parameter(g, states["mixed"][0])
            Op = diagnostics.order_parameter(g, states["fp64_perturbed"][0])
            c64, cm = diagnostics.countN(O64), diagnostics.countN(Om)
            out["checkpoints"]["t=%g" % (n * 0.02)] = {
                "steps": n, "err_vs_fp64": field_errors(states["mixed"], states["fp64"]),
                "err_vs_fp64_l2": field_errors_l2(states["mixed"], states["fp64"]),
                "O_abs_rel_err": float(np.max(np.abs(np.abs(Om) - np.abs(O64))) / np.max(np.abs(O64))),
                "O_abs_rel_err_l2": float(np.linalg.norm((np.abs(Om) - np.abs(O64)).ravel())
                                          / np.linalg.norm(np.abs(O64).ravel())),
                "condensate_mean_rel_err": float(abs(np.abs(Om).mean() - np.abs(O64).mean()) / np.abs(O64).mean()),
                "vortices_fp64": c64, "vortices_mixed": cm,
                "positions_equal": diagnostics.vortex_positions(Om) == diagnostics.vortex_positions(O64),
                "fp64_perturbed": {"err_vs_fp64": field_errors(states["fp64_perturbed"], states["fp64"]),
                                   "O_abs_rel_err": float(np.max(np.abs(np.abs(Op) - np.abs(O64))) / np.max(np.abs(O64)))}}
    finally:
        for st in sts.values():
            st.close()
    return out


def radial_ab(reps=200):
    """The radial contraction (u = Lz x, v = Lz u per column) in the stage kernel's FP64 MFMA form and as vector
    FMAs fed through LDS, on the padded boundary planes of 301^2 (nz = 20) and 64^2 (nz = 12): average launch time,
    algorithmic FP64 FLOPs (2 nz^2 per column per matvec) and HBM bytes (x in, u and v out) per second."""
    import ctypes
    import hqt
    out = []
    for ncol, nz in ((304 * 301, 20), (64 * 64, 12)):
        row = {"columns": ncol, "nz": nz}
        for form, name in ((0, "mfma"), (1, "valu_lds")):
            ms, ss = ctypes.c_double(0.0), ctypes.c_double(0.0)
            hqt._lib.check(hqt.lib().hqt_calibration_radial(ncol, nz, form, reps, ctypes.byref(ms), ctypes.byref(ss)))
            us = ms.value * 1e3
            row[name] = {"us": round(us, 2), "tflops": round(2 * 2 * nz * nz * ncol / (us * 1e-6) / 1e12, 3),
                         "gbps": round(3 * 8 * nz * ncol / (us * 1e-6) / 1e9, 1), "sumsq": ss.value}
        row["mfma_speedup"] = round(row["valu_lds"]["us"] / row["mfma"]["us"], 3)
        out.append(row)
    return out


def field_errors(a, b):
    names = ("Psi", "A_t", "A_x", "A_y", "rho")
    out = {}
    for nm, u, v in zip(names, a, b):
        out[nm] = float(np.max(np.abs(u - v)) / max(np.max(np.abs(v)), 1e-300))
    return out


def field_errors_l2(a, b):
    """Per-field relative L2 error ||a - b||_2 / ||b||_2 over the whole volume."""
    names = ("Psi", "A_t", "A_x", "A_y", "rho")
    return {nm: float(np.linalg.norm((u - v).ravel()) / max(np.linalg.norm(v.ravel()), 1e-300))
            for nm, u, v in zip(names, a, b)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=301)
    ap.add_argument("--nz", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-warmup", type=int, default=2, help="untimed oracle steps before the CPU timing")
    ap.add_argument("--cpu-steps", type=int, default=5, help="timed oracle steps (their median is reported)")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision leg")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched-ensemble leg")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="skip the mixed-precision accuracy leg (64^2 vortex lattice to t = 100)")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving (Poisson arrivals) leg")
    ap.add_argument("--ensemble-members", type=int, default=48)
    ap.add_argument("--ensemble-steps", type=int, default=500)
    ap.add_argument("--serving-rates", default="1,2,4,8",
                    help="Poisson arrival rates of the serving sweep, requests/s (comma-separated)")
    ap.add_argument("--serving-seconds", type=float, default=8.0, help="arrival span of each rate of the sweep")
    ap.add_argument("--folded", action="store_true",
                    help="RK-stage x/y derivatives as folded matrix products instead of the factored-DFT form")
    ap.add_argument("--dense-lines", action="store_true",
                    help="x/y operators as dense Toeplitz products (the reference's form) instead of the folded form")
    ap.add_argument("--radial-ab", action="store_true",
                    help="only the radial contraction A/B (FP64 MFMA slot matvec vs vector FMAs through LDS), one "
                         "JSON line")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MB",
                    help="also run the plain copy kernel on MB megabytes (traffic-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(independent evolutions would shard one per GPU)", file=sys.stderr)
        sys.exit(2)

    import ctypes

    import hqt
    from hqt import diagnostics, ic

    if args.calibrate_copy:
        ms = ctypes.c_double(0.0)
        hqt._lib.check(hqt.lib().hqt_calibration_copy(int(args.calibrate_copy) << 20, ctypes.byref(ms)))

    if args.radial_ab:
        print(json.dumps({"radial_ab": radial_ab()}), flush=True)
        return

    nx = ny = args.n
    nz = args.nz
    L = 50.0
    state = ic.initial_state(nx, ny, nz, L, L, mu=6.0, seed=1, n_vortices=4, phase_noise=1e-3)
    axb = ic.sinusoidal_drive(nx, ny, L, kk=7, amp=0.1)
    S = nx * ny * nz

    prof_steps = max(5, min(args.steps, 50))
    t64, fin64, stats = run_gpu(nx, ny, nz, L, "fp64", args.steps, args.warmup, state, axb, prof_steps, args.dense_lines,
                                args.folded)
    sps = args.steps / t64
    progress(f"fp64 leg: {sps:.1f} steps/s")

    # dominant kernel roofline (HIP events on the library's stream)
    kernels = {}
    for k, (cnt, ms) in stats.items():
        kernels[k] = {"launches": cnt, "avg_us": 1e3 * ms / cnt, "ms_per_step": ms / prof_steps}
    dom = max(stats.items(), key=lambda kv: kv[1][1])[0] if stats else None
    roof = None
    fact = {301: (7, 43), 32: (4, 8), 64: (8, 8), 96: (8, 12), 128: (8, 16), 192: (8, 24), 256: (8, 32)}
    factored_form = not args.dense_lines and not args.folded and nx == ny and nx in fact and nx >= 128
    if dom is not None and line_pass_fields(dom) is not None:
        cnt, ms = stats[dom]
        avg_s = ms / cnt / 1e3
        alg_bytes = line_pass_fields(dom) * S * 8
        ach = alg_bytes / avg_s / 1e9
        suffix = "_dense" if args.dense_lines else "_folded" if args.folded else ""
        tr = profiled_traffic(dom + suffix) if (nx, ny, nz) == (301, 301, 20) else None
        roof = {"kernel": dom, "bound": "hbm", "achieved": round(ach, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(ach / PEAK_HBM_GBS, 4), "traffic": tr, "avg_launch_us": round(avg_s * 1e6, 2),
                "algorithmic_bytes_per_launch": alg_bytes,
                "bytes_form": "fields read (sources, addends) + fields written, 8 B per point each"}
        # the streaming rate a plain copy kernel reaches on this GPU (1 GiB in, 1 GiB out; best of 3)
        cms = []
        for _ in range(3):
            m = ctypes.c_double(0.0)
            hqt._lib.check(hqt.lib().hqt_calibration_copy(1 << 30, ctypes.byref(m)))
            cms.append(m.value)
        copy_gbs = 2.0 * (1 << 30) / (min(cms) * 1e-3) / 1e9
        roof["copy_GBps"] = round(copy_gbs, 1)
        roof["frac_of_copy"] = round(ach / copy_gbs, 4)
        if factored_form and factored_transforms(dom) is not None:
            fl = factored_transforms(dom) * factored_flops_per_point(*fact[nx]) * S
            form = "factored DFT (%d x %d): each transform %.1f flops per point" % (fact[nx] + (factored_flops_per_point(*fact[nx]),))
        else:
            fl = line_pass_flops(dom, nx, ny, nz, folded=not args.dense_lines)
            form = ("dense operator, 2 N flops per point per application" if args.dense_lines else
                    "folded (even/odd) operator, 2 (N/2+1) flops per point per application")
        if fl is not None:
            roof["flops"] = {"achieved": round(fl / avg_s / 1e12, 3), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                             "frac": round(fl / avg_s / 1e12 / PEAK_FP64_TFLOPS, 4), "flops_per_launch": fl, "form": form}

    step_bytes = 368.0 * S          # compulsory FP64 bytes per step (4 stages + post-step)
    out = {
        "metric": "RK4 steps/sec @301^2x20 (FP64 vs mixed, stated err) + p50/p99 request latency",
        "value": round(sps, 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * t64 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded stationary profile + phase noise + 4 windings)",
        "config": {"workload": "driven flow, FP64, nx=ny=%d nz=%d L=50 mu=6 dt=0.02 kk=7 sinusoidal A_x drive" % (nx, nz),
                   "nx": nx, "ny": ny, "nz": nz, "L": L, "mu": 6.0, "dt": 0.02, "kk": 7, "drive_amp": 0.1},
        "point_updates_per_s": round(sps * S, 1),
        "step_hbm": {"algorithmic_bytes_per_step": step_bytes,
                     "achieved_GBps": round(step_bytes * sps / 1e9, 2),
                     "frac_of_peak": round(step_bytes * sps / 1e9 / PEAK_HBM_GBS, 4)},
        "step_flops": {"folded_form": step_flops(nx, ny, nz, True), "dense_form": step_flops(nx, ny, nz, False),
                       "factored_stages": (4 * (factored_transforms("line_x_stage") + factored_transforms("line_y_stage"))
                                           * factored_flops_per_point(*fact[nx]) * S
                                           + step_flops(nx, ny, nz, True)
                                           - 4 * (line_pass_flops("line_x_stage", nx, ny, nz) + line_pass_flops("line_y_stage", nx, ny, nz)))
                       if nx == ny and nx in fact else None},
        "roofline": roof,
        "kernels": kernels,
    }
    O64 = None
    try:
        # boundary order parameter and vortex count of the FP64 final state
        g = ic.Grid(nx, ny, nz, L, L)
        O64 = diagnostics.order_parameter(g, fin64[0])
        out["fp64_final"] = {"vortices": diagnostics.countN(O64), "O_abs_mean": float(np.mean(np.abs(O64)))}
    except Exception as e:  # diagnostics are reported, never fatal
        out["fp64_final"] = {"error": repr(e)}

    if not args.no_mixed:
        tm, finm, mstats = run_gpu(nx, ny, nz, L, "mixed", args.steps, args.warmup, state, axb, prof_steps,
                                   args.dense_lines)
        err = field_errors(finm, fin64)
        mixed = {"value": round(args.steps / tm, 3), "unit": "steps/s", "speedup_vs_fp64": round(t64 / tm, 3),
                 "err_vs_fp64_after_steps": err,
                 "kernels": {k: {"launches": c, "avg_us": 1e3 * ms / c, "ms_per_step": ms / prof_steps}
                             for k, (c, ms) in mstats.items()}}
        if O64 is not None:
            Om = diagnostics.order_parameter(g, finm[0])
            mixed["O_abs_rel_err"] = float(np.max(np.abs(np.abs(Om) - np.abs(O64))) / np.max(np.abs(O64)))
            mixed["vortices"] = diagnostics.countN(Om)
        out["mixed"] = mixed
        progress(f"mixed leg: {mixed['value']} steps/s")

    if not args.no_accuracy and not args.no_mixed:
        out["mixed_accuracy"] = mixed_accuracy_leg()
        progress("mixed accuracy leg done")

    if not args.no_ensemble:
        out["ensemble"] = ensemble_leg(args.ensemble_members, args.ensemble_steps)
        progress("ensemble leg done")

    if not args.no_serving:
        from hqt import loadgen
        rates = tuple(float(r) for r in args.serving_rates.split(",") if r.strip())
        out["serving"] = loadgen.sweep(rates=rates, duration=args.serving_seconds, seed=0, log=progress)

    if not args.no_cpu:
        cps, times = cpu_baseline(nx, ny, nz, L, state, axb, warmup=args.cpu_warmup, timed=args.cpu_steps)
        info = cpu_info()
        out["cpu_baseline"] = {"value": round(cps, 5), "unit": "steps/s", "cores": info["affinity_cpus"],
                               "kind": "port",
                               "sample": f"median of {len(times)} timed steps after {args.cpu_warmup} warm-up steps of "
                                         f"the same workload, NumPy/OpenBLAS dense-matrix restatement (oracle), "
                                         f"all threads of this process's CPU set",
                               "step_s": [round(t, 3) for t in times],
                               "projected_5000_steps_s": round(5000 / cps, 1), "cpu": info}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
