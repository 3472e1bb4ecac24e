"""Throughput benchmark of the MI355X BEM hot path.

Workload (one GPU): vertical cylinder with a hemispherical cap, radius 4 m,
draft 12 m, 4,096 panels (64 sectors x (40 side + 24 cap rows)), 64
frequencies uniform in 0.1-3.0 rad/s, 6 rigid dofs about the origin, heading
beta = 0, rho = 1025, g = 9.81, direct method.

A step is one sweep of all 64 frequencies, each an omega-solve: wave-term
assembly (the wavenumber-independent Rankine part is cached per body), one LU,
all right-hand sides (6 radiation + 1 diffraction) and pressure integration
into A, B and F_ex; value = 64 K / elapsed for K timed steps.  The mesh, dofs,
the cached Rankine part, the network weights and the batch buffers (allocated
by the warmup sweep at the same batch size) are resident on the device before
the timed region.

The headline path is the learned surrogate of the exact Guevel-Delhommeau term
(Wu fit + BF16-MFMA network residual, exact FP64 fallback near the free
surface) with the 64 frequencies of a sweep as one batch.  The plain GPU path
(FP64 throughout) is reported four ways on 4 of the frequencies: exact term one
frequency at a time (the denominator of speedup_vs_plain), the Wu term one
frequency at a time, and both with the reference's schedule (assembly and LU
per problem).  The headline path's coefficients are checked against the exact
path at those frequencies (relative error reported).  The CPU baseline is the
oracle on the host cores in two labelled modes (best CPU; reference-faithful,
sampled and extrapolated).

Prints one JSON line (last line of stdout).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "marinehydro.jl_amd"))

METRIC = "omega-solves/s at 4k panels: surrogate+batched vs plain GPU, at <=1e-3 rel err"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
BF16_PEAK_TFLOPS = 2500.0


WAVE_DESC = {"Surrogate": "learned surrogate (Wu + BF16 MFMA residual network, exact FP64 fallback)",
             "ExactGuevelDelhommeau": "exact FP64 Guevel-Delhommeau", "Wu": "FP64 Liang-Wu-Noblesse fit"}
DTYPE = {"Surrogate": "f64 (wave-term residual network: bf16 MFMA, f32 accumulate; S and D stored in FP32; "
                      "dense solve: FP32 LU with split-bf16 MFMA trailing updates, FP64-summed residuals and "
                      "refinement to 1e-12)",
         "ExactGuevelDelhommeau": "f64", "Wu": "f64"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload():
    from mhx import meshes
    mesh = meshes.capped_cylinder(4.0, 8.0, 64, 40, 24)
    omegas = np.linspace(0.1, 3.0, 64)
    return mesh, omegas


def host_cpu():
    """(logical CPUs the OS reports, CPU model) of the host."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return os.cpu_count() or 1, model


def cpu_baseline(mesh, omegas, nfreq, threads, sample_rows=32):
    """The oracle (C restatement of the reference) on the host cores, in the two modes of the
    CPU-baseline plan, both for complete omega-solves of the bench body (6 dofs, beta = 0):
      best CPU: nfreq complete omega-solves, exact assembly of S and D once per frequency on
        `threads` OpenMP threads, LAPACK LU and multi-RHS solve through numpy's OpenBLAS (timed in full);
      reference-faithful: the reference's schedule (solve.jl:75-78: assembly and LU per
        problem, 7 problems per frequency) with single-threaded assembly as the reference's serial
        loops; the single-thread assembly is timed on `sample_rows` seeded rows and scaled to n
        rows (the oracle evaluates every entry on its own, so the cost is linear in the rows), the
        LU and one-column solve are timed in full.
    Returns (best, faithful)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    import scipy.linalg as sl
    O.build()
    nproc, model = host_cpu()
    om = O.Mesh.from_mhx(mesh)
    n = mesh.nfaces
    ws = omegas[np.linspace(0, len(omegas) - 1, nfreq + 2).round().astype(int)[1:-1]]
    rng = np.random.default_rng(0)
    rhs = rng.standard_normal((n, 7)) + 1j * rng.standard_normal((n, 7))
    O.set_threads(threads)
    t_asm = t_lu = 0.0
    D = S = None
    for w in ws:
        t0 = time.perf_counter()
        S, D = O.assemble(om, w * w / 9.81, gf=O.GF_DELHOMMEAU, direct=True)
        t1 = time.perf_counter()
        lu = sl.lu_factor(D, check_finite=False)
        sl.lu_solve(lu, S @ rhs, check_finite=False)
        t2 = time.perf_counter()
        t_asm += t1 - t0
        t_lu += t2 - t1
    per = (t_asm + t_lu) / len(ws)
    host = f"host: {nproc} logical CPUs ({model}); {threads} threads used (the host's CPU share for one GPU)"
    best = {"value": 1.0 / per, "unit": "omega-solves/s", "cores": threads, "kind": "port", "mode": "best CPU",
            "sample": f"{len(ws)} complete omega-solves at omega = {np.round(ws, 3).tolist()} on {n} panels: "
                      f"exact assembly of S and D once per frequency (oracle, {threads} OpenMP threads, "
                      f"{t_asm:.2f} s) + LAPACK zgetrf/zgetrs and S.B for 7 right-hand sides (numpy's OpenBLAS, "
                      f"{t_lu:.2f} s); {host}"}
    # reference-faithful: single-thread assembly on seeded rows, per-problem LU
    O.set_threads(1)
    w = float(ws[0])
    rows = np.sort(np.random.default_rng(1).choice(n, size=sample_rows, replace=False))
    t0 = time.perf_counter()
    for i in rows:
        O.assemble_rows(om, w * w / 9.81, (int(i), int(i) + 1), gf=O.GF_DELHOMMEAU, direct=True)
    t_rows = time.perf_counter() - t0
    t_full_asm = t_rows * n / sample_rows
    t0 = time.perf_counter()
    lu = sl.lu_factor(D, check_finite=False)
    sl.lu_solve(lu, S @ rhs[:, :1], check_finite=False)
    t_lu1 = time.perf_counter() - t0
    per_f = 7 * (t_full_asm + t_lu1)
    faithful = {"value": 1.0 / per_f, "unit": "omega-solves/s", "cores": 1, "kind": "port",
                "mode": "reference-faithful (sampled, extrapolated)",
                "sample": f"single-thread exact assembly of {sample_rows} seeded rows at omega = {w:.3f} "
                          f"({t_rows:.2f} s) scaled to {n} rows ({t_full_asm:.1f} s per assembly), LAPACK "
                          f"zgetrf/zgetrs + S.b for one right-hand side ({t_lu1:.2f} s, numpy's OpenBLAS, "
                          f"{threads} threads), 7 problems per frequency each with its own assembly and LU "
                          f"(solve.jl:75-78); {host}"}
    return best, faithful


def cpu_baseline_fs(mesh, U, w, beta, threads):
    """The oracle on the host cores for one moving (U, omega, beta) group of the forward-speed
    workload: exact indirect assembly of S, K and of the x-derivative operator K_x at the
    encountered wavenumber (OpenMP), LAPACK LU of K, sigma for 7 right-hand sides, phi = S sigma and
    dphi/dx = K_x sigma (numpy's OpenBLAS)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    import scipy.linalg as sl
    O.build()
    O.set_threads(threads)
    om = O.Mesh.from_mhx(mesh)
    n = mesh.nfaces
    we, ke, _ = O.encountered(w, beta, U, 9.81)
    rng = np.random.default_rng(0)
    rhs = rng.standard_normal((n, 7)) + 1j * rng.standard_normal((n, 7))
    t0 = time.perf_counter()
    S, K = O.assemble(om, ke, gf=O.GF_DELHOMMEAU, direct=False)
    _, Kx = O.assemble_fixed_normal(om, ke, [1.0, 0.0, 0.0], gf=O.GF_DELHOMMEAU, direct=False)
    t1 = time.perf_counter()
    sigma = sl.lu_solve(sl.lu_factor(K, check_finite=False), rhs, check_finite=False)
    S @ sigma
    Kx @ sigma
    t2 = time.perf_counter()
    return {"value": 1.0 / (t2 - t0), "unit": "problem groups/s", "cores": threads, "kind": "port",
            "sample": f"one moving group (U={U}, omega={w}, beta={beta:.4f}) on {n} panels: exact assembly of S, K "
                      f"and K_x (oracle, {threads} OpenMP threads, {t1 - t0:.2f} s) + LAPACK zgetrf/zgetrs, "
                      f"S.sigma and K_x.sigma (numpy's OpenBLAS, {t2 - t1:.2f} s)"}


def surrogate_flops_per_pair():
    """Multiply-adds of one network evaluation (64 -> 4 x 128 -> 6), two flops each."""
    return 2.0 * (64 * 128 + 3 * 128 * 128 + 128 * 6)


def profiled_value(name, key="traffic_bytes_per_launch"):
    """A figure of the committed summary profiles/<name> (a line "key: value"): by default the HBM
    bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE (calibrated on a copy kernel)
    for the default workload."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key + ":"):
            return float(line.split(":")[1])
    return None


def lu_update_steps(n, nb=64, two_level=False, super_=4):
    """(K, KE) of the LU's timed trailing updates: one per 64-column panel, or (split updates) one
    per group of up to super_ panels that leaves a trailing matrix."""
    steps, K = [], 0
    while K < n:
        S = 1
        if two_level:
            while S < super_ and K + (S + 1) * nb < n:
                S += 1
        if S > 1:
            steps.append((K, K + S * nb))
            K += S * nb
        else:
            KE = min(n, K + nb)
            if KE < n:
                steps.append((K, KE))
            K = KE
    return steps


def lu_update_flops(n, nb=64, two_level=False, super_=4):
    """Algorithmic flops of the trailing updates of one blocked LU (complex multiply-add = 8 flops)."""
    return sum(8.0 * (n - KE) ** 2 * (KE - K) for K, KE in lu_update_steps(n, nb, two_level, super_))


def lu_update_bytes(n, nb=64, tile=128, super_=4):
    """Algorithmic HBM bytes of the split-bf16 trailing updates of one blocked LU (multi-panel
    steps): the trailing matrix read and written once per update (FP32 complex, 16 B per
    element) and the update's split panels (8 bf16 arrays of mpad x ks) read once."""
    b = 0.0
    for K, KE in lu_update_steps(n, nb, True, super_):
        m = n - KE
        mpad = (m + tile - 1) // tile * tile
        ks = (KE - K + nb - 1) // nb * nb
        b += 16.0 * m * m + 8.0 * mpad * ks * 2
    return b


def lu_kind(mode, gf):
    """Trailing-update kernel the dense solver uses: 'split' (bf16 hi/lo MFMA), 'fp32' or 'fp64'."""
    if mode == "mixed_bf16" or (mode == "auto" and gf == "Surrogate"):
        return "split"
    return "fp32" if mode == "mixed" else "fp64"


def measure(ctx, run, ws, n, gf, nlu=None, units=None, kind="fp64", calls=1):
    """`calls` timed calls of run(ws) with the library's phase timing off (the rate), then one
    call with it on: the phase table and the roofline of the dominant kernel (per-launch HIP
    events on the library's stream) come from that call.  nlu: factorizations per call (default
    one per frequency); units: what the rate counts per call (default the frequencies); kind: the
    trailing-update kernel (lu_kind)."""
    ctx.reset_timing()
    ctx.set_timing(False)
    t0 = time.perf_counter()
    for _ in range(calls):
        A, B, Fd, Fk = run(ws)
    elapsed = time.perf_counter() - t0
    assert np.all(np.isfinite(A)) and np.all(np.isfinite(Fd))
    rate = (len(ws) if units is None else units) * calls / elapsed
    # the instrumented call
    ctx.reset_timing()
    ctx.set_timing(True)
    run(ws)
    phases = {p: ctx.phase_time(p) for p in ("wave_table", "wave", "lu", "lu_update", "solve", "forces", "rankine",
                                             "forward_speed")}
    counts, flops = ctx.exact_work()
    tw, cw = phases["wave_table"]
    tu, cu = phases["lu_update"]
    tl, cl = phases["lu"]
    nw1 = len(ws) if nlu is None else nlu       # factorizations per call
    nw = nw1
    _, panels = ctx.lu_config()                  # panels per split update the library used
    sp = panels if panels else 4
    fu = lu_update_flops(n, two_level=kind == "split", super_=sp) * nw
    if kind == "split":
        # the split-bf16 update moves the trailing matrix once per step: HBM-bound
        bu = lu_update_bytes(n, super_=sp) * nw
        traffic = None
        items = nw / cu * len(lu_update_steps(n, two_level=True, super_=sp)) if cu else 0
        prof = "lu_update_split_4096x64_traffic.md"
        if n == 4096 and items and profiled_value(prof, "panels_per_update") == sp:
            # profiled at 32 matrices per launch and the same update width; the traffic scales with the matrices
            t32 = profiled_value(prof)
            traffic = t32 / 32 * items if t32 else None
        rl_update = {"bound": "hbm", "kernel": "zgemm_split_kernel (LU trailing update, split-bf16 MFMA)",
                     "achieved": bu / (tu * 1e-3) / 1e9 if tu else None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "traffic": traffic, "avg_launch_ms": tu / cu if cu else None, "launches": cu,
                     "matrices_per_launch": items,
                     "algorithmic_bytes_per_launch": bu / cu if cu else None,
                     "complex_tflops": fu / (tu * 1e-3) / 1e12 if tu else None}
        if rl_update["achieved"]:
            rl_update["frac"] = rl_update["achieved"] / HBM_PEAK_GBS
    else:
        fp32 = kind == "fp32"
        lu_peak = FP32_PEAK_TFLOPS if fp32 else FP64_PEAK_TFLOPS
        rl_update = {"bound": "mfma",
                     "kernel": "zgemm_mfma_kernel<%s> (LU trailing update)" % ("cplxf" if fp32 else "cplx"),
                     "achieved": fu / (tu * 1e-3) / 1e12 if tu else None, "peak": lu_peak, "unit": "TFLOP/s",
                     "traffic": None, "avg_launch_ms": tu / cu if cu else None, "launches": cu,
                     "algorithmic_flops_per_launch": fu / cu if cu else None}
        if rl_update["achieved"]:
            rl_update["frac"] = rl_update["achieved"] / lu_peak
    if gf == "Surrogate":
        pairs = n * (n + 1) / 2 * nw
        ach = pairs * surrogate_flops_per_pair() / (tw * 1e-3) / 1e12
        traffic = None
        if n == 4096 and cw and abs(nw / cw - 64) < 1e-9:   # the profiled configuration: 64 frequencies per launch
            traffic = profiled_value("surrogate_table_4096x64_traffic.md")
        rl_table = {"bound": "mfma", "kernel": "surr_table_kernel (+ exact fallback)", "achieved": ach,
                    "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": ach / BF16_PEAK_TFLOPS, "traffic": traffic,
                    "avg_launch_ms": tw / cw, "launches": cw, "frequencies_per_launch": nw / cw,
                    "algorithmic_flops_per_launch": pairs * surrogate_flops_per_pair() / cw}
    elif gf == "ExactGuevelDelhommeau":
        ach = flops / (tw * 1e-3) / 1e12
        rl_table = {"bound": "valu", "kernel": "delhommeau_table_kernel", "achieved": ach,
                    "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": ach / FP64_PEAK_TFLOPS, "traffic": None,
                    "avg_launch_ms": tw / cw, "launches": cw, "frequencies_per_launch": nw / cw,
                    "work_per_frequency": {"integrand_evals": int(counts[0] // nw), "series_steps": int(counts[1] // nw),
                                           "cf_steps": int(counts[2] // nw)}}
    else:
        rl_table = None
    # dominant kernel by time
    if rl_table is not None and tw >= tu:
        roof, other = rl_table, rl_update
    else:
        roof, other = rl_update, rl_table
    lu_tflops = 8.0 / 3.0 * n ** 3 * nw / (tl * 1e-3) / 1e12 if tl else None
    per_w = {k: v[0] / nw for k, v in phases.items()}
    return rate, elapsed, per_w, roof, lu_tflops, (A, B, Fd, Fk), other


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10,
                    help="timed steps; a step is one sweep of all 64 frequencies (64 omega-solves)")
    ap.add_argument("--warmup", type=int, default=1, help="untimed sweeps before the timed ones")
    ap.add_argument("--gf", default="Surrogate", choices=["Surrogate", "ExactGuevelDelhommeau", "Wu"])
    ap.add_argument("--schedule", default="batched", choices=["batched", "graph", "per_omega"])
    ap.add_argument("--cpu-freqs", type=int, default=2, help="complete omega-solves timed on the host for cpu_baseline")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain GPU comparison runs")
    ap.add_argument("--lu-mode", default="auto", choices=["auto", "fp64", "mixed", "mixed_bf16"],
                    help="dense solver precision (auto: mixed FP32 factor for the surrogate, FP64 otherwise)")
    ap.add_argument("--lu-streams", type=int, default=None,
                    help="streams a batched FP32 factorization is split over (1-4; default: the library's)")
    ap.add_argument("--freqs", type=int, default=64,
                    help="frequencies per step, spread over the 64 of the workload (fewer only for profiling)")
    ap.add_argument("--activation", default="silu", choices=["silu", "relu"],
                    help="surrogate network: the 4 x 128 SiLU network (the workload's), or its ReLU variant")
    ap.add_argument("--workload", default="headline", choices=["headline", "forward_speed", "multibody"],
                    help="forward_speed: the forward-speed grid on the same body (IndirectBEM, exact or Wu term); "
                         "multibody: the seeded uneven 8-body workload as one graph-captured sweep")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this benchmark measures one GPU; --gpus {args.gpus} is not supported "
              "(frequency items would shard across GPUs, one process per GPU)", file=sys.stderr)
        sys.exit(2)

    import mhx
    from mhx import _lib

    modes = {"auto": _lib.LU_AUTO, "fp64": _lib.LU_FP64, "mixed": _lib.LU_MIXED, "mixed_bf16": _lib.LU_MIXED_BF16}
    mesh, omegas = workload()
    n = mesh.nfaces
    ctx = _lib.Context(rho=1025.0, g=9.81, device=0, timing=True)
    mhx.api._contexts[0] = ctx
    mhx.set_rho(1025.0)
    mhx.set_g(9.81)
    names = ["Surge", "Sway", "Heave", "Roll", "Pitch", "Yaw"]
    body = mhx.FloatingBody(mesh, names)
    ctx.set_lu_mode(modes[args.lu_mode])
    if args.lu_streams:
        ctx.set_lu_streams(args.lu_streams)
    if args.workload == "forward_speed":
        return bench_forward_speed(args, mhx, _lib, ctx, body, names, mesh, omegas)
    if args.workload == "multibody":
        return bench_multibody(args, mhx, _lib, ctx, names)
    form = mhx.DirectBEM(args.gf)
    if args.freqs != len(omegas):
        omegas = omegas[np.linspace(0, len(omegas) - 1, max(1, args.freqs)).round().astype(int)]
    nw = len(omegas)
    mode = {"batched": _lib.MODE_BATCHED, "graph": _lib.MODE_GRAPH, "per_omega": _lib.MODE_PER_OMEGA}[args.schedule]
    ctx.set_batch(nw)

    def run_with(b, gf, direct, m):
        return lambda ws: mhx.api._body_coeffs(b, names, ws, [0.0], gf, direct, m, 1025.0, 9.81)

    if args.gf == "Surrogate":
        t0 = time.perf_counter()
        wpath = mhx.api.SURROGATE_WEIGHTS if args.activation == "silu" else mhx.api.SURROGATE_WEIGHTS_RELU
        if os.path.exists(wpath):
            mhx.load_surrogate(wpath)
            log(f"bench: loaded surrogate weights {os.path.basename(wpath)} ({time.perf_counter() - t0:.2f} s)")
        else:
            val = mhx.train_surrogate(activation=args.activation)
            log(f"bench: trained surrogate on the device ({time.perf_counter() - t0:.1f} s, val mse {val:.2e})")
    run = run_with(body, form.gf, form.direct, mode)
    log(f"bench: n={n}, {max(args.warmup, 1)} warmup sweep(s) of {nw} frequencies")
    for _ in range(max(args.warmup, 1)):
        run(omegas)   # body upload, Rankine part, batch buffers at the timed size: outside the timed region
    log(f"bench: timing {args.steps} sweeps of {nw} omega-solves ({args.schedule})")
    if args.schedule == "graph":
        # the captured sweep replayed: wall clock over the timed calls; the phase table and the
        # rooflines come from one eager pass of the same batched schedule (the same kernels on
        # the same inputs) right after
        ctx.reset_timing()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            res = run(omegas)
        elapsed = time.perf_counter() - t0
        assert np.all(np.isfinite(res[0])) and np.all(np.isfinite(res[2]))
        value = nw * args.steps / elapsed
        solved, steps, fallbacks, resid = ctx.solver_stats()
        eager = run_with(body, form.gf, form.direct, _lib.MODE_BATCHED)
        eager(omegas)
        _, _, phases, roof, lu_tflops, _, roof2 = measure(ctx, eager, omegas, n, args.gf,
                                                         kind=lu_kind(args.lu_mode, args.gf), calls=1)
    else:
        value, elapsed, phases, roof, lu_tflops, res, roof2 = measure(ctx, run, omegas, n, args.gf,
                                                                      kind=lu_kind(args.lu_mode, args.gf),
                                                                      calls=args.steps)
        solved, steps, fallbacks, resid = ctx.solver_stats()
    solver = {"mixed_precision_systems": solved, "refinement_steps": steps, "fp64_fallbacks": fallbacks,
              "max_relative_residual": resid}
    lu_streams, lu_panels = ctx.lu_config()
    if lu_streams:
        solver["lu_config"] = {"streams": lu_streams, "panels_per_split_update": lu_panels,
                               "chosen_by": "explicit --lu-streams" if args.lu_streams else "the library's LU tuner"}
    log(f"bench: {value:.1f} omega-solves/s; per-frequency phase ms:", phases, "solver:", solver)
    plain = None
    if not args.no_plain:
        # the plain GPU path on its own body handle (buffers for one frequency), FP64 LU as the reference
        pbody = mhx.FloatingBody(mesh, names)
        sub = np.linspace(0, nw - 1, 4).round().astype(int)
        ws4 = omegas[sub]
        ctx.set_lu_mode(_lib.LU_FP64)
        plain = {}
        for label, gf, m, ws in (("exact_per_omega", "ExactGuevelDelhommeau", _lib.MODE_PER_OMEGA, ws4),
                                 ("wu_per_omega", "Wu", _lib.MODE_PER_OMEGA, ws4),
                                 ("exact_reference_semantics", "ExactGuevelDelhommeau", _lib.MODE_PER_PROBLEM, ws4[1:2]),
                                 ("wu_reference_semantics", "Wu", _lib.MODE_PER_PROBLEM, ws4[1:2])):
            f = mhx.DirectBEM(gf)
            prun = run_with(pbody, f.gf, True, m)
            prun(ws[:1])  # buffers and the wave table of this Green function allocated
            log(f"bench: plain GPU {label} at omega = {np.round(ws, 3).tolist()}")
            pv, pel, pph, proof, plu, pres, proof2 = measure(ctx, prun, ws, n, gf, nlu=len(ws) * (7 if m == _lib.MODE_PER_PROBLEM else 1))
            entry = {"value": pv, "unit": "omega-solves/s", "omegas": np.round(ws, 4).tolist(),
                     "schedule": "assembly + FP64 LU per problem (7 per frequency, solve.jl:75-78)"
                     if m == _lib.MODE_PER_PROBLEM else "one assembly + FP64 LU + 7 right-hand sides per frequency",
                     "phases_ms_per_factorization": pph, "lu_tflops": plu}
            if label == "exact_per_omega":
                errs = {}
                pairs = (("added_mass", lambda r, q: r[0][:, :, q]), ("radiation_damping", lambda r, q: r[1][:, :, q]),
                         ("excitation_force", lambda r, q: (r[2] + r[3])[:, q]))
                for name, at in pairs:
                    errs[name] = max(float(np.linalg.norm(at(res, sub[q]) - at(pres, q)) / np.linalg.norm(at(pres, q)))
                                     for q in range(len(sub)))
                log("max relative error of the headline path vs exact FP64 per omega:", errs)
                entry.update(max_rel_err_vs_exact=errs, roofline=proof, roofline_next_kernel=proof2)
            plain[label] = entry
        ctx.set_lu_mode(modes[args.lu_mode])

    cpu = cpu_faithful = None
    if not args.no_cpu_baseline:
        log("bench: cpu baseline (oracle, both modes)")
        cpu, cpu_faithful = cpu_baseline(mesh, omegas, args.cpu_freqs, min(16, os.cpu_count() or 1))

    denom = plain["exact_per_omega"]["value"] if plain else None
    out = {
        "metric": METRIC, "value": value, "unit": "omega-solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": elapsed * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": DTYPE[args.gf], "data": "synthetic",
        "config": {"workload": f"{args.schedule} GPU path, wave term {WAVE_DESC[args.gf]}, direct method: "
                               f"capped cylinder R=4 draft=12, {n} panels, 6 dofs, beta=0, all {nw} frequencies "
                               f"in 0.1-3.0 rad/s per step (one step = one {nw}-frequency sweep)",
                   "panels": n, "dofs": 6, "headings": 1, "frequencies_per_step": nw, "schedule": args.schedule,
                   "parallelism": "single GPU",
                   **({"surrogate_network": "64 Fourier features -> 4 x 128 " +
                       ("SiLU" if args.activation == "silu" else "ReLU") + " -> 6"} if args.gf == "Surrogate" else {})},
        "phases_ms_per_omega": phases,
        **({"phases_source": "one eager pass of the batched schedule (same kernels and inputs) after the timed "
                             "graph replays"} if args.schedule == "graph" else {}),
        "lu_tflops": lu_tflops,
        "solver": solver,
        "plain_gpu": plain,
        "speedup_vs_plain": {"value": value / denom, "denominator": "plain_gpu.exact_per_omega"} if denom else None,
        "roofline": roof,
        "roofline_next_kernel": roof2,
        "cpu_baseline": cpu,
        "cpu_baseline_reference_faithful": cpu_faithful,
    }
    print(json.dumps(out), flush=True)


FS_SPEEDS = [0.0, 2.0]
FS_HEADINGS = [0.0, np.pi / 6, np.pi / 2, np.pi]


def bench_forward_speed(args, mhx, _lib, ctx, body, names, mesh, omegas):
    """The forward-speed grid: U in {0, 2} m/s, headings 0, pi/6, pi/2, pi, 16 of the 64
    frequencies, six rigid dofs, IndirectBEM.  A unit is one (U, omega, beta) problem group (six
    radiation problems and one diffraction problem); at U = 0 every heading shares one solve per
    frequency, each moving group has its own encountered frequency, assembly, LU and K_x."""
    gf = "ExactGuevelDelhommeau" if args.gf == "Surrogate" else args.gf
    form = mhx.IndirectBEM(gf)
    n = mesh.nfaces
    steps = 16
    ws = omegas[np.linspace(0, len(omegas) - 1, steps).round().astype(int)]
    nb, nu = len(FS_HEADINGS), len(FS_SPEEDS)
    ctx.set_batch(max(steps, 1))

    def run(w):
        return mhx.api._body_coeffs_fs(body, names, w, FS_HEADINGS, FS_SPEEDS, names, form.gf, False, 1025.0, 9.81)

    log(f"bench: forward speed, n={n}, warmup")
    run(ws[:1])
    groups = nu * len(ws) * nb
    nlu = len(ws) * (1 + nb * sum(1 for u in FS_SPEEDS if u != 0))
    log(f"bench: timing {groups} problem groups ({nlu} factorizations)")
    value, elapsed, phases, roof, lu_tflops, res, roof2 = measure(ctx, run, ws, n, gf, nlu=nlu, units=groups,
                                                                  kind=lu_kind(args.lu_mode, gf))
    log("per-factorization phase ms:", phases)
    cpu = None
    if not args.no_cpu_baseline:
        log("bench: cpu baseline (one moving group)")
        cpu = cpu_baseline_fs(mesh, 2.0, float(ws[len(ws) // 2]), np.pi / 6, min(16, os.cpu_count() or 1))
    out = {
        "metric": "forward-speed problem groups/s at 4k panels", "value": value, "unit": "problem groups/s",
        "n_gpus": 1, "steps": groups, "warmup": 1, "ms_per_step": elapsed * 1e3 / groups, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"forward-speed grid, IndirectBEM, wave term {WAVE_DESC[gf]}: capped cylinder R=4 "
                               f"draft=12, {n} panels, 6 dofs, U in {FS_SPEEDS} m/s, headings {len(FS_HEADINGS)}, "
                               f"{len(ws)} of the 64 frequencies in 0.1-3.0 rad/s",
                   "panels": n, "dofs": 6, "headings": nb, "speeds": nu, "frequencies": len(ws),
                   "factorizations": nlu, "parallelism": "single GPU"},
        "phases_ms_per_factorization": phases,
        "lu_tflops": lu_tflops,
        "roofline": roof,
        "roofline_next_kernel": roof2,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def bench_multibody(args, mhx, _lib, ctx, names):
    """The seeded uneven workload (8 bodies: spheres, cylinders, boxes of 256-4,096 panels, 16-64
    log-uniform frequencies in 0.1-3.0 rad/s each, 6 dofs, headings 0, pi/4, pi/2, pi, direct
    method) as one multi-body sweep captured into a HIP graph: the warmup captures it, each timed
    step replays it.  A unit is one (body, frequency) omega-solve with its 4 headings."""
    import math
    from mhx import meshes
    betas = [0.0, math.pi / 4, math.pi / 2, math.pi]
    if args.gf == "Surrogate":
        mhx.load_surrogate()
    jobs = [(mhx.FloatingBody(m, names), {"wave_frequencies": w, "radiating_dofs": names, "wave_directions": betas})
            for m, w, _ in meshes.random_bodies()]
    units = sum(len(par["wave_frequencies"]) for _, par in jobs)
    form = mhx.DirectBEM(args.gf)
    log(f"bench: multibody, {len(jobs)} bodies, panels {[b.mesh.nfaces for b, _ in jobs]}, {units} omega-solves")
    for _ in range(max(args.warmup, 1)):
        mhx.hydrodynamic_coefficients_many(jobs, form, graph=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res = mhx.hydrodynamic_coefficients_many(jobs, form, graph=True)
    elapsed = time.perf_counter() - t0
    assert all(np.all(np.isfinite(r["added_mass"])) for r in res)
    value = units * args.steps / elapsed
    log(f"bench: multibody {value:.1f} omega-solves/s")
    out = {
        "metric": "multi-body omega-solves/s (seeded 8-body workload, graph-captured sweep)", "value": value,
        "unit": "omega-solves/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": DTYPE[args.gf], "data": "synthetic",
        "config": {"workload": f"seeded uneven workload, 8 bodies of {sorted(b.mesh.nfaces for b, _ in jobs)} panels, "
                               f"{units} (body, frequency) items, 6 dofs, 4 headings, direct method, wave term "
                               f"{WAVE_DESC[args.gf]}; one step = the whole sweep replayed from its HIP graph",
                   "bodies": len(jobs), "omega_solves_per_step": units, "parallelism": "single GPU"},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
