"""Newton iterations/s of the arc-length (MSAL) solve of the 256x256 CST+DKT
cylindrical shell panel (131072 triangles, 66049 joints, 394752 equations) on
one MI355X.

A step is one arc-length increment with all its Newton iterations (predictor
included), tangent assembly and solver set-up included.  W warm-up increments
run first as an analysis of their own (untimed); then a fresh analysis from the
undeformed state runs and its first K increments are timed -- with the default
K = 40 that is the window the metric is defined on (increments 1..40,
the displacement-controlled first increment among them).  Time spent in the
exact FP64 pivot judge of a path that is not the FP64 factorization itself is
subtracted, as it would be from every variant.

Paths: "fp64" = FP64 band LDL^T on the GPU (the plain FP64 GPU path, the
baseline of the metric); "fp32" = the FP32 precision mode (the shell element
stiffness and the band LDL^T in FP32, one refinement step per solve against
an FP64 residual; internal forces, residual, arc-length constraint and
convergence tests in FP64), timed over the same increments with its
displacement error against the FP64 path; "mixed" = FGMRES with FP64 residual
preconditioned by an FP32 band LDL^T, timed over its own first
--compare-steps increments beside the plain path's rate over the same
increments.  Each reports its ratio to the plain path.  The learned
preconditioner is not run at this size: its FGMRES needs thousands of
iterations per solve here (it is measured on held-out panels up to 32x32 by
--workload learned).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cu-bens_amd"))

METRIC = "Newton iters/s at 396k-DOF DKT shell: learned-precond vs plain FP64 GPU"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload_file(n: int) -> str:
    from cubens import inputs
    d = os.environ.get("CUBENS_BENCH_DIR", "/tmp/cubens_bench")
    os.makedirs(d, exist_ok=True)
    path = os.path.join(d, f"panel{n}_c3.txt")
    if not os.path.exists(path):
        cfg = 3 if n == 256 else 2
        inputs.shell_panel(n=n, thick_sigma=0.1, imperfection=0.1, config_index=cfg).write(path)
    return path


PEAK_FP64_TFLOPS = 78.6
# the fused factorization's launches are timed as one HIP-event span per
# factorization and the split schedule's every 16th launch (the sweeps every
# time), which keeps the event overhead out of the measured window
TIMING_EVERY = 16
PEAK_HBM_GBS = 8000.0
TILE = 64


def kernel_model(info, sweeps: str = "chain") -> dict:
    """Algorithmic work of one launch of each band kernel family at this model.

    nt tile rows of 64, w sub-diagonal tiles (w = ceil(bandwidth/64)); column K
    has m = min(w, nt-1-K) sub-diagonal tiles:
      diag   : LDL^T of a 64x64 tile (2/3 T^3) + inverse of its unit factor (T^3/3)
      panel  : m tile products of 2 T^3
      update : m(m+1)/2 tile products of 2 T^3
      column : one fused launch = diag + panel + update of a column (average
               over the nt launches of a factorization); its time is the launch
               period (the nt launches of a factorization timed as one span)
      sweeps : every off-diagonal factor tile and every Linv tile read once (8 B
               entries), the right-hand side read and the solution written
      prep   : the sweeps' chain matrices, two tile products of 2 T^3 per
               tile row; with the look-ahead sweeps also W_I and V_I (one
               product each) and the two look-ahead bands (two products per
               sub-diagonal tile each)
    """
    nt = (info.neq + TILE - 1) // TILE
    w = min((info.bandwidth + TILE - 1) // TILE, nt - 1)
    t3 = TILE ** 3
    offd = sum(min(w, i) for i in range(nt))
    sweep_bytes = 8 * TILE * TILE * (offd + nt) + 3 * 8 * nt * TILE
    whole = sum(t3 + 2.0 * t3 * m + t3 * m * (m + 1) for m in (min(w, nt - 1 - k) for k in range(nt)))
    return {
        "band_column": ("mfma", whole / nt, "TFLOP/s"),
        "band_diag": ("mfma", float(t3), "TFLOP/s"),
        "band_panel": ("mfma", 2.0 * t3 * w, "TFLOP/s"),
        "band_update": ("mfma", 2.0 * t3 * w * (w + 1) / 2, "TFLOP/s"),
        "band_fwd_sweep": ("hbm", float(sweep_bytes), "GB/s"),
        "band_bwd_sweep": ("hbm", float(sweep_bytes), "GB/s"),
        "band_sweep_prep": ("mfma", 2.0 * t3 * (2 * nt + (2 * nt + 4 * offd if sweeps in ("look2", "look3") else 0)
                                                + (2 * (offd + nt) if sweeps == "look3" else 0)),
                            "TFLOP/s"),
    }


# HBM bytes per launch of the dominant kernels on the default workloads, from
# two rocprofv3 counter runs each (FETCH_SIZE, then WRITE_SIZE, KiB per
# dispatch), scaled by what the same counters read for an elementwise copy of
# 256 MiB (FETCH_SIZE x 2.000, WRITE_SIZE x 1.000): k_band_column over the
# interior launches 3000-3099 of the 256x256 panel, k_band_factor_models over
# the 13 rounds of 5 time steps of the 512 frames.
MEASURED_TRAFFIC = {
    "band_column": (8647.29 * 2.0 + 12001.31) * 1024.0,
    "band_factor_models": (1679014.64 * 2.0 + 3575412.50) * 1024.0,
}
TRAFFIC_SOURCE = ("rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs, KiB per dispatch scaled by a 256 MiB "
                  "copy (x2.000, x1.000)")


def skyline_work(kht: np.ndarray) -> dict:
    """Algorithmic work of the reference's skyline solver on this profile:
    skyfact sum(kht^2) flops (one multiply-add per column pair below the
    skyline, counted as 2 flops over the half-square, solve.c:539-653) and
    skysolve's read of the lss = sum(kht+1) factor entries once per sweep
    (solve.c:655-698)."""
    k = kht.astype(np.float64)
    return {"factor_flops": float((k * k).sum()), "sweep_bytes": 8.0 * float((k + 1).sum())}


def roofline(info, ktimes: dict, default_workload: bool, sky: dict | None = None, sweeps: str = "chain") -> dict:
    name = max(ktimes, key=lambda k: ktimes[k][0])
    ms, launches = ktimes[name]
    bound, work, unit = kernel_model(info, sweeps)[name]
    nt = (info.neq + TILE - 1) // TILE
    if sky is not None and name == "band_column":
        work = sky["factor_flops"] / nt          # a factorization is nt column launches
    elif sky is not None and name in ("band_fwd_sweep", "band_bwd_sweep"):
        work = sky["sweep_bytes"]
    avg_s = ms / 1000.0 / max(launches, 1)
    if bound == "hbm":
        achieved, peak = work / avg_s / 1e9, PEAK_HBM_GBS
    else:
        achieved, peak = work / avg_s / 1e12, PEAK_FP64_TFLOPS
    traffic = MEASURED_TRAFFIC.get(name) if default_workload else None
    out = {"kernel": name, "bound": bound, "achieved": achieved, "peak": peak, "unit": unit,
           "frac": achieved / peak, "traffic": traffic, "avg_launch_us": 1e6 * avg_s, "launches": launches,
           "share_of_kernel_time": ms / sum(v[0] for v in ktimes.values())}
    if sky is not None and name == "band_column":
        out["work_basis"] = "sum(kht^2) flops of the skyline factorization / nt launches"
    if traffic is not None:
        out["traffic_unit"] = "B per launch"
        out["traffic_source"] = TRAFFIC_SOURCE
    if sky is not None:
        # the two sweeps of every solve: algorithmic bytes (the skyline factor
        # read once per sweep) over their average duration, against HBM
        sw = [ktimes[k] for k in ("band_fwd_sweep", "band_bwd_sweep") if ktimes.get(k, (0, 0))[1]]
        if sw:
            per = sum(v[0] for v in sw) / 1000.0 / sum(v[1] for v in sw)
            out["sweeps"] = {"avg_sweep_ms": 1e3 * per, "bytes": sky["sweep_bytes"],
                             "achieved": sky["sweep_bytes"] / per / 1e9, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                             "frac": sky["sweep_bytes"] / per / 1e9 / PEAK_HBM_GBS}
    return out


def host_cpu() -> dict:
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "model": model}


def cpu_baseline(path: str, newton: int, increments: int, budget: float) -> dict:
    """Single-core timing of the oracle's serial skyline path on a bounded sample
    of the same workload, projected onto the GPU window's work."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    om = O.OracleModel(path)
    smp = om.time_sample(budget)
    per_incr = smp["t_tangent"] + smp["t_factor_est"] + smp["t_solve"]
    per_newton = smp["t_solve"] + smp["t_forces"]
    t = increments * per_incr + newton * per_newton
    return {"value": newton / t, "unit": "Newton iters/s", "cores": 1, "kind": "port", "host": host_cpu(),
            "sample": (f"oracle (C restatement of the serial skyline path) on 1 core, same model: one tangent "
                       f"{smp['t_tangent']:.2f}s, one internal-force pass {smp['t_forces']:.2f}s, one skyline "
                       f"back-substitution {smp['t_solve']:.2f}s, skyline LDL^T of the first {smp['cols_done']} "
                       f"columns {smp['t_factor_part']:.1f}s = {100 * smp['factor_fraction']:.2f}% of its "
                       f"sum(kht^2) work -> {smp['t_factor_est']:.0f}s per factorization; projected onto the "
                       f"GPU window ({increments} increments, {newton} Newton iterations)")}


FRAME_METRIC = "Newton iters/s, 512 space frames under nonlinear Newmark (batched)"


def frame_files(steps: int) -> list[str]:
    """The 512-frame workload (towers and lattice girders of U{200..5000}
    members, 200 time steps of average-acceleration Newmark), written once."""
    from cubens import inputs
    d = os.path.join(os.environ.get("CUBENS_BENCH_DIR", "/tmp/cubens_bench"), f"frames{steps}")
    os.makedirs(d, exist_ok=True)
    done = os.path.join(d, "complete")
    paths = [os.path.join(d, f"f{i:03d}.txt") for i in range(512)]
    if not os.path.exists(done):
        for p, m in zip(paths, inputs.frame_batch(steps=steps)):
            with open(p, "w") as fh:
                fh.write(m.text)
        open(done, "w").close()
    return paths


def frame_cpu_baseline(paths: list[str], steps: int, stride: int = 32) -> dict:
    """The oracle's serial Newmark driver on 1 core over every stride-th model
    of the batch for `steps` time steps: Newton iterations per second."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    t = 0.0
    newton = 0
    sample = paths[::stride]
    for p in sample:
        om = O.OracleModel(p)
        t0 = time.perf_counter()
        r = om.run(steps)
        t += time.perf_counter() - t0
        newton += int(r["n_newton"])
    return {"value": newton / t, "unit": "Newton iters/s", "cores": 1, "kind": "port", "host": host_cpu(),
            "sample": (f"oracle (C restatement of the serial Newmark driver with the skyline solver) on 1 core: "
                       f"{len(sample)} of the 512 models (one in {stride}), first {steps} time steps each, "
                       f"{newton} Newton iterations in {t:.1f}s")}


def run_frames(args) -> None:
    import cubens
    steps_total = 200
    log("[bench] frames: writing / reading the 512-model workload")
    paths = frame_files(steps_total)
    t0 = time.time()
    b = cubens.GpuBatch(paths)
    t_open = time.time() - t0
    neq = sum(i.neq for i in b.info)
    nfr = sum(i.ne_fr for i in b.info)
    log(f"[bench] frames: {b.count} models, {nfr} members, NEQ {neq}, opened in {t_open:.1f}s")
    # the warm-up run tunes the band factor's helper pairs on the initial
    # tangent; the timed runs use the count it keeps
    wst = b.run(max_steps=max(args.warmup, 1), keep_d=False, autotune=True)
    tuning = {"pairs_kept": int(wst["pairs"]), "pairs_max": int(wst["pairs_max"]),
              "trials_ms": [[p, round(ms, 3)] for p, ms in b.tuning()], "seconds": round(wst["tune_s"], 3)}
    log(f"[bench] frames: tuner kept {tuning['pairs_kept']} of {tuning['pairs_max']} helper pairs; "
        f"trials (pairs, factor ms) {tuning['trials_ms']} in {tuning['seconds']}s")
    k = min(args.steps, steps_total)
    # the same rounds three ways: launches one by one, each round's Newton
    # iteration replayed as one HIP graph (the timed value), and one by one
    # with HIP events around the band kernels (the kernel-level figures)
    e = b.run(max_steps=k, keep_d=False)
    g = b.run(max_steps=k, keep_d=False, graph=True)
    r = b.run(max_steps=k, keep_d=False, timing=True)
    # the member tangents in FP32 (rounded into the FP64 band), one graph per round
    g32 = b.run(max_steps=k, keep_d=False, graph=True, fp32_elements=True)
    rate32 = g32["newton"] / g32["t_total"]
    log(f"[bench] frames: {rate32:.0f} Newton it/s with FP32 member tangents ({g32['newton']} its, "
        f"{g32['rounds']} rounds in {g32['t_total']:.2f}s)")
    for x in (e, g, r):
        if x["failed"] or int(x["nsteps"].min()) < k:
            raise RuntimeError(f"frame batch stopped early: {x}")
    if not (g["newton"] == e["newton"] == r["newton"] and g["rounds"] == e["rounds"]):
        raise RuntimeError("the graph replay took other iterations than the launches one by one")
    rate_e = e["newton"] / e["t_total"]
    rate = g["newton"] / g["t_total"]
    log(f"[bench] frames: {rate:.0f} Newton it/s with the round's iteration as one graph "
        f"({g['newton']} its, {g['rounds']} rounds in {g['t_total']:.2f}s, {g['ops_per_round']:.1f} stream "
        f"operations per round); {rate_e:.0f} Newton it/s launching one by one ({e['ops_per_round']:.1f} "
        f"operations per round, {e['body_ops']} in the iteration)")
    log(f"[bench] frames: band factor {r['ms_factor']:.1f} ms, solve {r['ms_solve']:.1f} ms over the rounds")
    rounds = max(r["rounds"], 1)
    fam = {"band_factor_models": (r["ms_factor"], "mfma", r["factor_flops"] / 1e12, "TFLOP/s", PEAK_FP64_TFLOPS),
           "band_solve_models": (r["ms_solve"], "hbm", r["solve_bytes"] / 1e9, "GB/s", PEAK_HBM_GBS)}
    name = max(fam, key=lambda f: fam[f][0])
    ms, bound, work, unit, peak = fam[name]
    avg_s = ms / 1000.0 / rounds
    achieved = work / avg_s
    traffic = MEASURED_TRAFFIC.get(name)
    roof = {"kernel": name, "bound": bound, "achieved": achieved, "peak": peak, "unit": unit,
            "frac": achieved / peak, "traffic": traffic, "avg_launch_us": 1e6 * avg_s, "launches": rounds,
            "share_of_run": ms / 1000.0 / r["t_total"]}
    if traffic is not None:
        roof["traffic_unit"] = "B per launch"
        roof["traffic_source"] = TRAFFIC_SOURCE
    cpu = None
    if not args.no_cpu:
        log("[bench] cpu baseline: oracle on a sample of the models, 1 core")
        cpu = frame_cpu_baseline(paths, min(k, 10))
        log(f"[bench] cpu: {cpu['value']:.1f} Newton it/s; {cpu['sample']}")
    out = {
        "metric": FRAME_METRIC, "value": rate, "unit": "Newton iters/s", "n_gpus": 1, "steps": k,
        "warmup": args.warmup, "ms_per_step": 1000.0 * g["t_total"] / k, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded SplitMix64 topologies, sections and harmonic loads)",
        "config": {"workload": f"512 space frames (towers and lattice girders, 200-5000 members), {nfr} members, "
                               f"NEQ {neq}, average-acceleration Newmark, dt 1e-3",
                   "path": "batched fp64 band LDL^T, one workgroup per model; each round's Newton iteration "
                           "replayed as one captured HIP graph", "time_steps_timed": k,
                   "rounds": g["rounds"],
                   "launches": {"ops_per_round_graph": g["ops_per_round"], "ops_per_round_one_by_one":
                                e["ops_per_round"], "iteration_ops": g["body_ops"],
                                "newton_per_s_one_by_one": rate_e},
                   "autotune": tuning,
                   "fp32_element_tangents": {"newton_per_s": rate32, "newton": int(g32["newton"]),
                                             "rounds": int(g32["rounds"]), "seconds": g32["t_total"],
                                             "failed": int(g32["failed"])}},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    b.close()
    print(json.dumps(out))


def run_learned(args) -> None:
    """The learned preconditioner against block-Jacobi inside FGMRES (1e-8):
    trained by learned.fit on four 8x8 and two 12x12 seeded panels, selected
    on a 10x10 validation panel, evaluated on held-out panels of growing size
    over their first 2 arc-length steps."""
    import tempfile
    import cubens
    from cubens import inputs, learned
    tmp = tempfile.mkdtemp(prefix="cubens_learned_")

    def panel(n, seed):
        p = os.path.join(tmp, f"p{n}_{seed}.txt")
        inputs.shell_panel(n=n, thick_sigma=0.1, imperfection=0.1, seed=seed).write(p)
        return p

    rows = []
    with cubens.LearnedPrecond(0) as L:
        train = [panel(8, s) for s in (101, 102, 103, 104)] + [panel(12, s) for s in (201, 202)]
        t0 = time.time()
        hist = learned.fit(L, train, panel(10, 998), rounds=12, lr=2e-4)
        log(f"[bench] learned: fit in {time.time() - t0:.1f}s, validation iterations {hist['val_iters']}")
        for n in (int(v) for v in args.learned_sizes.split(",")):
            path = panel(n, 999)
            opt = dict(krylov_tol=1e-8, restart=30, maxit=40000, judge=1, max_steps=2)
            with cubens.GpuModel(path) as g:
                ref = g.run(cubens.options(cubens.SOLVER_BAND64, max_steps=2))
                t = time.time()
                bj = g.run(cubens.options(cubens.SOLVER_BJ_FGMRES, **opt))
                tb = time.time() - t
                L.attach(g)
                t = time.time()
                ln = g.run(cubens.options(cubens.SOLVER_LEARNED_FGMRES, **opt))
                tl = time.time() - t
                neq = g.neq
            err = lambda r: float(np.abs(r["d"] - ref["d"]).max() / np.abs(ref["d"]).max())  # noqa: E731
            row = {"panel": n, "neq": neq, "bj_iters": int(bj["krylov_iters"]),
                   "learned_iters": int(ln["krylov_iters"]), "bj_s": tb, "learned_s": tl,
                   "bj_rel_disp_err": err(bj), "learned_rel_disp_err": err(ln),
                   "fallbacks": int(bj["krylov_fallbacks"] + ln["krylov_fallbacks"])}
            log(f"[bench] learned: {json.dumps(row)}")
            rows.append(row)
    last = rows[-1]
    print(json.dumps({"metric": "FGMRES iterations at 1e-8, block-Jacobi / learned preconditioner",
                      "value": last["bj_iters"] / last["learned_iters"], "unit": "ratio", "n_gpus": 1,
                      "higher_is_better": True, "dtype": "bf16 network, f64 Krylov",
                      "data": "synthetic (seeded shell panels)",
                      "config": {"workload": f"held-out shell panels {args.learned_sizes} (n x n), 2 arc-length "
                                             "steps",
                                 "training": "four 8x8 and two 12x12 panels, 12 rounds, lr 2e-4",
                                 "selected_round": hist["best_round"]},
                      "panels": rows}))


def run_path(path: str, solver: int, warmup: int, steps: int, tol: float, timing: int = 0,
             schedule: int = 0, refine: int = 0, sweeps: int = 0) -> dict:
    """warmup increments of an untimed analysis, then the first `steps`
    increments of a fresh analysis, timed from its start."""
    import cubens
    t0 = time.time()
    with cubens.GpuModel(path) as g:
        t_open = time.time() - t0
        _, _, maxa = g.dofmaps()
        kht = np.diff(maxa) - 1
        # the FP32 precision mode decides det and psi on its own FP32 pivots
        judge = 0 if solver in (cubens.SOLVER_BAND64, cubens.SOLVER_BAND32) else 1
        common = dict(krylov_tol=tol, restart=30, maxit=400, judge=judge, factor_schedule=schedule,
                      refine=refine, sweep_schedule=sweeps)
        if warmup > 0:
            g.run(cubens.options(solver, max_steps=warmup, keep_d=0, **common))
        opt = cubens.options(solver, max_steps=steps, keep_d=1, kernel_timing=timing,
                             kernel_timing_every=TIMING_EVERY, **common)
        r = g.run(opt)
        info = g.info
        kt = None
        if timing and r["nsteps"] >= steps:
            prec = 64 if solver == cubens.SOLVER_BAND64 else 32
            kt = g.kernel_times(prec, steps - 1)
    n = r["nsteps"]
    if n < steps:
        raise RuntimeError(f"run stopped after {n} steps: {r['msg']}")
    tw, tj, nw = r["step_t"], r["step_judge"], r["step_newton"]
    t = tw[steps - 1] - tj[steps - 1]
    newton = int(nw[steps - 1])
    return {"rate": newton / t, "seconds": t, "newton": newton, "open_s": t_open, "stats": r,
            "neq": info.neq, "bandwidth": info.bandwidth, "lpf": r["lpf"], "d": r["d"], "info": info,
            "ktimes": kt, "kht": kht, "step_t": tw - tj, "step_newton": nw}


def path_error(ref: dict, got: dict) -> dict:
    """How far the converged states of a path lie from the FP64 one's
    equilibrium path.  The increments of two paths are not the same points (the
    arc length adapts to each increment's iteration count), so the load factor
    is compared at the same displacement of the DOF that moves most, on a
    cubic spline through the FP64 path's states, over the range both cover
    (before a limit point, where that displacement is monotone), with the
    spline's own error on the FP64 states (fitted through every other state,
    evaluated at the ones left out) for scale; and each converged
    state's FP64 equilibrium residual |lambda q - f| / |lambda q| is given
    beside the FP64 path's own."""
    k = int(np.argmax(np.abs(ref["d"][-1])))
    w_ref, l_ref = ref["d"][:, k], ref["lpf"]
    w_got, l_got = got["d"][:, k], got["lpf"]
    from scipy.interpolate import CubicSpline
    order = np.argsort(w_ref)
    ws, ls = w_ref[order], l_ref[order]
    keep = np.concatenate([[True], np.diff(ws) > 0])
    inside = (w_got >= ws.min()) & (w_got <= ws.max())
    ws, ls = ws[keep], ls[keep]
    lam_at = CubicSpline(ws, ls)(w_got[inside])
    scale = float(np.abs(l_ref).max())
    half = CubicSpline(ws[::2], ls[::2]) if len(ws) >= 8 else None
    self_err = (float(np.abs(half(ws[1:-1:2]) - ls[1:-1:2]).max() / scale) if half is not None else None)
    rel = lambda r: float(np.max(r["res"] / np.maximum(r["load"], 1e-300)))      # noqa: E731
    return {"lpf_err_at_same_displacement": float(np.abs(l_got[inside] - lam_at).max() / scale)
            if inside.any() else None,
            "states_compared": int(inside.sum()), "spline_self_error": self_err,
            "equilibrium_rel_residual": rel(got["stats"]), "equilibrium_rel_residual_fp64": rel(ref["stats"])}


def window_rate(res: dict, k: int) -> float:
    """Newton iterations/s of a timed path over its first k increments."""
    return float(res["step_newton"][k - 1]) / float(res["step_t"][k - 1])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40,
                    help="timed arc-length increments from the start of the analysis (the metric: 40)")
    ap.add_argument("--warmup", type=int, default=2, help="increments of an untimed warm-up analysis")
    ap.add_argument("--n", type=int, default=256, help="panel mesh (256 = the 396k-DOF workload)")
    ap.add_argument("--paths", default="fp64,fp32",
                    help="fp64 (the plain FP64 path, the headline), fp32 (the FP32 precision mode) "
                         "and/or mixed (FP32-factor FGMRES)")
    ap.add_argument("--compare-steps", type=int, default=6,
                    help="increments over which the mixed path is timed and compared with the plain one")
    ap.add_argument("--cpu-budget", type=float, default=8.0, help="seconds of partial CPU factorization")
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", type=float, default=0.0, metavar="MIB",
                    help="only run 20 elementwise copies of MIB MiB (the reference kernel for "
                         "scaling the FETCH_SIZE/WRITE_SIZE counters) and exit")
    ap.add_argument("--workload", default="panel", choices=("panel", "frames", "learned"),
                    help="panel: the 256x256 shell panel MSAL (the headline); frames: the 512-frame "
                         "nonlinear Newmark batch (--steps time steps); learned: the learned "
                         "preconditioner against block-Jacobi on held-out panels")
    ap.add_argument("--refine", type=int, default=1,
                    help="FP64-residual refinement steps per solve of the fp32 path")
    ap.add_argument("--sweeps", default="look2", choices=("chain", "rows", "look2", "look3"),
                    help="band solve sweeps: chain matrices (the library's default), the two-product "
                         "row step, or both sweeps on the look-ahead system (rows coupled two back; "
                         "its band formed after each factorization pays back over the ~13 solves "
                         "of an arc-length increment)")
    ap.add_argument("--learned-sizes", default="10,16,24,32",
                    help="held-out panel sizes n (n x n) of --workload learned")
    ap.add_argument("--schedule", default="fused", choices=("fused", "split"),
                    help="band factorization schedule: one fused launch per tile column, or "
                         "diag/panel/update launches")
    args = ap.parse_args()
    if args.calibrate_copy > 0:
        import torch
        n = int(args.calibrate_copy * (1 << 20)) // 8
        src = torch.rand(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(20):
            torch.mul(src, 1.0, out=dst)   # an elementwise kernel: reads and writes 8n bytes
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": 8 * n, "copies": 20}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if args.workload == "frames":
        run_frames(args)
        return
    if args.workload == "learned":
        run_learned(args)
        return
    import cubens
    path = workload_file(args.n)
    results = {}
    schedule = cubens.FACTOR_SPLIT if args.schedule == "split" else cubens.FACTOR_FUSED
    for name in args.paths.split(","):
        solver = {"fp64": cubens.SOLVER_BAND64, "fp32": cubens.SOLVER_BAND32,
                  "mixed": cubens.SOLVER_BAND32_FGMRES}[name]
        k = args.steps if name != "mixed" else min(args.compare_steps, args.steps)
        w = args.warmup if name != "mixed" else min(args.warmup, 1)
        log(f"[bench] path {name}: {w} warm-up increments, then increments 1..{k} timed")
        res = run_path(path, solver, w, k, args.tol, timing=127 if name == "fp64" else 0, schedule=schedule,
                       refine=args.refine if name == "fp32" else 0,
                       sweeps={"chain": 0, "rows": 1, "look2": 2, "look3": 3}[args.sweeps])
        st = res["stats"]
        log(f"[bench] {name}: {res['rate']:.2f} Newton it/s ({res['newton']} its in {res['seconds']:.2f}s); "
            f"tangent {st['t_tangent']:.2f}s factor {st['t_factor']:.2f}s solve {st['t_solve']:.2f}s "
            f"forces {st['t_forces']:.2f}s judge {st['t_judge']:.2f}s krylov {st['krylov_iters']} "
            f"fallbacks {st['krylov_fallbacks']} worst relres {st['worst_relres']:.2e}")
        results[name] = res
    plain = results.get("fp64")
    head = plain if plain is not None else results[next(iter(results))]
    extra = {}
    if plain is not None and "fp32" in results:
        m = results["fp32"]
        km = len(m["step_t"])
        base = window_rate(plain, km)
        err = path_error(plain, m)
        extra["precision"] = {
            "path": f"fp32: FP32 shell stiffness and band LDL^T, {args.refine} FP64-residual refinement "
                    "step(s) per solve; FP64 forces, residual, arc-length constraint and convergence tests",
            "refine": args.refine,
            "increments": km, "value": m["rate"], "plain_fp64_same_increments": base,
            "ratio_vs_plain_fp64": m["rate"] / base,
            "newton": int(m["newton"]), "newton_fp64": int(plain["step_newton"][km - 1]), **err}
        log(f"[bench] fp32 vs plain over increments 1..{km}: {m['rate']:.2f} / {base:.2f} = "
            f"{m['rate'] / base:.3f}; {err}")
    if plain is not None and "mixed" in results:
        m = results["mixed"]
        km = len(m["step_t"])
        base = window_rate(plain, km)
        extra["compare"] = {
            "path": "mixed: FGMRES with FP64 residual, FP32 band LDL^T preconditioner",
            "increments": km, "value": m["rate"], "plain_fp64_same_increments": base,
            "ratio_vs_plain_fp64": m["rate"] / base,
            "rel_disp_err_vs_fp64": float(np.abs(m["d"] - plain["d"][:km]).max() / np.abs(plain["d"][:km]).max()),
            "krylov_iters": int(m["stats"]["krylov_iters"]), "fallbacks": int(m["stats"]["krylov_fallbacks"]),
            "judge_s_excluded": float(m["stats"]["t_judge"]),
            "learned": "not run at this size (thousands of FGMRES iterations per solve); see --workload learned"}
        log(f"[bench] mixed vs plain over increments 1..{km}: {m['rate']:.2f} / {base:.2f} = "
            f"{m['rate'] / base:.3f}")
    sky = skyline_work(head["kht"]) if head is plain else None
    # the counters were read on the 256x256 panel's FP64 band with the fused schedule
    # kernel timings are taken on the FP64 path only; without it the line has no roofline
    roof = roofline(head["info"], head["ktimes"],
                    default_workload=(args.n == 256 and plain is not None and args.schedule == "fused"),
                    sky=sky, sweeps=args.sweeps) if head.get("ktimes") else None
    log(f"[bench] kernel times (ms, launches): {head['ktimes']}")
    cpu = None
    if not args.no_cpu:
        log("[bench] cpu baseline: oracle sample on 1 core")
        cpu = cpu_baseline(path, head["newton"], len(head["step_t"]), args.cpu_budget)
        log(f"[bench] cpu: {cpu['value']:.4f} Newton it/s; {cpu['sample']}")
    k = len(head["step_t"])
    out = {
        "metric": METRIC,
        "value": head["rate"],
        "unit": "Newton iters/s",
        "n_gpus": 1,
        "steps": k,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * head["seconds"] / k,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if head is plain else "f32 factor + f64 residual",
        "data": "synthetic (seeded SplitMix64 thickness field and imperfection)",
        "config": {"workload": f"cylindrical shell panel {args.n}x{args.n} MSAL, "
                               f"{2 * args.n * args.n} CST+DKT triangles, NEQ {head['neq']}",
                   "path": "plain fp64 band LDL^T (the metric's baseline path)" if head is plain
                           else "fp32-factor FGMRES",
                   "increments_timed": f"1..{k} of a fresh analysis after a {args.warmup}-increment warm-up run",
                   "newton_iterations": head["newton"], "factor_schedule": args.schedule,
                   "sweeps": args.sweeps},
        "roofline": roof,
        "cpu_baseline": cpu,
        **extra,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
