"""Newton iterations/s of the arc-length (MSAL) solve of the 256x256 CST+DKT
cylindrical shell panel (131072 triangles, 66049 joints, 394752 equations) on
one MI355X.

A step is one arc-length increment with all its Newton iterations (predictor
included), tangent assembly and solver set-up included.  The first W
increments (the displacement-controlled first increment among them) are warmup;
the next K increments are timed.  Time spent in the exact FP64 pivot judge of
a path that is not the FP64 factorization itself is subtracted, as it would be
from every variant.

Paths: "fp64" = FP64 band LDL^T on the GPU (the plain FP64 GPU path);
"mixed" = FGMRES with FP64 residual preconditioned by an FP32 band LDL^T.
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


def run_path(path: str, solver: int, warmup: int, steps: int, tol: float) -> dict:
    import cubens
    t0 = time.time()
    with cubens.GpuModel(path) as g:
        t_open = time.time() - t0
        opt = cubens.options(solver, krylov_tol=tol, restart=30, maxit=400,
                             judge=0 if solver == cubens.SOLVER_BAND64 else 1,
                             max_steps=warmup + steps, keep_d=1)
        r = g.run(opt)
        info = g.info
    n = r["nsteps"]
    if n < warmup + steps:
        raise RuntimeError(f"run stopped after {n} steps: {r['msg']}")
    tw, tj, nw = r["step_t"], r["step_judge"], r["step_newton"]
    a = warmup - 1
    t = (tw[-1] - tj[-1]) - ((tw[a] - tj[a]) if a >= 0 else 0.0)
    newton = int(nw[-1] - (nw[a] if a >= 0 else 0))
    return {"rate": newton / t, "seconds": t, "newton": newton, "open_s": t_open, "stats": r,
            "neq": info.neq, "bandwidth": info.bandwidth, "lpf": r["lpf"], "d": r["d"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=256, help="panel mesh (256 = the 396k-DOF workload)")
    ap.add_argument("--paths", default="fp64,mixed")
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import cubens
    path = workload_file(args.n)
    results = {}
    for name in args.paths.split(","):
        solver = {"fp64": cubens.SOLVER_BAND64, "mixed": cubens.SOLVER_BAND32_FGMRES}[name]
        log(f"[bench] path {name}: {args.warmup} warmup + {args.steps} timed increments")
        res = run_path(path, solver, args.warmup, args.steps, args.tol)
        st = res["stats"]
        log(f"[bench] {name}: {res['rate']:.2f} Newton it/s ({res['newton']} its in {res['seconds']:.2f}s); "
            f"tangent {st['t_tangent']:.2f}s factor {st['t_factor']:.2f}s solve {st['t_solve']:.2f}s "
            f"forces {st['t_forces']:.2f}s judge {st['t_judge']:.2f}s krylov {st['krylov_iters']} "
            f"fallbacks {st['krylov_fallbacks']} worst relres {st['worst_relres']:.2e}")
        results[name] = res
    head = results.get("mixed") or results["fp64"]
    plain = results.get("fp64")
    extra = {}
    if plain is not None and "mixed" in results:
        m = results["mixed"]
        extra["plain_fp64_value"] = plain["rate"]
        extra["ratio_vs_plain_fp64"] = m["rate"] / plain["rate"]
        extra["mixed_vs_fp64_rel_disp_err"] = float(
            np.abs(m["d"] - plain["d"]).max() / np.abs(plain["d"]).max())
    out = {
        "metric": METRIC,
        "value": head["rate"],
        "unit": "Newton iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * head["seconds"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if head is plain else "f32 factor + f64 residual",
        "data": "synthetic (seeded SplitMix64 thickness field and imperfection)",
        "config": {"workload": f"cylindrical shell panel {args.n}x{args.n} MSAL, "
                               f"{2 * args.n * args.n} CST+DKT triangles, NEQ {head['neq']}",
                   "path": "mixed" if head is not plain else "fp64",
                   "increments_timed": args.steps},
        **extra,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
