"""Training of the learned preconditioner with validation-based selection.

The network is trained online on the FP64 solves of arc-length analyses
(cb_learned_train: one Adam step per solve on the solve's (K, r, du) triple
plus random probes).  The training loss does not measure Krylov convergence
directly, so after every round the network is scored by the FGMRES
iterations it needs on a validation model, and the best parameters are kept.
Round 0 is the untrained network, which is exactly block-Jacobi: the selected
network is never worse than block-Jacobi on the validation model.
"""
from __future__ import annotations

import numpy as np

from . import GpuModel, LearnedPrecond, SOLVER_LEARNED_FGMRES, options


def krylov_iterations(L: LearnedPrecond, path: str, steps: int = 2, maxit: int = 20000) -> int:
    """FGMRES iterations (1e-8) of the learned preconditioner over the first
    `steps` converged steps of the model's analysis; a solve that misses its
    tolerance counts as maxit."""
    with GpuModel(path) as g:
        L.attach(g)
        r = g.run(options(SOLVER_LEARNED_FGMRES, krylov_tol=1e-8, restart=30, maxit=maxit, judge=1,
                          max_steps=steps, keep_d=0))
    return int(r["krylov_iters"] + maxit * r["krylov_fallbacks"])


def fit(L: LearnedPrecond, train_paths, val_path: str, rounds: int = 6, lr: float = 2e-4,
        probes: int = 6, train_steps: int = 6, val_steps: int = 2) -> dict:
    """Trains L for `rounds` rounds, one training model per round (cycling
    through train_paths), and keeps the parameters with the fewest
    validation iterations.  Returns the history."""
    best_p = L.params()
    best = krylov_iterations(L, val_path, val_steps)
    hist = {"val_iters": [best], "loss": [], "best_round": 0}
    for k in range(1, rounds + 1):
        with GpuModel(train_paths[(k - 1) % len(train_paths)]) as g:
            st = L.train(g, max_steps=train_steps, probes=probes, lr=lr)
        it = krylov_iterations(L, val_path, val_steps)
        hist["val_iters"].append(it)
        hist["loss"].append(float(st["loss_recent"]))
        if it < best:
            best, best_p, hist["best_round"] = it, L.params(), k
    L.set_params(best_p)
    hist["best_val_iters"] = best
    return hist
