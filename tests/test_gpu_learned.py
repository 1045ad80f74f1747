"""The learned preconditioner on the GPU: its blocks, loss and gradient against
the CPU oracle (oracle/learned_oracle.py), training on the triples of the
arc-length solves, and what the trained network gains over block-Jacobi
inside FGMRES.

Tolerances: the GPU rounds GEMM operands to bfloat16 and accumulates in
float32; the oracle rounds the same operands to bfloat16 and accumulates in
float64.  Blocks agree to 2e-2 of the largest entry, the loss to 1e-3
relative, the gradient to 5e-2 relative (L2) with cosine similarity > 0.995.
"""
import numpy as np
import pytest

import cubens
import learned_oracle as LO
from cubens import inputs
from helpers import model_file, relerr

pytestmark = pytest.mark.gpu


def _params(seed, scale=1.0):
    lay, n = LO.layout()
    rng = np.random.default_rng(seed)
    p = np.zeros(n)
    for off, k, m in lay:
        p[off:off + k * m] = scale * rng.uniform(-1, 1, k * m) * np.sqrt(6.0 / (k + m))
        p[off + k * m:off + k * m + m] = 0.05 * rng.standard_normal(m)
    return p.astype(np.float32)


def _oracle_graph(g):
    rp, ci = g.pattern()
    jc, _, _ = g.dofmaps()
    eq = cubens.eq_of_slot(jc, g.nj, 6)
    return LO.Graph(rp, ci), eq


def _panel(n, seed=None, **kw):
    txt = inputs.shell_panel(n=n, thick_sigma=0.1, imperfection=0.1, seed=seed, **kw).text
    return model_file(f"lpanel{n}_{seed}.txt", txt)


def test_blocks_loss_and_gradient_match_oracle():
    with cubens.GpuModel(_panel(4)) as g, cubens.LearnedPrecond(7) as L:
        p = _params(11)
        L.set_params(p)
        vals = g.tangent(cubens.GEN_BASE, cubens.GEN_BASE)
        gr, eq = _oracle_graph(g)
        f = LO.features(vals, gr, eq)
        net = LO.Net(p.astype(np.float64), gr, f)
        gd_o, go_o = net.forward()
        gd, go = L.blocks(g)
        assert relerr(gd, gd_o) < 2e-2
        assert relerr(go, go_o) < 2e-2
        # pairs with K u = r from the FP64 band solve
        g.factor(cubens.options(cubens.SOLVER_BAND64))
        rng = np.random.default_rng(5)
        R = rng.standard_normal((3, g.neq))
        U = np.stack([g.solve(r) for r in R])
        loss, grad = L.loss_grad(g, R, U)
        free = eq >= 0
        Rs = np.zeros((3, g.nj * 6)); Us = np.zeros_like(Rs)
        Rs[:, free] = R[:, eq[free]]
        Us[:, free] = U[:, eq[free]]
        loss_o, grad_o = net.loss_grad(Rs, Us)
        assert abs(loss / loss_o - 1) < 1e-3
        err = np.linalg.norm(grad - grad_o) / np.linalg.norm(grad_o)
        cos = grad @ grad_o / (np.linalg.norm(grad) * np.linalg.norm(grad_o))
        assert err < 5e-2 and cos > 0.995, (err, cos)


def test_untrained_network_is_exactly_block_jacobi():
    """Decoder initialised at zero: the learned FGMRES run takes the same
    Krylov iterations as block-Jacobi FGMRES."""
    path = _panel(6)
    opt = dict(krylov_tol=1e-8, restart=30, maxit=20000, judge=1, max_steps=3)
    with cubens.GpuModel(path) as g:
        rb = g.run(cubens.options(cubens.SOLVER_BJ_FGMRES, **opt))
    with cubens.GpuModel(path) as g, cubens.LearnedPrecond(3) as L:
        L.attach(g)
        rl = g.run(cubens.options(cubens.SOLVER_LEARNED_FGMRES, **opt))
    assert rb["status"] == 0 and rl["status"] == 0
    assert relerr(rl["d"], rb["d"]) < 1e-7
    assert abs(rl["krylov_iters"] - rb["krylov_iters"]) <= 0.02 * rb["krylov_iters"]


def test_trained_network_beats_block_jacobi_on_a_held_out_panel():
    """Trained on the arc-length solves of seeded 8x8 panels, selected on a
    validation panel, evaluated on a held-out panel: fewer FGMRES iterations
    than block-Jacobi at 1e-8, the same path as the plain FP64 solve to 1e-6."""
    from cubens import learned
    with cubens.LearnedPrecond(0) as L:
        train = [_panel(8, seed=s) for s in (101, 102, 103, 104)]
        hist = learned.fit(L, train, _panel(10, seed=998), rounds=6, lr=2e-4)
        assert hist["best_round"] > 0, hist
        assert hist["best_val_iters"] < hist["val_iters"][0], hist
        path = _panel(10, seed=999)
        opt = dict(krylov_tol=1e-8, restart=30, maxit=20000, judge=1, max_steps=4)
        with cubens.GpuModel(path) as g:
            ref = g.run(cubens.options(cubens.SOLVER_BAND64, max_steps=4))
            bj = g.run(cubens.options(cubens.SOLVER_BJ_FGMRES, **opt))
            L.attach(g)
            ln = g.run(cubens.options(cubens.SOLVER_LEARNED_FGMRES, **opt))
        for r in (bj, ln):
            assert r["status"] == 0, r["msg"]
            assert relerr(r["d"], ref["d"]) < 1e-6
        assert ln["krylov_fallbacks"] == 0
        print(f"held-out 10x10 panel: block-Jacobi {bj['krylov_iters']} FGMRES iterations, "
              f"learned {ln['krylov_iters']} (validation history {hist['val_iters']})")
        assert ln["krylov_iters"] < bj["krylov_iters"], (ln["krylov_iters"], bj["krylov_iters"])
