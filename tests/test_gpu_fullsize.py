"""The FP64 band path at the workload sizes, through size-independent checks.

- 256x256 panel (131072 CST+DKT triangles, NEQ 394752), the benchmark's
  model: the factorization of the initial tangent has only positive pivots
  (the hinged panel's tangent is positive definite); the solve has a normwise
  backward error at the FP64 round-off level (K applied on the host from the
  tangent the device assembled); the fused and split schedules give bitwise
  identical pivots and solutions.
- 64x64 panel (8192 triangles, NEQ 24960): the arc-length path of the plain
  FP64 GPU path follows the oracle's to 1e-9 with identical iteration counts.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import cubens
import oracle as O
from cubens import inputs
from helpers import model_file, panel, relerr

pytestmark = pytest.mark.gpu


def _tangent_operator(g):
    """K (free equations) as a scipy matrix from the device's joint-block tangent."""
    vals = g.tangent(cubens.GEN_BASE, cubens.GEN_BASE)
    rp, ci = g.pattern()
    bs = g.info.bs
    B = sp.bsr_matrix((vals.reshape(-1, bs, bs), ci, rp), shape=(g.nj * bs, g.nj * bs))
    jc, _, _ = g.dofmaps()
    eq = cubens.eq_of_slot(jc, g.nj, bs)
    free = np.flatnonzero(eq >= 0)
    P = sp.csr_matrix((np.ones(len(free)), (eq[free], free)), shape=(g.neq, g.nj * bs))
    return (P @ B.tocsr() @ P.T).tocsr()


def test_full_size_factor_and_solve():
    path = model_file("panel256_c3.txt",
                      inputs.shell_panel(n=256, thick_sigma=0.1, imperfection=0.1, config_index=3).text)
    with cubens.GpuModel(path) as g:
        assert g.neq == 394752
        K = _tangent_operator(g)
        b = np.random.default_rng(11).standard_normal(g.neq)
        out = []
        for sch in (cubens.FACTOR_FUSED, cubens.FACTOR_SPLIT):
            det, piv = g.factor(cubens.options(cubens.SOLVER_BAND64, factor_schedule=sch))
            out.append((det, piv.copy(), g.solve(b)))
        det, piv, x = out[0]
        assert det == 0 and np.all(piv > 0), (det, int(np.sum(piv <= 0)))
        r = K @ x - b
        res = np.linalg.norm(r) / np.linalg.norm(b)
        # normwise backward error (infinity norms): what a backward-stable
        # factorization bounds, independent of the conditioning of K
        knorm = np.abs(K).sum(axis=1).max()
        eta = np.abs(r).max() / (knorm * np.abs(x).max() + np.abs(b).max())
        print(f"NEQ {g.neq}: |Kx-b|/|b| = {res:.2e}, backward error {eta:.2e}, "
              f"pivots in [{piv.min():.3e}, {piv.max():.3e}]")
        assert eta < 1e-13
        assert out[1][0] == det
        assert np.array_equal(out[1][1], piv) and np.array_equal(out[1][2], x)


def test_panel64_msal_path_matches_oracle():
    path = panel(64)
    steps = 6
    g, om = cubens.GpuModel(path), O.OracleModel(path)
    gp = g.run(cubens.options(cubens.SOLVER_BAND64, max_steps=steps))
    op = om.run(steps)
    assert g.neq == 24960
    assert gp["nsteps"] == op["nsteps"] == steps
    assert np.array_equal(gp["itecnt"], op["itecnt"])
    assert relerr(gp["lpf"], op["lpf"]) < 1e-9
    assert relerr(gp["d"], op["d"]) < 1e-9
    g.close()
