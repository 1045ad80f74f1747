"""The band sweeps with precomputed chain matrices (CB_SWEEP_CHAIN: one
64x64 product with Linv_I L_{I,I-1} per chain step) against the sweeps with two
dependent products per step (CB_SWEEP_ROWS) and against the tangent itself:
both give a normwise backward error at the FP64 round-off level for random
right-hand sides, on models whose band shapes cover the kernels' edge cases --
a single tile row (nt = 1), odd and even numbers of tile rows, and 1, 2 and
more tiles below the diagonal -- and the arc-length paths of both schedules
agree with the oracle.  The look-ahead systems (CB_SWEEP_LOOK2, rows coupled
two and more back, and CB_SWEEP_LOOK3, three and more back) are held to the
same bars."""
import numpy as np
import pytest
import scipy.sparse as sp

import cubens
import oracle as O
from cubens import inputs
from helpers import model_file, relerr

pytestmark = pytest.mark.gpu


def _tangent_operator(g):
    vals = g.tangent(cubens.GEN_BASE, cubens.GEN_BASE)
    rp, ci = g.pattern()
    bs = g.info.bs
    B = sp.bsr_matrix((vals.reshape(-1, bs, bs), ci, rp), shape=(g.nj * bs, g.nj * bs))
    jc, _, _ = g.dofmaps()
    eq = cubens.eq_of_slot(jc, g.nj, bs)
    free = np.flatnonzero(eq >= 0)
    P = sp.csr_matrix((np.ones(len(free)), (eq[free], free)), shape=(g.neq, g.nj * bs))
    return (P @ B.tocsr() @ P.T).tocsr()


CASES = {
    "star3": lambda: inputs.star_truss(),                         # NEQ 3: one tile row
    "dome": lambda: inputs.truss_dome(),                          # NEQ 243: 4 tile rows
    "panel8": lambda: inputs.shell_panel(n=8),                    # 7 tile rows (odd)
    "panel16": lambda: inputs.shell_panel(n=16),                  # 26 tile rows
    "panel24": lambda: inputs.shell_panel(n=24, thick_sigma=0.1, imperfection=0.1),
    "panel40": lambda: inputs.shell_panel(n=40),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_chain_and_row_sweeps_solve_the_tangent(case):
    path = model_file(f"sweep_{case}.txt", CASES[case]().text)
    print(f"[sweeps] {case}", flush=True)
    with cubens.GpuModel(path) as g:
        K = _tangent_operator(g)
        nt = (g.neq + 63) // 64
        w = min((g.info.bandwidth + 63) // 64, nt - 1)
        rng = np.random.default_rng(nt * 31 + w)
        b = rng.standard_normal(g.neq)
        xs = {}
        for sch in (cubens.SWEEP_CHAIN, cubens.SWEEP_ROWS, cubens.SWEEP_LOOK2, cubens.SWEEP_LOOK3):
            det, piv = g.factor(cubens.options(cubens.SOLVER_BAND64, sweep_schedule=sch))
            xs[sch] = g.solve(b)
        knorm = np.abs(K).sum(axis=1).max()
        for sch, x in xs.items():
            r = K @ x - b
            eta = np.abs(r).max() / (knorm * np.abs(x).max() + np.abs(b).max())
            print(f"[sweeps] {case} nt {nt} w {w} schedule {sch}: backward error {eta:.2e}", flush=True)
            assert eta < 1e-13
        # the schedules differ only in rounding
        assert relerr(xs[cubens.SWEEP_CHAIN], xs[cubens.SWEEP_ROWS]) < 1e-8
        assert relerr(xs[cubens.SWEEP_LOOK2], xs[cubens.SWEEP_CHAIN]) < 1e-8
        assert relerr(xs[cubens.SWEEP_LOOK3], xs[cubens.SWEEP_CHAIN]) < 1e-8


@pytest.mark.parametrize("sch", [0, 1, 2, 3])
def test_panel_path_of_every_sweep_schedule_matches_the_oracle(sch):
    path = model_file("sweep_panel16_path.txt", inputs.shell_panel(n=16).text)
    with cubens.GpuModel(path) as g:
        gp = g.run(cubens.options(cubens.SOLVER_BAND64, max_steps=6, sweep_schedule=sch))
    op = O.OracleModel(path).run(6)
    assert np.array_equal(gp["itecnt"], op["itecnt"])
    assert relerr(gp["lpf"], op["lpf"]) < 1e-9
    assert relerr(gp["d"], op["d"]) < 1e-9
