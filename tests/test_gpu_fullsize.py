"""The FP64 band path at the workload sizes, through size-independent checks.

- 256x256 panel (131072 CST+DKT triangles, NEQ 394752), the benchmark's
  model: the factorization of the initial tangent has only positive pivots
  (the hinged panel's tangent is positive definite); the solve has a normwise
  backward error at the FP64 round-off level (K applied on the host from the
  tangent the device assembled) with the chain-matrix and with the two-product
  sweeps; the fused and split factorization schedules give bitwise
  identical pivots and solutions.
- 64x64 panel (8192 triangles, NEQ 24960): the arc-length path of the plain
  FP64 GPU path follows the oracle's to 1e-9 with identical iteration counts,
  live over 6 increments and against the oracle's fixture over the 60
  increments of the workload (tests/golden/panel64_path.npz);
- 256x256 panel: the first increment against the oracle's fixture
  (tests/golden/panel256_first.npz: one skyline factorization of the serial
  path, 3 minutes on one core), and over the 40 benchmarked increments the
  size-independent invariants of every converged step: the spherical
  arc-length constraint rebuilt from the path, and the out-of-balance force
  within the force criterion of test() (misc.c:215-232).

Fixture comparisons cover the whole displacement vector through its 2-norm
and a weighted checksum, and every 97th entry directly.
"""
import os
import sys

import numpy as np
import pytest
import scipy.sparse as sp

import cubens
import oracle as O
from cubens import inputs
from helpers import arc_residuals, model_file, panel, relerr

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
        # the two-product sweeps on the same factor: also at round-off
        g.factor(cubens.options(cubens.SOLVER_BAND64, sweep_schedule=cubens.SWEEP_ROWS))
        x2 = g.solve(b)
        r2 = K @ x2 - b
        eta2 = np.abs(r2).max() / (knorm * np.abs(x2).max() + np.abs(b).max())
        print(f"two-product sweeps: backward error {eta2:.2e}")
        assert eta2 < 1e-13


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


GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _weights(neq):
    sys.path.insert(0, GOLD)
    from make_fullsize import make_weights
    return make_weights(neq)


def _against_fixture(gp, fx):
    n = len(fx["lpf"])
    assert gp["status"] == int(fx["status"]) and gp["nsteps"] == n
    assert np.array_equal(gp["itecnt"], fx["itecnt"])
    # fixtures of elastic runs keep the first five event slots; the yield and
    # unload slots after them stay zero
    ne = fx["events"].shape[1]
    assert np.array_equal(gp["events"][:, :ne], fx["events"]) and not gp["events"][:, ne:].any()
    assert np.abs(gp["psi"] - fx["psi"]).max() <= 1e-6
    assert relerr(gp["lpf"], fx["lpf"]) <= 1e-9
    d = gp["d"]
    assert relerr(d[:, fx["sel"]], fx["d_sel"]) <= 1e-9
    assert relerr(np.linalg.norm(d, axis=1), fx["d_norm"]) <= 1e-9
    assert relerr(d @ _weights(d.shape[1]), fx["d_check"]) <= 1e-9


def _invariants(gp, q, tolforc=1e-3, msal=(1.0, 6, 0.1)):
    res = arc_residuals(msal, gp, q)
    assert len(res) >= gp["nsteps"] - 1 - int((gp["events"][:, 3] > 0).sum())
    worst_arc = max(res.values()) if res else 0.0
    lam = gp["lpf"]
    dl = np.abs(np.diff(np.concatenate([[0.0], lam])))
    # |lambda q - f| within the force criterion |qtot - f| <= tolforc |qtot - fp|
    # (qtot is formed before the last load-factor correction, main.c:2845-2855)
    ratio = gp["res"] / (dl * np.linalg.norm(q))
    print(f"[fullsize] worst arc-constraint residual {worst_arc:.2e}, "
          f"worst |lq-f|/(|dl||q|) {ratio.max():.3e}", flush=True)
    assert worst_arc <= 1e-8
    assert ratio.max() <= 2 * tolforc


def test_panel64_sixty_increments_against_fixture():
    fx = np.load(os.path.join(GOLD, "panel64_path.npz"))
    with cubens.GpuModel(panel(64)) as g:
        gp = g.run(cubens.options(cubens.SOLVER_BAND64, max_steps=len(fx["lpf"])))
        q = g.load()
    _against_fixture(gp, fx)
    _invariants(gp, q)


def _panel256():
    return model_file("panel256_c3.txt",
                      inputs.shell_panel(n=256, thick_sigma=0.1, imperfection=0.1, config_index=3).text)


def test_panel256_first_increment_against_fixture():
    fx = np.load(os.path.join(GOLD, "panel256_first.npz"))
    with cubens.GpuModel(_panel256()) as g:
        gp = g.run(cubens.options(cubens.SOLVER_BAND64, max_steps=1))
    _against_fixture(gp, fx)


def test_panel256_forty_increments_invariants():
    with cubens.GpuModel(_panel256()) as g:
        gp = g.run(cubens.options(cubens.SOLVER_BAND64, max_steps=40))
        q = g.load()
    assert gp["status"] == 0 and gp["nsteps"] == 40
    _invariants(gp, q)
