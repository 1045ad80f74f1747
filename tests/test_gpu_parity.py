"""GPU path against the oracle, through the C ABI (libcubens_amd.so).

Floating-point tolerances: per-kernel results (tangent entries, internal forces,
geometry) agree to 1e-12 relative to the largest entry; converged paths of the
plain FP64 GPU path agree with the oracle's to 1e-9 relative (the bar of the
FP64 path), the mixed FP32-preconditioned path to 1e-6.  Integer maps are
compared exactly.
"""
import numpy as np
import pytest

import cubens
import oracle as O
from helpers import panel, relerr, skyline_dense, truss_c1

pytestmark = pytest.mark.gpu


def _open_both(path):
    return cubens.GpuModel(path), O.OracleModel(path)


def _perturbed_state(om, seed=3, amp=2e-4):
    """A deformed configuration: random dd applied through the oracle's updatc."""
    rng = np.random.default_rng(seed)
    dd = amp * rng.standard_normal(om.neq)
    x, xip, defllen, deffarea, defslen, c1, c2, c3 = O.updatc(om, om.x, dd)
    return dd, x, deffarea, c1, c2, c3


@pytest.mark.parametrize("path_fn", [truss_c1, lambda: panel(8)])
def test_dofmaps_and_pattern(path_fn):
    path = path_fn()
    g, om = _open_both(path)
    jc, mc, mx = g.dofmaps()
    assert np.array_equal(jc, om.jcode)
    assert np.array_equal(mc, om.mcode)
    assert np.array_equal(mx, om.maxa)
    # scalar pattern of the joint blocks == union of element DOF pairs
    rp, ci = g.pattern()
    bs = g.info.bs
    eqs = cubens.eq_of_slot(jc, g.nj, bs)
    got = set()
    for a in range(g.nj):
        for k in range(rp[a], rp[a + 1]):
            b = ci[k]
            for p in range(bs):
                for q in range(bs):
                    i, j = eqs[a * bs + p], eqs[b * bs + q]
                    if i >= 0 and j >= 0:
                        got.add((i, j))
    want = set()
    w = 6 if om.ntr else 18
    for e in range(len(om.mcode) // w):
        m = om.mcode[e * w:(e + 1) * w]
        for i in m:
            for j in m:
                if i and j:
                    want.add((i - 1, j - 1))
    assert got == want
    # its envelope reproduces the skyline column heights
    top = np.full(om.neq, np.iinfo(np.int64).max)
    for i, j in got:
        top[j] = min(top[j], i)
    assert np.array_equal(np.arange(om.neq) - top, om.kht)
    g.close()


def _dense_from_gpu(g, vals):
    jc, _, _ = g.dofmaps()
    rp, ci = g.pattern()
    return cubens.bsr_to_dense(vals, rp, ci, g.info.bs, cubens.eq_of_slot(jc, g.nj, g.info.bs), g.neq)


def test_shell_tangent_initial_and_deformed():
    path = panel(8, thick_sigma=0.1, imperfection=0.1)
    g, om = _open_both(path)
    # initial state (MSAL call pattern: converged coordinates and triads)
    Kg = _dense_from_gpu(g, g.tangent(cubens.GEN_BASE, cubens.GEN_BASE))
    Ko = skyline_dense(O.stiff_sh(om, om.x, om.c1, om.c2, om.c3, om.farea), om.maxa)
    assert relerr(Kg, Ko) < 1e-12
    assert np.abs(Kg - Kg.T).max() == 0.0
    # deformed state with membrane forces (geometric stiffness active)
    dd, x, deffarea, c1, c2, c3 = _perturbed_state(om)
    g.set("x", x)
    g.set("c1", c1); g.set("c2", c2); g.set("c3", c3)
    g.set("deffarea", deffarea)
    Kg = _dense_from_gpu(g, g.tangent(cubens.GEN_BASE, cubens.GEN_BASE))
    Ko = skyline_dense(O.stiff_sh(om, x, c1, c2, c3, deffarea), om.maxa)
    assert relerr(Kg, Ko) < 1e-12
    g.close()


def test_truss_tangent_and_force():
    path = truss_c1()
    g, om = _open_both(path)
    rng = np.random.default_rng(5)
    dd = 1e-3 * rng.standard_normal(om.neq)
    x, xip, defllen, _, _, c1, c2, c3 = O.updatc(om, om.x, dd)
    ef = np.zeros(om.ntr * 2)
    ef[0::2] = 50 * rng.standard_normal(om.ntr)
    ef[1::2] = -ef[0::2]
    for name, v in (("c1_ip", c1), ("c2_ip", c2), ("c3_ip", c3), ("defllen_ip", defllen), ("ef_ip", ef),
                    ("x_temp", x)):
        g.set(name, v)
    Kg = _dense_from_gpu(g, g.tangent(cubens.GEN_I, cubens.GEN_IP))
    Ko = skyline_dense(O.stiff_tr(om, c1, c2, c3, defllen, ef), om.maxa)
    assert relerr(Kg, Ko) < 1e-12
    # internal force of the same state
    for name, v in (("c1_i", c1), ("c2_i", c2), ("c3_i", c3), ("defllen_i", defllen)):
        g.set(name, v)
    fg = g.internal_force(np.zeros(om.neq), cubens.GEN_I)
    fo, efo = O.forces_tr(om, c1, c2, c3, defllen)
    assert relerr(fg, fo) < 1e-12
    assert relerr(g.get("ef_i", om.ntr * 2), efo) < 1e-12
    g.close()


def test_updatc_matches_oracle():
    path = panel(8, imperfection=0.1)
    g, om = _open_both(path)
    rng = np.random.default_rng(7)
    dd = 1e-3 * rng.standard_normal(om.neq)
    g.set("x_temp", om.x)
    g.update_geometry(dd, cubens.GEN_I)
    x, xip, _, deffarea, defslen, c1, c2, c3 = O.updatc(om, om.x, dd)
    assert relerr(g.get("x_temp", om.nj * 3), x) < 1e-15
    assert relerr(g.get("x_ip", om.nj * 3), xip) == 0.0
    assert relerr(g.get("deffarea_i", om.nsh), deffarea) < 1e-13
    assert relerr(g.get("defslen_i", om.nsh * 3), defslen) < 1e-13
    for nm, v in (("c1_i", c1), ("c2_i", c2), ("c3_i", c3)):
        assert relerr(g.get(nm, len(v)), v) < 1e-12
    g.close()


def test_shell_internal_force_matches_oracle():
    path = panel(8, thick_sigma=0.1)
    g, om = _open_both(path)
    dd, x, deffarea, c1, c2, c3 = _perturbed_state(om, seed=11)
    ntri = 3 * om.nsh
    rng = np.random.default_rng(13)
    ef_ip = 10.0 * rng.standard_normal(18 * om.nsh)
    # previous triad = initial, current = after dd
    g.set("x_temp", x)
    for nm, v in (("c1_ip", om.c1), ("c2_ip", om.c2), ("c3_ip", om.c3), ("c1_i", c1), ("c2_i", c2),
                  ("c3_i", c3), ("ef_ip", ef_ip)):
        g.set(nm, v)
    fg = g.internal_force(dd, cubens.GEN_I)
    fo, efi, efp = O.forces_sh(om, dd, x, (om.c1, om.c2, om.c3), (c1, c2, c3), ef_ip)
    assert relerr(fg, fo) < 1e-12
    assert relerr(g.get("ef_i", 18 * om.nsh), efi) < 1e-12
    assert np.array_equal(g.get("ef_ip", 18 * om.nsh), efp)   # membrane entries cleared in place
    assert ntri == len(c1)
    g.close()


@pytest.mark.parametrize("n", [8, 16])
@pytest.mark.parametrize("schedule", [cubens.FACTOR_FUSED, cubens.FACTOR_SPLIT])
def test_band_factor_pivots_and_solve(n, schedule):
    path = panel(n, thick_sigma=0.1, imperfection=0.1)
    g, om = _open_both(path)
    g.tangent(cubens.GEN_BASE, cubens.GEN_BASE)
    det, piv = g.factor(cubens.options(cubens.SOLVER_BAND64, factor_schedule=schedule))
    ss = O.stiff_sh(om, om.x, om.c1, om.c2, om.c3, om.farea)
    rc, fac, opiv, odet = O.skyfact(om.maxa, ss)
    assert rc == 0 and det == odet == 0
    assert relerr(piv, opiv) < 1e-9
    rng = np.random.default_rng(17)
    b = rng.standard_normal(om.neq)
    x = g.solve(b)
    K = skyline_dense(ss, om.maxa)
    assert np.linalg.norm(K @ x - b) / np.linalg.norm(b) < 1e-10
    assert relerr(x, O.skysolve(om.maxa, fac, b)) < 1e-9
    g.close()


@pytest.mark.parametrize("n", [16, 96])
def test_band_factor_schedules_agree_bitwise(n):
    """The fused and split schedules apply the same tile operations in the
    same order per tile: identical pivots and solutions (the fused launch's
    flag hand-offs deliver every panel before its updates read it)."""
    g = cubens.GpuModel(panel(n, thick_sigma=0.1, imperfection=0.1))
    g.tangent(cubens.GEN_BASE, cubens.GEN_BASE)
    b = np.random.default_rng(3).standard_normal(g.neq)
    out = []
    for sch in (cubens.FACTOR_SPLIT, cubens.FACTOR_FUSED, cubens.FACTOR_SPLIT, cubens.FACTOR_FUSED):
        _, piv = g.factor(cubens.options(cubens.SOLVER_BAND64, factor_schedule=sch))
        out.append((piv.copy(), g.solve(b)))
    for piv, x in out[1:]:
        assert np.array_equal(piv, out[0][0]) and np.array_equal(x, out[0][1])
    g.close()


@pytest.mark.parametrize("solver,tol", [(cubens.SOLVER_BAND32_FGMRES, 1e-10), (cubens.SOLVER_BJ_FGMRES, 1e-10)])
def test_fgmres_solvers(solver, tol):
    path = panel(8)
    g, om = _open_both(path)
    g.tangent(cubens.GEN_BASE, cubens.GEN_BASE)
    g.factor(cubens.options(solver, krylov_tol=tol, restart=40, maxit=5000, judge=0))
    rng = np.random.default_rng(19)
    b = rng.standard_normal(om.neq)
    x = g.solve(b)
    K = skyline_dense(O.stiff_sh(om, om.x, om.c1, om.c2, om.c3, om.farea), om.maxa)
    assert np.linalg.norm(K @ x - b) / np.linalg.norm(b) < 20 * tol
    g.close()


def test_truss_dome_nr_path_matches_oracle():
    path = truss_c1()
    g, om = _open_both(path)
    gp = g.run(cubens.options(cubens.SOLVER_BAND64))
    op = om.run(40)
    assert gp["status"] == 0 and op["rc"] == 0
    assert gp["nsteps"] == op["nsteps"] == 20
    assert np.array_equal(gp["itecnt"], op["itecnt"])
    assert relerr(gp["lpf"], op["lpf"]) < 1e-12
    assert relerr(gp["d"], op["d"]) < 1e-9
    g.close()


@pytest.mark.parametrize("n,steps", [(8, 12), (16, 6)])
def test_panel_msal_path_matches_oracle(n, steps):
    path = panel(n)
    g, om = _open_both(path)
    gp = g.run(cubens.options(cubens.SOLVER_BAND64, max_steps=steps))
    op = om.run(steps)
    assert gp["nsteps"] == op["nsteps"] == steps
    assert np.array_equal(gp["itecnt"], op["itecnt"])
    assert relerr(gp["lpf"], op["lpf"]) < 1e-9
    assert relerr(gp["d"], op["d"]) < 1e-9
    g.close()


def test_panel_msal_mixed_precision_path():
    path = panel(8)
    g, om = _open_both(path)
    gp = g.run(cubens.options(cubens.SOLVER_BAND32_FGMRES, krylov_tol=1e-10, max_steps=8, judge=1))
    op = om.run(8)
    assert gp["nsteps"] == 8
    assert gp["krylov_fallbacks"] == 0
    assert relerr(gp["lpf"], op["lpf"]) < 1e-6
    assert relerr(gp["d"], op["d"]) < 1e-6
    g.close()


def test_reference_sample_5a_mnr_on_the_gpu():
    """The reference's own sample input (Sample_Input_Files/model_def_5a_truss.txt,
    modified Newton-Raphson, ALGFLAG 2) on the GPU against the path the
    reference's element routines give under the restated driver
    (tests/golden/sample_5a.npz)."""
    import os
    gold = np.load(os.path.join(os.path.dirname(__file__), "golden", "sample_5a.npz"))
    path = os.path.join(os.path.dirname(__file__), "golden", "model_def_5a_truss.txt")
    with cubens.GpuModel(path) as g:
        assert g.info.algflag == 2
        jc, _, _ = g.dofmaps()
        assert np.array_equal(jc, gold["jcode"])
        gp = g.run(cubens.options(cubens.SOLVER_BAND64, max_steps=50))
    assert gp["status"] == 0
    assert gp["nsteps"] == len(gold["path_lpf"])
    assert np.array_equal(gp["itecnt"], gold["path_itecnt"])
    assert relerr(gp["lpf"], gold["path_lpf"]) <= 1e-12
    assert relerr(gp["d"], gold["path_d"]) <= 1e-9
