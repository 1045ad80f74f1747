"""The FP32 precision mode (CB_SOLVER_BAND32) on the GPU through the C ABI:
the shell element stiffness in FP32 (k_shell_tangent<float>) against the FP64
kernel and the oracle's stiff_sh at the initial and a deformed state, and the
arc-length and Newton-Raphson paths whose corrections come from the FP32 band
LDL^T of that tangent while the internal forces, residual, arc-length
constraint and convergence tests stay FP64.

Bar: the FP32 tangent within 2e-6 of the FP64 one (relative to its largest
entry; FP32 rounds at 6e-8 and the DKT sums add a few roundings), fixed slots
exact.  The paths run the same converged steps and Newton iterations as the
FP64 path and the oracle.  Where the convergence tests are tight (the truss
dome, 1e-10) the displacements agree to 1e-6 (measured 6e-12); the shell
panels converge to a displacement criterion of 1e-3, inside which any change
of tangent settles elsewhere, and agree to 1e-5 (measured 8e-6 without
refinement, 3e-6 with one refinement step; the errors are printed per
path)."""
import numpy as np
import pytest

import cubens
import oracle as O
from helpers import model_file, panel, relerr, skyline_dense, truss_c1
from cubens import inputs

pytestmark = pytest.mark.gpu


def _dense(g, vals):
    jc, _, _ = g.dofmaps()
    rp, ci = g.pattern()
    return cubens.bsr_to_dense(np.asarray(vals, np.float64), rp, ci, g.info.bs,
                               cubens.eq_of_slot(jc, g.nj, g.info.bs), g.neq)


def _perturbed(om, seed=3, amp=2e-4):
    rng = np.random.default_rng(seed)
    dd = amp * rng.standard_normal(om.neq)
    x, xip, defllen, deffarea, defslen, c1, c2, c3 = O.updatc(om, om.x, dd)
    return x, deffarea, c1, c2, c3


def test_fp32_shell_tangent_against_fp64_and_the_oracle():
    path = panel(8, thick_sigma=0.1, imperfection=0.1)
    om = O.OracleModel(path)
    with cubens.GpuModel(path) as g:
        K32 = g.tangent_fp32(cubens.GEN_BASE, cubens.GEN_BASE)
        K64 = g.tangent(cubens.GEN_BASE, cubens.GEN_BASE)
        assert K32.dtype == np.float32
        # fixed slots are identity in both, exactly
        fixed = (K64 == 1.0) | (K64 == 0.0)
        assert np.array_equal(K32[fixed], K64[fixed].astype(np.float32))
        e0 = relerr(K32.astype(np.float64), K64)
        Ko = skyline_dense(O.stiff_sh(om, om.x, om.c1, om.c2, om.c3, om.farea), om.maxa)
        e0o = relerr(_dense(g, K32), Ko)
        x, deffarea, c1, c2, c3 = _perturbed(om)
        g.set("x", x)
        g.set("c1", c1); g.set("c2", c2); g.set("c3", c3)
        g.set("deffarea", deffarea)
        K32d = g.tangent_fp32(cubens.GEN_BASE, cubens.GEN_BASE)
        K64d = g.tangent(cubens.GEN_BASE, cubens.GEN_BASE)
        e1 = relerr(K32d.astype(np.float64), K64d)
        Kod = skyline_dense(O.stiff_sh(om, x, c1, c2, c3, deffarea), om.maxa)
        e1o = relerr(_dense(g, K32d), Kod)
    print(f"[precision] FP32 tangent error: initial {e0:.2e} (oracle {e0o:.2e}), deformed {e1:.2e} "
          f"(oracle {e1o:.2e})", flush=True)
    assert 0 < e0 < 2e-6 and 0 < e1 < 2e-6
    assert e0o < 2e-6 and e1o < 2e-6


@pytest.mark.parametrize("refine", [0, 1, 2])
@pytest.mark.parametrize("n,steps", [(8, 12), (16, 6)])
def test_fp32_precision_mode_panel_msal_path(n, steps, refine):
    path = panel(n)
    with cubens.GpuModel(path) as g:
        r64 = g.run(cubens.options(cubens.SOLVER_BAND64, max_steps=steps))
        r32 = g.run(cubens.options(cubens.SOLVER_BAND32, max_steps=steps, judge=0, refine=refine))
    o = O.OracleModel(path).run(steps)
    e_l, e_d = relerr(r32["lpf"], r64["lpf"]), relerr(r32["d"], r64["d"])
    print(f"[precision] panel {n}, refine {refine}: {steps} increments, Newton {int(np.sum(r32['itecnt']))} "
          f"(FP64 {int(np.sum(r64['itecnt']))}), lpf error {e_l:.2e}, d error {e_d:.2e}, vs oracle "
          f"{relerr(r32['d'], o['d']):.2e}", flush=True)
    assert r32["status"] == r64["status"] == 0, r32["msg"]
    assert r32["nsteps"] == r64["nsteps"] == o["nsteps"] == steps
    assert e_l < 1e-5 and e_d < 1e-5
    assert relerr(r32["d"], o["d"]) < 1e-5


@pytest.mark.parametrize("sweeps", [1, 2])
def test_fp32_precision_mode_under_every_sweep_schedule(sweeps):
    """The FP32 band's solves on the two-product row step and on the
    look-ahead system (both sweeps coupled two rows back) against the chain
    matrices: the same converged steps, the displacements to 1e-5 (the bar of
    the FP32 paths above; the schedules differ only in rounding)."""
    path = panel(8)
    with cubens.GpuModel(path) as g:
        a = g.run(cubens.options(cubens.SOLVER_BAND32, max_steps=12, judge=0, refine=1, sweep_schedule=0))
        b = g.run(cubens.options(cubens.SOLVER_BAND32, max_steps=12, judge=0, refine=1, sweep_schedule=sweeps))
    e = relerr(b["d"], a["d"])
    print(f"[precision] FP32 sweeps {sweeps} against the chain matrices: d error {e:.2e}", flush=True)
    assert a["status"] == b["status"] == 0, b["msg"]
    assert a["nsteps"] == b["nsteps"] == 12
    assert e < 1e-5 and relerr(b["lpf"], a["lpf"]) < 1e-5


def test_fp32_precision_mode_with_the_fp64_judge_pivots():
    """judge 1 takes det and psi from an FP64 factor of the FP64 tangent, as
    the other reduced-precision solvers do, and gives the same path."""
    path = panel(8)
    with cubens.GpuModel(path) as g:
        a = g.run(cubens.options(cubens.SOLVER_BAND32, max_steps=8, judge=0))
        b = g.run(cubens.options(cubens.SOLVER_BAND32, max_steps=8, judge=1))
    assert a["status"] == b["status"] == 0
    assert np.array_equal(a["itecnt"], b["itecnt"])
    assert np.array_equal(a["d"], b["d"])


def test_fp32_precision_mode_truss_dome_nr():
    """Bars take the FP64 tangent rounded once; NR load control to the end."""
    path = truss_c1()
    with cubens.GpuModel(path) as g:
        r64 = g.run(cubens.options(cubens.SOLVER_BAND64))
        r32 = g.run(cubens.options(cubens.SOLVER_BAND32, judge=0))
    print(f"[precision] dome: Newton {int(np.sum(r32['itecnt']))} (FP64 {int(np.sum(r64['itecnt']))}), "
          f"d error {relerr(r32['d'], r64['d']):.2e}", flush=True)
    assert r32["status"] == 0 and r32["nsteps"] == r64["nsteps"] == 20
    assert relerr(r32["lpf"], r64["lpf"]) < 1e-12
    assert relerr(r32["d"], r64["d"]) < 1e-6


def test_fp32_precision_mode_refuses_the_dynamic_drivers():
    path = model_file("prec_nm.txt", inputs.space_frame(40, kind="tower", steps=3).text)
    with cubens.GpuModel(path) as g:
        with pytest.raises(cubens.CubensError, match="static drivers"):
            g.run(cubens.options(cubens.SOLVER_BAND32, judge=0))


def test_unknown_solver_is_refused():
    with cubens.GpuModel(panel(4)) as g:
        with pytest.raises(cubens.CubensError, match="unknown solver"):
            g.run(cubens.options(7))
