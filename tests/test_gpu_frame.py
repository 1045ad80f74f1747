"""Frame members, lumped masses and the nonlinear Newmark driver on the GPU,
through the C ABI, against the oracle and the golden paths of the reference's
frame routines.

Tolerances: per-kernel results (tangent entries, internal forces, masses)
agree to 1e-12 relative to the largest entry; converged paths of the FP64
GPU path agree with the oracle's to 1e-9 relative with identical iteration
counts.
"""
import numpy as np
import pytest

import cubens
import oracle as O
from cubens import inputs
from helpers import model_file, relerr, skyline_dense

pytestmark = pytest.mark.gpu

GOLD = __import__("os").path.join(__import__("os").path.dirname(__file__), "golden")


def _dense_from_gpu(g, vals):
    jc, _, _ = g.dofmaps()
    rp, ci = g.pattern()
    return cubens.bsr_to_dense(vals, rp, ci, g.info.bs, cubens.eq_of_slot(jc, g.nj, g.info.bs), g.neq)


def _tower(n=60, **kw):
    return model_file(f"gtower{n}.txt", inputs.space_frame(n, kind="tower", **kw).text)


def _deformed(om, seed=5, amp=3e-3):
    rng = np.random.default_rng(seed)
    dd = amp * rng.standard_normal(om.neq)
    x, xip, defllen, deffarea, defslen, c1, c2, c3 = O.updatc(om, om.x, dd)
    return dd, x, defllen, c1, c2, c3


def test_frame_dofmaps_and_pattern_envelope():
    path = _tower(80, steps=5)
    g, om = cubens.GpuModel(path), O.OracleModel(path)
    jc, mc, mx = g.dofmaps()
    assert np.array_equal(jc, om.jcode) and np.array_equal(mc, om.mcode)
    assert np.array_equal(mx, om.maxa)
    g.close()


def test_frame_tangent_and_forces_at_a_deformed_state():
    """stiff_fr / forces_fr with non-zero previous end forces (geometric
    stiffness), deformed triads and lengths, against the oracle."""
    path = _tower(60, steps=5)
    g, om = cubens.GpuModel(path), O.OracleModel(path)
    dd, x, defllen, c1, c2, c3 = _deformed(om)
    rng = np.random.default_rng(11)
    ef_ip = 1e4 * rng.standard_normal(om.nfr * 14)
    # previous iterate: initial triads; current: deformed
    for name, v in (("c1_ip", om.c1), ("c2_ip", om.c2), ("c3_ip", om.c3), ("ef_ip", ef_ip),
                    ("defllen_ip", defllen), ("c1_i", c1), ("c2_i", c2), ("c3_i", c3),
                    ("x_temp", x)):
        g.set(name, v)
    Kg = _dense_from_gpu(g, g.tangent(cubens.GEN_I, cubens.GEN_IP))
    Ko = skyline_dense(O.stiff_fr(om, om.llength, defllen, om.c1, om.c2, om.c3, ef_ip), om.maxa)
    assert relerr(Kg, Ko) < 1e-12
    fg = g.internal_force(dd, cubens.GEN_I)
    fo, efo = O.forces_fr(om, dd, om.llength, defllen, (om.c1, om.c2, om.c3), (c1, c2, c3), ef_ip)
    assert relerr(fg, fo) < 1e-12
    assert relerr(g.get("ef_i", om.nfr * 14), efo) < 1e-12
    g.close()


def test_frame_geometry_update():
    path = _tower(60, steps=5)
    g, om = cubens.GpuModel(path), O.OracleModel(path)
    dd, x, defllen, c1, c2, c3 = _deformed(om, seed=8)
    g.update_geometry(dd, cubens.GEN_I)
    assert relerr(g.get("x_temp", om.nj * 3), x) < 1e-14
    assert relerr(g.get("defllen_i", om.nfr), defllen) < 1e-14
    for name, v in (("c1_i", c1), ("c2_i", c2), ("c3_i", c3)):
        assert relerr(g.get(name, om.nfr * 3), v) < 1e-13
    g.close()


def test_lumped_masses_against_oracle():
    path = _tower(60, steps=5)
    g, om = cubens.GpuModel(path), O.OracleModel(path)
    sm_o, llen_o, _, _ = O.mass(om, om.x)
    assert relerr(g.mass(), sm_o) < 1e-14
    assert relerr(g.get("llength", om.nfr), llen_o) < 1e-15
    g.close()


def test_shell_masses_against_oracle():
    path = model_file("p4m.txt", inputs.shell_panel(n=4, thick_sigma=0.1, imperfection=0.1).text)
    g, om = cubens.GpuModel(path), O.OracleModel(path)
    sm_o, _, fa, sl = O.mass(om, om.x)
    assert relerr(g.mass(), sm_o) < 1e-14
    assert relerr(g.get("farea", om.nsh), fa) < 1e-15
    assert relerr(g.get("slength", om.nsh * 3), sl) < 1e-15
    g.close()


def test_frame_newmark_path_against_reference_fixture():
    """The tower under harmonic lateral load (30 Newmark steps): the plain FP64
    GPU path follows the path the reference's frame routines produced."""
    z = np.load(__import__("os").path.join(GOLD, "frame_tower_newmark.npz"))
    path = model_file("gold_tower.txt", str(z["model"]))
    with cubens.GpuModel(path) as g:
        r = g.run(cubens.options(cubens.SOLVER_BAND64, max_steps=30))
    assert r["status"] == 0, r["msg"]
    assert np.array_equal(r["itecnt"], z["path_itecnt"])
    assert np.allclose(r["lpf"], z["path_lpf"], rtol=0, atol=1e-15)
    assert relerr(r["d"], z["path_d"]) < 1e-9


@pytest.mark.parametrize("kind,numopt,rho", [("grid", 0, 1.0), ("tower", 1, 0.8)])
def test_frame_newmark_paths_against_oracle(kind, numopt, rho):
    txt = inputs.space_frame(150, kind=kind, amp=8e4, steps=25, numopt=numopt, spectrds=rho).text
    path = model_file(f"nm_{kind}{numopt}.txt", txt)
    ro = O.OracleModel(path).run(25)
    assert ro["rc"] == 0, ro["msg"]
    with cubens.GpuModel(path) as g:
        r = g.run(cubens.options(cubens.SOLVER_BAND64, max_steps=25))
    assert r["status"] == 0, r["msg"]
    assert np.array_equal(r["itecnt"], ro["itecnt"])
    assert relerr(r["d"], ro["d"]) < 1e-9


def test_frame_static_nr_path_against_oracle():
    txt = inputs.space_frame(90, kind="grid", algflag=1, amp=2.5e4, tol=(1e-10, 1e-8, 1e-12)).text
    path = model_file("nr_grid.txt", txt)
    ro = O.OracleModel(path).run(25)
    with cubens.GpuModel(path) as g:
        r = g.run(cubens.options(cubens.SOLVER_BAND64, max_steps=25))
    assert r["status"] == 0, r["msg"]
    assert np.array_equal(r["itecnt"], ro["itecnt"])
    assert relerr(r["d"], ro["d"]) < 1e-9
