"""Member end bending releases (release frame.c:798-896, inverse
misc.c:284-342) on the GPU through the C ABI: paths against the reference
routines (tests/golden/releases.npz) and the oracle under NR, Newmark,
ANAFLAG 3 (NR and MSAL) and free warping; the clamped beam's textbook
deflections with one to four releases on a member and zero end moments at the
released DOFs; the tangent and internal force at a deformed state against the
oracle with releases and offsets together, 6 and 7 slots per joint.

Bar: FP64 paths 1e-9 with identical iteration counts, driver events and final
status; kernels 1e-12."""
import numpy as np
import pytest

import cubens
import oracle as O
from cubens import inputs
from helpers import model_file, relerr, skyline_dense
from test_oracle_releases import BEAMS, CASES, L, gold

pytestmark = pytest.mark.gpu


def _opts(steps=250):
    return cubens.options(cubens.SOLVER_BAND64, max_steps=steps)


@pytest.mark.parametrize("name", CASES)
def test_release_paths_match_reference_routines_and_oracle(name):
    g = gold()
    path = model_file(f"gpu_rel_{name}.txt", str(g[f"{name}_model"]))
    print(f"[releases] {name}", flush=True)
    with cubens.GpuModel(path) as m:
        r = m.run(_opts())
    o = O.OracleModel(path).run(250)
    assert r["status"] == int(g[f"{name}_status"]) == o["status"], (r["msg"], o["msg"])
    for ref in ({"itecnt": g[f"{name}_itecnt"], "lpf": g[f"{name}_lpf"], "d": g[f"{name}_d"],
                 "events": g[f"{name}_events"]}, o):
        assert np.array_equal(r["itecnt"], ref["itecnt"])
        assert np.array_equal(r["events"], ref["events"])
        assert relerr(r["lpf"], ref["lpf"]) < 1e-9
        assert relerr(r["d"], ref["d"]) < 1e-9


@pytest.mark.parametrize("f1,f2,strong,weak", BEAMS)
@pytest.mark.parametrize("direction", [3, 2])
def test_clamped_beam_with_releases_on_the_gpu(f1, f2, strong, weak, direction):
    mt = inputs.clamped_beam(n_elem=2, length=L, direction=direction)
    rows = {m: f for m, f in ((1, f1), (2, f2)) if any(f)}
    if rows:
        mt = mt.with_releases(rows)
    path = model_file("gpu_rel_beam.txt", mt.text)
    with cubens.GpuModel(path) as m:
        r = m.run(_opts(3))
        ef = m.get("ef", 28).reshape(2, 14)
        jc, _, _ = m.dofmaps()
    assert r["status"] == 0 and r["nsteps"] == 1, r["msg"]
    P, EI = mt.info["load"], mt.info["E"] * mt.info["I"]
    mid = jc.reshape(-1, 7)[1]
    want = (strong if direction == 3 else weak) * P * L ** 3 / EI
    assert abs(r["d"][0][mid[direction - 1] - 1] - want) <= 1e-11 * abs(want)
    for mem, f in ((0, f1), (1, f2)):
        for flag, dof in zip(f, (5, 4, 12, 11)):
            if flag:
                assert abs(ef[mem, dof]) <= 1e-9 * abs(P) * L


@pytest.mark.parametrize("warping", [False, True])
def test_kernels_with_releases_and_offsets_at_a_deformed_state(warping):
    mt = inputs.space_frame(60, kind="tower", steps=5, warping=warping)
    rng = np.random.default_rng(41)
    flags = [(1, 0, 0, 0), (0, 1, 1, 0), (1, 1, 1, 0), (1, 1, 1, 1), (0, 0, 0, 1), (1, 0, 1, 0)]
    rel = {m: flags[m % len(flags)] for m in range(1, mt.n_fr + 1) if m % 4 == 1}
    ofs = {m: tuple(0.04 * rng.uniform(-1, 1, 6)) for m in range(1, mt.n_fr + 1) if m % 3 == 0}
    path = model_file(f"gpu_rel_tower{int(warping)}.txt", mt.with_offsets(ofs).with_releases(rel).text)
    om = O.OracleModel(path)
    dd = 3e-3 * rng.standard_normal(om.neq)
    x, xip, defllen, deffarea, defslen, c1, c2, c3 = O.updatc(om, om.x, dd)
    ef_ip = 1e4 * rng.standard_normal(om.nfr * 14)
    print(f"[releases] kernels, warping {warping}", flush=True)
    with cubens.GpuModel(path) as g:
        bs = g.info.bs
        for name, v in (("c1_ip", om.c1), ("c2_ip", om.c2), ("c3_ip", om.c3), ("ef_ip", ef_ip),
                        ("defllen_ip", defllen), ("c1_i", c1), ("c2_i", c2), ("c3_i", c3), ("x_temp", x)):
            g.set(name, v)
        jc, _, _ = g.dofmaps()
        rp, ci = g.pattern()
        Kg = cubens.bsr_to_dense(g.tangent(cubens.GEN_I, cubens.GEN_IP), rp, ci, bs,
                                 cubens.eq_of_slot(jc, g.nj, bs), g.neq)
        fg = g.internal_force(dd, cubens.GEN_I)
        efg = g.get("ef_i", om.nfr * 14)
    Ko = skyline_dense(O.stiff_fr(om, om.llength, defllen, om.c1, om.c2, om.c3, ef_ip), om.maxa)
    fo, efo = O.forces_fr(om, dd, om.llength, defllen, (om.c1, om.c2, om.c3), (c1, c2, c3), ef_ip)
    assert relerr(Kg, Ko) < 1e-12
    assert relerr(fg, fo) < 1e-12
    assert relerr(efg, efo) < 1e-12


def test_release_flags_other_than_0_or_1_are_refused():
    text = inputs.clamped_beam().with_releases({1: (2, 0, 0, 0)}).text
    with pytest.raises(cubens.CubensError, match="release flags"):
        cubens.GpuModel(model_file("gpu_rel_bad.txt", text))
