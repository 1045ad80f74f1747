"""Member-end offsets (rigid links, frame.c:297-324) on the GPU through the C
ABI: paths against the reference routines (tests/golden/offsets.npz and
frame_girder_offsets_nr.npz) and the oracle under NR, MSAL, Newmark, ANAFLAG 3
and free warping; the tangent, internal force, geometry update and lumped
masses at a deformed state against the oracle, 6 and 7 slots per joint; the
rigid-zone cantilever's closed form and its support reactions; and the end
coordinates the Newmark checkpoint writes.

Bar: FP64 paths 1e-9 with identical iteration counts and driver events;
kernels 1e-12 (geometry 1e-13)."""
import os

import numpy as np
import pytest

import cubens
import oracle as O
from cubens import inputs
from helpers import model_file, relerr, skyline_dense
from test_oracle_offsets import CASES, _cantilever, gold

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _opts(steps=250, **kw):
    return cubens.options(cubens.SOLVER_BAND64, max_steps=steps, **kw)


@pytest.mark.parametrize("name", CASES)
def test_offset_paths_match_reference_routines_and_oracle(name):
    g = gold()
    path = model_file(f"gpu_ofs_{name}.txt", str(g[f"{name}_model"]))
    print(f"[offsets] {name}", flush=True)
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


def test_girder_nr_fixture():
    g = np.load(os.path.join(GOLD, "frame_girder_offsets_nr.npz"))
    path = model_file("gpu_ofs_girder.txt", str(g["model"]))
    print("[offsets] girder", flush=True)
    with cubens.GpuModel(path) as m:
        r = m.run(_opts(25))
    assert r["status"] == 0, r["msg"]
    assert np.array_equal(r["itecnt"], g["path_itecnt"])
    assert relerr(r["lpf"], g["path_lpf"]) < 1e-12
    assert relerr(r["d"], g["path_d"]) < 1e-9


def _tower_with_offsets(warping, n=60):
    mt = inputs.space_frame(n, kind="tower", steps=5, warping=warping)
    rng = np.random.default_rng(23)
    rows = {m: tuple(0.05 * rng.uniform(-1, 1, 6)) for m in range(1, mt.n_fr + 1) if m % 3 != 0}
    return model_file(f"gpu_ofs_tower{int(warping)}.txt", mt.with_offsets(rows).text)


@pytest.mark.parametrize("warping", [False, True])
def test_kernels_with_offsets_at_a_deformed_state(warping):
    path = _tower_with_offsets(warping)
    om = O.OracleModel(path)
    rng = np.random.default_rng(31)
    dd = 3e-3 * rng.standard_normal(om.neq)
    x, xip, defllen, deffarea, defslen, c1, c2, c3 = O.updatc(om, om.x, dd)
    ef_ip = 1e4 * rng.standard_normal(om.nfr * 14)
    print(f"[offsets] kernels, warping {warping}", flush=True)
    with cubens.GpuModel(path) as g:
        bs = g.info.bs
        assert bs == (7 if warping else 6)
        # geometry update from the initial state
        g.update_geometry(dd, cubens.GEN_I)
        assert relerr(g.get("x_temp", om.nj * 3), x) < 1e-14
        assert relerr(g.get("defllen_i", om.nfr), defllen) < 1e-13
        for name, v in (("c1_i", c1), ("c2_i", c2), ("c3_i", c3)):
            assert relerr(g.get(name, om.nfr * 3), v) < 1e-13
        for name, v in (("c1_ip", om.c1), ("c2_ip", om.c2), ("c3_ip", om.c3), ("ef_ip", ef_ip),
                        ("defllen_ip", defllen), ("c1_i", c1), ("c2_i", c2), ("c3_i", c3), ("x_temp", x)):
            g.set(name, v)
        jc, _, _ = g.dofmaps()
        rp, ci = g.pattern()
        Kg = cubens.bsr_to_dense(g.tangent(cubens.GEN_I, cubens.GEN_IP), rp, ci, bs,
                                 cubens.eq_of_slot(jc, g.nj, bs), g.neq)
        fg = g.internal_force(dd, cubens.GEN_I)
        efg = g.get("ef_i", om.nfr * 14)
        sm = g.mass()
        llen = g.get("llength", om.nfr)
    Ko = skyline_dense(O.stiff_fr(om, om.llength, defllen, om.c1, om.c2, om.c3, ef_ip), om.maxa)
    fo, efo = O.forces_fr(om, dd, om.llength, defllen, (om.c1, om.c2, om.c3), (c1, c2, c3), ef_ip)
    sm_o, llen_o, _, _ = O.mass(om, om.x)
    assert relerr(Kg, Ko) < 1e-12
    assert relerr(fg, fo) < 1e-12
    assert relerr(efg, efo) < 1e-12
    assert relerr(sm, sm_o) < 1e-14
    assert relerr(llen, llen_o) < 1e-15


@pytest.mark.parametrize("lateral", [(0.0, 0.0, 0.0, 0.0), (0.04, -0.03, 0.02, 0.05)])
def test_rigid_zone_cantilever_and_reactions(lateral):
    a1, a2 = 0.25, 0.15
    mt, n, L, P = _cantilever(a1, a2, lateral)
    path = model_file("gpu_ofs_cant.txt", mt.text)
    with cubens.GpuModel(path) as m:
        r = m.run(_opts(3))
        rx = m.reactions().reshape(-1, 7)
        jc, _, _ = m.dofmaps()
    o = O.OracleModel(path).run(3, keep_react=True)
    assert r["status"] == 0 == o["status"], r["msg"]
    assert relerr(r["d"], o["d"]) < 1e-9
    assert relerr(rx, o["react"][0]) < 1e-9
    assert np.allclose(rx[0, :6], [0.0, -P, 0.0, 0.0, 0.0, -P * L], rtol=0, atol=1e-9 * abs(P) * L)
    if not any(lateral):
        EI = mt.info["E"] * mt.info["I"]
        tip = jc.reshape(-1, 7)[-1]
        v_cf = P * ((L - a1) ** 3 - a2 ** 3) / (3 * EI)
        assert abs(r["d"][0][tip[1] - 1] - v_cf) <= 1e-10 * abs(v_cf)


def test_newmark_checkpoint_writes_the_offset_end_coordinates(tmp_path):
    g = gold()
    text = str(g["tower_nm_model"])
    path = model_file("gpu_ofs_ck.txt", text)
    om = O.OracleModel(path)
    # the same model with a checkpoint every 10 steps
    lines = text.split("\n")
    assert int(lines[1]) == 5 and lines[2] == "31,0"        # CHKPT, RFLAG
    lines[2] = "10,0"
    ckpath = model_file("gpu_ofs_ck10.txt", "\n".join(lines))
    print("[offsets] checkpoint", flush=True)
    ck, ock = str(tmp_path / "gpu.txt"), str(tmp_path / "oracle.txt")
    with cubens.GpuModel(ckpath) as m:
        r = m.run(_opts(12, checkpoint_path=ck, checkpoint_exact=1))
    assert r["status"] == 0, r["msg"]
    O.OracleModel(ckpath).run(12, ckpt_path=ock, ckpt_exact=True)
    mine, theirs = O.read_checkpoint(ck, om), O.read_checkpoint(ock, om)
    assert mine["tstep"] == theirs["tstep"] == 10
    assert relerr(mine["xfr"], theirs["xfr"]) < 1e-12
    assert relerr(mine["x"], theirs["x"]) < 1e-12
