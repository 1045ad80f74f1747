"""Member-end offsets (rigid links between a joint and the member end, fixed in
global axes: frame.c:297-324, misc.c:122-131, frame.c:1331-1345) in the
oracle, pinned by

  * the reference's own frame, geometry and mass routines under the restated
    drivers (tests/golden/offsets.npz, make_offsets.py): nonlinear Newmark,
    ANAFLAG 3 under NR and MSAL, and free warping DOFs; the NR girder of
    frame_girder_offsets_nr.npz is pinned in test_oracle_frame.py;
  * closed forms a dropped, transposed or mis-signed link fails: a
    cantilever whose first and last members end short of the joints in rigid
    zones a1 and a2 deflects P((L - a1)^3 - a2^3)/(3EI) and turns
    P((L - a1)^2 - a2^2)/(2EI) at the tip; its support reactions are (-P, -P L)
    whatever the offsets, out-of-plane ones included, since the reaction
    moment of an offset end is its end moment plus o x F (the warping slot
    takes the bimoment of the twist that lateral offsets bring in).
CPU only."""
import os

import numpy as np
import pytest

import oracle as O
from cubens import inputs
from helpers import model_file, relerr

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = ("tower_nm", "pc_nr", "pc_msal", "warp_nr")


def gold():
    return np.load(os.path.join(GOLD, "offsets.npz"))


@pytest.mark.parametrize("name", CASES)
def test_offset_paths_against_reference_routines(name):
    g = gold()
    r = O.OracleModel(model_file(f"ofs_{name}.txt", str(g[f"{name}_model"]))).run(250, keep_ef=True)
    assert r["status"] == int(g[f"{name}_status"]), r["msg"]
    assert np.array_equal(r["itecnt"], g[f"{name}_itecnt"])
    assert np.array_equal(r["events"], g[f"{name}_events"])
    assert relerr(r["lpf"], g[f"{name}_lpf"]) <= 1e-13
    assert relerr(r["d"], g[f"{name}_d"]) <= 1e-12
    assert relerr(r["ef"], g[f"{name}_ef"]) <= 1e-12


def test_fixture_models_are_the_generators():
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_offsets", os.path.join(GOLD, "make_offsets.py"))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    g = gold()
    for name, text in mk.models().items():
        assert str(g[f"{name}_model"]) == text, name
        om = O.OracleModel(model_file(f"ofs_gen_{name}.txt", text))
        rows = mk.PC_OFFSETS if name.startswith("pc") else (mk.TOWER_OFFSETS if name == "tower_nm"
                                                             else mk.GIRDER_OFFSETS)
        assert sorted(np.flatnonzero(om.osflag) + 1) == sorted(rows)
        for m, o in rows.items():
            assert np.array_equal(om.offset[(m - 1) * 6:m * 6], o)


def _cantilever(a1, a2, lateral=(0.0, 0.0, 0.0, 0.0)):
    n, L, P = 4, 3.0, -1000.0
    mt = inputs.cantilever_frame(n_elem=n, length=L, load=P)
    ly1, lz1, ly2, lz2 = lateral
    rows = {1: (a1, ly1, lz1, 0.0, 0.0, 0.0), n: (0.0, 0.0, 0.0, -a2, ly2, lz2)}
    return mt.with_offsets(rows), n, L, P


@pytest.mark.parametrize("a1,a2", [(0.0, 0.0), (0.2, 0.0), (0.0, 0.3), (0.25, 0.15)])
def test_cantilever_rigid_zones_closed_form(a1, a2):
    mt, n, L, P = _cantilever(a1, a2)
    om = O.OracleModel(model_file("ofs_cant.txt", mt.text))
    r = om.run(3, keep_react=True)
    assert r["status"] == 0 and r["nsteps"] == 1, r["msg"]
    EI = mt.info["E"] * mt.info["I"]
    tip = om.jcode.reshape(-1, 7)[-1]
    v, th = r["d"][0][tip[1] - 1], r["d"][0][tip[5] - 1]
    v_cf = P * ((L - a1) ** 3 - a2 ** 3) / (3 * EI)
    th_cf = P * ((L - a1) ** 2 - a2 ** 2) / (2 * EI)
    assert abs(v - v_cf) <= 1e-11 * abs(v_cf)
    assert abs(th - th_cf) <= 1e-11 * abs(th_cf)


@pytest.mark.parametrize("lateral", [(0.0, 0.0, 0.0, 0.0), (0.04, -0.03, 0.02, 0.05)])
def test_support_reactions_carry_the_offset_moments(lateral):
    mt, n, L, P = _cantilever(0.2, 0.1, lateral)
    om = O.OracleModel(model_file("ofs_react.txt", mt.text))
    r = om.run(3, keep_react=True)
    assert r["status"] == 0, r["msg"]
    root = r["react"][0][0][:6]       # slot 7 holds the restrained bimoment
    want = np.array([0.0, -P, 0.0, 0.0, 0.0, -P * L])
    assert np.allclose(root, want, rtol=0, atol=1e-9 * abs(P) * L)
