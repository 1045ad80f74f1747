"""OPTFLAG 2 joint renumbering (graph, optnum, updatenum; model.c:649-935) in
the oracle, pinned against the reference's own input sequence (fixtures of
tests/golden/make_renumber.py) and against what renumbering must preserve:

  * jnt is a permutation and the DOF maps, incidences, coordinates and load
    vector equal the reference's bit for bit;
  * the renumbered model is the same structure: its path equals the path of
    the same file read with OPTFLAG 1, equation for equation through jnt;
  * the band (largest column height) of a scrambled numbering shrinks --
    optnum minimises the band and uses the profile only to break ties;
  * load histories stay on the input joint numbers (model.c:1524), so the
    Newmark loads land on other joints than their jnt images.
CPU only."""
import os

import numpy as np
import pytest

import oracle as O
from helpers import model_file, relerr

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = ("panel", "dome", "tower", "girder", "mixed", "twin")
STATIC = ("panel", "dome", "tower", "mixed", "twin")       # no load histories


def gold():
    return np.load(os.path.join(GOLD, "renumber.npz"))


def with_optflag(text: str, v: int) -> str:
    """The model text with its OPTFLAG line (main.c:340-360) set to v."""
    lines = text.split("\n")
    k = 3 + (int(lines[1]) > 3)          # ANAFLAG, ALGFLAG, [CHKPT], SLVFLAG, OPTFLAG
    assert lines[k] in ("1", "2")
    lines[k] = str(v)
    return "\n".join(lines)


def model(name, optflag=2):
    text = str(gold()[f"{name}_model"])
    return O.OracleModel(model_file(f"renum_{name}_{optflag}.txt", with_optflag(text, optflag)))


@pytest.mark.parametrize("name", NAMES)
def test_renumbering_matches_the_reference(name):
    g = gold()
    om = model(name)
    jnt = O._arr(om.h.contents.jnt, om.nj, np.int64)
    assert np.array_equal(np.sort(jnt), np.arange(om.nj))
    assert not np.array_equal(jnt, np.arange(om.nj))      # the scramble gave it work to do
    assert np.array_equal(jnt, g[f"{name}_jnt"])
    for k in ("jcode", "mcode", "maxa", "minc"):
        assert np.array_equal(getattr(om, k), g[f"{name}_{k}"]), k
    assert np.array_equal(om.x, g[f"{name}_x"])
    assert np.array_equal(om.q, g[f"{name}_q"])


@pytest.mark.parametrize("name", NAMES)
def test_renumbered_path_against_reference_routines(name):
    g = gold()
    r = model(name).run(len(g[f"{name}_path_lpf"]))
    assert np.array_equal(r["itecnt"], g[f"{name}_path_itecnt"])
    assert relerr(r["lpf"], g[f"{name}_path_lpf"]) < 1e-10
    assert relerr(r["d"], g[f"{name}_path_d"]) < 1e-10


def _eq_maps(om1, om2):
    """Equation of each (input joint, direction) slot in both numberings."""
    jnt = O._arr(om2.h.contents.jnt, om2.nj, np.int64)
    j1 = om1.jcode.reshape(om1.nj, 7)
    j2 = om2.jcode.reshape(om2.nj, 7)[jnt]
    return j1, j2


@pytest.mark.parametrize("name", STATIC)
def test_renumbering_preserves_the_structure(name):
    om1, om2 = model(name, 1), model(name, 2)
    j1, j2 = _eq_maps(om1, om2)
    assert np.array_equal(j1 != 0, j2 != 0)
    e1, e2 = j1[j1 != 0] - 1, j2[j2 != 0] - 1
    assert np.array_equal(np.sort(e2), np.arange(om2.neq))
    assert np.array_equal(om1.q[e1], om2.q[e2])
    jnt = O._arr(om2.h.contents.jnt, om2.nj, np.int64)
    assert np.array_equal(om1.x.reshape(-1, 3), om2.x.reshape(-1, 3)[jnt])
    n = len(gold()[f"{name}_path_lpf"])
    r1, r2 = om1.run(n), om2.run(n)
    assert np.array_equal(r1["itecnt"], r2["itecnt"])
    # only the summation order differs between the two numberings
    assert relerr(r1["lpf"], r2["lpf"]) < 1e-9
    assert relerr(r1["d"][:, e1], r2["d"][:, e2]) < 1e-9
    if name != "twin":
        assert om2.kht.max() < om1.kht.max()


def test_newmark_histories_keep_the_input_joint_numbers():
    g = gold()
    om = model("girder")
    nt = om.h.contents.ntstpsinpt
    pin = O._arr(om.h.contents.pinpt, om.neq * nt, np.float64)
    assert np.array_equal(pin, g["girder_pinpt"])
    rows = set(np.flatnonzero(np.abs(pin.reshape(om.neq, nt)).max(axis=1) > 1e-20).tolist())
    text = str(g["girder_model"])
    jnt = O._arr(om.h.contents.jnt, om.nj, np.int64)
    jc = om.jcode.reshape(om.nj, 7)
    hist = {(int(a), int(b)) for a, b in
            (ln.split(",")[:2] for ln in text.split("\n") if ln.count(",") == 2 and ",1," in ln[:8])}
    as_read = {int(jc[jt - 1, d - 1]) - 1 for jt, d in hist if jc[jt - 1, d - 1] > 0}
    mapped = {int(jc[jnt[jt - 1], d - 1]) - 1 for jt, d in hist if jc[jnt[jt - 1], d - 1] > 0}
    assert rows == as_read
    assert rows != mapped
