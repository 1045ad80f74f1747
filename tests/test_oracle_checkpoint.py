"""Checkpoint and restart of the nonlinear Newmark driver in the oracle
(checkPoint / restartStep, misc.c:494-716; main.c:3378-3398, 3904-3907),
pinned against the reference's own reader and writer (fixtures of
tests/golden/make_checkpoint.py):

  * the file the oracle writes in the reference's format is, byte for byte,
    the file the reference's checkPoint writes for the same state;
  * the full-precision file reads back, through the reference's restartStep,
    to exactly the values the oracle's reader takes from it (so the reference
    can restart from it);
  * a run restarted from the reference-format file follows the reference
    routines' restarted path;
  * a run restarted from the full-precision file continues the uninterrupted
    run bit for bit (the extension carries the working lengths and the
    time-increment multiplier, which the reference's file lacks).
CPU only."""
import os

import numpy as np
import pytest

import oracle as O
from helpers import relerr

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "checkpoint")
BASE = os.path.join(GOLD, "tower_model.txt")
RESTART = os.path.join(GOLD, "tower_model_restart.txt")


def _copy(src, dst):
    with open(src) as a, open(dst, "w") as b:
        b.write(a.read())
    return dst


def test_reference_format_file_is_the_reference_writers(tmp_path):
    path = str(tmp_path / "results8.txt")
    r = O.OracleModel(BASE).run(10, ckpt_path=path)
    assert r["status"] == 0 and r["nsteps"] == 10
    with open(path) as fh, open(os.path.join(GOLD, "tower_results8.txt")) as ref:
        assert fh.read() == ref.read()


def test_exact_file_reads_back_through_the_reference_reader(tmp_path):
    path = str(tmp_path / "exact.txt")
    om = O.OracleModel(BASE)
    om.run(10, ckpt_path=path, ckpt_exact=True)
    mine = O.read_checkpoint(path, om)
    g = np.load(os.path.join(GOLD, "tower_ref_read.npz"))
    assert mine["tstep"] == int(g["tstep"]) == 8 and mine["has_ext"] == 1 and mine["ddt"] == 1.0
    for k in g.files:
        if k != "tstep":
            assert np.array_equal(mine[k], g[k]), k
    # the reference-format file holds the same values to the six digits it keeps
    e = str(tmp_path / "e.txt")
    O.OracleModel(BASE).run(10, ckpt_path=e)
    short = O.read_checkpoint(e, om)
    assert short["has_ext"] == 0
    for k in ("uc", "vc", "ac", "d", "f", "ef", "x", "c1", "ss", "sm"):
        assert np.allclose(short[k], mine[k], rtol=5e-6, atol=0), k


def test_restart_from_the_reference_file_follows_the_reference_routines(tmp_path):
    ck = _copy(os.path.join(GOLD, "tower_results8.txt"), str(tmp_path / "results8.txt"))
    r = O.OracleModel(RESTART).run(13, ckpt_path=ck)
    g = np.load(os.path.join(GOLD, "tower_restart.npz"))
    assert r["status"] == 0 and r["nsteps"] == 4
    assert np.array_equal(r["itecnt"], g["itecnt"])
    assert np.array_equal(r["lpf"], g["lpf"])          # times 9..12 dt
    assert relerr(r["d"], g["d"]) < 1e-12


def test_exact_restart_continues_bit_for_bit(tmp_path):
    full = O.OracleModel(BASE).run(13)
    ck = str(tmp_path / "exact.txt")
    O.OracleModel(BASE).run(10, ckpt_path=ck, ckpt_exact=True)
    r = O.OracleModel(RESTART).run(13, ckpt_path=ck)
    assert r["nsteps"] == 4
    assert np.array_equal(r["itecnt"], full["itecnt"][9:])
    assert np.array_equal(r["lpf"], full["lpf"][9:])
    assert np.array_equal(r["d"], full["d"][9:])
    # a six-digit file restarts within the error its digits allow
    e = str(tmp_path / "e.txt")
    O.OracleModel(BASE).run(10, ckpt_path=e)
    r6 = O.OracleModel(RESTART).run(13, ckpt_path=e)
    assert np.array_equal(r6["itecnt"], full["itecnt"][9:])
    assert relerr(r6["d"], full["d"][9:]) < 1e-5


def test_restart_edge_cases(tmp_path):
    # RFLAG 1 without a checkpoint file, and an unreadable file, stop with a message
    r = O.OracleModel(RESTART).run(4)
    assert r["status"] == 1 and "checkpoint" in r["msg"]
    r = O.OracleModel(RESTART).run(4, ckpt_path=str(tmp_path / "missing.txt"))
    assert r["status"] == 1 and "checkpoint" in r["msg"]
    # a checkpoint at the last step restarts to that step's row alone (main.c:3387-3392)
    ck = str(tmp_path / "last.txt")
    full = O.OracleModel(BASE).run(13, ckpt_path=ck)
    assert full["nsteps"] == 13 and O.read_checkpoint(ck, O.OracleModel(BASE))["tstep"] == 12
    r = O.OracleModel(RESTART).run(13, ckpt_path=ck)
    assert r["nsteps"] == 1 and r["itecnt"][0] == 0
    assert np.allclose(r["d"][0], full["d"][12], rtol=5e-6, atol=1e-20)


@pytest.mark.parametrize("exact", [False, True])
def test_write_read_round_trip(tmp_path, exact):
    om = O.OracleModel(BASE)
    g = dict(np.load(os.path.join(GOLD, "tower_ref_read.npz")))
    rng = np.random.default_rng(5)
    for k in ("llen_w", "farea_w", "slen_w"):
        g[k] = rng.standard_normal(O.ckpt_sizes(om)[k])
    g["ddt"] = 0.25
    p = str(tmp_path / "rt.txt")
    O.write_checkpoint(p, om, g, exact=exact)
    back = O.read_checkpoint(p, om)
    assert back["has_ext"] == int(exact)
    for k in g:
        if k in ("tstep", "ddt") or (k in ("llen_w", "farea_w", "slen_w") and not exact):
            continue
        if exact:
            assert np.array_equal(back[k], g[k][:len(back[k])]), k
        else:
            assert np.allclose(back[k], g[k][:len(back[k])], rtol=5e-6, atol=0), k
    if exact:
        assert back["ddt"] == 0.25
