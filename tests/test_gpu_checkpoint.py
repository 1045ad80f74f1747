"""Checkpoint and restart of the nonlinear Newmark driver on the GPU
(cb_options.checkpoint_path; checkPoint / restartStep, misc.c:494-716,
main.c:3378-3398, 3904-3907) through the C ABI, against the reference's own
checkpoint file and restarted path (tests/golden/checkpoint/) and the oracle:

  * the file written in the reference's format holds, line for line, the
    values of the file the reference's checkPoint writes for that state, to
    the six digits %e keeps (the GPU's state differs from the reference
    routines' in the last bits, which can move a sixth digit);
  * a restart from the reference's file follows the reference routines'
    restarted path to the FP64 bar (1e-9, iteration counts exact);
  * a restart from the full-precision file continues the uninterrupted GPU
    run bit for bit, and that file equals the oracle's to 1e-9;
  * a restart without a checkpoint file, or from a missing one, stops with
    a message."""
import os
import shutil

import numpy as np
import pytest

import cubens
import oracle as O
from helpers import relerr

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "checkpoint")
BASE = os.path.join(GOLD, "tower_model.txt")
RESTART = os.path.join(GOLD, "tower_model_restart.txt")


def _run(model, steps, path=None, exact=0):
    with cubens.GpuModel(model) as g:
        return g.run(cubens.options(cubens.SOLVER_BAND64, max_steps=steps, checkpoint_path=path,
                                    checkpoint_exact=exact))


def _close(a, b, rtol):
    if b.size == 0:
        return a.size == 0
    scale = max(np.abs(b).max(), 1e-300)
    return np.all(np.abs(a - b) <= rtol * np.abs(b) + 1e-9 * scale)


def test_reference_format_checkpoint_matches_the_reference_writer(tmp_path):
    path = str(tmp_path / "results8.txt")
    print("[checkpoint] reference format, 10 steps", flush=True)
    r = _run(BASE, 10, path)
    assert r["status"] == 0 and r["nsteps"] == 10, r["msg"]
    mine = open(path).read().split("\n")
    ref = open(os.path.join(GOLD, "tower_results8.txt")).read().split("\n")
    assert len(mine) == len(ref)
    for a, b in zip(mine, ref):
        assert a.count(",") == b.count(",")
        if a in ("0,0,0", "0,0") or b in ("0,0,0", "0,0"):
            assert a == b
    om = O.OracleModel(BASE)
    got = O.read_checkpoint(path, om)
    want = O.read_checkpoint(os.path.join(GOLD, "tower_results8.txt"), om)
    assert got["tstep"] == want["tstep"] == 8 and got["has_ext"] == 0
    for k in O.CKPT_ARRAYS[:19]:
        assert _close(got[k], want[k], 2e-6), k


def test_restart_from_the_reference_checkpoint_follows_the_reference(tmp_path):
    ck = str(tmp_path / "results8.txt")
    shutil.copy(os.path.join(GOLD, "tower_results8.txt"), ck)
    print("[checkpoint] restart from the reference's file", flush=True)
    r = _run(RESTART, 13, ck)
    g = np.load(os.path.join(GOLD, "tower_restart.npz"))
    assert r["status"] == 0 and r["nsteps"] == 4, r["msg"]
    assert np.array_equal(r["itecnt"], g["itecnt"])
    assert np.array_equal(r["lpf"], g["lpf"])
    assert relerr(r["d"], g["d"]) < 1e-9


def test_exact_restart_continues_bit_for_bit(tmp_path):
    print("[checkpoint] exact restart", flush=True)
    full = _run(BASE, 13)
    ck = str(tmp_path / "exact.txt")
    part = _run(BASE, 10, ck, exact=1)
    assert part["nsteps"] == 10
    om = O.OracleModel(BASE)
    mine = O.read_checkpoint(ck, om)
    ock = str(tmp_path / "oracle_exact.txt")
    om.run(10, ckpt_path=ock, ckpt_exact=True)
    theirs = O.read_checkpoint(ock, om)
    assert mine["has_ext"] == 1 and mine["tstep"] == theirs["tstep"] == 8 and mine["ddt"] == theirs["ddt"]
    for k in O.CKPT_ARRAYS:
        if theirs[k].size == 0 or not np.any(theirs[k]):
            assert np.array_equal(mine[k], theirs[k]), k
        elif k == "ss":
            assert _close(mine[k], theirs[k], 1e-9), k
        else:
            assert relerr(mine[k], theirs[k]) < 1e-9, k
    r = _run(RESTART, 13, ck)
    assert r["status"] == 0 and r["nsteps"] == 4, r["msg"]
    assert np.array_equal(r["itecnt"], full["itecnt"][9:])
    assert np.array_equal(r["lpf"], full["lpf"][9:])
    assert np.array_equal(r["d"], full["d"][9:])


def test_restart_needs_a_readable_checkpoint(tmp_path):
    r = _run(RESTART, 4)
    assert r["status"] == 1 and "checkpoint" in r["msg"]
    r = _run(RESTART, 4, str(tmp_path / "missing.txt"))
    assert r["status"] == 1 and "checkpoint" in r["msg"]
