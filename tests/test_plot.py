"""The BenPlot-style load-response plots (cubens.plot, BenPlot.py:35-174) on
the results files the reference's own writer produced (tests/golden/results/):
the joint found by its coordinates, one series per active DOF taken from the
results2.txt column of that DOF, static runs as lambda against displacement,
and the model read on the host alone through cb_model_info (no GPU).  CPU
only."""
import os

import numpy as np
import pytest

import cubens
import oracle as O
from cubens import plot

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "results")


def _files(name):
    return os.path.join(GOLD, name, "model_def.txt"), os.path.join(GOLD, name, "results2.txt")


@pytest.mark.parametrize("name", ["truss", "frame", "shell"])
def test_model_info_reads_the_model_on_the_host(name):
    model, _ = _files(name)
    mi = cubens.model_info(model)
    om = O.OracleModel(model)
    assert mi["neq"] == om.neq and mi["nj"] == om.nj
    assert np.array_equal(mi["jcode"], om.jcode)
    assert np.array_equal(mi["x"], om.x)
    assert np.array_equal(mi["jnt"], np.arange(om.nj))


@pytest.mark.parametrize("name", ["truss", "frame", "shell"])
def test_joint_series_are_the_results2_columns(name, tmp_path):
    model, res2 = _files(name)
    om = O.OracleModel(model)
    lpf, ite, d = plot.read_results2(res2)
    assert d.shape == (len(lpf), om.neq) and len(lpf) == 3
    jc = om.jcode.reshape(-1, 7)
    j = int(np.flatnonzero((jc[:, :6] != 0).any(axis=1))[-1])      # a joint with free DOFs
    xyz = om.x.reshape(-1, 3)[j]
    assert plot.find_joint(model, xyz) == j
    series = plot.joint_series(model, res2, xyz)
    want = [(k + 1, int(jc[j, k])) for k in range(6) if jc[j, k] != 0]
    assert [(s["direction"], s["dof"]) for s in series] == want
    for s in series:
        assert np.array_equal(s["y"], lpf)                  # static: lambda on the vertical axis
        assert np.array_equal(s["x"], d[:, s["dof"] - 1])
    paths = plot.plot_joint(model, res2, xyz, str(tmp_path))
    assert len(paths) == len(want) and all(os.path.getsize(p) > 0 for p in paths)


def test_unknown_coordinates_and_fixed_joints_are_refused():
    model, res2 = _files("truss")
    with pytest.raises(ValueError, match="Coordinates invalid"):
        plot.find_joint(model, (123.0, 456.0, 789.0))
    om = O.OracleModel(model)
    jc = om.jcode.reshape(-1, 7)
    j = int(np.flatnonzero((jc[:, :6] == 0).all(axis=1))[0])        # a pinned joint
    with pytest.raises(ValueError, match="does not have active DOFs"):
        plot.joint_series(model, res2, om.x.reshape(-1, 3)[j])
