"""Load-response plots of one joint, as the reference's BenPlot.py draws them
(BenPlot.py:35-174), from a model_def.txt and the results2.txt of its run.

The joint is found by its coordinates, compared as BenPlot compares them: the
three values printed with six decimals.  Each active DOF of the joint gets one
plot: for a static analysis (ALGFLAG 1-3) the load proportionality factor
against the DOF's displacement, for a dynamic one (ALGFLAG 4, 5) the
displacement against time.  BenPlot reads the joint table and DOF map from
the reference's log (results1.txt); here they come from the model itself
(cb_model_info), so no log is needed.  BenPlot plots directions 1-6 only
(BenPlot.py:139), and so does this.  BenPlot drops the last line of
results2.txt (BenPlot.py:132), which is the last converged step because
output() ends the file without a newline; every step is plotted here.
"""
from __future__ import annotations

import os

import numpy as np

from . import model_info

_DIR_LABEL = ("X-Translation", "Y-Translation", "Z-Translation", "X-Rotation", "Y-Rotation",
              "Z-Rotation")


def read_results2(path: str):
    """(lpf or time [nsteps], itecnt [nsteps], d [nsteps, NEQ]) from a results2.txt
    (misc.c:350-417: a header line, a DOF line, one row per step)."""
    lpf, ite, rows = [], [], []
    with open(path) as fh:
        lines = fh.read().split("\n")
    for ln in lines[2:]:
        f = [v for v in ln.split("\t") if v != ""]
        if not f:
            continue
        lpf.append(float(f[0]))
        ite.append(int(f[1]))
        rows.append([float(v) for v in f[2:]])
    n = max((len(r) for r in rows), default=0)
    d = np.array([r + [0.0] * (n - len(r)) for r in rows]).reshape(len(rows), n)
    return np.array(lpf), np.array(ite, np.int64), d


def find_joint(model_path: str, xyz) -> int:
    """0-based number of the joint at coordinates xyz (matched to six decimals,
    as BenPlot.py:36-40 prints them); ValueError when none is there."""
    mi = model_info(model_path)
    key = tuple("%.6f" % float(v) for v in xyz)
    x = mi["x"].reshape(-1, 3)
    for j in range(mi["nj"]):
        if tuple("%.6f" % v for v in x[j]) == key:
            return j
    raise ValueError("Coordinates invalid")


def joint_series(model_path: str, results2_path: str, xyz) -> list[dict]:
    """The plots of the joint at xyz: per active DOF (directions 1-6), a dict
    with the label, the axis titles and the x and y series."""
    mi = model_info(model_path)
    j = find_joint(model_path, xyz)
    lpf, _, d = read_results2(results2_path)
    dyn = mi["algflag"] > 3
    out = []
    for k in range(6):
        dof = int(mi["jcode"][j * 7 + k])
        if dof == 0:
            continue
        resp = d[:, dof - 1]
        name = "Displacement" if k < 3 else "Rotation"
        if dyn:
            out.append({"label": f"DOF {dof}: {_DIR_LABEL[k]}", "x": lpf, "y": resp,
                        "xlabel": "Time (sec)", "ylabel": name, "dof": dof, "direction": k + 1})
        else:
            out.append({"label": f"DOF {dof}: {_DIR_LABEL[k]}", "x": resp, "y": lpf, "xlabel": name,
                        "ylabel": "Load Proportionality Factor (LPF)", "dof": dof, "direction": k + 1})
    if not out:
        raise ValueError("Chosen Coordinate does not have active DOFs")
    return out


def plot_joint(model_path: str, results2_path: str, xyz, out_dir: str) -> list[str]:
    """Writes one PNG per active DOF of the joint at xyz into out_dir and
    returns their paths."""
    import matplotlib
    matplotlib.use("Agg")
    import matplotlib.pyplot as plt
    os.makedirs(out_dir, exist_ok=True)
    paths = []
    for s in joint_series(model_path, results2_path, xyz):
        fig = plt.figure()
        plt.plot(s["x"], s["y"])
        plt.xlabel(s["xlabel"])
        plt.ylabel(s["ylabel"])
        plt.title(s["label"])
        p = os.path.join(out_dir, f"dof{s['dof']}.png")
        fig.savefig(p)
        plt.close(fig)
        paths.append(p)
    return paths
