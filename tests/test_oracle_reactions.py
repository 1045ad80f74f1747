"""Support reactions of the oracle (orc_reactions): the element forces T^T ef
summed at the constrained slots.  The reference computes none, so they are
pinned by statics:

* star truss: each support carries -N c of its bar (truss.c:349-372), the
  closed form N(L) of the bar force at the apex position of the path;
* every model: the element end forces are self-equilibrated, so per
  translational direction the reactions plus the internal force at the free
  DOFs sum to zero (round-off), and the reactions balance the applied load
  lambda q up to the residual the convergence test left.
"""
import math
import os
import tempfile

import numpy as np

import oracle as O
from cubens import inputs


def _run(mt, steps, name):
    d = tempfile.mkdtemp(prefix="cb_react_")
    path = mt.write(os.path.join(d, name))
    om = O.OracleModel(path)
    return om, om.run(steps, keep_f=True, keep_react=True)


def direction_sums(om, v_eq):
    """Sum of an NEQ vector per joint direction 1..6 (via jcode)."""
    jc = om.jcode.reshape(om.nj, 7)
    out = np.zeros(6)
    for k in range(6):
        eq = jc[:, k]
        eq = eq[eq > 0] - 1
        out[k] = v_eq[eq].sum()
    return out


def test_star_truss_reactions_follow_the_bar_forces():
    p = dict(bars=4, radius=1.0, rise=0.1, emod=2e11, area=1e-4)
    om, r = _run(inputs.star_truss(**p), 400, "star.txt")
    assert r["status"] == 0
    L0 = math.hypot(p["radius"], p["rise"])
    for s in range(r["nsteps"]):
        w = r["d"][s, 2]
        H = p["rise"] + w
        L = math.hypot(p["radius"], H)
        e = (L - L0) / L0
        N = p["emod"] * p["area"] * (e + 0.5 * e * e) * L / L0
        R = r["react"][s]
        for k in range(p["bars"]):
            t = 2 * math.pi * k / p["bars"]
            c = np.array([-p["radius"] * math.cos(t), -p["radius"] * math.sin(t), H]) / L
            want = -N * c
            assert np.abs(R[k + 1, :3] - want).max() <= 1e-9 * abs(N)
        assert np.abs(R[0]).max() == 0.0            # the apex is free
        assert np.abs(R[1:, 3:]).max() == 0.0       # no moments in a truss
        # vertical equilibrium with the applied load lambda * q (q_z = -1)
        assert abs(R[:, 2].sum() - r["lpf"][s]) <= 1e-9 * max(1.0, abs(r["lpf"][s]))


def _check_balance(om, r, steps, tol_res):
    for s in steps:
        R = r["react"][s]
        f = r["f"][s]
        fr = direction_sums(om, f)
        rs = R[:, :6].sum(axis=0)
        scale = max(np.abs(fr).max(), np.abs(rs).max(), 1e-300)
        # self-equilibrium of the element forces (translations)
        assert np.abs(rs[:3] + fr[:3]).max() <= 1e-11 * scale, s
        # balance of the applied load, to the converged residual
        lq = direction_sums(om, r["lpf"][s] * om.q)
        res = np.linalg.norm(r["lpf"][s] * om.q - f, 1)
        assert np.abs(rs[:3] + lq[:3]).max() <= res + 1e-11 * scale, s
        assert np.abs(rs[:3] + lq[:3]).max() <= tol_res * np.abs(lq).max(), s


def test_truss_dome_reactions_balance_the_load():
    om, r = _run(inputs.truss_dome(), 40, "dome.txt")
    assert r["nsteps"] == 20
    _check_balance(om, r, range(20), 1e-6)


def test_shell_panel_reactions_balance_the_load():
    om, r = _run(inputs.shell_panel(n=8), 12, "panel8.txt")
    assert r["nsteps"] == 12
    _check_balance(om, r, range(12), 2e-2)


def test_frame_tower_reactions_balance_the_load():
    om, r = _run(inputs.space_frame(n_elem=200, kind="tower", algflag=1, amp=2e3), 20, "tower.txt")
    assert r["nsteps"] >= 10, r["msg"]
    _check_balance(om, r, range(r["nsteps"]), 1e-6)
