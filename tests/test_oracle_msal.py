"""Pins the oracle's arc-length (MSAL) driver, main.c:2158-3141 with
arc.c:41-157, by what the mathematics fixes rather than by the restatement:

* a shallow star truss traced through both limit points and the flat
  position: the load factor at every converged increment equals the closed
  form lambda(w) of the bar equilibrium (truss.c:327-331 for the bar force)
  to the force tolerance times the step's load increment (the criterion of
  test(), misc.c:215-232), the apex stays on the axis and the path never
  reverses;
* the pivot-sign flag det (solve.c:608-621) of every increment equals the sign
  pattern of the closed-form apex tangent (truss.c:106-166) at the previous
  converged state, and the recorded psi (main.c:2632-2642) equals the
  closed-form pivot ratio (capped at 1, its starting value);
* every converged increment satisfies the spherical constraint
  |d_s - d_{s-1}|^2 + [psi >= psi_thresh] (lambda_s - lambda_{s-1})^2 q.q
  = arc_s^2, with arc_s rebuilt from the recorded path alone: the previous
  increment, beta = sqrt(iteopt/itecnt) alpha |d_0| / |d_{s-1} - d_{s-2}|
  (main.c:2625-2630, 3048-3049) and one halving per quad retry or itemax
  exceedance (main.c:2995-3036);
* the first increment holds the control DOF at dk exactly (arc.c:41-68).
"""
import math
import os
import tempfile

import numpy as np
import pytest

import oracle as O
from cubens import inputs
from helpers import arc_residuals


def closed_form_gap(lam, lc):
    """|lambda - lambda_closed| over each step's load increment."""
    dl = np.abs(np.diff(np.concatenate([[0.0], lam])))
    return np.abs(lam - lc) / dl


def _run(mt, steps=400):
    d = tempfile.mkdtemp(prefix="cb_msal_")
    path = mt.write(os.path.join(d, "model.txt"))
    om = O.OracleModel(path)
    return om, om.run(steps, keep_f=True)


def star_tangent(w, bars, radius, rise, emod, area):
    """Diagonal of the closed-form apex tangent of star_truss at apex
    displacement w (x, y, z): sum over bars of E A L^2/L0^3 c c^T + N/L I,
    with c the bar direction (truss.c:106-111, 155-166)."""
    L0 = math.hypot(radius, rise)
    H = rise + w
    L = math.hypot(radius, H)
    e = (L - L0) / L0
    N = emod * area * (e + 0.5 * e * e) * L / L0
    ka = emod * area * L * L / L0 ** 3
    kxx = ka * (bars / 2.0) * radius * radius / (L * L) + bars * N / L
    kzz = ka * bars * H * H / (L * L) + bars * N / L
    return np.array([kxx, kxx, kzz])


STAR = dict(bars=4, radius=1.0, rise=0.1, emod=2e11, area=1e-4)
# convergence tolerances (displacement, force, energy) with every test decision
# of the traced paths well clear of round-off (see the module docstring)
TOL = (1e-8, 1e-6, 1.0)


def test_star_truss_follows_the_closed_form_through_both_limit_points():
    mt = inputs.star_truss(**STAR, tol=TOL)
    om, r = _run(mt)
    assert r["status"] == 0, r["msg"]
    n = r["nsteps"]
    w = r["d"][:, 2]
    lam = r["lpf"]
    lc = inputs.star_truss_lambda(w, **STAR)
    # equilibrium along the whole path, to the force tolerance
    assert closed_form_gap(lam, lc).max() <= TOL[1]
    # the apex stays on the axis
    assert np.abs(r["d"][:, :2]).max() <= 1e-12
    # first increment: the control DOF is held at dk (arc.c:41-68)
    assert abs(w[0] - (-0.005)) <= 1e-15
    # the path never reverses, and it covers rise, limit, flat, inverted and
    # the stiffening branch past the second limit point
    assert np.all(np.diff(w) < 0)
    assert w[-1] < -2 * STAR["rise"]
    wl = np.linspace(-0.3, 0.0, 30001)
    lmax = inputs.star_truss_lambda(wl, **STAR)
    w_top = wl[np.argmax(np.where(wl > -0.1, lmax, -np.inf))]
    w_bot = wl[np.argmin(np.where(wl > -0.2, lmax, np.inf))]
    assert (w > w_top).any() and ((w < w_top) & (w > w_bot)).any() and (w < w_bot).any()
    assert (lam > 0).any() and (lam < 0).any()
    assert n > 40


def test_star_truss_det_and_psi_match_the_closed_form_tangent():
    mt = inputs.star_truss(**STAR, tol=TOL)
    om, r = _run(mt)
    w = r["d"][:, 2]
    ev, psi = r["events"], r["psi"]
    k0 = star_tangent(0.0, **STAR)
    states = [0.0] + list(w)          # converged apex positions, undeformed first
    checked_det = checked_psi = 0
    for s in range(r["nsteps"]):
        # the increment s factors the tangent at converged state s-1
        k = star_tangent(states[s], **STAR)
        if np.all(np.abs(k) > 1e-9 * np.abs(k0)):
            assert ev[s, O.EV_DET] == int((k <= 0).any()), s
            checked_det += 1
        # psi of increment s comes from the pivots of increment s-1
        # (main.c:3121-3127): the tangent at converged state s-2
        if s >= 1:
            kp = star_tangent(states[s - 1], **STAR)
            assert abs(psi[s] - min(1.0, np.min(np.abs(kp / k0)))) <= 1e-9, s
            checked_psi += 1
    assert checked_det >= r["nsteps"] - 2 and checked_psi == r["nsteps"] - 1
    # both branches of the predictor and of the psi switch are taken
    assert ev[:, O.EV_DET].sum() > 10 and (ev[:, O.EV_DET] == 0).sum() > 10
    assert (psi < 0.1).sum() >= 2 and (psi >= 0.1).sum() >= 2


def test_star_truss_converged_increments_satisfy_the_arc_constraint():
    mt = inputs.star_truss(**STAR, tol=TOL)
    om, r = _run(mt)
    res = arc_residuals((1.0, 6, 0.1), r, om.q)
    assert len(res) == r["nsteps"] - 1
    assert max(res.values()) <= 1e-9
    ev = r["events"]
    # quad's imaginary-root and negative-gamma returns and the itemax
    # halving all occur, and the rebuilt arc accounts for them
    assert ev[:, O.EV_IMAG].sum() >= 1 and ev[:, O.EV_NEG].sum() >= 1
    assert ev[:, O.EV_HALVE].sum() >= 1


def test_soft_star_truss_dnorm_restart_and_subdivision_stop():
    """A soft truss (E A = 2 N) makes the arc displacement-dominated, and a large
    iteopt lets a converged increment exceed 100 alpha |d_0|, which restarts it
    with a smaller beta (main.c:2936-2953); the run later stops with the
    reference's 'maximum allowable number of subdivisions exceeded'
    (main.c:3038-3043).  In a displacement-only arc a converged increment is
    sqrt(iteopt/itecnt) alpha |d_0| long, so iteopt is kept off the multiples
    of 1e4 itecnt at which that length would sit exactly on the dnorm test's
    threshold."""
    kw = dict(STAR, emod=2e4)
    mt = inputs.star_truss(**kw, alpha=0.01, iteopt=25000, dk=-0.005, tol=TOL)
    om, r = _run(mt)
    assert r["status"] == 1 and "subdivisions" in r["msg"]
    ev = r["events"]
    assert ev[:, O.EV_DNORM].sum() >= 2 and ev[:, O.EV_DET].sum() >= 1
    lc = inputs.star_truss_lambda(r["d"][:, 2], **kw)
    assert closed_form_gap(r["lpf"], lc).max() <= TOL[1]
    res = arc_residuals((0.01, 25000, 0.1), r, om.q)
    assert res and max(res.values()) <= 1e-9


def test_truss_dome_crosses_its_limit_point_under_arc_length():
    """The 10x10 dome under arc-length control: past the crown's limit point
    (about 92 N for the load level of the default tolerances) the pivots turn
    negative (det = 1) and the load factor falls; every increment whose arc the
    path determines satisfies the constraint."""
    mt = inputs.truss_dome(algflag=3, load=-1.0, tol=(1e-6, 1e-4, 1.0),
                           msal=dict(alpha=0.4, dk=-0.0015))
    om, r = _run(mt, steps=60)
    assert r["nsteps"] == 60, r["msg"]
    ev = r["events"]
    lam = r["lpf"]
    assert ev[:, O.EV_DET].sum() >= 5 and ev[:, O.EV_IMAG].sum() >= 1
    first_neg = int(np.argmax(ev[:, O.EV_DET] > 0))
    assert lam[:first_neg].max() > lam[first_neg:first_neg + 3].min()
    res = arc_residuals((0.4, 6, 0.1), r, om.q)
    assert len(res) >= 50 and max(res.values()) <= 1e-8
    # residual equilibrium at each converged increment: the force criterion
    # of test() holds to its tolerance relative to the applied load
    f = r["f"]
    rel = [np.linalg.norm(lam[s] * om.q - f[s]) / np.linalg.norm(lam[s] * om.q) for s in range(60)]
    assert max(rel) <= 2e-3


@pytest.mark.parametrize("n", [8])
def test_shell_panel_stops_at_its_limit_point(n):
    """The reference's arc-length method on its CST+DKT panel reaches the limit
    load and then cannot place a converged increment past it: quad finds
    imaginary roots and the arc is halved until the subdivision limit stops
    the run (main.c:3038-3043).  The stop is part of the path the GPU must
    reproduce."""
    mt = inputs.shell_panel(n=n, dk=-1e-3)
    om, r = _run(mt, steps=80)
    assert r["status"] == 1 and "subdivisions" in r["msg"]
    lam = r["lpf"]
    assert 2150 < lam.max() < 2250
    assert r["events"][:, O.EV_IMAG].sum() >= 5


def test_star_hub_of_70_bars_follows_the_closed_form():
    kw = dict(STAR, bars=70)
    om, r = _run(inputs.star_truss(**kw, tol=TOL))
    assert r["status"] == 0 and r["nsteps"] > 40
    lc = inputs.star_truss_lambda(r["d"][:, 2], **kw)
    assert closed_form_gap(r["lpf"], lc).max() <= TOL[1]
    assert r["events"][:, O.EV_DET].sum() > 10


def test_shell_panel_increments_satisfy_the_invariants():
    """The invariants the GPU full-size tests apply at 256x256, here on the
    oracle's 16x16 panel with the same thickness field and imperfection: the
    rebuilt arc-length constraint at every increment, and the out-of-balance
    force within the force criterion of test() (tolforc 1e-3 of the load
    increment, up to the last load-factor correction)."""
    mt = inputs.shell_panel(n=16, thick_sigma=0.1, imperfection=0.1, config_index=3)
    om, r = _run(mt, steps=30)
    assert r["nsteps"] == 30
    res = arc_residuals((1.0, 6, 0.1), r, om.q)
    assert len(res) == 29 and max(res.values()) <= 1e-8
    lam = r["lpf"]
    dl = np.abs(np.diff(np.concatenate([[0.0], lam])))
    ratio = r["res"] / (dl * np.linalg.norm(om.q))
    assert ratio.max() <= 2e-3
    # the residual norm the driver records is |lambda q - f| of its converged f
    assert np.allclose(r["res"], np.linalg.norm(lam[:, None] * om.q[None, :] - r["f"], axis=1), rtol=1e-12)
