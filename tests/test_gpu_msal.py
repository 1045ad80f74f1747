"""The HIP arc-length path (cb_run, ALGFLAG 3) through limit points, against
the oracle on the same inputs, through the C ABI.

Each case runs the branches of main.c:2158-3141 that only a limit point
reaches: the predictor sign flip on det = 1 (main.c:2765-2769), the psi switch
of the constraint (main.c:2636-2642, 2758-2762, 2875-2889), quad's returns 2
and 3 (arc.c:134, 145) with their arc halvings, the dnorm restart
(main.c:2936-2953) and the stop on too many subdivisions (main.c:3038-3043).
The bar: identical status, message, step count, iteration counts and event
record (so every branch is taken at the same place), load factors and
displacements to 1e-9 of their largest value, the pivot-ratio psi to 1e-6.  The oracle's
driver itself is pinned by the closed forms of test_oracle_msal.py.

The convergence tolerances of these models keep every test() decision of the
traced paths at least 0.2% away from its threshold while the converged
residual is about 1e-4 of the load increment, so the round-off difference
between the GPU's and the oracle's internal forces (about 1e-12 of the bar
forces) cannot flip a decision.  At tolerances near round-off (force 1e-10
with a displacement criterion of 1e-12) decisions land within 0.1% of the
threshold while the residual is only 1e-8 of the forces, and the two
correct implementations can then legitimately take different iteration counts.
"""
import numpy as np
import pytest

import cubens
import oracle as O
from cubens import inputs
from helpers import model_file, relerr, skyline_dense

pytestmark = pytest.mark.gpu

STAR = dict(bars=4, radius=1.0, rise=0.1, emod=2e11, area=1e-4)
# convergence tolerances (displacement, force, energy) with every test decision
# of the traced paths well clear of round-off (see the module docstring)
TOL = (1e-8, 1e-6, 1.0)


def _compare(name, mt, steps):
    path = model_file(name, mt.text)
    print(f"[msal] {name}: {steps} steps", flush=True)
    with cubens.GpuModel(path) as g:
        gp = g.run(cubens.options(cubens.SOLVER_BAND64, max_steps=steps))
    op = O.OracleModel(path).run(steps)
    m = min(gp["nsteps"], op["nsteps"])
    for s in range(m):
        if gp["itecnt"][s] != op["itecnt"][s] or not np.array_equal(gp["events"][s], op["events"][s]):
            print(f"[msal] first difference at step {s}: itecnt {gp['itecnt'][s]} / {op['itecnt'][s]}, "
                  f"events {gp['events'][s]} / {op['events'][s]}, lpf {gp['lpf'][s]!r} / {op['lpf'][s]!r}, "
                  f"psi {gp['psi'][s]!r} / {op['psi'][s]!r}", flush=True)
            break
    assert gp["status"] == op["status"], (gp["msg"], op["msg"])
    assert gp["msg"] == op["msg"]
    assert gp["nsteps"] == op["nsteps"]
    assert np.array_equal(gp["itecnt"], op["itecnt"])
    assert np.array_equal(gp["events"], op["events"])
    assert relerr(gp["lpf"], op["lpf"]) <= 1e-9
    assert relerr(gp["d"], op["d"]) <= 1e-9
    # psi is the smallest pivot ratio: next to a limit point that pivot is a
    # near-cancelling difference and carries the tangent's conditioning; psi
    # enters the driver only through psi >= psi_thresh, whose outcomes the
    # identical event records already show to agree
    assert np.abs(gp["psi"] - op["psi"]).max() <= 1e-6
    return gp, op


def test_star_truss_through_both_limit_points():
    gp, _ = _compare("star_msal.txt", inputs.star_truss(**STAR, tol=TOL), 400)
    ev = gp["events"]
    assert gp["status"] == 0 and gp["nsteps"] > 40
    assert ev[:, cubens.EV_DET].sum() > 10
    assert (gp["psi"] < 0.1).sum() >= 2
    assert ev[:, cubens.EV_IMAG].sum() >= 1 and ev[:, cubens.EV_NEG].sum() >= 1
    # the GPU path itself against the closed form of the bar equilibrium
    lc = inputs.star_truss_lambda(gp["d"][:, 2], **STAR)
    dl = np.abs(np.diff(np.concatenate([[0.0], gp["lpf"]])))
    assert (np.abs(gp["lpf"] - lc) <= TOL[1] * dl).all()
    assert np.all(np.diff(gp["d"][:, 2]) < 0)


def test_soft_star_truss_dnorm_restart():
    """The soft star truss of test_oracle_msal.py through its first dnorm
    restart (main.c:2936-2953), in step 10.  The comparison stops there: the
    restarts shrink the arc to 4e-5 m (step 10) and then to 4e-13 m (step 11)
    while |d| is 0.05 m, and quad's c, assembled from expanded dot products of
    size |d|^2 (main.c:2873-2889), keeps about 10 significant digits in step 10
    but none in step 11, where the imaginary-root decisions of the reference's
    own formulation are round-off."""
    kw = dict(STAR, emod=2e4)
    mt = inputs.star_truss(**kw, alpha=0.01, iteopt=25000, dk=-0.005, tol=TOL)
    gp, _ = _compare("star_soft_msal.txt", mt, 11)
    assert gp["status"] == 0 and gp["nsteps"] == 11
    assert gp["events"][:, cubens.EV_DNORM].sum() >= 1


def test_truss_dome_past_its_limit_point():
    mt = inputs.truss_dome(algflag=3, load=-1.0, tol=(1e-6, 1e-4, 1.0),
                           msal=dict(alpha=0.4, dk=-0.0015))
    gp, _ = _compare("dome_msal.txt", mt, 60)
    assert gp["nsteps"] == 60
    assert gp["events"][:, cubens.EV_DET].sum() >= 5
    assert gp["events"][:, cubens.EV_IMAG].sum() >= 1


def test_shell_panel_stop_at_its_limit_point():
    gp, _ = _compare("panel8_limit.txt", inputs.shell_panel(n=8, dk=-1e-3), 80)
    assert gp["status"] == 1 and "subdivisions" in gp["msg"]
    assert 2150 < gp["lpf"].max() < 2250


def test_hub_of_70_bars_gets_70_colours():
    """70 bars meet at the apex, so every bar needs its own colour for the
    atomic-free scatter; the tangent and the arc-length path match the oracle
    and the closed form."""
    kw = dict(STAR, bars=70)
    mt = inputs.star_truss(**kw, tol=TOL)
    path = model_file("star70_msal.txt", mt.text)
    with cubens.GpuModel(path) as g:
        assert g.info.ncolours_tr == 70
        om = O.OracleModel(path)
        kt = g.tangent()
        bs = g.info.bs
        rp, ci = g.pattern()
        # the apex (joint 1) diagonal block against the skyline of stiff_tr
        ss = O.stiff_tr(om, om.c1, om.c2, om.c3, om.llength, np.zeros(om.ntr * 2))
        K = skyline_dense(ss, om.maxa)
        blk = kt[rp[0] * bs * bs:(rp[0] + 1) * bs * bs].reshape(bs, bs)[:3, :3]
        assert np.abs(blk - K).max() <= 1e-12 * np.abs(K).max()
    gp, _ = _compare("star70_msal.txt", mt, 400)
    lc = inputs.star_truss_lambda(gp["d"][:, 2], **kw)
    dl = np.abs(np.diff(np.concatenate([[0.0], gp["lpf"]])))
    assert (np.abs(gp["lpf"] - lc) <= TOL[1] * dl).all()
