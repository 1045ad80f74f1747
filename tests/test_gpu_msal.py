"""The HIP arc-length path (cb_run, ALGFLAG 3) through limit points, against
the oracle on the same inputs, through the C ABI.

Each case runs the branches of main.c:2158-3141 that only a limit point
reaches: the predictor sign flip on det = 1 (main.c:2765-2769), the psi switch
of the constraint (main.c:2636-2642, 2758-2762, 2875-2889), quad's returns 2
and 3 (arc.c:134, 145) with their arc halvings, the dnorm restart
(main.c:2936-2953) and the stop on too many subdivisions (main.c:3038-3043).
The bar: identical status, message, step count, iteration counts and event
record (so every branch is taken at the same place), pivot-ratio psi to 1e-9,
load factors and displacements to 1e-9 of their largest value.  The oracle's
driver itself is pinned by the closed forms of test_oracle_msal.py.
"""
import numpy as np
import pytest

import cubens
import oracle as O
from cubens import inputs
from helpers import model_file, relerr, skyline_dense

pytestmark = pytest.mark.gpu

STAR = dict(bars=4, radius=1.0, rise=0.1, emod=2e11, area=1e-4)


def _compare(name, mt, steps):
    path = model_file(name, mt.text)
    print(f"[msal] {name}: {steps} steps", flush=True)
    with cubens.GpuModel(path) as g:
        gp = g.run(cubens.options(cubens.SOLVER_BAND64, max_steps=steps))
    op = O.OracleModel(path).run(steps)
    assert gp["status"] == op["status"], (gp["msg"], op["msg"])
    assert gp["msg"] == op["msg"]
    assert gp["nsteps"] == op["nsteps"]
    assert np.array_equal(gp["itecnt"], op["itecnt"])
    assert np.array_equal(gp["events"], op["events"])
    assert np.abs(gp["psi"] - op["psi"]).max() <= 1e-9
    assert relerr(gp["lpf"], op["lpf"]) <= 1e-9
    assert relerr(gp["d"], op["d"]) <= 1e-9
    return gp, op


def test_star_truss_through_both_limit_points():
    gp, _ = _compare("star_msal.txt", inputs.star_truss(**STAR), 400)
    ev = gp["events"]
    assert gp["status"] == 0 and gp["nsteps"] > 40
    assert ev[:, cubens.EV_DET].sum() > 10
    assert (gp["psi"] < 0.1).sum() >= 2
    assert ev[:, cubens.EV_IMAG].sum() >= 1 and ev[:, cubens.EV_NEG].sum() >= 1
    # the GPU path itself against the closed form of the bar equilibrium
    lc = inputs.star_truss_lambda(gp["d"][:, 2], **STAR)
    assert np.abs(gp["lpf"] - lc).max() <= 1e-9 * np.abs(lc).max()
    assert np.all(np.diff(gp["d"][:, 2]) < 0)


def test_soft_star_truss_dnorm_restart_and_stop():
    kw = dict(STAR, emod=2e4)
    gp, _ = _compare("star_soft_msal.txt", inputs.star_truss(**kw, alpha=0.01, iteopt=30000), 400)
    assert gp["status"] == 1 and "subdivisions" in gp["msg"]
    assert gp["events"][:, cubens.EV_DNORM].sum() >= 2


def test_truss_dome_past_its_limit_point():
    mt = inputs.truss_dome(algflag=3, load=-1.0, tol=(1e-6, 1e-4, 1.0),
                           msal=dict(alpha=0.3, dk=-0.002))
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
    mt = inputs.star_truss(**kw)
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
    assert np.abs(gp["lpf"] - lc).max() <= 1e-9 * np.abs(lc).max()
