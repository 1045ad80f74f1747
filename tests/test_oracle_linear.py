"""The linear analyses of the oracle: linear static (ANAFLAG 1 under ALGFLAG
1/2, main.c:1714-1798) and linear Newmark (ALGFLAG 4, main.c:3143-3303,
solve.c:139-468), pinned by

  * the reference's own element and mass routines under the restated drivers
    (tests/golden/linear.npz, make_linear.py), including its sample input
    model_def_5c_shell.txt;
  * closed forms: the star truss apex under a static load (K_zz = n EA H^2/L0^3,
    bar force EA/L0 (H/L0) w), the Euler-Bernoulli cantilever (P L^3/3EI,
    P L^2/2EI, root shear P and moment P L), and the average-acceleration
    Newmark solution of a free one-DOF oscillator started in equilibrium
    (a0 = -w^2 u0; the reference takes the input's initial acceleration as
    given), u_k = u0 cos((k+1) theta) with cos theta = (1 - (w dt)^2/4)/(1 + (w dt)^2/4);
  * the generalized-alpha equilibrium every step of numopt 1-3 must satisfy,
    (1-am) M a_new + am M a + (1-af) K u_new + af K u = (1-af) P_k + af P_{k-1},
    which a dropped or mis-signed term of Keff or Reff breaks;
  * skymult (solve.c:700-760) against the dense product.
CPU only."""
import math
import os

import numpy as np
import pytest

import oracle as O
from cubens import inputs
from helpers import model_file, relerr, skyline_dense

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = ("star_static", "cant_static", "panel_static", "sample_5c", "star_nm1", "cant_nm3", "cant_nm2")


def gold():
    return np.load(os.path.join(GOLD, "linear.npz"))


def model(name):
    return O.OracleModel(model_file(f"lin_{name}.txt", str(gold()[f"{name}_model"])))


@pytest.mark.parametrize("name", CASES)
def test_linear_paths_against_reference_routines(name):
    g = gold()
    r = model(name).run(200, keep_ef=True)
    assert r["status"] == 0, r["msg"]
    assert np.array_equal(r["itecnt"], g[f"{name}_itecnt"])
    assert np.array_equal(r["lpf"], g[f"{name}_lpf"])
    shells = "panel" in name or "5c" in name
    tol = 1e-12 if shells else 0.0          # DKT by quadrature here, explicit there
    assert relerr(r["d"], g[f"{name}_d"]) <= tol
    assert relerr(r["ef"], g[f"{name}_ef"]) <= tol or not np.any(g[f"{name}_ef"])


def test_sample_5c_runs_its_sixteen_steps_and_the_closing_row():
    g = gold()
    t = g["sample_5c_lpf"]
    # 16 steps of 0.26934/16 from time 0, then (lpfmax, 0, d = 0) (main.c:3277-3301)
    assert len(t) == 18 and np.allclose(t[:17], np.arange(17) * 0.26934 / 16, rtol=0, atol=1e-15)
    assert t[17] == 0.001 and not np.any(g["sample_5c_d"][17])
    assert np.all(g["sample_5c_itecnt"] == 0)


def test_star_truss_static_closed_form():
    E, A, R, H, P, lam = 2e11, 1e-4, 1.0, 0.1, -1000.0, 2.0
    p = model_file("lin_star_cf.txt", inputs.star_truss(analysis="static", load=P, lpfmax=lam).text)
    r = O.OracleModel(p).run(3, keep_ef=True)
    L0 = math.hypot(R, H)
    w = lam * P / (4 * E * A * H * H / L0 ** 3)
    assert r["nsteps"] == 1 and r["lpf"][0] == lam
    assert abs(r["d"][0][2] - w) <= 1e-13 * abs(w)
    assert np.abs(r["d"][0][:2]).max() <= 1e-13 * abs(w)
    N = E * A / L0 * (H / L0) * w          # every bar, apex end; the support end opposite
    ef = r["ef"][0].reshape(4, 2)
    assert np.allclose(ef[:, 1], N, rtol=1e-12, atol=0) and np.allclose(ef[:, 0], -N, rtol=1e-12, atol=0)


def test_cantilever_static_closed_form():
    mt = inputs.cantilever_frame(n_elem=4, load=-1000.0)
    r = O.OracleModel(model_file("lin_cant_cf.txt", mt.text)).run(3, keep_ef=True)
    i = mt.info
    E, I, L, P = i["E"], i["I"], i["length"], i["load"]
    om = O.OracleModel(model_file("lin_cant_cf.txt", mt.text))
    tip = om.jcode.reshape(-1, 7)[-1]
    assert abs(r["d"][0][tip[1] - 1] - P * L ** 3 / (3 * E * I)) <= 1e-12 * abs(P * L ** 3 / (3 * E * I))
    assert abs(r["d"][0][tip[5] - 1] - P * L ** 2 / (2 * E * I)) <= 1e-12 * abs(P * L ** 2 / (2 * E * I))
    root = r["ef"][0][:14]
    assert np.isclose(np.abs(root).max(), abs(P) * L, rtol=1e-12)          # root moment P L
    assert np.isclose(np.sort(np.abs(root))[-2], abs(P) * (L - L / 4), rtol=1e-12) or \
        np.isclose(np.sort(np.abs(root))[-2], abs(P), rtol=1e-12)


def _star_sdof(numopt, rho, steps, dt, P=0.0, u0=1e-4, consistent=False):
    E, A, R, H, dens = 2e11, 1e-4, 1.0, 0.1, 7850.0
    L0 = math.hypot(R, H)
    K = 4 * E * A * H * H / L0 ** 3
    M = 4 * dens * A * L0 / 3              # lumped: rho A L/3 per bar at its free end
    a0 = (P - K * u0) / M if consistent else 0.0
    mt = inputs.star_truss(analysis="newmark", load=P, u0=u0, a0=a0, steps=steps, dt=dt, numopt=numopt,
                           spectrds=rho)
    r = O.OracleModel(model_file(f"lin_sdof_{numopt}.txt", mt.text)).run(steps + 5)
    return r, K, M


def test_average_acceleration_free_vibration_closed_form():
    dt, steps, u0 = 2e-4, 40, 1e-4
    r, K, M = _star_sdof(0, 1.0, steps, dt, u0=u0, consistent=True)
    assert r["nsteps"] == steps + 2        # time 0..steps dt, then the closing row
    om2 = (K / M) * dt * dt
    theta = math.acos((1 - om2 / 4) / (1 + om2 / 4))
    k = np.arange(steps + 1)
    want = u0 * np.cos((k + 1) * theta)
    assert np.max(np.abs(r["d"][:steps + 1, 2] - want)) <= 1e-11 * u0
    assert np.max(np.abs(r["d"][:steps + 1, :2])) <= 1e-12 * u0     # round-off of K's coupling


@pytest.mark.parametrize("numopt,rho", [(1, 0.8), (2, 0.6), (3, 0.7)])
def test_generalized_alpha_equilibrium_every_step(numopt, rho):
    dt, steps, P, u0 = 2e-4, 30, -500.0, 1e-4
    r, K, M = _star_sdof(numopt, rho, steps, dt, P=P, u0=u0)
    am, af = {1: ((2 * rho - 1) / (rho + 1), rho / (rho + 1)), 2: (0.0, (1 - rho) / (1 + rho)),
              3: ((rho - 1) / (rho + 1), 0.0)}[numopt]
    alpha, delta = (1 - am + af) ** 2 / 4, 0.5 - am + af
    a0, a2, a3 = 1 / (alpha * dt * dt), 1 / (alpha * dt), 1 / (2 * alpha) - 1
    a6, a7 = dt * (1 - delta), delta * dt
    u, v, a = u0, 0.0, 0.0
    worst = 0.0
    for k in range(steps + 1):
        un = r["d"][k, 2]
        an = a0 * (un - u) - a2 * v - a3 * a
        vn = v + a6 * a + a7 * an
        lhs = (1 - am) * M * an + am * M * a + (1 - af) * K * un + af * K * u
        rhs = (1 - af) * P + (af * P if k > 0 else 0.0)
        worst = max(worst, abs(lhs - rhs) / (abs(K * u0) + abs(P)))
        u, v, a = un, vn, an
    assert worst < 1e-10


@pytest.mark.parametrize("seed", [1, 2])
def test_skymult_is_the_symmetric_product(seed):
    rng = np.random.default_rng(seed)
    n = 40
    kht = np.minimum(np.arange(n), rng.integers(0, 7, n))
    maxa = np.concatenate([[1], 1 + np.cumsum(kht + 1)]).astype(np.int64)
    ss = rng.standard_normal(maxa[-1] - 1)
    v = rng.standard_normal(n)
    out = np.zeros(n)
    O.lib().orc_skymult(n, O.ptr(maxa, O.PL), O.ptr(ss), O.ptr(v), O.ptr(out))
    Kd = skyline_dense(ss, maxa)
    assert np.allclose(out, Kd @ v, rtol=1e-13, atol=1e-13 * np.abs(Kd).max() * np.abs(v).max())


def test_linear_newmark_checkpoint_and_restart(tmp_path):
    kw = dict(analysis="newmark", load=-500.0, steps=25, dt=1e-3, numopt=2, spectrds=0.6, u0=1e-3, chkpt=5)
    base = model_file("lin_ck.txt", inputs.cantilever_frame(**kw).text)
    rst = model_file("lin_ck_r.txt", inputs.cantilever_frame(**kw, rflag=1).text)
    full = O.OracleModel(base).run(40)
    ck = str(tmp_path / "exact.txt")
    part = O.OracleModel(base).run(13, ckpt_path=ck, ckpt_exact=True)     # rows of steps 0..12
    assert part["nsteps"] == 13
    head = open(ck).read().split("\n")
    assert head[0] == "10" and head[-2] == "0,0,0"
    r = O.OracleModel(rst).run(40, ckpt_path=ck)
    assert np.array_equal(r["lpf"], full["lpf"][11:])
    assert np.array_equal(r["d"], full["d"][11:])
    e = str(tmp_path / "e.txt")
    O.OracleModel(base).run(13, ckpt_path=e)
    r6 = O.OracleModel(rst).run(40, ckpt_path=e)
    assert relerr(r6["d"], full["d"][11:]) < 1e-5


@pytest.mark.parametrize("name", ["star_static", "cant_static", "panel_static"])
def test_linear_static_reactions_balance_the_load(name):
    """The reactions of the linear end forces balance lambda q per direction;
    the linear truss branch's end forces carry the opposite sign to the
    nonlinear ones (truss.c:300-305 against 349-350), which the reactions undo."""
    om = model(name)
    r = om.run(3, keep_react=True)
    jc = om.jcode.reshape(-1, 7)
    applied = np.array([r["lpf"][0] * om.q[jc[:, k][jc[:, k] > 0] - 1].sum() for k in range(3)])
    assert np.allclose(r["react"][0][:, :3].sum(axis=0) + applied, 0, atol=1e-12 * np.abs(applied).max())
