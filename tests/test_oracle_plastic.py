"""Inelastic analyses (ANAFLAG 3) of the oracle: plastic hinges at frame
member ends (stiffm_fr frame.c:580-797, the yield checks of forces_fr
frame.c:1153-1270, regula_falsi frame.c:1397-1455, unload frame.c:1457-1658)
and bars capped at their squash load (stiffm_tr truss.c:206-229, forces_tr
truss.c:336-347), under the NR/MNR, MSAL and nonlinear Newmark drivers with
their restart rules for forces_fr's codes (main.c:2033-2057, 2954-2970,
3773-3795).  Pinned by

  * the reference's own element routines under the restated drivers
    (tests/golden/plastic.npz, make_plastic.py), among them its sample input
    model_def_5b_frame.txt;
  * closed forms of small-displacement plastic theory: the three-bar truss
    (vertical bar at P / (1 + 2 cos^3 t) until it yields at
    P1 = Py (1 + 2 cos^3 t), then capped at Py with the inclined bars taking
    (P - Py) / (2 cos t)), and the propped cantilever under a central load
    (first hinge at the clamp at P1 = 16 Mp / 3L, collapse at Pc = 6 Mp / L);
  * Orbison's surface through the squash load and the plastic moments
    (phi(1,0,0) = phi(0,1,0) = phi(0,0,1) = 1), its gradient against central
    differences, the plastic tangent annihilating the yield gradients of the
    ends it reduces (k_p G = 0), and the consistency of the hinge: once the
    clamp has yielded under NR it stays on the surface within the tolerance.
CPU only."""
import math
import os

import numpy as np
import pytest

import oracle as O
from cubens import inputs
from helpers import model_file, relerr

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = ("tb_nr", "pc_nr", "pc_mnr", "pc_msal", "pc_nm", "pc_nm_fail", "sample_5b")
PHITOL = 1e-4


def gold():
    return np.load(os.path.join(GOLD, "plastic.npz"))


def model(name):
    return O.OracleModel(model_file(f"pl_{name}.txt", str(gold()[f"{name}_model"])))


def end_phis(ef, Py, Mpy, Mpz):
    """phi at both ends of every member of every row: [rows, members, 2]."""
    e = ef.reshape(len(ef), -1, 14)
    out = np.zeros(e.shape[:2] + (2,))
    for s in range(e.shape[0]):
        for n in range(e.shape[1]):
            for a in range(2):
                out[s, n, a] = O.orbison(e[s, n, 7 * a] / Py, e[s, n, 7 * a + 4] / Mpy,
                                         e[s, n, 7 * a + 5] / Mpz)
    return out


@pytest.mark.parametrize("name", CASES)
def test_plastic_paths_against_reference_routines(name):
    g = gold()
    r = model(name).run(250, keep_ef=True)
    assert r["status"] == int(g[f"{name}_status"]), r["msg"]
    assert np.array_equal(r["itecnt"], g[f"{name}_itecnt"])
    assert np.array_equal(r["events"], g[f"{name}_events"])
    assert relerr(r["lpf"], g[f"{name}_lpf"]) <= 1e-13
    assert relerr(r["d"], g[f"{name}_d"]) <= 1e-12
    assert relerr(r["ef"], g[f"{name}_ef"]) <= 1e-12


def test_fixtures_cover_yield_and_unload_restarts():
    g = gold()
    ev = {n: g[f"{n}_events"].sum(axis=0) for n in CASES}
    assert ev["pc_nr"][O.EV_YIELD] == 1 and ev["pc_mnr"][O.EV_YIELD] == 1
    assert ev["pc_msal"][O.EV_YIELD] > 0 and ev["pc_msal"][O.EV_UNLOAD] > 0
    assert ev["pc_nm"][O.EV_YIELD] > 0
    assert ev["pc_nm_fail"][O.EV_UNLOAD] > 0 and int(g["pc_nm_fail_status"]) == 3


def test_three_bar_truss_closed_form():
    mt = inputs.three_bar_truss()
    i = mt.info
    r = O.OracleModel(model_file("pl_tb.txt", mt.text)).run(50, keep_ef=True)
    assert r["status"] == 0
    c = math.cos(i["theta"])
    P = r["lpf"] * abs(i["load"])
    N = r["ef"][:, 0::2]                    # axial force of bars 1 (vertical), 2, 3
    below, above = P < 0.999 * i["P1"], P > 1.001 * i["P1"]
    assert below.sum() >= 3 and above.sum() >= 2
    # tension is positive; geometric nonlinearity at strains near 1e-3 is the slack
    assert np.allclose(N[below, 0], P[below] / (1 + 2 * c ** 3), rtol=5e-3)
    assert np.allclose(N[below, 1], P[below] * c * c / (1 + 2 * c ** 3), rtol=5e-3)
    assert np.all(N[above, 0] == i["Py"])   # capped exactly
    assert np.allclose(N[above, 1:], ((P[above] - i["Py"]) / (2 * c))[:, None], rtol=5e-3)
    assert P.max() < i["Pc"]


def test_propped_cantilever_first_hinge_and_surface_consistency():
    g = gold()
    info = inputs.propped_cantilever().info
    Py, Mp = info["Py"], info["Mp"]
    lam = g["pc_nr_lpf"]
    phi = end_phis(g["pc_nr_ef"], Py, Mp, Mp)
    root = phi[:, 0, 0]
    s = int(np.argmax(root >= 1 - PHITOL))
    assert root[s] >= 1 - PHITOL and np.all(root[:s] < 1 - PHITOL)
    # the regula-falsi restart lands the first hinge at P1 (second-order terms aside)
    assert abs(lam[s] * 1000.0 - info["P1"]) <= 2e-4 * info["P1"]
    assert g["pc_nr_events"][s, O.EV_YIELD] == 1
    # the hinge stays on the surface and nothing else crosses it
    assert np.all(np.abs(root[s:] - 1) <= PHITOL)
    assert phi.max() <= 1 + PHITOL
    assert np.sum(phi >= 1 - PHITOL) == len(lam) - s


def test_propped_cantilever_collapse_plateau_under_msal():
    g = gold()
    info = inputs.propped_cantilever().info
    lam = g["pc_msal_lpf"] * 1000.0
    # the arc-length path climbs to the mechanism load of plastic theory and
    # stalls there (the subdivisions run out on the plateau)
    assert int(g["pc_msal_status"]) == 1
    assert 0.97 * info["Pc"] <= lam.max() <= 1.01 * info["Pc"]
    phi = end_phis(g["pc_msal_ef"], info["Py"], info["Mp"], info["Mp"])
    s = int(np.argmax(phi[:, 0, 0] >= 1 - PHITOL))
    assert abs(lam[s] - info["P1"]) <= 2e-4 * info["P1"]


def test_orbison_surface_and_gradient():
    for p, my, mz in [(1, 0, 0), (0, 1, 0), (0, 0, 1), (-1, 0, 0), (0, -1, 0), (0, 0, -1)]:
        assert O.orbison(p, my, mz) == 1.0
    assert O.orbison(0.5, 0, 0) == 0.25 and O.orbison(0, 0.5, 0) == 0.0625
    rng = np.random.default_rng(3)
    Py, Mpy, Mpz = 2.0e6, 1.1e5, 1.3e5
    for _ in range(20):
        x = rng.uniform(-0.9, 0.9, 3)
        g = O.orbison_grad(*x, Py, Mpy, Mpz)
        h = 1e-6
        fd = []
        for k, scale in enumerate((Py, Mpy, Mpz)):
            e = np.zeros(3)
            e[k] = h
            fd.append((O.orbison(*(x + e)) - O.orbison(*(x - e))) / (2 * h * scale))
        assert np.allclose(g, fd, rtol=1e-7, atol=1e-12 * np.abs(g).max())


def _member_k(seed=0):
    info = inputs.propped_cantilever().info
    rng = np.random.default_rng(seed)
    k = O.frame_ke(info["E"], info["E"] / 2.6, info["A"], 0.5, info["I"], info["I"], 2 * info["I"], 0.0)
    f = rng.standard_normal(14) * 1e3
    k = k + O.frame_kg(f, 0.5, info["A"], 2 * info["I"])
    return k, info


@pytest.mark.parametrize("ends", [(0,), (1,), (0, 1)])
def test_plastic_tangent_annihilates_the_active_gradients(ends):
    k, info = _member_k(1)
    Py, Mp = info["Py"], info["Mp"]
    f = np.zeros(14)
    for a in (0, 1):
        # on the surface (phi = 1) for the ends that yield, inside for the others
        t = 1.0 if a in ends else 0.5
        v = np.array([0.3, 0.2, 0.5])
        v *= t / O.orbison(*v) ** 0.25 if a in ends else 0.5
        while a in ends and abs(O.orbison(*v) - 1) > 1e-12:
            v *= (1 / O.orbison(*v)) ** 0.25
        f[7 * a], f[7 * a + 4], f[7 * a + 5] = v[0] * Py, v[1] * Mp, v[2] * Mp
    kp = O.frame_plastic(k, f, [0, 0], Py, Mp, Mp)
    assert np.allclose(kp, kp.T, rtol=0, atol=1e-9 * np.abs(k).max())
    for a in ends:
        G = np.zeros(14)
        G[[7 * a, 7 * a + 4, 7 * a + 5]] = O.orbison_grad(f[7 * a] / Py, f[7 * a + 4] / Mp,
                                                          f[7 * a + 5] / Mp, Py, Mp, Mp)
        assert np.abs(kp @ G).max() <= 1e-9 * np.abs(k @ G).max()
    # an end flagged 2 (unloading) is not reduced
    if ends == (0,):
        assert np.array_equal(O.frame_plastic(k, f, [2, 0], Py, Mp, Mp), k)


def test_unload_reads_the_sign_of_the_plastic_multipliers():
    Py, Mpy, Mpz = 1.0, 1.0, 1.0
    f = np.zeros(14)
    f[0], f[7] = 1.0, -1.0                   # both ends at the squash load
    k = np.eye(14)
    g0, g1 = np.zeros(14), np.zeros(14)
    g0[0], g1[7] = 2.0, -2.0                 # d phi / dP at each end
    both = np.array([1.0, 1.0])
    assert O.frame_unload(both, f, Py, Mpy, Mpz, k, -(g0 + g1)) == 1
    assert O.frame_unload(both, f, Py, Mpy, Mpz, k, -g0 + g1) == 2
    assert O.frame_unload(both, f, Py, Mpy, Mpz, k, g0 - g1) == 3
    assert O.frame_unload(both, f, Py, Mpy, Mpz, k, g0 + g1) == 0
    assert O.frame_unload(np.array([1.0, 0.5]), f, Py, Mpy, Mpz, k, -g0) == 2
    assert O.frame_unload(np.array([0.5, 1.0]), f, Py, Mpy, Mpz, k, -g1) == 3
    assert O.frame_unload(np.array([0.5, 0.5]), f, Py, Mpy, Mpz, k, -g0) == 0


def test_regula_falsi_brackets_the_surface():
    rng = np.random.default_rng(5)
    for _ in range(20):
        v = rng.uniform(-1, 1, 3)
        while abs(O.orbison(*v) - 1) > 1e-13:
            v *= (1 / O.orbison(*v)) ** 0.25
        # an increment through the surface: from 0.1% inside to 0.1% outside
        x, dx = 0.999 * v, 0.002 * v
        t = O.regula_falsi(x[0], dx[0], x[1], dx[1], x[2], dx[2])
        assert 0 < t < 1
        # false position is exact to second order on so short a chord
        assert abs(O.orbison(*(x + t * dx)) - 1) < 1e-7
        # on a long chord through a curved surface it stays inside the bracket
        t = O.regula_falsi(*(w for pair in zip(0.5 * v, v) for w in pair))
        assert 0 < t < 1


def test_sample_5b_stays_elastic_and_runs_its_sixteen_steps():
    g = gold()
    t = g["sample_5b_lpf"]
    assert len(t) == 17 and np.allclose(t, np.arange(17) * 0.11538 / 16, rtol=0, atol=1e-15)
    assert not g["sample_5b_events"].any()
    om = model("sample_5b")
    assert om.anaflag == 3


def test_anaflag3_shells_are_rejected():
    mt = inputs.shell_panel(n=4)
    text = "3" + mt.text[1:]
    with pytest.raises(ValueError, match="ANAFLAG 3 shells"):
        O.OracleModel(model_file("pl_shell.txt", text))
