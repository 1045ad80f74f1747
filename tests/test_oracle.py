"""The oracle (CPU restatement) pinned against the reference's own routines
(golden fixtures written by tests/golden/make_golden.py) and against closed
forms and invariants chosen so that plausible mistakes fail: a transposed
triad, a dropped drilling term, a wrong sign in the bending interpolation, a
shifted skyline address, or the wrong root in the arc-length quadratic.
CPU only."""
import os

import numpy as np
import pytest

import oracle as O
from cubens import inputs
from helpers import model_file, relerr, skyline_dense

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def gold(name):
    return np.load(os.path.join(GOLD, name))


# --------------------------------------------------------------- golden vectors
def test_truss_dome_against_reference_routines():
    g = gold("truss_dome.npz")
    om = O.OracleModel(model_file("c1.txt", inputs.truss_dome().text))
    assert np.array_equal(om.jcode, g["jcode"])
    assert np.array_equal(om.mcode, g["mcode"])
    assert np.array_equal(om.maxa, g["maxa"])
    assert np.array_equal(om.x, g["x"]) and np.array_equal(om.q, g["q"])
    assert np.array_equal(om.llength, g["llength"])
    r = om.run(40)
    assert r["nsteps"] == len(g["path_lpf"]) == 20
    assert np.array_equal(r["itecnt"], g["path_itecnt"])
    # the truss routines are restated operation for operation: bit-identical
    assert np.array_equal(r["lpf"], g["path_lpf"])
    assert np.array_equal(r["d"], g["path_d"])


def test_reference_sample_5a_against_reference_routines():
    g = gold("sample_5a.npz")
    om = O.OracleModel(os.path.join(GOLD, "model_def_5a_truss.txt"))
    assert np.array_equal(om.jcode, g["jcode"]) and np.array_equal(om.q, g["q"])
    r = om.run(50)
    assert np.array_equal(r["itecnt"], g["path_itecnt"])
    assert np.array_equal(r["d"], g["path_d"])


def _panel8():
    return O.OracleModel(model_file("p8g.txt", inputs.shell_panel(n=8, thick_sigma=0.1, imperfection=0.1).text))


def test_shell_model_input_against_reference():
    g = gold("panel8.npz")
    om = _panel8()
    for k in ("jcode", "mcode", "maxa"):
        assert np.array_equal(getattr(om, k), g[k]), k
    for k in ("x", "c1", "c2", "c3", "xlocal", "farea", "slength", "q"):
        assert np.array_equal(getattr(om, k), g[k]), k


def test_shell_routines_against_reference():
    g = gold("panel8.npz")
    om = _panel8()
    x, xip, _, deffarea, defslen, c1, c2, c3 = O.updatc(om, om.x, g["dd"])
    assert np.array_equal(x, g["x_new"])
    assert np.array_equal(deffarea, g["deffarea"]) and np.array_equal(defslen, g["defslen"])
    assert np.array_equal(c1, g["c1_new"]) and np.array_equal(c3, g["c3_new"])
    ss = O.stiff_sh(om, x, c1, c2, c3, deffarea)
    # DKT by quadrature here, by the explicit product form in the reference
    assert relerr(ss, g["ss"]) < 1e-13
    f, efi, efp = O.forces_sh(om, g["dd"], x, (om.c1, om.c2, om.c3), (c1, c2, c3), g["ef_ip_in"])
    assert relerr(f, g["f"]) < 1e-12
    assert relerr(efi, g["ef_i"]) < 1e-12
    assert np.array_equal(efp, g["ef_ip_out"])


def test_shell_path_against_reference_routines():
    g = gold("panel8.npz")
    r = _panel8().run(12)
    assert r["nsteps"] == 12
    assert np.array_equal(r["itecnt"], g["path_itecnt"])
    assert relerr(r["lpf"], g["path_lpf"]) < 1e-10
    assert relerr(r["d"], g["path_d"]) < 1e-10


# --------------------------------------------------------------- skyline solver
def _random_skyline(n, seed, spd=True):
    rng = np.random.default_rng(seed)
    kht = np.minimum(np.arange(n), rng.integers(0, 6, n))
    maxa = np.concatenate([[1], 1 + np.cumsum(kht + 1)]).astype(np.int64)
    A = np.zeros((n, n))
    for j in range(n):
        for i in range(j - kht[j], j + 1):
            A[i, j] = A[j, i] = rng.standard_normal()
    A += (n if spd else -0.5) * np.eye(n) if spd else np.diag(rng.standard_normal(n) * 3)
    ss = np.zeros(maxa[-1] - 1)
    for j in range(n):
        for a in range(maxa[j], maxa[j + 1]):
            ss[a - 1] = A[j - (a - maxa[j]), j]
    return A, ss, maxa


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_skyline_ldlt_pivots_are_leading_minor_ratios(seed):
    A, ss, maxa = _random_skyline(24, seed)
    rc, fac, piv, det = O.skyfact(maxa, ss)
    assert rc == 0 and det == 0
    minors = [np.linalg.det(A[:k, :k]) for k in range(1, 25)]
    want = np.array([minors[0]] + [minors[k] / minors[k - 1] for k in range(1, 24)])
    assert relerr(piv, want) < 1e-10
    b = np.random.default_rng(seed).standard_normal(24)
    assert relerr(O.skysolve(maxa, fac, b), np.linalg.solve(A, b)) < 1e-10


def test_skyline_det_flag_on_indefinite_matrix():
    A, ss, maxa = _random_skyline(20, 5, spd=False)
    rc, fac, piv, det = O.skyfact(maxa, ss)
    eig = np.linalg.eigvalsh(A)
    assert rc == 0
    # Sylvester: the number of negative pivots is the number of negative eigenvalues
    assert (piv < 0).sum() == (eig < 0).sum()
    assert det == int((eig < 0).any())
    # the non-arc-length branch refuses a non-positive pivot
    rc2, _, _, _ = O.skyfact(maxa, ss, msal=0)
    assert rc2 == 1


# --------------------------------------------------------------- arc length
def test_quad_satisfies_the_constraint_and_picks_the_forward_root():
    import ctypes as C
    rng = np.random.default_rng(3)
    n = 30
    dp = rng.standard_normal(n)
    dt = dp + 0.1 * rng.standard_normal(n)
    ddr = 1e-3 * rng.standard_normal(n)
    ddq = rng.standard_normal(n)
    arc = 1.5 * np.linalg.norm(dt - dp + ddr)
    a = ddq @ ddq
    b = 2 * (dt @ ddq - dp @ ddq + ddr @ ddq)
    c = 2 * (dt @ ddr - dp @ ddr - dt @ dp) + dt @ dt + dp @ dp + ddr @ ddr - arc * arc
    d_t = dt.copy()
    dd = np.zeros(n)
    dl = C.c_double()
    lt = C.c_double(0.0)
    rc = O.lib().orc_quad(n, a, b, c, O.ptr(d_t), O.ptr(dp), O.ptr(ddr), O.ptr(ddq), O.ptr(dd),
                          C.byref(dl), C.byref(lt))
    assert rc == 0
    # |d - dp|^2 = arc^2 after the update (cylindrical constraint)
    assert abs(np.linalg.norm(d_t - dp) - arc) < 1e-12
    assert np.allclose(dd, ddr + dl.value * ddq)
    roots = np.roots([a, b, c])
    other = roots[np.argmax(np.abs(roots - dl.value))]
    g_pick = (dt - dp) @ (dt - dp + ddr + dl.value * ddq)
    g_other = (dt - dp) @ (dt - dp + ddr + other * ddq)
    assert g_pick > 0 and g_pick > g_other
    # imaginary roots are reported
    assert O.lib().orc_quad(n, 1.0, 0.0, 1.0, O.ptr(d_t), O.ptr(dp), O.ptr(ddr), O.ptr(ddq), O.ptr(dd),
                            C.byref(dl), C.byref(lt)) == 2


# --------------------------------------------------------------- element invariants
def _elem_K(E=2e9, nu=0.3, t=0.01, X=((0, 0, 0), (1.2, 0.1, 0.05), (0.3, 0.9, -0.1)), geometric=False, dm=None):
    """Global 18x18 tangent of one triangle via the oracle."""
    import ctypes as C
    X = np.array(X, float)
    e12, e31 = X[1] - X[0], X[2] - X[0]
    nrm = np.cross(e12, e31)
    A = 0.5 * np.linalg.norm(nrm)
    c1 = e12 / np.linalg.norm(e12)
    c3 = nrm / (2 * A)
    c2 = np.cross(c3, c1)
    c2 /= np.linalg.norm(c2)
    xl = np.array([c1 @ e12, c1 @ e31, c2 @ e31])
    sl = np.array([np.linalg.norm(e12), np.linalg.norm(X[2] - X[1]), np.linalg.norm(e31)])
    K = np.zeros(324)
    dmv = np.zeros(6) if dm is None else np.asarray(dm, float)
    O.lib().orc_shell_elem_tangent(E, nu, t, O.ptr(xl), A, O.ptr(sl), A, O.ptr(dmv), O.ptr(c1), O.ptr(c2),
                                   O.ptr(c3), int(geometric), O.ptr(K))
    return K.reshape(18, 18), X, (c1, c2, c3), xl, A


def test_shell_element_symmetric_and_rigid_body_free():
    K, X, (c1, c2, c3), _, _ = _elem_K()
    assert np.abs(K - K.T).max() < 1e-9 * np.abs(K).max()
    # rigid translations
    for d in np.eye(3):
        u = np.zeros(18)
        for a in range(3):
            u[6 * a:6 * a + 3] = d
        assert np.abs(K @ u).max() < 1e-8 * np.abs(K).max()
    # rigid rotation about the element normal is resisted only by the drilling
    # term, about an in-plane axis not at all: rotation omega about c1
    om = c1
    u = np.zeros(18)
    for a in range(3):
        u[6 * a:6 * a + 3] = np.cross(om, X[a])
        u[6 * a + 3:6 * a + 6] = om
    assert np.abs(K @ u).max() < 1e-8 * np.abs(K).max()


def test_dkt_constant_curvature_energy():
    """w = kappa x^2 / 2 (in the element plane) stores D kappa^2 A / 2."""
    E, nu, t = 2e9, 0.3, 0.01
    K, X, (c1, c2, c3), xl, A = _elem_K(E, nu, t, X=((0, 0, 0), (1.0, 0, 0), (0.4, 0.8, 0)))
    kappa = 1e-3
    u = np.zeros(18)
    for a in range(3):
        x = X[a, 0]
        u[6 * a + 2] = 0.5 * kappa * x * x       # w
        u[6 * a + 4] = -kappa * x                  # theta_y = -dw/dx
    D = E * t ** 3 / (12 * (1 - nu ** 2))
    assert abs(0.5 * u @ K @ u - 0.5 * D * kappa ** 2 * A) < 1e-10 * D * kappa ** 2 * A


def test_cst_constant_strain_energy():
    E, nu, t = 2e9, 0.3, 0.01
    K, X, _, _, A = _elem_K(E, nu, t, X=((0, 0, 0), (1.0, 0, 0), (0.4, 0.8, 0)))
    ex, ey, g = 1e-4, -3e-5, 2e-5
    u = np.zeros(18)
    for a in range(3):
        x, y = X[a, :2]
        u[6 * a] = ex * x + 0.5 * g * y
        u[6 * a + 1] = ey * y + 0.5 * g * x
    C_ = E / (1 - nu ** 2) * np.array([[1, nu, 0], [nu, 1, 0], [0, 0, (1 - nu) / 2]])
    eps = np.array([ex, ey, g])
    assert abs(0.5 * u @ K @ u - 0.5 * t * A * eps @ C_ @ eps) < 1e-10 * t * A * eps @ C_ @ eps


def test_drilling_is_1e4_of_the_bending_rotation_term():
    K, X, R, _, _ = _elem_K(X=((0, 0, 0), (1.0, 0, 0), (0.4, 0.8, 0)))
    # flat element in the global xy plane: local == global
    assert abs(K[5, 5] - K[3, 3] / 1e4) < 1e-12 * K[3, 3]


def test_simply_supported_plate_centre_deflection():
    """Square plate a = 1, simply supported (translations fixed on the edges),
    centre load P: w = 0.01160 P a^2 / D (Timoshenko & Woinowsky-Krieger)."""
    import ctypes as C
    n = 16
    E, nu, t, P = 1e9, 0.3, 0.01, 1.0
    m = n + 1
    node = lambda i, j: j * m + i + 1  # noqa: E731
    tris = []
    for j in range(n):
        for i in range(n):
            a, b, c, d = node(i, j), node(i + 1, j), node(i + 1, j + 1), node(i, j + 1)
            tris += [(a, b, c), (a, c, d)]
    lines = [f"2\n3\n0\n1\n{m * m}\n0,0,{len(tris)},0,0\n"]
    lines += [f"{a},{b},{c}\n" for a, b, c in tris]
    for j in range(m):
        for i in range(m):
            if i in (0, n) or j in (0, n):
                lines += [f"{node(i, j)},{k}\n" for k in (1, 2, 3)]
    lines.append("0,0\n")
    lines += [f"{i / n},{j / n},0.0\n" for j in range(m) for i in range(m)]
    lines += [f"{E},{nu},{t},1.0,1e30\n"] * len(tris)
    ctr = node(n // 2, n // 2)
    lines.append(f"{ctr},3,{-P}\n0,0,0\n{ctr},3,-1e-6\n1\n0.1\n6\n1e30,1e30\n10,1,1,1\n1,1e-3,1\n")
    om = O.OracleModel(model_file("plate.txt", "".join(lines)))
    ss = O.stiff_sh(om, om.x, om.c1, om.c2, om.c3, om.farea)
    rc, fac, piv, det = O.skyfact(om.maxa, ss)
    w = O.skysolve(om.maxa, fac, om.q)[om.dkdof]
    D = E * t ** 3 / (12 * (1 - nu ** 2))
    exact = -0.01160 * P / D
    assert abs(w / exact - 1) < 0.02


def test_truss_force_closed_form_and_tangent():
    """Single bar: N = EA (eps + eps^2/2) L/L0 and its tangent EA L^2/L0^3 + N/L
    (truss.c:108-111, 155-166, 329-331) against a finite difference."""
    txt = ("2\n1\n0\n1\n2\n1,0,0,0,0\n1,2\n1,1\n1,2\n1,3\n2,2\n2,3\n0,0\n"
           "0,0,0\n2.0,0.5,0\n2e11,1e-4,7850,1e30\n2,1,1000\n0,0,0\n"
           "1.0,1.0,1.0,1.0,1.0\n20,5,10\n1e-12,1e-12,1e-12\n")
    om = O.OracleModel(model_file("bar.txt", txt))
    assert om.neq == 1
    r = om.run(5)
    u = r["d"][0, 0]
    L0 = np.hypot(2.0, 0.5)
    L = np.hypot(2.0 + u, 0.5)
    eps = (L - L0) / L0
    N = 2e11 * 1e-4 * (eps + eps * eps / 2) * L / L0
    assert abs(N * (2.0 + u) / L - 1000) < 1e-6 * 1000   # equilibrium in x
