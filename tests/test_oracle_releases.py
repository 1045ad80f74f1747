"""Member end bending releases in the oracle (static condensation of the
released end rotations out of the local tangent: release frame.c:798-896,
inverse misc.c:284-342, applied by stiff_fr and forces_fr after the plastic
reduction, frame.c:280-283), pinned by

  * the reference's own routines under the restated drivers
    (tests/golden/releases.npz, make_releases.py): NR and Newmark frames with
    one to four releases per member, ANAFLAG 3 under NR and MSAL, free warping;
  * the textbook mid-span deflections of a beam clamped at both ends and
    loaded at mid-span, which one, two, three and four releases on a member
    turn into a propped cantilever (7 P L^3/768 E I), a simply supported beam
    (P L^3/48 E I) and a cantilever of half the span propping a link
    (P a^3/3 E I); a release dropped, put on the wrong DOF or in the wrong
    plane fails one of them;
  * zero end moments at the released DOFs, and the Gauss-Jordan inverse
    against numpy's.
CPU only."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle as O
from cubens import inputs
from helpers import model_file, relerr

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = ("girder_nr", "tower_nm", "pc_nr", "pc_msal", "warp_nr")


def gold():
    return np.load(os.path.join(GOLD, "releases.npz"))


def _maker():
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_releases", os.path.join(GOLD, "make_releases.py"))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    return mk


@pytest.mark.parametrize("name", CASES)
def test_release_paths_against_reference_routines(name):
    g = gold()
    r = O.OracleModel(model_file(f"rel_{name}.txt", str(g[f"{name}_model"]))).run(250, keep_ef=True)
    assert r["status"] == int(g[f"{name}_status"]), r["msg"]
    assert np.array_equal(r["itecnt"], g[f"{name}_itecnt"])
    assert np.array_equal(r["events"], g[f"{name}_events"])
    assert relerr(r["lpf"], g[f"{name}_lpf"]) <= 1e-13
    assert relerr(r["d"], g[f"{name}_d"]) <= 1e-12
    assert relerr(r["ef"], g[f"{name}_ef"]) <= 1e-12


def test_fixture_models_are_the_generators():
    g = gold()
    for name, text in _maker().models().items():
        assert str(g[f"{name}_model"]) == text, name
    # every release count occurs in the girder
    rows = _maker().girder_releases(inputs.space_frame(60, kind="grid", algflag=1).n_fr)
    assert {sum(f) for f in rows.values()} == {1, 2, 3, 4}


L, A = 4.0, 2.0      # span and member length of the two-member clamped beam
# (flags of member 1, flags of member 2) -> mid-span deflection / (P L^3 / E I)
# for a load bending about the strong (local z) axis, direction 3, and about
# the weak (local y) axis, direction 2
BEAMS = [
    ((0, 0, 0, 0), (0, 0, 0, 0), 1 / 192, 1 / 192),                     # clamped-clamped
    ((1, 0, 0, 0), (0, 0, 0, 0), 7 / 768, 1 / 192),                     # propped (strong plane)
    ((0, 1, 0, 0), (0, 0, 0, 0), 1 / 192, 7 / 768),                     # propped (weak plane)
    ((1, 1, 0, 0), (0, 0, 1, 1), 1 / 48, 1 / 48),                       # simply supported
    ((1, 1, 1, 0), (0, 0, 0, 0), A ** 3 / 3 / L ** 3, 7 / 768),         # link / propped
    ((1, 1, 1, 1), (0, 0, 0, 0), A ** 3 / 3 / L ** 3, A ** 3 / 3 / L ** 3),   # link in both planes
]


@pytest.mark.parametrize("f1,f2,strong,weak", BEAMS)
@pytest.mark.parametrize("direction", [3, 2])
def test_clamped_beam_with_releases_closed_form(f1, f2, strong, weak, direction):
    mt = inputs.clamped_beam(n_elem=2, length=L, direction=direction)
    rows = {m: f for m, f in ((1, f1), (2, f2)) if any(f)}
    if rows:
        mt = mt.with_releases(rows)
    om = O.OracleModel(model_file("rel_beam.txt", mt.text))
    r = om.run(3, keep_ef=True)
    assert r["status"] == 0 and r["nsteps"] == 1, r["msg"]
    P, EI = mt.info["load"], mt.info["E"] * mt.info["I"]
    mid = om.jcode.reshape(-1, 7)[1]
    v = r["d"][0][mid[direction - 1] - 1]
    want = (strong if direction == 3 else weak) * P * L ** 3 / EI
    assert abs(v - want) <= 1e-11 * abs(want)
    # no end moment at a released DOF (local 5, 4, 12, 11 for flags 1-4)
    ef = r["ef"][0].reshape(2, 14)
    for m, f in ((0, f1), (1, f2)):
        for flag, dof in zip(f, (5, 4, 12, 11)):
            if flag:
                assert abs(ef[m, dof]) <= 1e-9 * abs(P) * L


@pytest.mark.parametrize("n", [1, 2, 3, 4])
def test_gauss_jordan_inverse(n):
    lib = O.lib()
    lib.orc_gj_inverse.argtypes = [O.PD, C.c_int]
    rng = np.random.default_rng(n)
    M = rng.standard_normal((n, n))
    M = M @ M.T + n * np.eye(n)
    a = M.copy().ravel()
    lib.orc_gj_inverse(O.ptr(a), n)
    assert np.allclose(a.reshape(n, n), np.linalg.inv(M), rtol=1e-13, atol=1e-14)
    if n >= 2:
        # a zero diagonal: the row search swaps in the row of the first non-zero column
        P = np.eye(n)[::-1] * 2.0
        a = P.copy().ravel()
        lib.orc_gj_inverse(O.ptr(a), n)
        assert np.allclose(a.reshape(n, n), np.linalg.inv(P), rtol=0, atol=1e-15)
