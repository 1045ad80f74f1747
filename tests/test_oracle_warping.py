"""Frame models with free warping DOFs in the oracle: the warping DOF of a
joint shared by its members (codes(), model.c:951-957) and Vlasov's
non-uniform torsion in stiffe_fr / stiffg_fr (frame.c:354-579), pinned by

  * the reference's own element routines under the restated drivers
    (tests/golden/warping.npz, make_warping.py): a thin-walled cantilever in
    torsion (linear and NR), a tower under Newmark and a girder under NR;
  * Vlasov's closed forms for a cantilever under a tip torque T with its root
    unable to warp: tip twist T/(GJ) (L - tanh(kL)/k), k^2 = GJ/(E Cw),
    approached at the h^4 rate of the cubic elements, and the root bimoment
    reaction (T/k) tanh(kL); with the root free to warp, Saint-Venant's
    T L/(GJ) exactly.
CPU only."""
import math
import os

import numpy as np
import pytest

import oracle as O
from cubens import inputs
from helpers import model_file, relerr

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = ("tors_static", "tors_nr", "tower_nm", "girder_nr")


def gold():
    return np.load(os.path.join(GOLD, "warping.npz"))


@pytest.mark.parametrize("name", CASES)
def test_warping_paths_against_reference_routines(name):
    g = gold()
    r = O.OracleModel(model_file(f"w_{name}.txt", str(g[f"{name}_model"]))).run(100, keep_ef=True)
    assert r["status"] == 0, r["msg"]
    assert np.array_equal(r["itecnt"], g[f"{name}_itecnt"])
    assert np.array_equal(r["lpf"], g[f"{name}_lpf"])
    assert relerr(r["d"], g[f"{name}_d"]) <= 1e-13
    assert relerr(r["ef"], g[f"{name}_ef"]) <= 1e-13


def test_warping_dofs_are_numbered_per_joint():
    om = O.OracleModel(model_file("w_tc.txt", inputs.torsion_cantilever(n_elem=4).text))
    jc = om.jcode.reshape(-1, 7)
    assert jc[0, 6] == 0 and np.all(jc[1:, 6] > 0)
    mc = om.mcode.reshape(-1, 14)
    # a joint's warping DOF is the one both of its members carry
    assert np.array_equal(mc[:-1, 13], mc[1:, 6])


@pytest.mark.parametrize("n,tol", [(4, 2e-3), (16, 1e-5), (64, 1e-7)])
def test_vlasov_tip_twist_and_root_bimoment(n, tol):
    mt = inputs.torsion_cantilever(n_elem=n)
    i = mt.info
    om = O.OracleModel(model_file(f"w_tc{n}.txt", mt.text))
    r = om.run(3, keep_react=True)
    tip = om.jcode.reshape(-1, 7)[-1]
    theta = r["d"][0][tip[3] - 1]
    assert abs(theta / i["theta_fixed"] - 1) <= tol
    rx = r["react"][0][0]
    assert abs(rx[3] + i["torque"]) <= 1e-10 * i["torque"]
    B = i["torque"] / i["k"] * math.tanh(i["k"] * i["length"])
    assert abs(-rx[6] / B - 1) <= 8 * tol


def test_saint_venant_with_the_root_free_to_warp():
    mt = inputs.torsion_cantilever(n_elem=8, root_warping="free")
    om = O.OracleModel(model_file("w_tcf.txt", mt.text))
    r = om.run(3)
    tip = om.jcode.reshape(-1, 7)[-1]
    assert abs(r["d"][0][tip[3] - 1] / mt.info["theta_free"] - 1) <= 1e-13
