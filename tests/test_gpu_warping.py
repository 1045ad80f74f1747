"""Frame models with free warping DOFs on the GPU through the C ABI: the
device layout then carries 7 slots per joint (the shared warping DOF in slot
6, codes() model.c:951-957).  Paths against the reference routines
(tests/golden/warping.npz) and the oracle, the kernels at a deformed state,
Vlasov's closed form, and the reference's own sample model_def_5b_frame.txt
as given (its warping DOFs are free).

Bar: FP64 paths 1e-9 with identical iteration counts; kernels 1e-12."""
import math

import numpy as np
import pytest

import cubens
import oracle as O
from cubens import inputs
from helpers import model_file, relerr, skyline_dense
from test_oracle_warping import CASES, gold

pytestmark = pytest.mark.gpu


def _opts(steps=100):
    return cubens.options(cubens.SOLVER_BAND64, max_steps=steps)


@pytest.mark.parametrize("name", CASES)
def test_warping_paths_match_reference_routines_and_oracle(name):
    g = gold()
    path = model_file(f"gpu_w_{name}.txt", str(g[f"{name}_model"]))
    print(f"[warping] {name}", flush=True)
    with cubens.GpuModel(path) as m:
        assert m.info.bs == 7
        r = m.run(_opts())
    o = O.OracleModel(path).run(100)
    assert r["status"] == 0 == o["status"], r["msg"]
    for ref in ({"itecnt": g[f"{name}_itecnt"], "lpf": g[f"{name}_lpf"], "d": g[f"{name}_d"]}, o):
        assert np.array_equal(r["itecnt"], ref["itecnt"])
        assert relerr(r["lpf"], ref["lpf"]) < 1e-12
        assert relerr(r["d"], ref["d"]) < 1e-9


def test_warping_tangent_and_forces_at_a_deformed_state():
    path = model_file("gpu_w_tower.txt", inputs.space_frame(60, kind="tower", steps=5, warping=True).text)
    om = O.OracleModel(path)
    rng = np.random.default_rng(17)
    dd = 3e-3 * rng.standard_normal(om.neq)
    x, xip, defllen, deffarea, defslen, c1, c2, c3 = O.updatc(om, om.x, dd)
    ef_ip = 1e4 * rng.standard_normal(om.nfr * 14)
    print("[warping] kernels", flush=True)
    with cubens.GpuModel(path) as g:
        assert g.info.bs == 7
        for name, v in (("c1_ip", om.c1), ("c2_ip", om.c2), ("c3_ip", om.c3), ("ef_ip", ef_ip),
                        ("defllen_ip", defllen), ("c1_i", c1), ("c2_i", c2), ("c3_i", c3), ("x_temp", x)):
            g.set(name, v)
        jc, _, _ = g.dofmaps()
        rp, ci = g.pattern()
        Kg = cubens.bsr_to_dense(g.tangent(cubens.GEN_I, cubens.GEN_IP), rp, ci, 7,
                                 cubens.eq_of_slot(jc, g.nj, 7), g.neq)
        fg = g.internal_force(dd, cubens.GEN_I)
        efg = g.get("ef_i", om.nfr * 14)
    Ko = skyline_dense(O.stiff_fr(om, om.llength, defllen, om.c1, om.c2, om.c3, ef_ip), om.maxa)
    fo, efo = O.forces_fr(om, dd, om.llength, defllen, (om.c1, om.c2, om.c3), (c1, c2, c3), ef_ip)
    assert relerr(Kg, Ko) < 1e-12
    assert relerr(fg, fo) < 1e-12
    assert relerr(efg, efo) < 1e-12


@pytest.mark.parametrize("root", ["fixed", "free"])
def test_vlasov_torsion_on_the_gpu(root):
    mt = inputs.torsion_cantilever(n_elem=64, root_warping=root)
    path = model_file(f"gpu_w_tc_{root}.txt", mt.text)
    with cubens.GpuModel(path) as m:
        r = m.run(_opts(3))
        jc, _, _ = m.dofmaps()
        rx = m.reactions().reshape(-1, 7)
    tip = jc.reshape(-1, 7)[-1]
    i = mt.info
    theta = r["d"][0][tip[3] - 1]
    if root == "fixed":
        assert abs(theta / i["theta_fixed"] - 1) <= 1e-7
        B = i["torque"] / i["k"] * math.tanh(i["k"] * i["length"])
        assert abs(-rx[0, 6] / B - 1) <= 1e-6
    else:
        assert abs(theta / i["theta_free"] - 1) <= 1e-10      # round-off of 64 members


def test_reference_sample_5b_as_given():
    """model_def_5b_frame.txt (ANAFLAG 3, Newmark, warping free) against the
    reference routines' path of tests/golden/plastic.npz."""
    g = np.load(__import__("os").path.join(__import__("os").path.dirname(__file__), "golden", "plastic.npz"))
    path = model_file("gpu_5b.txt", str(g["sample_5b_model"]))
    print("[warping] sample 5b", flush=True)
    with cubens.GpuModel(path) as m:
        assert m.info.bs == 7
        r = m.run(_opts())
    assert r["status"] == 0, r["msg"]
    assert np.array_equal(r["itecnt"], g["sample_5b_itecnt"])
    assert np.array_equal(r["events"], g["sample_5b_events"])
    assert relerr(r["lpf"], g["sample_5b_lpf"]) < 1e-12
    assert relerr(r["d"], g["sample_5b_d"]) < 1e-9
