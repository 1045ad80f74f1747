"""Inelastic analyses (ANAFLAG 3) on the GPU through the C ABI: plastic hinges
at frame member ends and bars capped at their squash load, under the NR/MNR,
MSAL and nonlinear Newmark drivers, against the reference routines' paths
(tests/golden/plastic.npz, among them the reference's sample input
model_def_5b_frame.txt) and the oracle; the kernels against the oracle's
stiff_fr / forces_fr at yield, plastic flow and unloading states; and the
yield flags through the Newmark checkpoint.

Bar: the FP64 bar (1e-9 relative on paths, 1e-12 on kernels), identical
iteration counts, driver events (yield and unload restarts included) and
final status, failures included."""
import os

import numpy as np
import pytest

import cubens
import oracle as O
from cubens import inputs
from helpers import model_file, relerr
from test_oracle_plastic import CASES, gold

pytestmark = pytest.mark.gpu


def _opts(steps=250, **kw):
    return cubens.options(cubens.SOLVER_BAND64, max_steps=steps, **kw)


# the reference's sample leaves the warping DOFs free (7 slots per joint); it
# runs as given in test_gpu_warping.py, and here with the warping fixed
GPU_CASES = tuple(c for c in CASES if c != "sample_5b")


def _warping_fixed(text: str) -> str:
    """A model text with "j,7,0" (warping restrained) for every joint added
    to its joint constraints."""
    lines = text.split("\n")
    at = 5 if int(lines[1]) > 3 else 4          # a checkpoint line follows ALGFLAG > 3
    nj = int(lines[at])
    k = next(i for i in range(at + 2, len(lines)) if lines[i].replace(" ", "") == "0,0")
    return "\n".join(lines[:k] + [f"{j},7,0" for j in range(1, nj + 1)] + lines[k:])


@pytest.mark.parametrize("name", GPU_CASES)
def test_plastic_paths_match_reference_routines_and_oracle(name):
    g = gold()
    path = model_file(f"gpu_pl_{name}.txt", str(g[f"{name}_model"]))
    print(f"[plastic] {name}", flush=True)
    with cubens.GpuModel(path) as m:
        r = m.run(_opts())
    o = O.OracleModel(path).run(250)
    want = int(g[f"{name}_status"])
    assert r["status"] == want == o["status"], (r["msg"], o["msg"])
    for ref in ({"itecnt": g[f"{name}_itecnt"], "lpf": g[f"{name}_lpf"], "d": g[f"{name}_d"],
                 "events": g[f"{name}_events"]}, o):
        assert np.array_equal(r["itecnt"], ref["itecnt"])
        assert np.array_equal(r["events"], ref["events"])
        assert relerr(r["lpf"], ref["lpf"]) < 1e-9
        assert relerr(r["d"], ref["d"]) < 1e-9


def test_sample_5b_with_warping_fixed_matches_the_oracle():
    g = gold()
    text = _warping_fixed(str(g["sample_5b_model"]))
    path = model_file("gpu_pl_5b_w.txt", text)
    print("[plastic] sample 5b, warping fixed", flush=True)
    with cubens.GpuModel(path) as m:
        r = m.run(_opts())
    o = O.OracleModel(path).run(250)
    assert r["status"] == o["status"] == 0, r["msg"]
    assert r["nsteps"] == o["nsteps"] == 17
    assert np.array_equal(r["itecnt"], o["itecnt"]) and np.array_equal(r["events"], o["events"])
    assert relerr(r["lpf"], o["lpf"]) < 1e-12 and relerr(r["d"], o["d"]) < 1e-9


def _cantilever_state():
    """The propped cantilever of the NR fixture: the step before the clamp
    yields (s - 1), the first yielded step (s) and the one after."""
    g = gold()
    info = inputs.propped_cantilever().info
    ef = g["pc_nr_ef"]
    e = ef.reshape(len(ef), -1, 14)
    phi = np.array([O.orbison(r[0, 0] / info["Py"], r[0, 4] / info["Mp"], r[0, 5] / info["Mp"]) for r in e])
    s = int(np.argmax(phi >= 1 - 1e-4))
    path = model_file("gpu_pl_state.txt", str(g["pc_nr_model"]))
    return path, g["pc_nr_d"], ef, s


def _element_case(path, ef_ip, dd, yld):
    om = O.OracleModel(path)
    with cubens.GpuModel(path) as m:
        for name, v in (("c1_ip", om.c1), ("c2_ip", om.c2), ("c3_ip", om.c3), ("c1_i", om.c1),
                        ("c2_i", om.c2), ("c3_i", om.c3), ("ef_ip", ef_ip), ("ef_i", np.zeros_like(ef_ip)),
                        ("defllen_ip", om.llength)):
            m.set(name, v)
        m.set_yield_flags(yld)
        Kg = m.tangent(cubens.GEN_I, cubens.GEN_IP)
        jc, _, _ = m.dofmaps()
        rp, ci = m.pattern()
        Kd = cubens.bsr_to_dense(Kg, rp, ci, m.info.bs, cubens.eq_of_slot(jc, m.nj, m.info.bs), m.neq)
        fg, code = m.internal_force(dd, cubens.GEN_I, with_code=True)
        yg, tau = m.yield_flags()
        efg = m.get("ef_i", om.nfr * 14)
    from helpers import skyline_dense
    Ko = skyline_dense(O.stiff_fr(om, om.llength, om.llength, om.c1, om.c2, om.c3, ef_ip, yld), om.maxa)
    fo, efo, co, yo, dlo = O.forces_fr(om, dd, om.llength, om.llength, (om.c1, om.c2, om.c3),
                                       (om.c1, om.c2, om.c3), ef_ip, yld=yld, dlpf=1.0)
    return (Kd, fg, code, yg, tau, efg), (Ko, fo, co, yo, dlo, efo)


def test_kernels_at_a_yield_overshoot():
    """An increment that carries the clamp through the surface: code 1 at the
    first member, the regula-falsi fraction, and the forces of the members
    before it only (none: it is member 1)."""
    path, d, ef, s = _cantilever_state()
    dd = 1.5 * (d[s] - d[s - 1])
    yld = np.zeros(16, np.int32)
    (Kd, fg, code, yg, tau, efg), (Ko, fo, co, yo, dlo, efo) = _element_case(path, ef[s - 1], dd, yld)
    print("[plastic] overshoot codes", code, co, flush=True)
    assert co == 1 and code == 1
    assert abs(tau - dlo) <= 1e-12 and 0 < tau < 1
    assert np.array_equal(yg, yo) and yg[0] == 1
    assert relerr(Kd, Ko) < 1e-12
    assert np.array_equal(fg, fo) or relerr(fg, fo) < 1e-12
    assert relerr(efg[:14], efo[:14]) < 1e-12       # the returning member's ef_i is written
    assert np.array_equal(efg[14:], np.zeros(efg.size - 14))   # the rest untouched


@pytest.mark.parametrize("sign,want", [(1.0, 0), (-1.0, 2)])
def test_kernels_at_plastic_flow_and_unloading(sign, want):
    """The clamp yielded (flag 1, end forces on the surface): loading on goes
    through the plastic tangent (code 0, every member's forces), reversing
    the increment unloads the hinge (code 2, flag 2)."""
    path, d, ef, s = _cantilever_state()
    dd = sign * (d[s + 1] - d[s])
    yld = np.zeros(16, np.int32)
    yld[0] = 1
    (Kd, fg, code, yg, tau, efg), (Ko, fo, co, yo, dlo, efo) = _element_case(path, ef[s], dd, yld)
    print("[plastic] flow/unload codes", code, co, flush=True)
    assert co == want and code == want
    assert np.array_equal(yg, yo)
    assert relerr(Kd, Ko) < 1e-12
    assert relerr(fg, fo) < 1e-12
    n = 16 * 14 // 2 if want == 0 else 14     # all members, or up to the returning one
    assert relerr(efg[:n], efo[:n]) < 1e-12


def test_newmark_checkpoint_carries_the_yield_flags(tmp_path):
    kw = dict(analysis="newmark", load=-1.62e5, steps=40, dt=1e-3, tol=(1e-8, 1e-8, 1e-8), chkpt=10)
    base = model_file("gpu_pl_ck.txt", inputs.propped_cantilever(**kw).text)
    rst = model_file("gpu_pl_ck_r.txt", inputs.propped_cantilever(**kw, rflag=1).text)
    print("[plastic] checkpoint", flush=True)
    with cubens.GpuModel(base) as m:
        full = m.run(_opts(60))
    ck = str(tmp_path / "exact.txt")
    with cubens.GpuModel(base) as m:
        part = m.run(_opts(31, checkpoint_path=ck, checkpoint_exact=1))
    assert full["status"] == 0 and part["nsteps"] == 31
    om = O.OracleModel(base)
    ock = str(tmp_path / "oracle.txt")
    om.run(31, ckpt_path=ock, ckpt_exact=True)
    mine, theirs = O.read_checkpoint(ck, om), O.read_checkpoint(ock, om)
    assert mine["tstep"] == theirs["tstep"] == 30
    assert np.array_equal(mine["yldflag"], theirs["yldflag"]) and mine["yldflag"].any()
    with cubens.GpuModel(rst) as m:
        r = m.run(_opts(60, checkpoint_path=ck))
    assert r["status"] == 0, r["msg"]
    assert np.array_equal(r["itecnt"], full["itecnt"][31:])
    assert np.array_equal(r["lpf"], full["lpf"][31:])
    assert np.array_equal(r["d"], full["d"][31:])


def test_anaflag3_shells_are_refused():
    text = "3" + inputs.shell_panel(n=4).text[1:]
    with pytest.raises(cubens.CubensError, match="ANAFLAG 3 shells"):
        cubens.GpuModel(model_file("gpu_pl_shell.txt", text))
