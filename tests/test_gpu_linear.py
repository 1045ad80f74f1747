"""The linear analyses on the GPU through the C ABI: linear static (ANAFLAG 1
under ALGFLAG 1/2, main.c:1714-1798) and linear Newmark (ALGFLAG 4,
main.c:3143-3303, solve.c:139-468), against the reference routines' paths
(tests/golden/linear.npz, among them the reference's sample input
model_def_5c_shell.txt) and the oracle at the FP64 bar (1e-9 relative, rows
and iteration counts exact); the static element forces and the reactions
balancing the load; and the linear Newmark checkpoint (a full-precision
restart continues the run bit for bit, a restart from the oracle's
six-digit file follows the oracle's)."""
import os
import shutil

import numpy as np
import pytest

import cubens
import oracle as O
from cubens import inputs
from helpers import model_file, relerr
from test_oracle_linear import CASES, gold

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", CASES)
def test_linear_paths_match_reference_routines_and_oracle(name):
    g = gold()
    path = model_file(f"gpu_lin_{name}.txt", str(g[f"{name}_model"]))
    print(f"[linear] {name}", flush=True)
    with cubens.GpuModel(path) as m:
        r = m.run(cubens.options(cubens.SOLVER_BAND64, max_steps=200))
        ef = m.get("ef", len(g[f"{name}_ef"][-1]))
    assert r["status"] == 0, r["msg"]
    ro = O.OracleModel(path).run(200)
    for want in ({"itecnt": g[f"{name}_itecnt"], "lpf": g[f"{name}_lpf"], "d": g[f"{name}_d"]}, ro):
        assert np.array_equal(r["itecnt"], want["itecnt"])
        assert np.array_equal(r["lpf"], want["lpf"])
        assert relerr(r["d"], want["d"]) < 1e-9
    if name.endswith("_static"):
        assert relerr(ef, g[f"{name}_ef"][-1]) < 1e-9


@pytest.mark.parametrize("name", ["star_static", "cant_static", "panel_static"])
def test_linear_static_reactions_balance_the_load(name):
    g = gold()
    path = model_file(f"gpu_lin_{name}.txt", str(g[f"{name}_model"]))
    with cubens.GpuModel(path) as m:
        r = m.run(cubens.options(cubens.SOLVER_BAND64))
        rx = m.reactions().reshape(-1, 7)
        q = m.load()
        jc, _, _ = m.dofmaps()
    lam = r["lpf"][-1]
    jc = jc.reshape(-1, 7)
    applied = np.zeros(3)
    for k in range(3):
        eq = jc[:, k][jc[:, k] > 0] - 1
        applied[k] = lam * q[eq].sum()
    scale = np.abs(applied).max()
    assert np.allclose(rx[:, :3].sum(axis=0) + applied, 0, atol=1e-9 * scale)


def test_linear_newmark_checkpoint_on_the_gpu(tmp_path):
    kw = dict(analysis="newmark", load=-500.0, steps=25, dt=1e-3, numopt=1, spectrds=0.7, u0=1e-3, chkpt=5)
    base = model_file("gpu_lin_ck.txt", inputs.cantilever_frame(**kw).text)
    rst = model_file("gpu_lin_ck_r.txt", inputs.cantilever_frame(**kw, rflag=1).text)

    def run(p, steps, ck=None, exact=0):
        with cubens.GpuModel(p) as m:
            return m.run(cubens.options(cubens.SOLVER_BAND64, max_steps=steps, checkpoint_path=ck,
                                        checkpoint_exact=exact))
    print("[linear] checkpoint", flush=True)
    full = run(base, 40)
    ck = str(tmp_path / "exact.txt")
    assert run(base, 13, ck, 1)["nsteps"] == 13
    r = run(rst, 40, ck)
    assert r["status"] == 0, r["msg"]
    assert np.array_equal(r["lpf"], full["lpf"][11:]) and np.array_equal(r["d"], full["d"][11:])
    e = str(tmp_path / "oracle_e.txt")
    O.OracleModel(base).run(13, ckpt_path=e)
    # a restarted run writes its own checkpoints: each restart reads a copy
    e1, e2 = str(tmp_path / "e1.txt"), str(tmp_path / "e2.txt")
    for c in (e1, e2):
        shutil.copy(e, c)
    want = O.OracleModel(rst).run(40, ckpt_path=e1)
    got = run(rst, 40, e2)
    assert np.array_equal(got["lpf"], want["lpf"]) and relerr(got["d"], want["d"]) < 1e-9
