"""OPTFLAG 2 joint renumbering (model.c:649-935) in the library's model reader:
the renumbering, DOF maps and load vector equal the reference's
(tests/golden/renumber.npz), the band of the device factorization narrows,
and the arc-length, Newton-Raphson and Newmark paths of renumbered models
follow the reference's path (fixture) and the oracle's to the FP64 path's
bar (1e-9 relative; iteration counts exact)."""
import os

import numpy as np
import pytest

import cubens
import oracle as O
from helpers import model_file, relerr
from test_oracle_renumber import NAMES, with_optflag

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def gold():
    return np.load(os.path.join(GOLD, "renumber.npz"))


def _path(name, optflag=2):
    text = str(gold()[f"{name}_model"])
    return model_file(f"gpu_renum_{name}_{optflag}.txt", with_optflag(text, optflag))


@pytest.mark.parametrize("name", NAMES)
def test_renumbered_model_maps_match_the_reference(name):
    g = gold()
    print(f"[renumber] maps {name}", flush=True)
    with cubens.GpuModel(_path(name)) as m:
        assert np.array_equal(m.joint_numbers(), g[f"{name}_jnt"])
        jc, mc, mx = m.dofmaps()
        assert np.array_equal(jc, g[f"{name}_jcode"])
        assert np.array_equal(mc, g[f"{name}_mcode"])
        assert np.array_equal(mx, g[f"{name}_maxa"])
        assert np.array_equal(m.load(), g[f"{name}_q"])
        assert np.array_equal(m.get("x", m.nj * 3), g[f"{name}_x"])
        bw2 = m.info.bandwidth
    with cubens.GpuModel(_path(name, 1)) as m1:
        assert np.array_equal(m1.joint_numbers(), np.arange(m1.nj))
        bw1 = m1.info.bandwidth
    print(f"[renumber] {name}: half-bandwidth {bw1} -> {bw2}", flush=True)
    if name != "twin":
        assert bw2 < bw1


@pytest.mark.parametrize("name", NAMES)
def test_renumbered_paths_match_reference_and_oracle(name):
    g = gold()
    steps = len(g[f"{name}_path_lpf"])
    path = _path(name)
    print(f"[renumber] path {name} ({steps} steps)", flush=True)
    with cubens.GpuModel(path) as m:
        r = m.run(cubens.options(cubens.SOLVER_BAND64, max_steps=steps))
    assert r["status"] == 0, r["msg"]
    ro = O.OracleModel(path).run(steps)
    for want in ({"itecnt": g[f"{name}_path_itecnt"], "lpf": g[f"{name}_path_lpf"], "d": g[f"{name}_path_d"]}, ro):
        assert np.array_equal(r["itecnt"], want["itecnt"])
        assert relerr(r["lpf"], want["lpf"]) < 1e-9
        assert relerr(r["d"], want["d"]) < 1e-9
