"""cb_write_results against the files the reference's own writer output()
(misc.c:345-451) produced for the same three models and steps
(tests/golden/results/): the headers byte for byte (including the tabs the
frame and shell headers leave in results3.txt), the rows to the 6
significant digits the files keep; and the averaged element forces of
cb_get_path_forces against the oracle's element forces at 1e-9."""
import os
import tempfile

import numpy as np
import pytest

import cubens
import oracle as O
from helpers import parse_results, relerr, rows_close
from test_oracle_results import averaged

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "results")


@pytest.mark.parametrize("name", ["truss", "frame", "shell"])
def test_results_files_match_the_reference_writer(name):
    d = os.path.join(GOLD, name)
    path = os.path.join(d, "model_def.txt")
    out = tempfile.mkdtemp(prefix=f"cubens_results_{name}_")
    with cubens.GpuModel(path) as g:
        gp = g.run(cubens.options(cubens.SOLVER_BAND64, max_steps=3, keep_forces=1))
        g.write_results(out)
    assert gp["nsteps"] == 3
    for k in (2, 3, 4, 5):
        want_head, want_rows = parse_results(os.path.join(d, f"results{k}.txt"))
        got_head, got_rows = parse_results(os.path.join(out, f"results{k}.txt"))
        assert got_head == want_head, k
        assert rows_close(got_rows, want_rows), k
    with open(os.path.join(d, "results3.txt")) as a, open(os.path.join(out, "results3.txt")) as b:
        if name != "truss":
            assert a.read() == b.read() == "\t"
    om = O.OracleModel(path)
    r = om.run(3, keep_ef=True)
    fa = np.array([averaged(om, r["ef"][s]) for s in range(3)])
    assert relerr(gp["forces"], fa) <= 1e-9


def test_results_need_the_kept_paths():
    path = os.path.join(GOLD, "truss", "model_def.txt")
    with cubens.GpuModel(path) as g:
        g.run(cubens.options(cubens.SOLVER_BAND64, max_steps=2))
        with pytest.raises(cubens.CubensError):
            g.write_results(tempfile.mkdtemp())
        g.run(cubens.options(cubens.SOLVER_BAND64, max_steps=2, keep_d=0, keep_forces=1))
        with pytest.raises(cubens.CubensError):
            g.write_results2(os.path.join(tempfile.mkdtemp(), "results2.txt"))
