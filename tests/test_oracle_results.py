"""The results files the reference's own writer output() (misc.c:345-451)
produced for three small models (tests/golden/results/, made by
make_results.py from oracle/_ref) against the oracle's path of the same
models: displacements (results2.txt) and the averaged element forces of
results3-5.txt -- (|N1| + |N2|)/2 per truss, (|f_j| + |f_{7+j}|)/2 for the 7
entries of a frame member, (|f_j| + |f_{6+j}| + |f_{12+j}|)/3 for the 6 of a
shell -- to the 6 significant digits the files keep."""
import os

import numpy as np
import pytest

import oracle as O
from helpers import parse_results, rows_close

GOLD = os.path.join(os.path.dirname(__file__), "golden", "results")


def averaged(om, ef):
    ntr, nfr, nsh = om.ntr, om.nfr, om.nsh
    out = []
    for e in range(ntr):
        out.append((abs(ef[2 * e]) + abs(ef[2 * e + 1])) / 2)
    for e in range(nfr):
        f = ef[2 * ntr + 14 * e: 2 * ntr + 14 * e + 14]
        out += [(abs(f[j]) + abs(f[7 + j])) / 2 for j in range(7)]
    for e in range(nsh):
        f = ef[2 * ntr + 14 * nfr + 18 * e: 2 * ntr + 14 * nfr + 18 * e + 18]
        out += [(abs(f[j]) + abs(f[6 + j]) + abs(f[12 + j])) / 3 for j in range(6)]
    return out


@pytest.mark.parametrize("name", ["truss", "frame", "shell"])
def test_oracle_path_against_the_reference_writer(name):
    d = os.path.join(GOLD, name)
    om = O.OracleModel(os.path.join(d, "model_def.txt"))
    r = om.run(3, keep_ef=True)
    assert r["nsteps"] == 3
    _, rows2 = parse_results(os.path.join(d, "results2.txt"))
    mine2 = [[r["lpf"][s], r["itecnt"][s]] + list(r["d"][s]) for s in range(3)]
    assert rows_close(mine2, rows2)
    fa = [averaged(om, r["ef"][s]) for s in range(3)]
    for k, (n, lo, hi) in enumerate([(om.ntr, 0, om.ntr), (om.nfr, om.ntr, om.ntr + 7 * om.nfr),
                                     (om.nsh, om.ntr + 7 * om.nfr, om.ntr + 7 * om.nfr + 6 * om.nsh)]):
        head, rows = parse_results(os.path.join(d, f"results{k + 3}.txt"))
        if n == 0:
            assert rows == []
            continue
        mine = [[r["lpf"][s], r["itecnt"][s]] + fa[s][lo:hi] for s in range(3)]
        assert rows_close(mine, rows), k + 3
