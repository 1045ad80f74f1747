"""Shared helpers of the test-suite: workload files and comparisons."""
from __future__ import annotations

import os
import tempfile

import numpy as np

from cubens import inputs

_TMP = tempfile.mkdtemp(prefix="cubens_tests_")


def model_file(name: str, text: str) -> str:
    path = os.path.join(_TMP, name)
    with open(path, "w") as fh:
        fh.write(text)
    return path


def truss_c1() -> str:
    return model_file("c1_truss.txt", inputs.truss_dome().text)


def panel(n: int, **kw) -> str:
    tag = "_".join(f"{k}{v}" for k, v in sorted(kw.items()))
    return model_file(f"panel{n}_{tag}.txt", inputs.shell_panel(n=n, **kw).text)


def relerr(a: np.ndarray, b: np.ndarray) -> float:
    """max |a-b| / max |b|"""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    s = np.abs(b).max()
    return float(np.abs(a - b).max() / (s if s > 0 else 1.0))


def skyline_entries(ss: np.ndarray, maxa: np.ndarray):
    """(rows, cols, vals) of the upper triangle held by a skyline array."""
    rows, cols, vals = [], [], []
    n = len(maxa) - 1
    for j in range(n):
        lo, hi = maxa[j], maxa[j + 1]
        for a in range(lo, hi):
            rows.append(j - (a - lo))
            cols.append(j)
            vals.append(ss[a - 1])
    return np.array(rows), np.array(cols), np.array(vals)


def skyline_dense(ss: np.ndarray, maxa: np.ndarray) -> np.ndarray:
    n = len(maxa) - 1
    r, c, v = skyline_entries(ss, maxa)
    K = np.zeros((n, n))
    K[r, c] = v
    K[c, r] = v
    return K
