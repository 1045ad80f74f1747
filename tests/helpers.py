"""Shared helpers of the test-suite: workload files and comparisons."""
from __future__ import annotations

import os
import tempfile

import numpy as np

import math

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


class Hip:
    """Device buffers and streams through the HIP runtime the library itself
    links (libamdhip64.so.7 under /opt/rocm); torch bundles a runtime of its
    own, which cannot share a process with this one."""

    def __init__(self):
        import ctypes as C
        self.C = C
        self.rt = C.CDLL("/opt/rocm/lib/libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.rt.hipStreamDestroy.argtypes = [C.c_void_p]
        self.rt.hipStreamSynchronize.argtypes = [C.c_void_p]
        self.rt.hipDeviceSynchronize.argtypes = []
        self.live = []

    def _ok(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed: hipError {rc}")

    def alloc(self, nbytes: int) -> int:
        p = self.C.c_void_p()
        self._ok(self.rt.hipMalloc(self.C.byref(p), max(int(nbytes), 8)), "hipMalloc")
        self.live.append(p.value)
        return p.value

    def upload(self, a: np.ndarray) -> int:
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        self._ok(self.rt.hipMemcpy(p, a.ctypes.data, a.nbytes, 1), "hipMemcpy H2D")
        return p

    def download(self, p: int, n: int, dtype=np.float64) -> np.ndarray:
        out = np.empty(n, dtype)
        self._ok(self.rt.hipMemcpy(out.ctypes.data, p, out.nbytes, 2), "hipMemcpy D2H")
        return out

    def stream(self) -> int:
        s = self.C.c_void_p()
        self._ok(self.rt.hipStreamCreate(self.C.byref(s)), "hipStreamCreate")
        return s.value

    def sync(self, s=None):
        if s is None:
            self._ok(self.rt.hipDeviceSynchronize(), "hipDeviceSynchronize")
        else:
            self._ok(self.rt.hipStreamSynchronize(s), "hipStreamSynchronize")

    def free_all(self):
        for p in self.live:
            self.rt.hipFree(p)
        self.live = []


def arc_residuals(om_params, r, q):
    """Relative residual of the spherical arc-length constraint
    |d_s - d_{s-1}|^2 + [psi_s >= psi_thresh] (lambda_s - lambda_{s-1})^2 q.q = arc_s^2
    of each converged increment s >= 1 whose arc the recorded path determines
    (no dnorm restart within the step), with arc_s rebuilt from the path alone:
    beta = sqrt(iteopt/itecnt) alpha |d_0| / |d_{s-1} - d_{s-2}| times the
    previous increment (main.c:2625-2630, 2655-2664, 3048-3049), halved once per
    quad retry, itemax exceedance or yield restart (main.c:2961-3036).  r: a
    path with keys d, lpf, itecnt, events (EV_* slots), psi; returns
    {s: residual}."""
    alpha, iteopt, thr = om_params
    d, lam, ite, ev, psi = r["d"], r["lpf"], r["itecnt"], r["events"], r["psi"]
    EV_IMAG, EV_NEG, EV_DNORM, EV_HALVE, EV_YIELD = 1, 2, 3, 4, 5
    qq = float(q @ q)
    dn0 = np.linalg.norm(d[0])
    out = {}
    for s in range(1, r["nsteps"]):
        if ev[s, EV_DNORM]:
            continue
        dprev = d[s - 1] - (d[s - 2] if s >= 2 else 0.0)
        lprev = lam[s - 1] - (lam[s - 2] if s >= 2 else 0.0)
        beta = math.sqrt(iteopt / ite[s - 1]) * alpha * dn0 / np.linalg.norm(dprev)
        load = psi[s] >= thr
        arc = beta * math.sqrt(dprev @ dprev + (lprev * lprev * qq if load else 0.0))
        arc *= 0.5 ** int(ev[s, EV_IMAG] + ev[s, EV_NEG] + ev[s, EV_HALVE] + ev[s, EV_YIELD])
        dd = d[s] - d[s - 1]
        dl = lam[s] - lam[s - 1]
        lhs = dd @ dd + (dl * dl * qq if load else 0.0)
        out[s] = abs(lhs - arc * arc) / (arc * arc)
    return out


def parse_results(path: str):
    """A results file of output() (misc.c:345-451): (header text, rows), each
    row [lambda, itecnt, values...] from its tab-separated %e / %d fields."""
    with open(path) as fh:
        text = fh.read()
    head, rows = [], []
    for line in text.split("\n"):
        fields = [t for t in line.split("\t") if t]
        try:
            vals = [float(t) for t in fields]
            if fields and len(fields) >= 2 and line.startswith("\t"):
                rows.append(vals)
                continue
        except ValueError:
            pass
        head.append(line)
    return "\n".join(head), rows


def rows_close(a, b, rel=1.1e-5):
    """Rows of two results files equal to the 6 significant digits %e keeps:
    itecnt exactly, every value within one unit of its last printed digit
    (relative to the file's largest value for entries at round-off level)."""
    if len(a) != len(b):
        return False
    big = max((abs(v) for r in b for v in r[2:]), default=0.0)
    for ra, rb in zip(a, b):
        if len(ra) != len(rb) or ra[1] != rb[1]:
            return False
        for x, y in zip([ra[0]] + ra[2:], [rb[0]] + rb[2:]):
            if abs(x - y) > rel * max(abs(y), 1e-9 * big):
                return False
    return True
