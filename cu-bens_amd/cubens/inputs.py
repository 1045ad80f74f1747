"""Synthetic CU-BENs inputs (``model_def.txt``) for the benchmark workloads.

The text follows the positional format read by the reference
(main.c:340-427, model.c:95-264, main.c:1394-1396, truss.c:66, frame.c:62-220,
shell.c:61, model.c:1313-1606, arc.c:48, main.c:1809-1812, 2240-2245,
3320-3326).  Random fields come from SplitMix64 so every workload is
reproducible from its seed without numpy's generator versions mattering.
"""
from __future__ import annotations

import io
import math
from dataclasses import dataclass, field

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


class SplitMix64:
    """SplitMix64 stream (Steele, Lea, Flood 2014); uniform doubles in [0,1)."""

    def __init__(self, seed: int):
        self.state = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)

    def next_u64(self, n: int) -> np.ndarray:
        with np.errstate(over="ignore"):
            inc = np.uint64(0x9E3779B97F4A7C15)
            z = self.state + inc * np.arange(1, n + 1, dtype=np.uint64)
            self.state = self.state + inc * np.uint64(n)
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            return z ^ (z >> np.uint64(31))

    def uniform(self, n: int, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
        u = (self.next_u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u

    def normal(self, n: int) -> np.ndarray:
        """Box-Muller pairs from two uniform streams."""
        m = (n + 1) // 2
        u1 = 1.0 - self.uniform(m)          # (0, 1]
        u2 = self.uniform(m)
        r = np.sqrt(-2.0 * np.log(u1))
        z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
        return z[:n]


def workload_seed(config_index: int, instance: int = 0) -> int:
    return 20261202 + config_index * 1000 + instance


def _fmt(v: float) -> str:
    return repr(float(v))


@dataclass
class ModelText:
    """A model_def.txt image plus the facts tests need about it."""
    text: str
    nj: int
    n_tr: int = 0
    n_fr: int = 0
    n_sh: int = 0
    info: dict = field(default_factory=dict)

    def write(self, path: str) -> str:
        with open(path, "w") as fh:
            fh.write(self.text)
        return path


# --------------------------------------------------------------------------
# C1: shallow space-truss dome, Newton-Raphson load control
# --------------------------------------------------------------------------
def truss_dome(n: int = 10, rise: float = 0.5, seed: int | None = None,
               load: float = -80.0, steps: int = 20, tol=(1e-9, 1e-6, 1e-12),
               algflag: int = 1) -> ModelText:
    """n x n bays at 1 m, paraboloid crown rise, boundary pinned, centre load.

    The crown of a 0.5 m paraboloid over 10 m is locally flat (bar slope 0.01),
    so the loaded joint snaps through near 100 N under load control; the default
    total load of 80 N keeps the Newton-Raphson path on its stable branch while
    softening the crown 34% against the linear response.  The force and
    displacement criteria are set above their round-off floors (the bar force
    comes from L - L0, which cancels).  lpfmax is a hair above 1 so the twentieth
    step runs although 20 * 0.05 rounds above 1.
    """
    rng = SplitMix64(workload_seed(1) if seed is None else seed)
    m = n + 1
    ii, jj = np.meshgrid(np.arange(m), np.arange(m), indexing="xy")
    x = ii.ravel().astype(float)
    y = jj.ravel().astype(float)
    c = n / 2.0
    r2max = 2 * c * c
    z = rise * (1 - ((x - c) ** 2 + (y - c) ** 2) / r2max) * (1 + rng.uniform(m * m, -0.02, 0.02))
    node = lambda i, j: j * m + i + 1          # noqa: E731  joint id, i fastest
    bars = []
    for j in range(m):
        for i in range(n):
            bars.append((node(i, j), node(i + 1, j)))
    for j in range(n):
        for i in range(m):
            bars.append((node(i, j), node(i, j + 1)))
    for j in range(n):
        for i in range(n):
            bars.append((node(i, j), node(i + 1, j + 1)))
            bars.append((node(i + 1, j), node(i, j + 1)))
    area = 1e-4 * (1 + rng.uniform(len(bars), -0.05, 0.05))
    out = io.StringIO()
    w = out.write
    w(f"2\n{algflag}\n0\n1\n{m * m}\n{len(bars)},0,0,0,0\n")
    for a, b in bars:
        w(f"{a},{b}\n")
    for j in range(m):
        for i in range(m):
            if i in (0, n) or j in (0, n):
                for d in (1, 2, 3):
                    w(f"{node(i, j)},{d}\n")
    w("0,0\n")
    for k in range(m * m):
        w(f"{_fmt(x[k])},{_fmt(y[k])},{_fmt(z[k])}\n")
    for a in area:
        w(f"200000000000.0,{_fmt(a)},7850.0,1e30\n")
    ctr = node(n // 2, n // 2)
    w(f"{ctr},3,{_fmt(load)}\n0,0,0\n")
    dl = 1.0 / steps
    w(f"{_fmt(1.0 + 1e-9)},{_fmt(dl)},{_fmt(dl)},{_fmt(dl)},{_fmt(dl)}\n")
    w("50,10,1000\n")
    w(f"{tol[0]},{tol[1]},{tol[2]}\n")
    return ModelText(out.getvalue(), m * m, n_tr=len(bars),
                     info={"centre_joint": ctr, "neq": 3 * (n - 1) ** 2})


# --------------------------------------------------------------------------
# C2 / C3 / C5: hinged cylindrical shell panel, arc-length snap-through
# --------------------------------------------------------------------------
def shell_panel(n: int = 64, radius: float = 2.54, angle: float = 0.2, length: float = 0.508,
                thick: float = 0.0127, emod: float = 3.10275e9, nu: float = 0.3,
                thick_sigma: float = 0.0, imperfection: float = 0.0,
                seed: int | None = None, config_index: int = 2, dk: float = -1e-4,
                alpha: float = 1.0, psi_thresh: float = 0.1, iteopt: int = 6,
                itemax: int = 40, submax: int = 10, imagmax: int = 10, negmax: int = 10,
                tol=(1.0, 1e-3, 1.0), load: float = -1.0, lpfmax: float = 1e30,
                dkimax: float = 1e30) -> ModelText:
    """n x n quads, each split on one fixed diagonal into two CST+DKT triangles.

    Straight edges (i = 0 and i = n) are hinged (translations fixed); the curved
    edges are free; a unit inward point load sits on the centre joint and the
    arc-length control DOF is that joint's direction 3.  ``thick_sigma`` > 0
    gives per-element thickness t exp(sigma g) with g the smooth unit-variance
    field of smooth_field(); ``imperfection`` is
    the amplitude (as a fraction of t) of a sin(pi i/n) sin(pi j/n) radial
    offset.

    Tolerances (displacement, force, energy): the reference's shell internal
    force carries no drilling term while its tangent does (shell.c:482-484), so
    the modified Newton iteration of the arc-length method converges linearly
    (about 0.45 per iteration); the defaults are levels that iteration reaches.
    """
    rng = SplitMix64(workload_seed(config_index) if seed is None else seed)
    m = n + 1
    ii, jj = np.meshgrid(np.arange(m), np.arange(m), indexing="xy")
    ii = ii.ravel()
    jj = jj.ravel()
    phi = -angle / 2 + angle * ii / n
    w0 = imperfection * thick * np.sin(np.pi * ii / n) * np.sin(np.pi * jj / n)
    x = (radius + w0) * np.sin(phi)
    z = (radius + w0) * np.cos(phi) - radius * math.cos(angle / 2)
    y = length * jj / n
    node = lambda i, j: j * m + i + 1          # noqa: E731
    tris = []
    for j in range(n):
        for i in range(n):
            a, b, c, d = node(i, j), node(i + 1, j), node(i + 1, j + 1), node(i, j + 1)
            tris.append((a, b, c))
            tris.append((a, c, d))
    ne = len(tris)
    if thick_sigma > 0:
        t_e = thick * np.exp(thick_sigma * smooth_field(rng, tris, ii, jj, n))
    else:
        t_e = np.full(ne, thick)
    out = io.StringIO()
    w = out.write
    w(f"2\n3\n0\n1\n{m * m}\n0,0,{ne},0,0\n")
    for t in tris:
        w(f"{t[0]},{t[1]},{t[2]}\n")
    for j in range(m):
        for i in (0, n):
            for dd in (1, 2, 3):
                w(f"{node(i, j)},{dd}\n")
    w("0,0\n")
    for k in range(m * m):
        w(f"{_fmt(x[k])},{_fmt(y[k])},{_fmt(z[k])}\n")
    for t in t_e:
        w(f"{_fmt(emod)},{_fmt(nu)},{_fmt(t)},1200.0,1e30\n")
    ctr = node(n // 2, n // 2)
    w(f"{ctr},3,{_fmt(load)}\n0,0,0\n")
    w(f"{ctr},3,{_fmt(dk)}\n")
    w(f"{_fmt(alpha)}\n{_fmt(psi_thresh)}\n{iteopt}\n")
    w(f"{_fmt(lpfmax)},{_fmt(dkimax)}\n")
    w(f"{itemax},{submax},{imagmax},{negmax}\n")
    w(f"{tol[0]},{tol[1]},{tol[2]}\n")
    neq = 6 * m * m - 3 * 2 * m
    return ModelText(out.getvalue(), m * m, n_sh=ne,
                     info={"centre_joint": ctr, "neq": neq, "n": n})


def smooth_field(rng: SplitMix64, tris, ii, jj, n: int, modes: int = 16, kmax: int = 4) -> np.ndarray:
    """Zero-mean, unit-variance random field at the element centroids: a sum of
    ``modes`` cosines with integer wavenumbers up to ``kmax`` across the panel
    and uniform random phases, scaled by sqrt(2/modes).  Element-to-element
    independent thicknesses stall the arc-length iteration of this element
    formulation, so the thickness field is spatially correlated instead."""
    t = np.asarray(tris) - 1
    u = (ii[t].mean(axis=1)) / n
    v = (jj[t].mean(axis=1)) / n
    kx = np.floor(rng.uniform(modes, 0, kmax + 1))
    ky = np.floor(rng.uniform(modes, 0, kmax + 1))
    ph = rng.uniform(modes, 0, 2 * np.pi)
    g = np.zeros(len(t))
    for a, b, p in zip(kx, ky, ph):
        g += np.cos(2 * np.pi * (a * u + b * v) + p)
    return g * np.sqrt(2.0 / modes)


def config_inputs(index: int, **kw) -> ModelText:
    """The workloads by index: 1 truss dome, 2 panel 64^2, 3 panel 256^2."""
    if index == 1:
        return truss_dome(**kw)
    if index == 2:
        return shell_panel(n=64, config_index=2, **kw)
    if index == 3:
        return shell_panel(n=256, thick_sigma=0.1, imperfection=0.1, config_index=3, **kw)
    raise ValueError(f"no generator for workload {index}")
