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


def _joint_ids(nj: int, scramble: int | None) -> np.ndarray:
    """Output id (1-based) of each generator joint: the identity, or a seeded
    permutation (SplitMix64 keys) that scatters the numbering so that the
    input's own joint order gives a wide band."""
    if scramble is None:
        return np.arange(1, nj + 1)
    keys = SplitMix64(scramble).next_u64(nj)
    order = np.argsort(keys, kind="stable")
    ids = np.empty(nj, np.int64)
    ids[order] = np.arange(1, nj + 1)
    return ids


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

    def with_offsets(self, rows: dict) -> "ModelText":
        """The same frame model with member-end offsets: ``rows`` maps a member
        (1-based) to its six offsets (end 1 x, y, z, end 2 x, y, z, global
        axes), written before the zero line that ends the offsets, the first
        all-zero seven-field line of every frame model written here."""
        assert self.n_fr > 0 and self.n_tr == 0
        lines = "".join(f"{m}," + ",".join(_fmt(v) for v in o) + "\n" for m, o in sorted(rows.items()))
        end = "0,0,0,0,0,0,0\n"
        assert end in self.text
        return ModelText(self.text.replace(end, lines + end, 1), self.nj, self.n_tr, self.n_fr, self.n_sh,
                         dict(self.info, offsets=dict(rows)))

    def with_releases(self, rows: dict) -> "ModelText":
        """The same frame model with member end bending releases: ``rows`` maps
        a member (1-based) to its four flags (strong axis end 1, weak axis end
        1, strong axis end 2, weak axis end 2; 1 released), written before the
        zero line that ends the releases, the first line "0,0,0,0,0"."""
        assert self.n_fr > 0
        lines = self.text.split("\n")
        at = lines.index("0,0,0,0,0")
        add = [f"{m}," + ",".join(str(int(v)) for v in f) for m, f in sorted(rows.items())]
        return ModelText("\n".join(lines[:at] + add + lines[at:]), self.nj, self.n_tr, self.n_fr, self.n_sh,
                         dict(self.info, releases=dict(rows)))


# --------------------------------------------------------------------------
# C1: shallow space-truss dome, Newton-Raphson load control
# --------------------------------------------------------------------------
def truss_dome(n: int = 10, rise: float = 0.5, seed: int | None = None,
               load: float = -80.0, steps: int = 20, tol=(1e-9, 1e-6, 1e-12),
               algflag: int = 1, msal: dict | None = None, optflag: int = 1,
               scramble: int | None = None) -> ModelText:
    """n x n bays at 1 m, paraboloid crown rise, boundary pinned, centre load.

    The crown of a 0.5 m paraboloid over 10 m is locally flat (bar slope 0.01),
    so the loaded joint snaps through near 100 N under load control; the default
    total load of 80 N keeps the Newton-Raphson path on its stable branch while
    softening the crown 34% against the linear response.  The force and
    displacement criteria are set above their round-off floors (the bar force
    comes from L - L0, which cancels).  lpfmax is a hair above 1 so the twentieth
    step runs although 20 * 0.05 rounds above 1.

    algflag 3 runs the arc-length method instead, with ``load`` as the
    reference load and the centre joint's direction 3 as the control DOF;
    ``msal`` overrides its parameters (dk, alpha, psi_thresh, iteopt, lpfmax,
    dkimax, itemax, submax, imagmax, negmax).  ``scramble`` and ``optflag`` as
    for shell_panel.
    """
    rng = SplitMix64(workload_seed(1) if seed is None else seed)
    m = n + 1
    ii, jj = np.meshgrid(np.arange(m), np.arange(m), indexing="xy")
    x = ii.ravel().astype(float)
    y = jj.ravel().astype(float)
    c = n / 2.0
    r2max = 2 * c * c
    z = rise * (1 - ((x - c) ** 2 + (y - c) ** 2) / r2max) * (1 + rng.uniform(m * m, -0.02, 0.02))
    ids = _joint_ids(m * m, scramble)
    node = lambda i, j: int(ids[j * m + i])     # noqa: E731  joint id, i fastest
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
    w(f"2\n{algflag}\n0\n{optflag}\n{m * m}\n{len(bars)},0,0,0,0\n")
    for a, b in bars:
        w(f"{a},{b}\n")
    for j in range(m):
        for i in range(m):
            if i in (0, n) or j in (0, n):
                for d in (1, 2, 3):
                    w(f"{node(i, j)},{d}\n")
    w("0,0\n")
    at = np.empty(m * m, np.int64)
    at[ids - 1] = np.arange(m * m)
    for k in at:
        w(f"{_fmt(x[k])},{_fmt(y[k])},{_fmt(z[k])}\n")
    for a in area:
        w(f"200000000000.0,{_fmt(a)},7850.0,1e30\n")
    ctr = node(n // 2, n // 2)
    w(f"{ctr},3,{_fmt(load)}\n0,0,0\n")
    if algflag == 3:
        p = {"dk": -0.01, "alpha": 1.0, "psi_thresh": 0.1, "iteopt": 6, "lpfmax": 1e30,
             "dkimax": 1.2, "itemax": 30, "submax": 10, "imagmax": 10, "negmax": 10}
        p.update(msal or {})
        w(f"{ctr},3,{_fmt(p['dk'])}\n")
        w(f"{_fmt(p['alpha'])}\n{_fmt(p['psi_thresh'])}\n{p['iteopt']}\n")
        w(f"{_fmt(p['lpfmax'])},{_fmt(p['dkimax'])}\n")
        w(f"{p['itemax']},{p['submax']},{p['imagmax']},{p['negmax']}\n")
    else:
        dl = 1.0 / steps
        w(f"{_fmt(1.0 + 1e-9)},{_fmt(dl)},{_fmt(dl)},{_fmt(dl)},{_fmt(dl)}\n")
        w("50,10,1000\n")
    w(f"{tol[0]},{tol[1]},{tol[2]}\n")
    return ModelText(out.getvalue(), m * m, n_tr=len(bars),
                     info={"centre_joint": ctr, "neq": 3 * (n - 1) ** 2})


def _linear_tail(w, analysis: str, lpfmax: float, load_lines, steps: int, dt: float,
                 numopt: int, spectrds: float, u0_lines, chkpt: int | None, rflag: int) -> None:
    """Load section and driver parameters of the linear analyses: static
    (ANAFLAG 1 under ALGFLAG 1, main.c:1717) or linear Newmark (ALGFLAG 4,
    model.c:1492-1606, main.c:3184-3197).  load_lines are (joint, dir, value)
    triples: static loads, or constant histories of the Newmark run."""
    if analysis == "static":
        for j, d, v in load_lines:
            w(f"{j},{d},{_fmt(v)}\n")
        w("0,0,0\n")
    else:
        w(f"{steps},{_fmt(steps * dt)}\n")
        for j, d, v in load_lines:
            for _ in range(steps + 1):
                w(f"{j},{d},{_fmt(v)}\n")
        w("0,0,0\n0,0,0\n")
        for j, d, *uva in u0_lines:
            u, v, a = (list(uva) + [0.0, 0.0])[:3]
            w(f"{j},{d},{_fmt(u)},{_fmt(v)},{_fmt(a)}\n")
        w("0,0,0,0,0\n")
        w(f"{numopt},{_fmt(spectrds)}\n")
    w(f"{_fmt(lpfmax)}\n")


def _linear_head(analysis: str, chkpt: int | None, rflag: int, steps: int) -> str:
    if analysis == "static":
        return "1\n1\n"
    return f"1\n4\n{steps + 1 if chkpt is None else chkpt},{rflag}\n"


def star_truss(bars: int = 4, radius: float = 1.0, rise: float = 0.1, emod: float = 2e11,
               area: float = 1e-4, load: float = -1.0, dk: float = -0.005, alpha: float = 1.0,
               psi_thresh: float = 0.1, iteopt: int = 6, dkimax: float | None = None,
               lpfmax: float = 1e30, itemax: int = 30, submax: int = 10, imagmax: int = 10,
               negmax: int = 10, tol=(1e-12, 1e-10, 1.0), tilt: float = 0.0,
               analysis: str = "msal", steps: int = 20, dt: float = 1e-3, numopt: int = 0,
               spectrds: float = 1.0, u0: float = 0.0, dens: float = 7850.0,
               chkpt: int | None = None, rflag: int = 0, a0: float = 0.0) -> ModelText:
    """A shallow star truss under arc-length control: ``bars`` equal bars from
    pinned supports on a circle of ``radius`` (at height 0, angles 2 pi k/bars)
    to one apex at height ``rise``, loaded by ``load`` in direction 3 at the
    apex, which is also the arc-length control DOF.  The apex has its three
    translations free (NEQ = 3).

    Under a vertical load the apex stays on the axis and the load factor
    follows the closed form (truss.c:327-331 for the bar force)
        lambda(w) = bars * N(L) * H / (L * load),   H = rise + w,
        L = sqrt(radius^2 + H^2),  N = E A (e + e^2/2) L / L0,  e = (L - L0)/L0,
    which rises to a limit point, falls through zero at the flat position
    (H = 0), reaches a minimum and rises again once the apex has snapped
    through.  ``tilt`` > 0 moves the apex sideways by tilt * radius so the
    three DOFs couple (no closed form then).

    ``analysis`` "static" writes a linear elastic analysis (ANAFLAG 1,
    ALGFLAG 1) of ``load`` times ``lpfmax``; "newmark" a linear Newmark run
    (ALGFLAG 4) of ``steps`` steps of ``dt`` under the constant history
    ``load`` from the initial apex displacement ``u0`` (and acceleration
    ``a0``) in direction 3, with Newmark option ``numopt``/``spectrds`` and
    checkpoint ``chkpt``/``rflag``.
    """
    h = float(rise)
    if dkimax is None:
        dkimax = 2.6 * h
    out = io.StringIO()
    w = out.write
    nj = bars + 1
    head = "2\n3\n" if analysis == "msal" else _linear_head(analysis, chkpt, rflag, steps)
    w(f"{head}0\n1\n{nj}\n{bars},0,0,0,0\n")
    for k in range(bars):
        w(f"{k + 2},1\n")             # support -> apex (apex is joint 1)
    for k in range(bars):
        for d in (1, 2, 3):
            w(f"{k + 2},{d}\n")
    w("0,0\n")
    if analysis == "newmark":
        w("0,0\n")                    # no prescribed displacement DOFs
    w(f"{_fmt(tilt * radius)},0.0,{_fmt(h)}\n")
    for k in range(bars):
        t = 2 * math.pi * k / bars
        w(f"{_fmt(radius * math.cos(t))},{_fmt(radius * math.sin(t))},0.0\n")
    for _ in range(bars):
        w(f"{_fmt(emod)},{_fmt(area)},{_fmt(dens)},1e30\n")
    if analysis != "msal":
        _linear_tail(w, analysis, lpfmax, [(1, 3, load)], steps, dt, numopt, spectrds,
                     [(1, 3, u0, 0.0, a0)] if (u0 or a0) else [], chkpt, rflag)
        return ModelText(out.getvalue(), nj, n_tr=bars,
                         info={"bars": bars, "radius": radius, "rise": h, "emod": emod, "area": area,
                               "load": load, "neq": 3, "dens": dens})
    w(f"1,3,{_fmt(load)}\n0,0,0\n")
    w(f"1,3,{_fmt(dk)}\n")
    w(f"{_fmt(alpha)}\n{_fmt(psi_thresh)}\n{iteopt}\n")
    w(f"{_fmt(lpfmax)},{_fmt(dkimax)}\n")
    w(f"{itemax},{submax},{imagmax},{negmax}\n")
    w(f"{tol[0]},{tol[1]},{tol[2]}\n")
    return ModelText(out.getvalue(), nj, n_tr=bars,
                     info={"bars": bars, "radius": radius, "rise": h, "emod": emod, "area": area,
                           "load": load, "neq": 3})


def star_truss_lambda(w, bars: int = 4, radius: float = 1.0, rise: float = 0.1,
                      emod: float = 2e11, area: float = 1e-4, load: float = -1.0):
    """Closed-form load factor of star_truss at apex displacement w (direction 3)."""
    w = np.asarray(w, dtype=np.float64)
    L0 = math.sqrt(radius * radius + rise * rise)
    H = rise + w
    L = np.sqrt(radius * radius + H * H)
    e = (L - L0) / L0
    N = emod * area * (e + 0.5 * e * e) * L / L0
    return bars * N * H / (L * load)


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
                dkimax: float = 1e30, optflag: int = 1, scramble: int | None = None,
                edge_beam: float = 0.0, tie_area: float = 0.0, analysis: str = "msal") -> ModelText:
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

    ``scramble`` (a seed) permutes the joint numbers of the input file;
    ``optflag`` 2 asks the reader to renumber the joints (model.c:649-935).

    A mixed model: ``edge_beam`` > 0 adds frame members along both curved
    edges (circular hollow sections of that outer radius, the shell's E, warping
    fixed at their joints), ``tie_area`` > 0 adds truss bars of that area on the
    second diagonal of the quads of the middle row.

    ``analysis`` "static" writes a linear elastic analysis (ANAFLAG 1,
    ALGFLAG 1) of the centre load times ``lpfmax``.
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
    ids = _joint_ids(m * m, scramble)
    node = lambda i, j: int(ids[j * m + i])     # noqa: E731
    tris = []
    for j in range(n):
        for i in range(n):
            a, b, c, d = node(i, j), node(i + 1, j), node(i + 1, j + 1), node(i, j + 1)
            tris.append((a, b, c))
            tris.append((a, c, d))
    ne = len(tris)
    bars = [(node(i + 1, n // 2), node(i, n // 2 + 1)) for i in range(n)] if tie_area > 0 else []
    beams = ([(node(i, j), node(i + 1, j)) for j in (0, n) for i in range(n)] if edge_beam > 0 else [])
    if thick_sigma > 0:
        t_e = thick * np.exp(thick_sigma * smooth_field(rng, tris, ii, jj, n))
    else:
        t_e = np.full(ne, thick)
    out = io.StringIO()
    w = out.write
    head = "2\n3\n" if analysis == "msal" else _linear_head(analysis, None, 0, 0)
    w(f"{head}0\n{optflag}\n{m * m}\n{len(bars)},{len(beams)},{ne},0,0\n")
    for a, b in bars + beams:
        w(f"{a},{b}\n")
    for t in tris:
        w(f"{t[0]},{t[1]},{t[2]}\n")
    for j in range(m):
        for i in (0, n):
            for dd in (1, 2, 3):
                w(f"{node(i, j)},{dd}\n")
    for jt in sorted({v for b in beams for v in b}):
        w(f"{jt},7,0\n")
    w("0,0\n")
    at = np.empty(m * m, np.int64)
    at[ids - 1] = np.arange(m * m)                # generator joint of each output id
    for k in at:
        w(f"{_fmt(x[k])},{_fmt(y[k])},{_fmt(z[k])}\n")
    for _ in bars:
        w(f"{_fmt(emod)},{_fmt(tie_area)},1200.0,1e30\n")
    if beams:
        sec = chs_section(edge_beam, E=emod)
        for _ in beams:
            w(f"{_fmt(sec['E'])},{_fmt(sec['G'])},1200.0,{_fmt(sec['A'])},{_fmt(sec['I'])},"
              f"{_fmt(sec['I'])},{_fmt(sec['J'])},0.0\n")
        w("0,0,0,0,0,0,0\n")
        for a, _b in beams:
            k = at[a - 1]                          # aux point off the member, across the panel
            w(f"{_fmt(x[k])},{_fmt(y[k] + 1.0)},{_fmt(z[k])}\n")
        for _ in beams:
            w("1e+30,1.0,1.0\n")
        w("0,0,0,0,0\n")
    for t in t_e:
        w(f"{_fmt(emod)},{_fmt(nu)},{_fmt(t)},1200.0,1e30\n")
    ctr = node(n // 2, n // 2)
    w(f"{ctr},3,{_fmt(load)}\n0,0,0\n")
    if beams:
        w("0,0,0\n")                              # no distributed member loads
    neq = 6 * m * m - 3 * 2 * m
    if analysis == "static":
        w(f"{_fmt(lpfmax)}\n")
        return ModelText(out.getvalue(), m * m, n_tr=len(bars), n_fr=len(beams), n_sh=ne,
                         info={"centre_joint": ctr, "neq": neq, "n": n})
    w(f"{ctr},3,{_fmt(dk)}\n")
    w(f"{_fmt(alpha)}\n{_fmt(psi_thresh)}\n{iteopt}\n")
    w(f"{_fmt(lpfmax)},{_fmt(dkimax)}\n")
    w(f"{itemax},{submax},{imagmax},{negmax}\n")
    w(f"{tol[0]},{tol[1]},{tol[2]}\n")
    neq = 6 * m * m - 3 * 2 * m
    return ModelText(out.getvalue(), m * m, n_tr=len(bars), n_fr=len(beams), n_sh=ne,
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


# --------------------------------------------------------------------------
# C4: 3-D corotational space frames under harmonic lateral load, Newmark
# --------------------------------------------------------------------------
def chs_section(r_o: float, E: float = 2e11):
    """Circular hollow section with wall r_o/10: A, I, J = 2I, warping 0."""
    r_i = 0.9 * r_o
    A = math.pi * (r_o ** 2 - r_i ** 2)
    I = math.pi * (r_o ** 4 - r_i ** 4) / 4
    return {"E": E, "G": E / 2.6, "A": A, "I": I, "J": 2 * I, "Iw": 0.0}


def space_frame(n_elem: int = 400, kind: str = "tower", seed: int | None = None,
                instance: int = 0, steps: int = 200, dt: float = 1e-3, amp: float | None = None,
                tol=(1e-8, 1e-8, 1e-8), algflag: int = 5, numopt: int = 0,
                spectrds: float = 1.0, itemax: int = 20, slice_shape: tuple | None = None,
                optflag: int = 1, scramble: int | None = None, chkpt: int | None = None,
                rflag: int = 0, warping: bool = False) -> ModelText:
    """A random tower or lattice girder of about ``n_elem`` frame members.

    tower : square storeys of 4 columns, 4 beams and 8 face X-braces; plan
            width and storey height U(1, 4) m; base joints fixed.
    grid  : a long 3-D lattice girder of ny x nz joints per slice and nx
            slices, members along the three lattice directions at spacings
            U(1, 4) m; the joints of the first slice are fixed.
    Every joint has its warping DOF fixed (restraint 0), so each member has 12
    active DOF.  Circular hollow sections r_o ~ U(0.05, 0.2) m, E 200 GPa,
    density 7850.  The top joints (tower) or the free-end slice (grid) carry a
    lateral load amp * sin(2 pi f t) with f ~ U(2, 10) Hz over ``steps`` steps
    of ``dt`` (average acceleration Newmark with numopt 0, spectral radius 1).
    Joints are numbered level by level so the skyline band stays narrow.
    slice_shape = (ny, nz) replaces the drawn girder cross-section (the draws
    are still made, so the rest of the model is the same); a wide slice gives
    a wide band.  ``scramble`` (a seed) permutes the joint numbers of the
    input file; ``optflag`` 2 asks the reader to renumber the joints
    (model.c:649-935).  The reference applies that renumbering to static loads
    but not to load histories (model.c:1524), so a scrambled Newmark model with
    optflag 2 loads other joints than the same model with optflag 1.
    ``chkpt`` and ``rflag`` are the Newmark checkpoint interval and restart flag
    (main.c:347-371); by default the interval lies beyond the last step.
    ``warping`` leaves the warping DOFs free (shared by the members at each
    joint, model.c:951-957) except at the fixed joints, and gives the sections
    a warping constant Cw = I r_o^2 / 4.
    """
    rng = SplitMix64(workload_seed(4, instance) if seed is None else seed)
    joints, members, fixed, loaded = [], [], [], []
    if kind == "tower":
        b = float(rng.uniform(1, 1.0, 4.0)[0])
        h = float(rng.uniform(1, 1.0, 4.0)[0])
        ns = max(1, int(math.ceil(n_elem / 16)))
        plan = [(0.0, 0.0), (b, 0.0), (b, b), (0.0, b)]
        for lv in range(ns + 1):
            for (px, py) in plan:
                joints.append((px, py, lv * h))
        jid = lambda lv, c: lv * 4 + c + 1      # noqa: E731
        fixed = [jid(0, c) for c in range(4)]
        loaded = [jid(ns, c) for c in range(4)]
        for lv in range(ns):
            for c in range(4):
                members.append((jid(lv, c), jid(lv + 1, c)))                 # column
            for c in range(4):
                members.append((jid(lv + 1, c), jid(lv + 1, (c + 1) % 4)))   # beam
            for c in range(4):
                members.append((jid(lv, c), jid(lv + 1, (c + 1) % 4)))       # X-brace
                members.append((jid(lv, (c + 1) % 4), jid(lv + 1, c)))
    elif kind == "grid":
        ny = 2 + int(rng.uniform(1, 0, 3)[0])
        nz = 2 + int(rng.uniform(1, 0, 3)[0])
        if slice_shape is not None:
            ny, nz = int(slice_shape[0]), int(slice_shape[1])
        per_slice = ny * nz * 3 - ny - nz          # members between and within slices
        nx = max(2, int(math.ceil(n_elem / per_slice)))
        sx, sy, sz = (float(v) for v in rng.uniform(3, 1.0, 4.0))
        jid = lambda a, b_, c: a * ny * nz + b_ * nz + c + 1   # noqa: E731
        for a in range(nx):
            for b_ in range(ny):
                for c in range(nz):
                    joints.append((a * sx, b_ * sy, c * sz))
        fixed = [jid(0, b_, c) for b_ in range(ny) for c in range(nz)]
        loaded = [jid(nx - 1, b_, c) for b_ in range(ny) for c in range(nz)]
        for a in range(nx):
            for b_ in range(ny):
                for c in range(nz):
                    if a + 1 < nx:
                        members.append((jid(a, b_, c), jid(a + 1, b_, c)))
                    if b_ + 1 < ny:
                        members.append((jid(a, b_, c), jid(a, b_ + 1, c)))
                    if c + 1 < nz:
                        members.append((jid(a, b_, c), jid(a, b_, c + 1)))
    else:
        raise ValueError(kind)
    nj, ne = len(joints), len(members)
    r_o = rng.uniform(ne, 0.05, 0.2)
    freq = float(rng.uniform(1, 2.0, 10.0)[0])
    if amp is None:
        amp = float(rng.uniform(1, 1e3, 5e3)[0])
    X = np.array(joints)
    ids = _joint_ids(nj, scramble)
    jo = lambda j: int(ids[j - 1])              # noqa: E731  output id of generator joint j
    out = io.StringIO()
    w = out.write
    w(f"2\n{algflag}\n")
    if algflag > 3:
        # default checkpoint interval beyond the last step: no checkpoint is
        # written (CHKPT 0 would divide by zero in the reference's k % CHKPT)
        w(f"{steps + 1 if chkpt is None else chkpt},{rflag}\n")
    w(f"0\n{optflag}\n{nj}\n0,{ne},0,0,0\n")
    for a, b_ in members:
        w(f"{jo(a)},{jo(b_)}\n")
    for j in fixed:
        for d in range(1, 7):
            w(f"{jo(j)},{d}\n")
    for j in (sorted(jo(f) for f in fixed) if warping else range(1, nj + 1)):
        w(f"{j},7,0\n")
    w("0,0\n")
    if algflag > 3:
        w("0,0\n")
    at = np.empty(nj, np.int64)
    at[ids - 1] = np.arange(nj)
    for k in at:
        p = X[k]
        w(f"{_fmt(p[0])},{_fmt(p[1])},{_fmt(p[2])}\n")
    secs = [chs_section(r) for r in r_o]
    for s_, r in zip(secs, r_o):
        cw = s_["I"] * r * r / 4 if warping else 0.0
        w(f"{_fmt(s_['E'])},{_fmt(s_['G'])},7850.0,{_fmt(s_['A'])},{_fmt(s_['I'])},"
          f"{_fmt(s_['I'])},{_fmt(s_['J'])},{_fmt(cw)}\n")
    w("0,0,0,0,0,0,0\n")
    for a, b_ in members:
        pa, pb = X[a - 1], X[b_ - 1]
        e = (pb - pa) / np.linalg.norm(pb - pa)
        aux = pa + (np.array([1.0, 0.0, 0.0]) if abs(e[2]) > 0.9 else np.array([0.0, 0.0, 1.0]))
        w(f"{_fmt(aux[0])},{_fmt(aux[1])},{_fmt(aux[2])}\n")
    for _ in members:
        w("1e+30,1.0,1.0\n")
    w("0,0,0,0,0\n")
    t = np.arange(steps + 1) * dt
    if algflag > 3:
        w(f"{steps},{_fmt(steps * dt)}\n")
        hist = amp * np.sin(2 * np.pi * freq * t)
        for j in loaded:
            for v in hist:
                w(f"{jo(j)},1,{_fmt(v)}\n")
        w("0,0,0\n0,0,0\n0,0,0,0,0\n")
        w(f"{numopt},{_fmt(spectrds)}\n1.0,1.0,1.0,1.0,1.0\n{itemax},10,1000\n")
        w(f"{tol[0]},{tol[1]},{tol[2]}\n")
    else:
        for j in loaded:
            w(f"{jo(j)},1,{_fmt(amp)}\n")
        w("0,0,0\n0,0,0\n")
        w(f"1.0,0.05,0.05,0.05,0.05\n{itemax},10,1000\n{tol[0]},{tol[1]},{tol[2]}\n")
    return ModelText(out.getvalue(), nj, n_fr=ne,
                     info={"kind": kind, "freq": freq, "amp": amp, "fixed": [jo(j) for j in fixed],
                           "loaded": [jo(j) for j in loaded], "neq": 6 * (nj - len(fixed))})


def cantilever_frame(n_elem: int = 4, length: float = 3.0, r_o: float = 0.1, load: float = -1000.0,
                     lpfmax: float = 1.0, analysis: str = "static", steps: int = 20, dt: float = 1e-3,
                     numopt: int = 0, spectrds: float = 1.0, u0: float = 0.0,
                     chkpt: int | None = None, rflag: int = 0) -> ModelText:
    """A straight cantilever of ``n_elem`` frame members along x, clamped at
    joint 1 (warping fixed at every joint), with ``load`` in direction 2 at the
    tip: a linear elastic analysis (static or linear Newmark, as star_truss
    writes them).  The members are Euler-Bernoulli beams, so the static tip
    deflection is load lpfmax L^3 / (3 E I) exactly for any n_elem."""
    sec = chs_section(r_o)
    nj = n_elem + 1
    out = io.StringIO()
    w = out.write
    w(f"{_linear_head(analysis, chkpt, rflag, steps)}0\n1\n{nj}\n0,{n_elem},0,0,0\n")
    for e in range(n_elem):
        w(f"{e + 1},{e + 2}\n")
    for d in range(1, 7):
        w(f"1,{d}\n")
    for j in range(1, nj + 1):
        w(f"{j},7,0\n")
    w("0,0\n")
    if analysis == "newmark":
        w("0,0\n")
    for j in range(nj):
        w(f"{_fmt(length * j / n_elem)},0.0,0.0\n")
    for _ in range(n_elem):
        w(f"{_fmt(sec['E'])},{_fmt(sec['G'])},7850.0,{_fmt(sec['A'])},{_fmt(sec['I'])},"
          f"{_fmt(sec['I'])},{_fmt(sec['J'])},0.0\n")
    w("0,0,0,0,0,0,0\n")
    for e in range(n_elem):
        w(f"{_fmt(length * e / n_elem)},0.0,1.0\n")     # aux point above the member
    for _ in range(n_elem):
        w("1e+30,1.0,1.0\n")
    w("0,0,0,0,0\n")
    if analysis == "static":
        w(f"{nj},2,{_fmt(load)}\n0,0,0\n0,0,0\n{_fmt(lpfmax)}\n")
    else:
        _linear_tail(w, analysis, lpfmax, [(nj, 2, load)], steps, dt, numopt, spectrds,
                     [(nj, 2, u0)] if u0 else [], chkpt, rflag)
    return ModelText(out.getvalue(), nj, n_fr=n_elem,
                     info={"E": sec["E"], "I": sec["I"], "A": sec["A"], "length": length,
                           "load": load, "neq": 6 * n_elem})


def clamped_beam(n_elem: int = 2, length: float = 4.0, r_o: float = 0.1, load: float = -1000.0,
                 direction: int = 2) -> ModelText:
    """A straight beam of ``n_elem`` frame members along x, clamped at both end
    joints (warping fixed everywhere), with ``load`` in ``direction`` at the
    middle joint: a linear static analysis.  The auxiliary points above the
    members make local y = global z and local z = -global y, so a load in
    direction 3 bends the members about their local z (strong) axis and one
    in direction 2 about local y (weak).  Euler-Bernoulli members and a
    circular section (equal second moments), so with end releases the
    mid-span deflection takes the textbook forms, e.g. P L^3/(192 E I)
    without releases and P L^3/(48 E I) with both supports released in the
    load's plane."""
    assert n_elem % 2 == 0 and direction in (2, 3)
    sec = chs_section(r_o)
    nj = n_elem + 1
    out = io.StringIO()
    w = out.write
    w(f"{_linear_head('static', None, 0, 1)}0\n1\n{nj}\n0,{n_elem},0,0,0\n")
    for e in range(n_elem):
        w(f"{e + 1},{e + 2}\n")
    for j in (1, nj):
        for d in range(1, 7):
            w(f"{j},{d}\n")
    for j in range(1, nj + 1):
        w(f"{j},7,0\n")
    w("0,0\n")
    for j in range(nj):
        w(f"{_fmt(length * j / n_elem)},0.0,0.0\n")
    for _ in range(n_elem):
        w(f"{_fmt(sec['E'])},{_fmt(sec['G'])},7850.0,{_fmt(sec['A'])},{_fmt(sec['I'])},"
          f"{_fmt(sec['I'])},{_fmt(sec['J'])},0.0\n")
    w("0,0,0,0,0,0,0\n")
    for e in range(n_elem):
        w(f"{_fmt(length * e / n_elem)},0.0,1.0\n")
    for _ in range(n_elem):
        w("1e+30,1.0,1.0\n")
    w("0,0,0,0,0\n")
    w(f"{n_elem // 2 + 1},{direction},{_fmt(load)}\n0,0,0\n0,0,0\n1.0\n")
    return ModelText(out.getvalue(), nj, n_fr=n_elem,
                     info={"E": sec["E"], "I": sec["I"], "length": length, "load": load,
                           "direction": direction})


# --------------------------------------------------------------------------
# Inelastic analyses (ANAFLAG 3): plastic hinges at frame member ends, bars
# capped at their squash load
# --------------------------------------------------------------------------
def chs_plastic_modulus(r_o: float) -> float:
    """Plastic section modulus of the circular hollow section of chs_section
    (wall r_o/10): Z = 4/3 (r_o^3 - r_i^3)."""
    r_i = 0.9 * r_o
    return 4.0 / 3.0 * (r_o ** 3 - r_i ** 3)


def _nonlinear_tail(w, analysis: str, load_lines, nr, msal, nm, frames: bool) -> None:
    """Loads and driver parameters of an NR/MNR (main.c:1808-1811), MSAL
    (main.c:2172-2186) or nonlinear Newmark (main.c:3312-3330) analysis.
    load_lines are (joint, dir, value) static loads, or (joint, dir, values)
    load histories under Newmark."""
    if analysis == "newmark":
        steps, dt = nm["steps"], nm["dt"]
        w(f"{steps},{_fmt(steps * dt)}\n")
        for j, d, vals in load_lines:
            for v in vals:
                w(f"{j},{d},{_fmt(v)}\n")
        w("0,0,0\n0,0,0\n0,0,0,0,0\n")
        w(f"{nm['numopt']},{_fmt(nm['spectrds'])}\n")
        w(f"1.0,{_fmt(nm['dlpf'])},{_fmt(nm['dlpf'])},{_fmt(nm['dlpfmax'])},{_fmt(nm['dlpfmin'])}\n")
        w(f"{nm['itemax']},{nm['submax']},{nm['solmin']}\n")
        t = nm["tol"]
        w(f"{t[0]},{t[1]},{t[2]}\n")
        return
    for j, d, v in load_lines:
        w(f"{j},{d},{_fmt(v)}\n")
    w("0,0,0\n")
    if frames:
        w("0,0,0\n")                 # no member distributed loads
    if analysis == "msal":
        j, d, dk = msal["ctrl"]
        w(f"{j},{d},{_fmt(dk)}\n")
        w(f"{_fmt(msal['alpha'])}\n{_fmt(msal['psi_thresh'])}\n{msal['iteopt']}\n")
        w(f"{_fmt(msal['lpfmax'])},{_fmt(msal['dkimax'])}\n")
        w(f"{msal['itemax']},{msal['submax']},{msal['imagmax']},{msal['negmax']}\n")
        t = msal["tol"]
    else:
        w(f"{_fmt(nr['lpfmax'])},{_fmt(nr['lpf'])},{_fmt(nr['dlpf'])},{_fmt(nr['dlpfmax'])},"
          f"{_fmt(nr['dlpfmin'])}\n")
        w(f"{nr['itemax']},{nr['submax']},{nr['solmin']}\n")
        t = nr["tol"]
    w(f"{t[0]},{t[1]},{t[2]}\n")


def propped_cantilever(n_elem: int = 8, length: float = 4.0, r_o: float = 0.1, fy: float = 345e6,
                       load: float = -1000.0, axial: float = 0.0, analysis: str = "nr",
                       algflag: int | None = None, lpfmax: float | None = None,
                       dlpf: float | None = None, dlpfmin: float | None = None, solmin: int = 1000,
                       itemax: int = 30, submax: int = 20, tol=(1e-10, 1e-10, 1e-10),
                       dk: float | None = None, alpha: float = 1.0, iteopt: int = 6,
                       steps: int = 20, dt: float = 2e-3, history=None, numopt: int = 0,
                       spectrds: float = 1.0, chkpt: int | None = None, rflag: int = 0) -> ModelText:
    """A propped cantilever of ``n_elem`` circular hollow members (chs_section
    with the plastic modulus chs_plastic_modulus and yield stress ``fy``)
    along x, clamped at joint 1, on a roller at the far end (translations y, z
    and the twist restrained, the axial slide free), warping fixed
    everywhere, with ``load`` in direction 2 at the middle joint and
    ``axial`` in direction 1 at the roller: an inelastic analysis (ANAFLAG 3).

    Small-displacement plastic theory for the central load P: the clamped end
    carries the largest moment, 3 P L/16, and yields first at
    P1 = 16 Mp / (3 L); a second hinge under the load makes the mechanism at
    Pc = 6 Mp / L, with Mp = Z fy.

    ``analysis`` "nr" (ALGFLAG 1, or ``algflag`` 2 for modified NR) steps the
    load factor by ``dlpf`` up to ``lpfmax``; "msal" drives the loaded joint's
    direction 2 by ``dk`` first (ALGFLAG 3); "newmark" runs ``steps`` steps of
    ``dt`` (ALGFLAG 5) under the load factor history ``history`` (default: a
    ramp to 1 over the first half, then constant) times ``load``.
    """
    assert n_elem % 2 == 0
    sec = chs_section(r_o)
    Z = chs_plastic_modulus(r_o)
    Mp = Z * fy
    P1 = 16 * Mp / (3 * length)
    Pc = 6 * Mp / length
    nj = n_elem + 1
    mid = n_elem // 2 + 1
    if algflag is None:
        algflag = {"nr": 1, "msal": 3, "newmark": 5}[analysis]
    if lpfmax is None:
        lpfmax = 0.95 * Pc / abs(load) if analysis == "nr" else (10 * Pc / abs(load) if analysis == "msal" else 1.0)
    if dlpf is None:
        dlpf = 0.08 * Pc / abs(load) if analysis == "nr" else 0.1
    if dlpfmin is None:
        dlpfmin = dlpf / 1024 if analysis == "nr" else 1e-3
    out = io.StringIO()
    w = out.write
    w(f"3\n{algflag}\n")
    if algflag > 3:
        w(f"{steps + 1 if chkpt is None else chkpt},{rflag}\n")
    w(f"0\n1\n{nj}\n0,{n_elem},0,0,0\n")
    for e in range(n_elem):
        w(f"{e + 1},{e + 2}\n")
    for d in range(1, 7):
        w(f"1,{d}\n")
    for d in (2, 3, 4):
        w(f"{nj},{d}\n")
    for j in range(1, nj + 1):
        w(f"{j},7,0\n")
    w("0,0\n")
    if algflag > 3:
        w("0,0\n")
    for j in range(nj):
        w(f"{_fmt(length * j / n_elem)},0.0,0.0\n")
    for _ in range(n_elem):
        w(f"{_fmt(sec['E'])},{_fmt(sec['G'])},7850.0,{_fmt(sec['A'])},{_fmt(sec['I'])},"
          f"{_fmt(sec['I'])},{_fmt(sec['J'])},0.0\n")
    w("0,0,0,0,0,0,0\n")
    for e in range(n_elem):
        w(f"{_fmt(length * e / n_elem)},0.0,1.0\n")
    for _ in range(n_elem):
        w(f"{_fmt(fy)},{_fmt(Z)},{_fmt(Z)}\n")
    w("0,0,0,0,0\n")
    if analysis == "newmark":
        if history is None:
            t = np.arange(steps + 1)
            history = np.minimum(1.0, t / max(1, steps // 2))
        loads = [(mid, 2, load * np.asarray(history, float))]
        if axial:
            loads.append((nj, 1, axial * np.asarray(history, float)))
    else:
        loads = [(mid, 2, load)] + ([(nj, 1, axial)] if axial else [])
    nr = dict(lpfmax=lpfmax, lpf=dlpf, dlpf=dlpf, dlpfmax=dlpf, dlpfmin=dlpfmin, itemax=itemax,
              submax=submax, solmin=solmin, tol=tol)
    if dk is None:
        dk = -length / 400 if load < 0 else length / 400
    msal = dict(ctrl=(mid, 2, dk), alpha=alpha, psi_thresh=0.1, iteopt=iteopt, lpfmax=lpfmax,
                dkimax=length / 8, itemax=itemax, submax=submax, imagmax=10, negmax=10, tol=tol)
    nm = dict(steps=steps, dt=dt, numopt=numopt, spectrds=spectrds, dlpf=dlpf, dlpfmax=10.0,
              dlpfmin=dlpfmin, itemax=itemax, submax=submax, solmin=solmin, tol=tol)
    _nonlinear_tail(w, analysis, loads, nr, msal, nm, True)
    return ModelText(out.getvalue(), nj, n_fr=n_elem,
                     info={"E": sec["E"], "I": sec["I"], "A": sec["A"], "Z": Z, "fy": fy, "Mp": Mp,
                           "Py": sec["A"] * fy, "P1": P1, "Pc": Pc, "length": length, "load": load,
                           "mid": mid, "neq": 6 * (nj - 1) - 3})


def three_bar_truss(height: float = 1.0, theta: float = math.pi / 4, emod: float = 2e11,
                    area: float = 1e-4, fy: float = 250e6, load: float = -1000.0,
                    lpfmax: float | None = None, dlpf: float | None = None,
                    dlpfmin: float | None = None, itemax: int = 60, submax: int = 20,
                    tol=(1e-10, 1e-10, 1e-10), algflag: int = 1) -> ModelText:
    """The classic three-bar truss: a vertical bar of length ``height`` and two
    bars at +-``theta`` from it, hung from supports in the plane y = 0 and
    meeting at joint 1, which carries ``load`` in direction 3 (its y
    translation restrained); inelastic (ANAFLAG 3), NR load control.

    Small-displacement elastic-plastic theory with Py = fy A per bar: the
    vertical bar carries P / (1 + 2 cos^3 theta) and yields first at
    P1 = Py (1 + 2 cos^3 theta); the inclined bars then take the rest,
    (P - Py) / (2 cos theta) each, and the truss collapses at
    Pc = Py (1 + 2 cos theta).  The reference's bar force is capped at +-Py
    with no plastic strain (truss.c:336-347), which under a monotonic load
    follows the same path."""
    Py = fy * area
    c = math.cos(theta)
    P1 = Py * (1 + 2 * c ** 3)
    Pc = Py * (1 + 2 * c)
    if lpfmax is None:
        lpfmax = 0.97 * Pc / abs(load)
    if dlpf is None:
        dlpf = 0.1 * Pc / abs(load)
    if dlpfmin is None:
        dlpfmin = dlpf / 1024
    out = io.StringIO()
    w = out.write
    w(f"3\n{algflag}\n0\n1\n4\n3,0,0,0,0\n")
    for k in range(3):
        w(f"1,{k + 2}\n")
    w("1,2\n")
    for k in range(3):
        for d in (1, 2, 3):
            w(f"{k + 2},{d}\n")
    w("0,0\n")
    s = math.tan(theta) * height
    w(f"0.0,0.0,{_fmt(-height)}\n")
    for xk in (0.0, -s, s):
        w(f"{_fmt(xk)},0.0,0.0\n")
    for _ in range(3):
        w(f"{_fmt(emod)},{_fmt(area)},7850.0,{_fmt(fy)}\n")
    nr = dict(lpfmax=lpfmax, lpf=dlpf, dlpf=dlpf, dlpfmax=dlpf, dlpfmin=dlpfmin, itemax=itemax,
              submax=submax, solmin=1000, tol=tol)
    _nonlinear_tail(w, "nr", [(1, 3, load)], nr, None, None, False)
    return ModelText(out.getvalue(), 4, n_tr=3,
                     info={"Py": Py, "P1": P1, "Pc": Pc, "height": height, "theta": theta,
                           "emod": emod, "area": area, "load": load, "neq": 2})


def torsion_cantilever(n_elem: int = 16, length: float = 2.0, E: float = 2e11, G: float = 7.7e10,
                       J: float = 2e-7, Cw: float = 4e-9, A: float = 5e-3, I: float = 2e-5,
                       torque: float = 100.0, root_warping: str = "fixed", analysis: str = "static",
                       lpfmax: float = 1.0, dlpf: float | None = None,
                       tol=(1e-10, 1e-10, 1e-10)) -> ModelText:
    """A straight cantilever of ``n_elem`` thin-walled members along x under a
    tip torque (direction 4), clamped at joint 1 with its warping DOF
    ``root_warping`` ("fixed" or "free") and free to warp at every other
    joint (the warping DOFs shared by the members at a joint, model.c:951-957):
    Vlasov's non-uniform torsion.  With k^2 = GJ / (E Cw), the tip twist of a
    cantilever whose root cannot warp is (Vlasov, Timoshenko & Gere)
        theta(L) = T / (GJ) (L - tanh(kL) / k),
    and with the root free to warp it is Saint-Venant's T L / (GJ).
    ``analysis`` "static" is linear (ANAFLAG 1); "nr" geometrically nonlinear
    NR load control (ANAFLAG 2) to ``lpfmax`` in steps of ``dlpf``."""
    nj = n_elem + 1
    out = io.StringIO()
    w = out.write
    w("1\n1\n" if analysis == "static" else "2\n1\n")
    w(f"0\n1\n{nj}\n0,{n_elem},0,0,0\n")
    for e in range(n_elem):
        w(f"{e + 1},{e + 2}\n")
    for d in range(1, 7):
        w(f"1,{d}\n")
    if root_warping == "fixed":
        w("1,7,0\n")
    w("0,0\n")
    for j in range(nj):
        w(f"{_fmt(length * j / n_elem)},0.0,0.0\n")
    for _ in range(n_elem):
        w(f"{_fmt(E)},{_fmt(G)},7850.0,{_fmt(A)},{_fmt(I)},{_fmt(I)},{_fmt(J)},{_fmt(Cw)}\n")
    w("0,0,0,0,0,0,0\n")
    for e in range(n_elem):
        w(f"{_fmt(length * e / n_elem)},0.0,1.0\n")
    for _ in range(n_elem):
        w("1e+30,1.0,1.0\n")
    w("0,0,0,0,0\n")
    if analysis == "static":
        w(f"{nj},4,{_fmt(torque)}\n0,0,0\n0,0,0\n{_fmt(lpfmax)}\n")
    else:
        step = lpfmax / 5 if dlpf is None else dlpf
        nr = dict(lpfmax=lpfmax, lpf=step, dlpf=step, dlpfmax=step, dlpfmin=step / 64, itemax=30,
                  submax=10, solmin=1000, tol=tol)
        _nonlinear_tail(w, "nr", [(nj, 4, torque)], nr, None, None, True)
    k = math.sqrt(G * J / (E * Cw))
    return ModelText(out.getvalue(), nj, n_fr=n_elem,
                     info={"E": E, "G": G, "J": J, "Cw": Cw, "k": k, "length": length, "torque": torque,
                           "theta_fixed": torque / (G * J) * (length - math.tanh(k * length) / k),
                           "theta_free": torque * length / (G * J)})


def frame_batch(count: int = 512, lo: int = 200, hi: int = 5000, **kw) -> list[ModelText]:
    """``count`` frames of U{lo..hi} members, towers and girders alternating."""
    rng = SplitMix64(workload_seed(4, 100000))
    sizes = lo + np.floor(rng.uniform(count, 0, hi - lo + 1)).astype(int)
    return [space_frame(int(sz), kind=("tower", "grid")[i % 2], instance=i, **kw)
            for i, sz in enumerate(sizes)]


def config_inputs(index: int, **kw) -> ModelText:
    """The workloads by index: 1 truss dome, 2 panel 64^2, 3 panel 256^2,
    4 one space frame of the frame batch."""
    if index == 1:
        return truss_dome(**kw)
    if index == 2:
        return shell_panel(n=64, config_index=2, **kw)
    if index == 3:
        return shell_panel(n=256, thick_sigma=0.1, imperfection=0.1, config_index=3, **kw)
    if index == 4:
        return space_frame(**kw)
    raise ValueError(f"no generator for workload {index}")
