"""CPU tests of the oracle's frame element, lumped masses and nonlinear Newmark
driver: golden paths from the reference's own frame and mass routines, and
closed forms that fix what the golden paths leave open."""
from __future__ import annotations

import os

import numpy as np
import pytest
import scipy.linalg as sla

import oracle as O
from cubens import inputs
from helpers import model_file

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def gold(name):
    return np.load(os.path.join(GOLD, name))


# --------------------------------------------------------------- golden paths
@pytest.mark.parametrize("name,steps", [("frame_tower_newmark.npz", 30),
                                        ("frame_girder_offsets_nr.npz", 25)])
def test_frame_paths_against_reference_routines(name, steps):
    """stiff_fr / forces_fr / mass_fr of the reference (frame.c:226-1393) drove
    these paths; the oracle's restatement reproduces them bit for bit."""
    g = gold(name)
    om = O.OracleModel(model_file(name + ".txt", str(g["model"])))
    r = om.run(steps, keep_f=True)
    assert r["rc"] == 0, r["msg"]
    assert np.array_equal(r["itecnt"], g["path_itecnt"])
    assert np.array_equal(r["lpf"], g["path_lpf"])
    assert np.array_equal(r["d"], g["path_d"])
    assert np.array_equal(r["f"], g["path_f"])


def test_frame_generator_matches_fixture():
    g = gold("frame_tower_newmark.npz")
    assert inputs.space_frame(100, kind="tower", amp=1.5e5, steps=30).text == str(g["model"])


# ------------------------------------------------------------- closed forms
E, G, A, Is, Iw, J = 2e11, 2e11 / 2.6, 3.0e-3, 4.0e-6, 2.5e-6, 8.0e-6


def _cantilever_tip(k, dofs):
    """Flexibility at end 2 with end 1 fixed (rows/cols 7..13)."""
    kk = k[np.ix_(dofs, dofs)]
    return np.linalg.inv(kk)


def test_frame_elastic_stiffness_closed_forms():
    """Cantilever tip flexibilities PL^3/3EI, PL^2/2EI, L/EI in both planes,
    axial L/EA, and the Vlasov torsion term 6GJ/(5L) + 12 E Cw/L^3 that the
    reference uses in place of GJ/L (frame.c:377-379)."""
    L, Cw = 2.5, 1e-7
    k = O.frame_ke(E, G, A, L, Is, Iw, J, Cw)
    assert np.allclose(k, k.T, rtol=0, atol=1e-9 * np.abs(k).max())
    assert np.isclose(k[7, 7], E * A / L) and np.isclose(k[0, 7], -E * A / L)
    fz = _cantilever_tip(k, [8, 12])               # v and theta_z at end 2
    assert np.isclose(fz[0, 0], L ** 3 / (3 * E * Is), rtol=1e-12)
    assert np.isclose(fz[0, 1], L ** 2 / (2 * E * Is), rtol=1e-12)
    assert np.isclose(fz[1, 1], L / (E * Is), rtol=1e-12)
    fy = _cantilever_tip(k, [9, 11])               # w and theta_y at end 2
    assert np.isclose(fy[0, 0], L ** 3 / (3 * E * Iw), rtol=1e-12)
    assert np.isclose(fy[0, 1], -L ** 2 / (2 * E * Iw), rtol=1e-12)
    assert np.isclose(k[10, 10], 6 * G * J / (5 * L) + 12 * E * Cw / L ** 3, rtol=1e-14)
    assert np.isclose(k[6, 6], 2 * G * J * L / 15 + 4 * E * Cw / L, rtol=1e-14)
    # rigid-body modes: translation and rotation about the element axis-free
    # rotation (y rotation with matching w) produce no force
    x = np.zeros(14); x[[0, 7]] = 1.0
    assert np.abs(k @ x).max() < 1e-6 * np.abs(k).max()
    x = np.zeros(14); x[[4, 11]] = 1.0; x[9] = -L       # rotate about y through end 1
    assert np.abs(k @ x).max() < 1e-6 * np.abs(k).max()


@pytest.mark.parametrize("nel", [8, 16])
def test_frame_geometric_stiffness_euler_buckling(nel):
    """Pinned-pinned column of nel elements under axial compression P: the
    smallest lambda with det(Ke + lambda Kg(P=-1)) = 0 approaches the Euler
    load pi^2 E I / L^2 in both bending planes (stiffg_fr, frame.c:410-579)."""
    Ltot = 6.0
    L = Ltot / nel
    n = 7 * (nel + 1)
    Ke = np.zeros((n, n)); Kg = np.zeros((n, n))
    f = np.zeros(14); f[7] = -1.0; f[0] = 1.0            # unit compression
    ke = O.frame_ke(E, G, A, L, Is, Iw, J, 0.0)
    kg = O.frame_kg(f, L, A, J)
    for e in range(nel):
        idx = np.r_[7 * e:7 * e + 14]
        Ke[np.ix_(idx, idx)] += ke
        Kg[np.ix_(idx, idx)] += kg
    fixed = [0, 1, 2, 3, 6] + [7 * nel + 1, 7 * nel + 2] + [7 * j + 6 for j in range(nel + 1)]
    free = np.setdiff1d(np.arange(n), fixed)
    Kf, Gf = Ke[np.ix_(free, free)], -Kg[np.ix_(free, free)]
    # generalised eigenproblem Kf x = lambda Gf x restricted to Gf's range
    w = sla.eigvals(Kf, Gf)
    w = np.real(w[np.isfinite(w) & (np.abs(np.imag(w)) < 1e-6 * np.abs(w))])
    lam = np.sort(w[w > 0])
    pe_weak = np.pi ** 2 * E * Iw / Ltot ** 2
    pe_strong = np.pi ** 2 * E * Is / Ltot ** 2
    assert abs(lam[0] / pe_weak - 1) < 2e-3
    assert np.any(np.abs(lam / pe_strong - 1) < 2e-3)


def _sdof_model(numopt, rho, amp, steps, dt, freq, L=3.0, dens=7850.0):
    lines = ["2", "5", f"{steps + 1},0", "0", "1", "2", "0,1,0,0,0", "1,2"]
    lines += [f"1,{d}" for d in range(1, 7)] + ["1,7,0"]
    # two free DOFs (axial u and transverse v at end 2, uncoupled): with a
    # single equation (lss = 1) the reference divides by K alone and ignores
    # the mass (main.c:3622-3625), which diverges under Newmark
    lines += [f"2,{d}" for d in range(3, 7)] + ["2,7,0", "0,0", "0,0"]
    lines += ["0.0,0.0,0.0", f"{L},0.0,0.0"]
    lines += [f"{E},{G},{dens},{A},{Is},{Iw},{J},0.0", "0,0,0,0,0,0,0", "0.0,0.0,1.0",
              "1e+30,1.0,1.0", "0,0,0,0,0", f"{steps},{repr(steps * dt)}"]
    t = np.arange(steps + 1) * dt
    p = amp * np.sin(2 * np.pi * freq * t)
    lines += [f"2,1,{repr(float(v))}" for v in p]
    lines += ["0,0,0", "0,0,0", "0,0,0,0,0", f"{numopt},{rho}", "1.0,1.0,1.0,1.0,1.0",
              "50,10,1000", "1e-11,1e-11,1.0"]
    return "\n".join(lines) + "\n", np.where(p == 0, 1e-30, p), E * A / L, dens * A * L / 2


@pytest.mark.parametrize("numopt,rho", [(0, 1.0), (1, 0.8), (2, 0.6), (3, 0.5)])
def test_newmark_single_dof_recurrence(numopt, rho):
    """One axial DOF (k = EA/L, lumped m = rho A L/2) under a small harmonic
    load: the converged steps follow the textbook Newmark recurrence with
    beta = (1 - am + af)^2/4, gamma = 1/2 - am + af and equilibrium at the new
    time (main.c:3330-3357, 3553-3653; solve.c:189-207, 469-533), starting from
    rest one step before t = 0 as the reference's step 0 does."""
    dt, steps, freq = 2e-4, 40, 40.0
    txt, p, k, m = _sdof_model(numopt, rho, 1.0, steps, dt, freq)
    om = O.OracleModel(model_file(f"sdof{numopt}.txt", txt))
    assert om.neq == 2
    r = om.run(steps + 1)
    assert r["rc"] == 0, r["msg"]
    am, af = {0: (0, 0), 1: ((2 * rho - 1) / (rho + 1), rho / (rho + 1)),
              2: (0, (1 - rho) / (1 + rho)), 3: ((rho - 1) / (rho + 1), 0)}[numopt]
    beta, gamma = (1 - am + af) ** 2 / 4, 0.5 - am + af
    u = v = a = 0.0
    us = []
    for pk in p:
        a_new = (pk - k * (u + dt * v + dt * dt * (0.5 - beta) * a)) / (m + k * dt * dt * beta)
        u_new = u + dt * v + dt * dt * ((0.5 - beta) * a + beta * a_new)
        v = v + dt * ((1 - gamma) * a + gamma * a_new)
        u, a = u_new, a_new
        us.append(u)
    us = np.array(us)
    assert r["nsteps"] == steps + 1
    assert np.allclose(r["lpf"], np.arange(steps + 1) * dt)
    # geometric stiffness adds N/L ~ 1e-11 relative at this load
    assert np.abs(r["d"][:, 0] - us).max() < 1e-8 * np.abs(us).max()
    assert np.abs(r["d"][:, 1]).max() < 1e-12 * np.abs(us).max()


def test_lumped_masses_closed_form():
    """Single frame: translations rho A L/2, rotations rho A L^3/24, warping 0
    (frame.c:1357-1363); the free end of the one-DOF model carries rho A L/2."""
    txt, _, _, m = _sdof_model(0, 1.0, 1.0, 4, 1e-3, 10.0)
    om = O.OracleModel(model_file("mass1.txt", txt))
    r = om.run(1)
    assert r["rc"] == 0
    # K_eff = k + a0 m with a0 = 4/dt^2: the first step from rest gives
    # u0 = p0/(k + 4m/dt^2) exactly for a linear DOF
    k = E * A / 3.0
    assert np.isclose(r["d"][0, 0], 1e-30 / (k + 4 * m / 1e-6), rtol=1e-9)
