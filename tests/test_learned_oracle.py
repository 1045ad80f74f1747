"""CPU checks of the learned-preconditioner oracle: its hand-written reverse
mode against central finite differences (float64), and the closed form of the
untrained network (decoder at zero gives exactly block-Jacobi)."""
import numpy as np
import pytest

import learned_oracle as LO
import oracle as O
from cubens import inputs
from helpers import model_file, skyline_dense


def _system(n=2, seed=0):
    """Scaled-block test system from the oracle's shell tangent of an n x n
    panel, laid out in 6-slot joint blocks (fixed slots identity)."""
    om = O.OracleModel(model_file(f"lp{n}.txt", inputs.shell_panel(n=n, thick_sigma=0.1).text))
    K = skyline_dense(O.stiff_sh(om, om.x, om.c1, om.c2, om.c3, om.farea), om.maxa)
    nj = om.nj
    eq = np.full(nj * 6, -1)
    for j in range(nj):
        for k in range(6):
            if om.jcode[j * 7 + k]:
                eq[j * 6 + k] = om.jcode[j * 7 + k] - 1
    # joint adjacency through shells
    adj = [set([j]) for j in range(nj)]
    mc = om.minc.reshape(-1, 3) - 1
    for tri in mc:
        for a in tri:
            adj[a].update(tri)
    rowptr, colidx = [0], []
    for j in range(nj):
        colidx += sorted(adj[j])
        rowptr.append(len(colidx))
    vals = np.zeros((len(colidx), 6, 6))
    for a in range(nj):
        for k in range(rowptr[a], rowptr[a + 1]):
            b = colidx[k]
            for p in range(6):
                for q in range(6):
                    i, j = eq[a * 6 + p], eq[b * 6 + q]
                    if i >= 0 and j >= 0:
                        vals[k, p, q] = K[i, j]
                    elif a == b and p == q:
                        vals[k, p, q] = 1.0
    g = LO.Graph(np.array(rowptr), np.array(colidx))
    Ks = np.zeros((nj * 6, nj * 6))
    for a in range(nj):
        for k in range(rowptr[a], rowptr[a + 1]):
            Ks[a * 6:a * 6 + 6, colidx[k] * 6:colidx[k] * 6 + 6] = vals[k]
    return g, vals.reshape(-1), eq, Ks


def _params(seed, scale=1.0):
    _, n = LO.layout()
    rng = np.random.default_rng(seed)
    p = np.zeros(n)
    for off, k, m in LO.layout()[0]:
        p[off:off + k * m] = scale * rng.uniform(-1, 1, k * m) * np.sqrt(6.0 / (k + m))
        p[off + k * m:off + k * m + m] = 0.05 * rng.standard_normal(m)
    return p


def test_untrained_network_is_block_jacobi():
    g, vals, eq, Ks = _system()
    f = LO.features(vals, g, eq)
    p = _params(1)
    for off, k, m in LO.layout()[0][8:]:       # decoder weights and biases at zero
        p[off:off + k * m + m] = 0.0
    net = LO.Net(p, g, f)
    net.forward()
    rng = np.random.default_rng(2)
    v = rng.standard_normal(g.nv * 6)
    free = eq >= 0
    v[~free] = 0.0
    want = np.zeros_like(v)
    for i in range(g.nv):
        blk = Ks[i * 6:i * 6 + 6, i * 6:i * 6 + 6]
        want[i * 6:i * 6 + 6] = np.linalg.solve(blk, v[i * 6:i * 6 + 6])
    got = net.apply(v)
    assert np.abs(got - want)[free].max() < 1e-5 * np.abs(want).max()


@pytest.mark.parametrize("objective", [0, 1, 2])
def test_backward_matches_finite_differences(objective):
    g, vals, eq, Ks = _system()
    f = LO.features(vals, g, eq, r16=False)
    p = _params(3, scale=1.0)
    rng = np.random.default_rng(4)
    free = eq >= 0
    Kf = Ks[np.ix_(free, free)]
    R = np.zeros((2, g.nv * 6))
    U = np.zeros_like(R)
    for b in range(2):
        r = rng.standard_normal(free.sum())
        R[b, free] = r
        U[b, free] = np.linalg.solve(Kf, r)
    net = LO.Net(p, g, f, r16=False)
    net.forward()
    loss, grad = net.loss_grad(R, U, objective)
    assert np.isclose(loss, LO.loss_only(p, g, f, R, U, objective), rtol=1e-12)
    idx = rng.choice(len(p), 60, replace=False)
    # include decoder and first-layer entries explicitly
    lay = LO.layout()[0]
    idx = np.concatenate([idx, [lay[8][0] + 5, lay[9][0] + 7, lay[0][0] + 3, lay[2][0] + 11]])
    for i in idx:
        h = 1e-6 * max(1.0, abs(p[i]))
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (LO.loss_only(pp, g, f, R, U, objective) - LO.loss_only(pm, g, f, R, U, objective)) / (2 * h)
        # central differences carry ~eps * loss / h of round-off
        assert abs(fd - grad[i]) <= 1e-5 * abs(fd) + 1e-9 * max(1.0, loss), (i, fd, grad[i], loss)
