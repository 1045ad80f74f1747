"""CPU oracle of the learned preconditioner (numpy).

TEST INFRASTRUCTURE ONLY: the checker the GPU network is compared against; it
is never loaded by the product library.

It restates, from the formulation (not from the GPU code), the features, the
message-passing network, the preconditioner blocks M = S G^T G S, the training
loss on the error t = M~ r~ - u~ (K~ = S K S, r~ = S r, u~ = u / S, on free
slots): objective 1 (default) the energy norm t.K~t / r~.u~, 0 |K~ t|^2/|r~|^2,
2 |t|^2/|u~|^2, averaged over the pairs, and its gradient by hand-written
reverse mode.  With ``bf16=True`` the operands of every matrix product are
rounded to bfloat16 where the GPU rounds them (products accumulate in float64
here, float32 on the GPU); with ``bf16=False`` everything is float64, which is
what the finite-difference check of the backward pass uses.
"""
from __future__ import annotations

import numpy as np

H, FIN, FOUT, NL, ALPHA = 64, 48, 48, 3, 0.1
# weights in order, each followed by its bias: Wn, We, Wm1..3, Wu1..3, Wo, Wd
DIMS = [(FIN, H), (FIN, H)] + [(3 * H, H)] * 3 + [(2 * H, H)] * 3 + [(3 * H, FOUT), (H, FOUT)]


def layout():
    out, off = [], 0
    for k, n in DIMS:
        out.append((off, k, n))
        off += k * n + n
    return out, off


def unpack(p):
    lay, _ = layout()
    return [(p[o:o + k * n].reshape(k, n).astype(np.float64), p[o + k * n:o + k * n + n].astype(np.float64))
            for o, k, n in lay]


def bf16(x):
    """Round to bfloat16 (nearest even) through float32, returned as float64."""
    f = np.ascontiguousarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


class Graph:
    """Edges = off-diagonal joint blocks of the pattern, grouped by row."""

    def __init__(self, rowptr, colidx):
        nv = len(rowptr) - 1
        er, ec, eb, ep = [], [], [], [0]
        diag = np.zeros(nv, np.int64)
        for i in range(nv):
            for k in range(rowptr[i], rowptr[i + 1]):
                if colidx[k] == i:
                    diag[i] = k
                else:
                    er.append(i); ec.append(colidx[k]); eb.append(k)
            ep.append(len(er))
        self.nv, self.ne = nv, len(er)
        self.erow, self.ecol, self.eblk = np.array(er, np.int64), np.array(ec, np.int64), np.array(eb, np.int64)
        self.eptr, self.diag = np.array(ep, np.int64), diag
        pos = {(i, j): e for e, (i, j) in enumerate(zip(er, ec))}
        self.erev = np.array([pos[(j, i)] for i, j in zip(er, ec)], np.int64)


def features(Kvals, g: Graph, eq, r16=True):
    """Scale, mask, node/edge inputs and the block-Jacobi factors C_i."""
    Kb = Kvals.reshape(-1, 6, 6)
    nv = g.nv
    free = (np.asarray(eq).reshape(nv, 6) >= 0)
    kdiag = np.abs(np.einsum("bpp->bp", Kb[g.diag]))
    scale = np.where(free & (kdiag > 0), (1.0 / np.sqrt(np.where(kdiag > 0, kdiag, 1.0))).astype(np.float32), 1.0)
    scale = scale.astype(np.float32).astype(np.float64)
    mask = free.astype(np.float64)
    A = scale[:, :, None] * Kb[g.diag] * scale[:, None, :]
    mm = mask[:, :, None] * mask[:, None, :]
    eye = np.broadcast_to(np.eye(6), A.shape)
    A = np.where(mm > 0, A, eye)
    xn = np.zeros((nv, FIN))
    xn[:, :36] = A.reshape(nv, 36).astype(np.float32)
    xn[:, 36:42] = np.where(free, (0.1 * np.log10(kdiag + 1e-300)).astype(np.float32), 0.0)
    xn[:, 42:48] = mask
    C = np.zeros((nv, 6, 6))
    for i in range(nv):
        try:
            L = np.linalg.cholesky(A[i])
            if np.min(np.diag(L)) ** 2 <= 1e-12:
                raise np.linalg.LinAlgError
            C[i] = np.linalg.inv(L)
        except np.linalg.LinAlgError:
            C[i] = np.eye(6)
    C = C.astype(np.float32).astype(np.float64)
    xe = np.zeros((g.ne, FIN))
    if g.ne:
        B = scale[g.erow][:, :, None] * Kb[g.eblk] * scale[g.ecol][:, None, :]
        me = mask[g.erow][:, :, None] * mask[g.ecol][:, None, :]
        xe[:, :36] = np.where(me > 0, B, 0.0).reshape(g.ne, 36).astype(np.float32)
        xe[:, 36:42] = mask[g.erow]
        xe[:, 42:48] = mask[g.ecol]
    rnd = bf16 if r16 else (lambda x: x)
    return {"scale": scale.reshape(-1), "mask": mask, "xn": rnd(xn), "xe": rnd(xe), "C": C, "K": Kb}


class Net:
    def __init__(self, params, g: Graph, feats, r16=True):
        self.W = unpack(np.asarray(params, np.float64))
        self.g, self.f, self.r16 = g, feats, r16

    def rb(self, x):
        return bf16(x) if self.r16 else x

    def f32(self, x):
        return x.astype(np.float32).astype(np.float64) if self.r16 else x

    def lin(self, X, i, relu):
        W, b = self.W[i]
        y = self.f32(X @ self.rb(W) + b)
        return np.maximum(y, 0.0) if relu else y

    def forward(self):
        g, c = self.g, {}
        c["h"] = [self.lin(self.f["xn"], 0, True)]
        c["e"] = [self.lin(self.f["xe"], 1, True)]
        c["m"], c["u"], c["a"] = [None], [None], [None]
        for l in range(1, NL + 1):
            hb, eb = self.rb(c["h"][-1]), self.rb(c["e"][-1])
            Xm = np.concatenate([hb[g.erow], hb[g.ecol], eb], axis=1)
            m = self.lin(Xm, 1 + l, True)
            e = self.f32(c["e"][-1] + m)
            a = np.zeros((g.nv, H))
            np.add.at(a, g.erow, m)
            a = self.f32(a)
            Xu = np.concatenate([hb, self.rb(a)], axis=1)
            u = self.lin(Xu, 4 + l, True)
            h = self.f32(c["h"][-1] + u)
            c["m"].append(m); c["u"].append(u); c["a"].append(a)
            c["e"].append(e); c["h"].append(h)
        hb, eb = self.rb(c["h"][-1]), self.rb(c["e"][-1])
        c["do"] = self.lin(np.concatenate([hb[g.erow], hb[g.ecol], eb], axis=1), 8, False)
        c["dd"] = self.lin(hb, 9, False)
        self.c = c
        mask = self.f["mask"]
        gd = self.f["C"] + ALPHA * (mask[:, :, None] * mask[:, None, :]) * c["dd"][:, :36].reshape(-1, 6, 6)
        lower = (g.erow > g.ecol).astype(np.float64)[:, None, None]
        go = ALPHA * lower * (mask[g.erow][:, :, None] * mask[g.ecol][:, None, :]) * c["do"][:, :36].reshape(-1, 6, 6)
        self.gd = self.f32(gd)
        self.go = self.f32(go)
        return self.gd, self.go

    # y = G x ;  z = G^T y  (slot vectors reshaped [nv,6])
    def gx(self, x):
        g = self.g
        y = np.einsum("ipq,iq->ip", self.gd, x)
        low = g.erow > g.ecol
        np.add.at(y, g.erow[low], np.einsum("epq,eq->ep", self.go[low], x[g.ecol[low]]))
        return y

    def gtx(self, y):
        g = self.g
        z = np.einsum("ipq,ip->iq", self.gd, y)
        low = g.erow > g.ecol
        np.add.at(z, g.ecol[low], np.einsum("epq,ep->eq", self.go[low], y[g.erow[low]]))
        return z

    def kt(self, v):
        """mask S K S v on [nv,6] slot vectors (fixed slots give 0)."""
        g, Kb = self.g, self.f["K"]
        s = self.f["scale"].reshape(-1, 6)
        w = s * v
        out = np.einsum("ipq,iq->ip", Kb[g.diag], w)
        if g.ne:
            np.add.at(out, g.erow, np.einsum("epq,eq->ep", Kb[g.eblk], w[g.ecol]))
        return np.where(self.f["mask"] > 0, s * out, 0.0)

    def apply(self, v):
        s = self.f["scale"].reshape(-1, 6)
        return (s * self.gtx(self.gx(s * v.reshape(-1, 6)))).reshape(-1)

    def err_loss(self, x, uu, B, objective):
        """Loss term of one pair and dL/dt for t = M~ x - u~."""
        mask = self.f["mask"]
        z = self.gtx(self.gx(x))
        t = np.where(mask > 0, z - uu, 0.0)
        if objective == 0:
            e = self.kt(t)
            den = np.sum(np.where(mask > 0, x * x, 0.0))
            den = den if den > 0 else 1.0
            return np.sum(e * e) / den / B, self.kt(2.0 / (B * den) * e)
        if objective == 1:
            w = self.kt(t)
            den = np.sum(np.where(mask > 0, x * uu, 0.0))
            den = den if den > 0 else 1.0
            return np.sum(t * w) / den / B, 2.0 / (B * den) * w
        den = np.sum(np.where(mask > 0, uu * uu, 0.0))
        den = den if den > 0 else 1.0
        return np.sum(t * t) / den / B, 2.0 / (B * den) * t

    def loss_grad(self, R, U, objective=1):
        """R, U: [B, nslot] with K u = r; returns (loss, flat gradient)."""
        g, mask = self.g, self.f["mask"]
        s = self.f["scale"].reshape(-1, 6)
        B = R.shape[0]
        dgd = np.zeros_like(self.gd)
        dgo = np.zeros_like(self.go)
        loss = 0.0
        for b in range(B):
            x = s * R[b].reshape(-1, 6)
            uu = U[b].reshape(-1, 6) / s
            y = self.gx(x)
            lb, gz = self.err_loss(x, uu, B, objective)
            loss += lb
            gy = self.gx(gz)
            dgd += np.einsum("ip,iq->ipq", y, gz) + np.einsum("ip,iq->ipq", gy, x)
            low = g.erow > g.ecol
            dgo[low] += (np.einsum("ep,eq->epq", y[g.erow[low]], gz[g.ecol[low]])
                         + np.einsum("ep,eq->epq", gy[g.erow[low]], x[g.ecol[low]]))
        dgd, dgo = self.f32(dgd), self.f32(dgo)
        mm = mask[:, :, None] * mask[:, None, :]
        ddd = np.zeros((g.nv, FOUT))
        ddd[:, :36] = (ALPHA * mm * dgd).reshape(g.nv, 36)
        ddo = np.zeros((g.ne, FOUT))
        lower = (g.erow > g.ecol).astype(np.float64)[:, None, None]
        me = mask[g.erow][:, :, None] * mask[g.ecol][:, None, :]
        ddo[:, :36] = (ALPHA * lower * me * dgo).reshape(g.ne, 36)
        ddd, ddo = self.f32(ddd), self.f32(ddo)
        return loss, self.backward(ddd, ddo)

    def backward(self, ddd, ddo):
        g, c = self.g, self.c
        grads = [None] * len(self.W)

        def wgrad(i, X, D):
            grads[i] = (self.f32(self.rb(X).T @ self.rb(D)), self.f32(D.sum(axis=0)))

        def rowsum(X):
            out = np.zeros((g.nv, X.shape[1]))
            np.add.at(out, g.erow, X)
            return out

        def colsum(X):
            out = np.zeros((g.nv, X.shape[1]))
            np.add.at(out, g.ecol, X)
            return out

        hb, eb = self.rb(c["h"][NL]), self.rb(c["e"][NL])
        wgrad(8, np.concatenate([hb[g.erow], hb[g.ecol], eb], axis=1), ddo)
        wgrad(9, hb, ddd)
        Wo, Wd = self.rb(self.W[8][0]), self.rb(self.W[9][0])
        dh = self.f32(self.rb(ddd) @ Wd.T)
        dxe = self.f32(self.rb(ddo) @ Wo.T)
        dh = dh + rowsum(dxe[:, :H]) + colsum(dxe[:, H:2 * H])
        de = dxe[:, 2 * H:]
        for l in range(NL, 0, -1):
            du = np.where(c["u"][l] > 0, dh, 0.0)
            hbp = self.rb(c["h"][l - 1])
            wgrad(4 + l, np.concatenate([hbp, self.rb(c["a"][l])], axis=1), du)
            dxn = self.f32(self.rb(du) @ self.rb(self.W[4 + l][0]).T)
            dh = dh + dxn[:, :H]
            da = dxn[:, H:]
            dq = np.where(c["m"][l] > 0, de + da[g.erow], 0.0)
            ebp = self.rb(c["e"][l - 1])
            wgrad(1 + l, np.concatenate([hbp[g.erow], hbp[g.ecol], ebp], axis=1), dq)
            dxm = self.f32(self.rb(dq) @ self.rb(self.W[1 + l][0]).T)
            dh = dh + rowsum(dxm[:, :H]) + colsum(dxm[:, H:2 * H])
            de = de + dxm[:, 2 * H:]
        dh0 = np.where(c["h"][0] > 0, dh, 0.0)
        wgrad(0, self.f["xn"], dh0)
        de0 = np.where(c["e"][0] > 0, de, 0.0)
        wgrad(1, self.f["xe"], de0)
        return np.concatenate([np.concatenate([gw.reshape(-1), gb]) for gw, gb in grads])


def loss_only(params, g, feats, R, U, objective=1):
    net = Net(params, g, feats, r16=False)
    net.forward()
    s = feats["scale"].reshape(-1, 6)
    tot = 0.0
    for b in range(R.shape[0]):
        tot += net.err_loss(s * R[b].reshape(-1, 6), U[b].reshape(-1, 6) / s, R.shape[0], objective)[0]
    return tot
