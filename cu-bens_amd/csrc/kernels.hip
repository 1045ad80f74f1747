// Element, geometry-update and vector kernels of the nonlinear solve path.
//
// One thread per element.  The shell tangent is never formed as an 18x18
// array: the local matrix is kept as its membrane (CST), bending (DKT) and
// geometric parts, each joint-pair 6x6 block is rotated with the element triad
// (block-diagonal T, shell.c:285-305 + misc.c:41-69) and added into the
// joint-block CSR.  Elements of one colour share no joint, so the adds of one
// launch never collide (no atomics); colours run one after another.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>

#include "dev.h"

namespace cb {
namespace {

constexpr int TPB = 256;

inline int nblk(long n, int tpb = TPB) { return int((n + tpb - 1) / tpb); }

// ---------------------------------------------------------------- shell math
// Local coordinates (x2', x3', y3') of the vertices in triad rows c1, c2
// (mem_coord, shell.c:2402-2446).
__device__ inline void local_xy(const double* p1, const double* p2, const double* p3,
                                const double* c1, const double* c2, double out[3]) {
    double e2x = p2[0] - p1[0], e2y = p2[1] - p1[1], e2z = p2[2] - p1[2];
    double e3x = p3[0] - p1[0], e3y = p3[1] - p1[1], e3z = p3[2] - p1[2];
    out[0] = c1[0] * e2x + c1[1] * e2y + c1[2] * e2z;
    out[1] = c1[0] * e3x + c1[1] * e3y + c1[2] * e3z;
    out[2] = c2[0] * e3x + c2[1] * e3y + c2[2] * e3z;
}

// CST shape-function gradients / (2A): vertex 1 at the origin, 2 at (x2,0),
// 3 at (x3,y3).
template <class R>
__device__ inline void cst_grad(const R xl[3], R twoA, R gx[3], R gy[3]) {
    gx[0] = -(xl[2] / twoA);
    gx[1] = xl[2] / twoA;
    gx[2] = R(0);
    gy[0] = (xl[1] - xl[0]) / twoA;
    gy[1] = -(xl[1] / twoA);
    gy[2] = xl[0] / twoA;
}

// R: the arithmetic of the element stiffness (double, or float in the
// FP32 precision mode)
template <class R>
struct ShellLocal {
    R km[6][6];   // membrane over (u1,v1,u2,v2,u3,v3)
    R kb[9][9];   // bending over (w,tx,ty) per vertex
};

// Plane-stress membrane stiffness t A Bm^T C Bm (shell.c:487-531).
template <class R>
__device__ inline void shell_membrane(R E, R nu, R t, const R xl[3], R A, R km[6][6]) {
    R gx[3], gy[3];
    cst_grad<R>(xl, 2 * A, gx, gy);
    R c0 = E / (1 - nu * nu);
    R C00 = c0, C01 = c0 * nu, C22 = c0 * (1 - nu) / 2;
    // columns of Bm: u_a -> (gx_a, 0, gy_a), v_a -> (0, gy_a, gx_a)
    R B[3][6];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        B[0][2 * a] = gx[a]; B[1][2 * a] = R(0); B[2][2 * a] = gy[a];
        B[0][2 * a + 1] = R(0); B[1][2 * a + 1] = gy[a]; B[2][2 * a + 1] = gx[a];
    }
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        R bc0 = B[0][i] * C00 + B[1][i] * C01;
        R bc1 = B[0][i] * C01 + B[1][i] * C00;
        R bc2 = B[2][i] * C22;
#pragma unroll
        for (int j = 0; j < 6; ++j) km[i][j] = t * A * (bc0 * B[0][j] + bc1 * B[1][j] + bc2 * B[2][j]);
    }
}

// DKT bending stiffness by exact mid-side quadrature of the Batoz-Bathe-Ho
// curvature field; squared side lengths from the 3-D side lengths sl.
template <class R>
__device__ inline void shell_bending(R E, R nu, R t, const R xl[3], R A, const R sl[3], R kb[9][9]) {
    const R x2 = xl[0], x3 = xl[1], y3 = xl[2];
    const R xs[3] = {x2 - x3, x3, -x2};
    const R ys[3] = {-y3, y3, R(0)};
    const R l2[3] = {sl[1] * sl[1], sl[2] * sl[2], sl[0] * sl[0]};
    R a[3], b[3], c[3], dd[3], e[3];
#pragma unroll
    for (int s = 0; s < 3; ++s) {
        a[s] = -xs[s] / l2[s];
        b[s] = R(0.75) * xs[s] * ys[s] / l2[s];
        c[s] = (R(0.25) * xs[s] * xs[s] - R(0.5) * ys[s] * ys[s]) / l2[s];
        dd[s] = -ys[s] / l2[s];
        e[s] = (R(0.25) * ys[s] * ys[s] - R(0.5) * xs[s] * xs[s]) / l2[s];
    }
    const R D = E * t * t * t / (12 * (1 - nu * nu));
    const R D01 = D * nu, D22 = D * (1 - nu) / 2;
    const R w = R(1.0) / (R(6.0) * 2 * A);
#pragma unroll
    for (int i = 0; i < 9; ++i)
#pragma unroll
        for (int j = 0; j < 9; ++j) kb[i][j] = R(0);
#pragma unroll
    for (int g = 0; g < 3; ++g) {
        const R xi = (g == 2) ? R(0) : R(0.5), eta = (g == 0) ? R(0) : R(0.5), L = 1 - xi - eta;
        R n1[2] = {-(4 * L - 1), -(4 * L - 1)}, n2[2] = {4 * xi - 1, R(0)}, n3[2] = {R(0), 4 * eta - 1};
        R n4[2] = {4 * eta, 4 * xi}, n5[2] = {-4 * eta, 4 * L - 4 * eta}, n6[2] = {4 * L - 4 * xi, -4 * xi};
        R Hx[9][2], Hy[9][2];
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            Hx[0][r] = R(1.5) * (a[2] * n6[r] - a[1] * n5[r]);
            Hx[1][r] = b[1] * n5[r] + b[2] * n6[r];
            Hx[2][r] = n1[r] - c[1] * n5[r] - c[2] * n6[r];
            Hx[3][r] = R(1.5) * (a[0] * n4[r] - a[2] * n6[r]);
            Hx[4][r] = b[2] * n6[r] + b[0] * n4[r];
            Hx[5][r] = n2[r] - c[2] * n6[r] - c[0] * n4[r];
            Hx[6][r] = R(1.5) * (a[1] * n5[r] - a[0] * n4[r]);
            Hx[7][r] = b[0] * n4[r] + b[1] * n5[r];
            Hx[8][r] = n3[r] - c[0] * n4[r] - c[1] * n5[r];
            Hy[0][r] = R(1.5) * (dd[2] * n6[r] - dd[1] * n5[r]);
            Hy[1][r] = -n1[r] + e[1] * n5[r] + e[2] * n6[r];
            Hy[2][r] = -Hx[1][r];
            Hy[3][r] = R(1.5) * (dd[0] * n4[r] - dd[2] * n6[r]);
            Hy[4][r] = -n2[r] + e[2] * n6[r] + e[0] * n4[r];
            Hy[5][r] = -Hx[4][r];
            Hy[6][r] = R(1.5) * (dd[1] * n5[r] - dd[0] * n4[r]);
            Hy[7][r] = -n3[r] + e[0] * n4[r] + e[1] * n5[r];
            Hy[8][r] = -Hx[7][r];
        }
        R B0[9], B1[9], B2[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            B0[i] = y3 * Hx[i][0];
            B1[i] = -x3 * Hy[i][0] + x2 * Hy[i][1];
            B2[i] = -x3 * Hx[i][0] + x2 * Hx[i][1] + y3 * Hy[i][0];
        }
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            R d0 = D * B0[i] + D01 * B1[i], d1 = D01 * B0[i] + D * B1[i], d2 = D22 * B2[i];
#pragma unroll
            for (int j = i; j < 9; ++j) kb[i][j] += w * (d0 * B0[j] + d1 * B1[j] + d2 * B2[j]);
        }
    }
#pragma unroll
    for (int i = 0; i < 9; ++i)
#pragma unroll
        for (int j = 0; j < i; ++j) kb[i][j] = kb[j][i];
}

// Entry (r,s) of the local 6x6 block for vertices (a,b): slots (u,v,w,tx,ty,tz).
template <class R>
__device__ inline R local_entry(const ShellLocal<R>& L, const R g[3][3], const R drill[3], int a, int b, int r,
                                int s) {
    R v = R(0);
    if (r < 2 && s < 2) v += L.km[2 * a + r][2 * b + s];
    if (r >= 2 && r < 5 && s >= 2 && s < 5) v += L.kb[3 * a + r - 2][3 * b + s - 2];
    if (r == 5 && s == 5 && a == b) v += drill[a];
    if (r < 3 && r == s) v += g[a][b];
    return v;
}

template <class T>
__device__ inline void add_block(T* Kb, const T Kab[6][6], bool transpose, bool upper_only) {
#pragma unroll
    for (int p = 0; p < 6; ++p)
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            T v;
            if (upper_only) v = (p <= q) ? Kab[p][q] : Kab[q][p];
            else v = transpose ? Kab[q][p] : Kab[p][q];
            Kb[p * 6 + q] += T(v);
        }
}

// stiff_sh (shell.c:110-840) for the shells of one colour, with the
// element stiffness in the arithmetic T: double, or float for the FP32
// precision mode.  The geometry stays FP64 in both: the deformed local
// coordinates, the strains of the deformed area and the membrane forces are
// formed in double and rounded once to T.
template <class T>
__global__ void __launch_bounds__(128) k_shell_tangent(
    const int32_t* __restrict__ els, int n, const int32_t* __restrict__ minc,
    const int32_t* __restrict__ blk, const double* __restrict__ x, const double* __restrict__ c1,
    const double* __restrict__ c2, const double* __restrict__ c3, const double* __restrict__ emod,
    const double* __restrict__ nu, const double* __restrict__ thick, const double* __restrict__ xlocal,
    const double* __restrict__ farea, const double* __restrict__ slength,
    const double* __restrict__ deffarea, T* __restrict__ K, long ioff, long toff, long poff,
    int geo) {
    int it = blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= n) return;
    const long e = els[it];
    const double E = emod[poff + e], v = nu[e], t = thick[e], A0 = farea[e], Aip = deffarea[e];
    const double xl[3] = {xlocal[e * 3], xlocal[e * 3 + 1], xlocal[e * 3 + 2]};
    const double sl[3] = {slength[e * 3], slength[e * 3 + 1], slength[e * 3 + 2]};
    double Rd[3][3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        Rd[0][k] = c1[toff + e * 3 + k];
        Rd[1][k] = c2[toff + e * 3 + k];
        Rd[2][k] = c3[toff + e * 3 + k];
    }
    const int32_t v0 = minc[ioff + e * 3] - 1, v1 = minc[ioff + e * 3 + 1] - 1, v2 = minc[ioff + e * 3 + 2] - 1;
    double P[3][3];
#pragma unroll
    for (int k = 0; k < 3; ++k) { P[0][k] = x[v0 * 3 + k]; P[1][k] = x[v1 * 3 + k]; P[2][k] = x[v2 * 3 + k]; }
    double xc[3];
    local_xy(P[0], P[1], P[2], Rd[0], Rd[1], xc);
    const double du2 = xc[0] - xl[0], du3 = xc[1] - xl[1], dv3 = xc[2] - xl[2];

    const T Et = T(E), vt = T(v), tt = T(t), A0t = T(A0);
    const T xlt[3] = {T(xl[0]), T(xl[1]), T(xl[2])};
    const T slt[3] = {T(sl[0]), T(sl[1]), T(sl[2])};
    ShellLocal<T> L;
    shell_membrane<T>(Et, vt, tt, xlt, A0t, L.km);
    shell_bending<T>(Et, vt, tt, xlt, A0t, slt, L.kb);
    const T drill[3] = {L.kb[1][1] / 10000, L.kb[4][4] / 10000, L.kb[7][7] / 10000};
    // membrane force of the deformed-area strain (stiffg_sh, shell.c:660-840)
    double gxd[3], gyd[3];
    cst_grad<double>(xl, 2 * Aip, gxd, gyd);
    const double c0 = E / (1 - v * v);
    const double ex = gxd[1] * du2 + gxd[2] * du3;
    const double ey = gyd[2] * dv3;
    const double gxy = gyd[1] * du2 + gyd[2] * du3 + gxd[2] * dv3;
    const T Nx = T(t * (c0 * ex + c0 * v * ey)), Ny = T(t * (c0 * v * ex + c0 * ey));
    const T Nxy = T(t * (c0 * (1 - v) / 2 * gxy));
    const T gx[3] = {T(gxd[0]), T(gxd[1]), T(gxd[2])}, gy[3] = {T(gyd[0]), T(gyd[1]), T(gyd[2])};
    const T Aipt = T(Aip);
    // linear elastic analysis (ANAFLAG 1) has no geometric term (shell.c:150-153)
    T g[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            T r0 = gx[a] * Nx + gy[a] * Nxy, r1 = gx[a] * Nxy + gy[a] * Ny;
            g[a][b] = geo ? Aipt * (r0 * gx[b] + r1 * gy[b]) : T(0);
        }
    T R[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) R[r][c] = T(Rd[r][c]);
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = a; b < 3; ++b) {
            // K_ab = blockdiag(R,R)^T k_ab blockdiag(R,R)
            T tmp[6][6], Kab[6][6];
#pragma unroll
            for (int p = 0; p < 6; ++p)
#pragma unroll
                for (int s = 0; s < 6; ++s) {
                    int bp = p / 3 * 3, cp = p % 3;
                    tmp[p][s] = R[0][cp] * local_entry<T>(L, g, drill, a, b, bp, s) +
                                R[1][cp] * local_entry<T>(L, g, drill, a, b, bp + 1, s) +
                                R[2][cp] * local_entry<T>(L, g, drill, a, b, bp + 2, s);
                }
#pragma unroll
            for (int p = 0; p < 6; ++p)
#pragma unroll
                for (int q = 0; q < 6; ++q) {
                    int bq = q / 3 * 3, cq = q % 3;
                    Kab[p][q] = tmp[p][bq] * R[0][cq] + tmp[p][bq + 1] * R[1][cq] + tmp[p][bq + 2] * R[2][cq];
                }
            if (a == b) {
                add_block<T>(K + (long)blk[e * 9 + a * 3 + a] * 36, Kab, false, true);
            } else {
                add_block<T>(K + (long)blk[e * 9 + a * 3 + b] * 36, Kab, false, false);
                add_block<T>(K + (long)blk[e * 9 + b * 3 + a] * 36, Kab, true, false);
            }
        }
}

// Internal force of the shells of one colour (forces_sh ANAFLAG 2,
// shell.c:1675-1786, 2305-2400).
__global__ void __launch_bounds__(128) k_shell_forces(
    const int32_t* __restrict__ els, int n, const int32_t* __restrict__ minc,
    const int32_t* __restrict__ mcode, const double* __restrict__ x, const double* __restrict__ dd,
    const double* __restrict__ c1p, const double* __restrict__ c2p, const double* __restrict__ c3p,
    const double* __restrict__ c1c, const double* __restrict__ c2c, const double* __restrict__ c3c,
    const double* __restrict__ emod, const double* __restrict__ nu, const double* __restrict__ thick,
    const double* __restrict__ xlocal, const double* __restrict__ farea,
    const double* __restrict__ slength, double* __restrict__ ef_ip, double* __restrict__ ef_i,
    double* __restrict__ f, long ioff, long toff, long poff, long moff, long eoff) {
    int it = blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= n) return;
    const long e = els[it];
    const double E = emod[poff + e], v = nu[e], t = thick[e], A0 = farea[e];
    const double xl[3] = {xlocal[e * 3], xlocal[e * 3 + 1], xlocal[e * 3 + 2]};
    const double sl[3] = {slength[e * 3], slength[e * 3 + 1], slength[e * 3 + 2]};
    double Rp[3][3], Rc[3][3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        Rp[0][k] = c1p[toff + e * 3 + k]; Rp[1][k] = c2p[toff + e * 3 + k]; Rp[2][k] = c3p[toff + e * 3 + k];
        Rc[0][k] = c1c[toff + e * 3 + k]; Rc[1][k] = c2c[toff + e * 3 + k]; Rc[2][k] = c3c[toff + e * 3 + k];
    }
    const int32_t v0 = minc[ioff + e * 3] - 1, v1 = minc[ioff + e * 3 + 1] - 1, v2 = minc[ioff + e * 3 + 2] - 1;
    double P[3][3];
#pragma unroll
    for (int k = 0; k < 3; ++k) { P[0][k] = x[v0 * 3 + k]; P[1][k] = x[v1 * 3 + k]; P[2][k] = x[v2 * 3 + k]; }
    double xc[3];
    local_xy(P[0], P[1], P[2], Rc[0], Rc[1], xc);
    const double du2 = xc[0] - xl[0], du3 = xc[1] - xl[1], dv3 = xc[2] - xl[2];

    const int32_t* mc = mcode + moff + e * 18;
    double DD[18];
#pragma unroll
    for (int i = 0; i < 18; ++i) {
        int32_t q = mc[i];
        DD[i] = q ? dd[q - 1] : 0.0;
    }
    double ef[18];
    double* efp = ef_ip + eoff + e * 18;
#pragma unroll
    for (int i = 0; i < 18; ++i) ef[i] = efp[i];
    double km[6][6];
    shell_membrane(E, v, t, xl, A0, km);
    double memb[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) memb[i] = km[i][2] * du2 + km[i][4] * du3 + km[i][5] * dv3;
    double kb[9][9];
    shell_bending(E, v, t, xl, A0, sl, kb);
    // bending increment (w, tx, ty) per vertex in the previous triad
    double ddb[9];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double* tr = &DD[6 * a];
        const double* rt = &DD[6 * a + 3];
        ddb[3 * a] = Rp[2][0] * tr[0] + Rp[2][1] * tr[1] + Rp[2][2] * tr[2];
        ddb[3 * a + 1] = Rp[0][0] * rt[0] + Rp[0][1] * rt[1] + Rp[0][2] * rt[2];
        ddb[3 * a + 2] = Rp[1][0] * rt[0] + Rp[1][1] * rt[1] + Rp[1][2] * rt[2];
    }
    // previous force + increment, membrane entries of the previous force dropped
    double pe[18];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        pe[6 * a + 0] = 0.0;
        pe[6 * a + 1] = 0.0;
        pe[6 * a + 5] = ef[6 * a + 5];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            double s = 0;
#pragma unroll
            for (int j = 0; j < 9; ++j) s += kb[3 * a + k][j] * ddb[j];
            pe[6 * a + 2 + k] = s + ef[6 * a + 2 + k];
        }
    }
    // rotate previous -> current triad, add membrane; Q = Rc Rp^T
    double Q[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) Q[a][b] = Rc[a][0] * Rp[b][0] + Rc[a][1] * Rp[b][1] + Rc[a][2] * Rp[b][2];
    double efc[18];
#pragma unroll
    for (int blk = 0; blk < 6; ++blk)
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            double s = Q[r][0] * pe[3 * blk] + Q[r][1] * pe[3 * blk + 1] + Q[r][2] * pe[3 * blk + 2];
            int i = 3 * blk + r;
            int vtx = i / 6, loc = i % 6;
            double m = (loc < 2) ? memb[2 * vtx + loc] : 0.0;
            efc[i] = m + s;
        }
    double* efo = ef_i + eoff + e * 18;
#pragma unroll
    for (int i = 0; i < 18; ++i) efo[i] = efc[i];
#pragma unroll
    for (int a = 0; a < 3; ++a) { efp[6 * a] = 0.0; efp[6 * a + 1] = 0.0; }
    // global force T_i^T ef_i scattered by mcode
#pragma unroll
    for (int blk = 0; blk < 6; ++blk)
#pragma unroll
        for (int col = 0; col < 3; ++col) {
            int i = 3 * blk + col;
            double s = Rc[0][col] * efc[3 * blk] + Rc[1][col] * efc[3 * blk + 1] + Rc[2][col] * efc[3 * blk + 2];
            int32_t q = mc[i];
            if (q) f[q - 1] += s;
        }
}

// ---------------------------------------------------------------- truss math
// truss.c:82-204: EA L^2/L0^3 on the axis plus N/L on the translation blocks;
// EA/L0 alone in linear elastic analysis (ANAFLAG 1, truss.c:104-106).
template <class T>
__global__ void k_truss_tangent(const int32_t* __restrict__ els, int n, const int32_t* __restrict__ blk,
                                const double* __restrict__ c1, const double* __restrict__ c2,
                                const double* __restrict__ c3, const double* __restrict__ emod,
                                const double* __restrict__ carea, const double* __restrict__ llen,
                                const double* __restrict__ defllen, const double* __restrict__ ef,
                                const double* __restrict__ yield, T* __restrict__ K, int bs, int ana) {
    int it = blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= n) return;
    const long e = els[it];
    const double L0 = llen[e], L = defllen[e];
    const bool geo = ana != 1;
    const double ka = geo ? emod[e] * carea[e] * (L * L) / (L0 * L0 * L0) : emod[e] * carea[e] / L0;
    // local 2x2 [k00 k01; k10 k11] of the bar
    double kl[2][2] = {{ka, -ka}, {-ka, ka}};
    double gN = geo ? ef[e * 2] / L : 0.0;
    if (ana == 3) {
        // stiffm_tr (truss.c:206-229) beyond the surface, (N/Py)^2 > 1 + tol:
        // k - (k g)(k g)^T / g^T k g with g = 2 (N1, N2) / Py^2; the geometric
        // term only at the squash load, (N/Py)^2 >= 1 - tol (truss.c:155-166)
        const double Py = carea[e] * yield[e], r = ef[e * 2] / Py;
        if (r * r > 1 + 1e-4) {
            const double g0 = 2 * ef[e * 2] / (Py * Py), g1 = 2 * ef[e * 2 + 1] / (Py * Py);
            const double u0 = kl[0][0] * g0 + kl[0][1] * g1, u1 = kl[1][0] * g0 + kl[1][1] * g1;
            const double gkg = u0 * g0 + u1 * g1;
            const double r00 = u0 * u0 / gkg, r01 = u0 * u1 / gkg, r11 = u1 * u1 / gkg;
            kl[0][0] -= r00; kl[0][1] -= r01; kl[1][0] -= r01; kl[1][1] -= r11;
        }
        if (!(r * r >= 1 - 1e-4)) gN = 0.0;
    }
    const double c[3] = {c1[e], c2[e], c3[e]};
    const int bb = bs * bs;
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) {
            T* Kb = K + (long)blk[e * 4 + a * 2 + b] * bb;
            const double kab = kl[a][b], sgn = (a == b) ? 1.0 : -1.0;
            for (int p = 0; p < 3; ++p)
                for (int q = 0; q < 3; ++q) {
                    double vv = kab * c[p] * c[q];
                    if (p == q) vv += sgn * gN;
                    Kb[p * bs + q] += T(vv);
                }
        }
}

// truss.c:327-376: N = EA (eps + eps^2/2) L/L0; f -= N c at end 1, += at end 2;
// under ANAFLAG 3 (yield non-null) N at or beyond the squash load,
// (N/Py)^2 >= 1 - tol, is set to +-Py with no plastic strain kept
// (truss.c:336-347).
__global__ void k_truss_forces(const int32_t* __restrict__ els, int n, const int32_t* __restrict__ mcode,
                               const double* __restrict__ c1, const double* __restrict__ c2,
                               const double* __restrict__ c3, const double* __restrict__ emod,
                               const double* __restrict__ carea, const double* __restrict__ llen,
                               const double* __restrict__ defllen, const double* __restrict__ yield,
                               double* __restrict__ ef_i, double* __restrict__ f) {
    int it = blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= n) return;
    const long e = els[it];
    const double L0 = llen[e], L = defllen[e];
    const double eps = (L - L0) / L0;
    double N = emod[e] * carea[e] * (eps + 0.5 * (eps * eps)) * L / L0;
    if (yield) {
        const double Py = yield[e] * carea[e], r = N / Py;
        if (r * r >= 1 - 1e-4) N = N < 0 ? -Py : Py;
    }
    ef_i[e * 2] = N;
    ef_i[e * 2 + 1] = -N;
    const double c[3] = {c1[e], c2[e], c3[e]};
    for (int j = 0; j < 6; ++j) {
        int32_t q = mcode[e * 6 + j];
        if (q) f[q - 1] -= ((j < 3) ? N : -N) * c[j % 3];
    }
}

// ---------------------------------------------------------------- linear forces
// Linear elastic analysis (ANAFLAG 1) from the total displacement D.
// truss.c:243-326: d = T D, ef = EA/L0 [1 -1; -1 1] d, f -= ef c at each end.
__global__ void k_truss_forces_lin(const int32_t* __restrict__ els, int n, const int32_t* __restrict__ mcode,
                                   const double* __restrict__ D, const double* __restrict__ c1,
                                   const double* __restrict__ c2, const double* __restrict__ c3,
                                   const double* __restrict__ emod, const double* __restrict__ carea,
                                   const double* __restrict__ llen, double* __restrict__ ef_i,
                                   double* __restrict__ f) {
    int it = blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= n) return;
    const long e = els[it];
    const double c[3] = {c1[e], c2[e], c3[e]};
    double De[6];
    for (int j = 0; j < 6; ++j) {
        int32_t q = mcode[e * 6 + j];
        De[j] = q ? D[q - 1] : 0.0;
    }
    const double d0 = c[0] * De[0] + c[1] * De[1] + c[2] * De[2];
    const double d1 = c[0] * De[3] + c[1] * De[4] + c[2] * De[5];
    const double ka = carea[e] * emod[e] / llen[e];
    const double e0 = ka * d0 + (-ka) * d1, e1 = (-ka) * d0 + ka * d1;
    ef_i[e * 2] = e0;
    ef_i[e * 2 + 1] = e1;
    for (int j = 0; j < 6; ++j) {
        int32_t q = mcode[e * 6 + j];
        if (q) f[q - 1] -= ((j < 3) ? e0 : e1) * c[j % 3];
    }
}

// shell.c:1695-1727 with the scatter of shell.c:2385-2397: ef = k (T_ip D) with
// the elastic matrix of stiffe_sh (membrane, bending and the drilling term),
// f += T_i^T ef.
__global__ void __launch_bounds__(128) k_shell_forces_lin(
    const int32_t* __restrict__ els, int n, const int32_t* __restrict__ mcode, const double* __restrict__ D,
    const double* __restrict__ c1p, const double* __restrict__ c2p, const double* __restrict__ c3p,
    const double* __restrict__ c1c, const double* __restrict__ c2c, const double* __restrict__ c3c,
    const double* __restrict__ emod, const double* __restrict__ nu, const double* __restrict__ thick,
    const double* __restrict__ xlocal, const double* __restrict__ farea,
    const double* __restrict__ slength, double* __restrict__ ef_i, double* __restrict__ f, long toff,
    long poff, long moff, long eoff) {
    int it = blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= n) return;
    const long e = els[it];
    const double E = emod[poff + e], v = nu[e], t = thick[e], A0 = farea[e];
    const double xl[3] = {xlocal[e * 3], xlocal[e * 3 + 1], xlocal[e * 3 + 2]};
    const double sl[3] = {slength[e * 3], slength[e * 3 + 1], slength[e * 3 + 2]};
    double Rp[3][3], Rc[3][3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        Rp[0][k] = c1p[toff + e * 3 + k]; Rp[1][k] = c2p[toff + e * 3 + k]; Rp[2][k] = c3p[toff + e * 3 + k];
        Rc[0][k] = c1c[toff + e * 3 + k]; Rc[1][k] = c2c[toff + e * 3 + k]; Rc[2][k] = c3c[toff + e * 3 + k];
    }
    const int32_t* mc = mcode + moff + e * 18;
    double dl[18];
#pragma unroll
    for (int b = 0; b < 6; ++b) {
        double Db[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            int32_t q = mc[3 * b + k];
            Db[k] = q ? D[q - 1] : 0.0;
        }
#pragma unroll
        for (int r = 0; r < 3; ++r) dl[3 * b + r] = Rp[r][0] * Db[0] + Rp[r][1] * Db[1] + Rp[r][2] * Db[2];
    }
    ShellLocal<double> L;
    shell_membrane(E, v, t, xl, A0, L.km);
    shell_bending(E, v, t, xl, A0, sl, L.kb);
    const double drill[3] = {L.kb[1][1] / 10000, L.kb[4][4] / 10000, L.kb[7][7] / 10000};
    const double g0[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    double ef[18];
    for (int i = 0; i < 18; ++i) {
        const int a = i / 6, r = i % 6;
        double s = 0;
        for (int j = 0; j < 18; ++j) s += local_entry(L, g0, drill, a, j / 6, r, j % 6) * dl[j];
        ef[i] = s;
    }
    double* efo = ef_i + eoff + e * 18;
#pragma unroll
    for (int i = 0; i < 18; ++i) efo[i] = ef[i];
#pragma unroll
    for (int b = 0; b < 6; ++b)
#pragma unroll
        for (int col = 0; col < 3; ++col) {
            int i = 3 * b + col;
            double s = Rc[0][col] * ef[3 * b] + Rc[1][col] * ef[3 * b + 1] + Rc[2][col] * ef[3 * b + 2];
            int32_t q = mc[i];
            if (q) f[q - 1] += s;
        }
}

// ---------------------------------------------------------------- updatc
__global__ void k_update_joints(const int32_t* __restrict__ jcode, const double* __restrict__ dd,
                                 double* __restrict__ x, double* __restrict__ xip, long nj) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= nj * 3) return;
    long j = i / 3;
    int k = int(i % 3);
    int32_t q = jcode[j * 7 + k];
    double old = x[i];
    xip[i] = old;
    if (q) x[i] = old + dd[q - 1];
}

__global__ void k_update_truss(const int32_t* __restrict__ minc, const double* __restrict__ x,
                               double* __restrict__ defllen, double* __restrict__ c1,
                               double* __restrict__ c2, double* __restrict__ c3, long n) {
    long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (e >= n) return;
    const double* a = x + (minc[e * 2] - 1) * 3;
    const double* b = x + (minc[e * 2 + 1] - 1) * 3;
    double el0 = b[0] - a[0], el1 = b[1] - a[1], el2 = b[2] - a[2];
    double L = sqrt(el0 * el0 + el1 * el1 + el2 * el2);
    defllen[e] = L;
    c1[e] = el0 / L;
    c2[e] = el1 / L;
    c3[e] = el2 / L;
}

// misc.c:145-184: side lengths, area and triad of each shell.
__global__ void k_update_shell(const int32_t* __restrict__ minc, const double* __restrict__ x,
                               double* __restrict__ defslen, double* __restrict__ deffarea,
                               double* __restrict__ c1, double* __restrict__ c2,
                               double* __restrict__ c3, long n, long ioff, long toff) {
    long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (e >= n) return;
    const double* p1 = x + (minc[ioff + e * 3] - 1) * 3;
    const double* p2 = x + (minc[ioff + e * 3 + 1] - 1) * 3;
    const double* p3 = x + (minc[ioff + e * 3 + 2] - 1) * 3;
    double e12[3], e23[3], e31[3];
    for (int k = 0; k < 3; ++k) {
        e23[k] = p3[k] - p2[k];
        e31[k] = p3[k] - p1[k];
        e12[k] = p2[k] - p1[k];
    }
    double l23 = sqrt(e23[0] * e23[0] + e23[1] * e23[1] + e23[2] * e23[2]);
    double l31 = sqrt(e31[0] * e31[0] + e31[1] * e31[1] + e31[2] * e31[2]);
    double l12 = sqrt(e12[0] * e12[0] + e12[1] * e12[1] + e12[2] * e12[2]);
    defslen[e * 3 + 1] = l23;
    defslen[e * 3 + 2] = l31;
    defslen[e * 3] = l12;
    double nx = e12[1] * e31[2] - e12[2] * e31[1];
    double ny = e12[2] * e31[0] - e12[0] * e31[2];
    double nz = e12[0] * e31[1] - e12[1] * e31[0];
    double A = 0.5 * sqrt(nx * nx + ny * ny + nz * nz);
    deffarea[e] = A;
    double lx[3] = {e12[0] / l12, e12[1] / l12, e12[2] / l12};
    double lz[3] = {nx / (2 * A), ny / (2 * A), nz / (2 * A)};
    double ly[3] = {lz[1] * lx[2] - lz[2] * lx[1], lz[2] * lx[0] - lz[0] * lx[2], lz[0] * lx[1] - lz[1] * lx[0]};
    double ll = sqrt(ly[0] * ly[0] + ly[1] * ly[1] + ly[2] * ly[2]);
    for (int k = 0; k < 3; ++k) {
        c1[toff + e * 3 + k] = lx[k];
        c2[toff + e * 3 + k] = ly[k] / ll;
        c3[toff + e * 3 + k] = lz[k];
    }
}

// Rows/columns of fixed slots become identity rows/columns.
template <class T>
__global__ void k_fix_slots(T* __restrict__ K, const int32_t* __restrict__ rowptr,
                            const int32_t* __restrict__ colidx, const int32_t* __restrict__ eq,
                            long nj, int bs) {
    long r = blockIdx.x * (long)blockDim.x + threadIdx.x;   // slot row
    if (r >= nj * bs) return;
    long A = r / bs;
    int p = int(r % bs);
    bool rowfix = eq[r] < 0;
    for (int32_t k = rowptr[A]; k < rowptr[A + 1]; ++k) {
        long B = colidx[k];
        for (int q = 0; q < bs; ++q) {
            bool colfix = eq[B * bs + q] < 0;
            if (rowfix || colfix) K[(long)k * bs * bs + p * bs + q] = (A == B && p == q) ? T(1) : T(0);
        }
    }
}

__global__ void k_cast32(float* __restrict__ y, const double* __restrict__ x, long n) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) y[i] = float(x[i]);
}

// ---------------------------------------------------------------- vectors
constexpr int DOT_BLOCKS = DOTS_BLOCKS;
constexpr int MAXDOT = DOTS_MAX;

struct DotArgs {
    const double* a[MAXDOT];
    const double* b[MAXDOT];
};

__global__ void __launch_bounds__(TPB) k_dot_partial(DotArgs args, int nd, long n, double* __restrict__ part) {
    __shared__ double sh[MAXDOT][TPB / 64];
    double acc[MAXDOT];
#pragma unroll
    for (int k = 0; k < MAXDOT; ++k) acc[k] = 0.0;
    for (long i = blockIdx.x * (long)TPB + threadIdx.x; i < n; i += (long)gridDim.x * TPB) {
#pragma unroll
        for (int k = 0; k < MAXDOT; ++k)
            if (k < nd) acc[k] += args.a[k][i] * args.b[k][i];
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < MAXDOT; ++k) {
        if (k >= nd) break;
        double v = acc[k];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
        if (lane == 0) sh[k][wv] = v;
    }
    __syncthreads();
    if (threadIdx.x < nd) {
        double s = 0.0;
        for (int w = 0; w < TPB / 64; ++w) s += sh[threadIdx.x][w];
        part[threadIdx.x * gridDim.x + blockIdx.x] = s;
    }
}

__global__ void k_dot_final(const double* __restrict__ part, int nblocks, int nd, double* __restrict__ out) {
    __shared__ double sh[TPB];
    for (int k = 0; k < nd; ++k) {
        double s = 0.0;
        for (int i = threadIdx.x; i < nblocks; i += TPB) s += part[k * nblocks + i];
        sh[threadIdx.x] = s;
        __syncthreads();
        for (int o = TPB / 2; o > 0; o >>= 1) {
            if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[k] = sh[0];
        __syncthreads();
    }
}

__global__ void k_fill(double* y, double v, long n) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) y[i] = v;
}
__global__ void k_axpby(double* y, double a, const double* x, double b, long n) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) y[i] = a * x[i] + b * y[i];
}
__global__ void k_residual(double* r, double lam, const double* q, const double* f, long n) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) r[i] = q[i] * lam - f[i];
}
__global__ void k_to_slots(double* ext, const double* v, const int32_t* eq, long ns) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < ns) {
        int32_t q = eq[i];
        ext[i] = q >= 0 ? v[q] : 0.0;
    }
}
__global__ void k_from_slots(double* v, const double* ext, const int32_t* eq, long ns) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < ns) {
        int32_t q = eq[i];
        if (q >= 0) v[q] = ext[i];
    }
}

// y = K x, one thread per slot row; bs x bs row-major blocks.
template <class T, int BS>
__global__ void __launch_bounds__(TPB) k_bsr_spmv(const T* __restrict__ K, const int32_t* __restrict__ rowptr,
                                                  const int32_t* __restrict__ colidx,
                                                  const double* __restrict__ x, double* __restrict__ y, long nj) {
    long r = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (r >= nj * BS) return;
    long A = r / BS;
    int p = int(r % BS);
    double s = 0.0;
    for (int32_t k = rowptr[A]; k < rowptr[A + 1]; ++k) {
        const T* Kb = K + (long)k * BS * BS + p * BS;
        const double* xb = x + (long)colidx[k] * BS;
#pragma unroll
        for (int q = 0; q < BS; ++q) s += double(Kb[q]) * xb[q];
    }
    y[r] = s;
}

}  // namespace

// ============================================================== launchers
void launch_tangent(Dev& d, int gx, int g, bool want_fp32) {
    const long nb = d.nnzb * d.bs * d.bs;
    if (d.nfr && !d.ntr && !d.nsh && !d.ofs && !d.relmask) {
        // frame-only: every block is summed and written once, fixed slots included
        launch_frame_tangent_gather(d, g);
        if (want_fp32) {
            hipLaunchKernelGGL(k_cast32, dim3(nblk(nb)), dim3(TPB), 0, d.s, d.K32, d.K, nb);
            CB_CHECK(hipGetLastError());
        }
        return;
    }
    CB_CHECK(hipMemsetAsync(d.K, 0, sizeof(double) * nb, d.s));
    const long to = d.ntr + d.nfr * 3, io = d.ntr * 2 + d.nfr * 2, po = d.ntr + d.nfr;
    for (size_t c = 0; c + 1 < d.col_ptr_tr.size(); ++c) {
        int n = d.col_ptr_tr[c + 1] - d.col_ptr_tr[c];
        if (!n) continue;
        hipLaunchKernelGGL(k_truss_tangent<double>, dim3(nblk(n, 128)), dim3(128), 0, d.s,
                           d.col_el_tr + d.col_ptr_tr[c], n, d.blk_tr, d.c1[g], d.c2[g], d.c3[g],
                           d.emod, d.carea, d.llen, d.defllen[g], d.ef[g], d.yield, d.K, d.bs,
                           d.anaflag);
    }
    if (d.nfr) launch_frame_tangent(d, g);
    for (size_t c = 0; c + 1 < d.col_ptr_sh.size(); ++c) {
        int n = d.col_ptr_sh[c + 1] - d.col_ptr_sh[c];
        if (!n) continue;
        hipLaunchKernelGGL(k_shell_tangent<double>, dim3(nblk(n, 128)), dim3(128), 0, d.s,
                           d.col_el_sh + d.col_ptr_sh[c], n, d.minc, d.blk_sh, d.x[gx], d.c1[g],
                           d.c2[g], d.c3[g], d.emod, d.nu, d.thick, d.xlocal, d.farea, d.slength,
                           d.deffarea[g], d.K, io, to, po, int(d.anaflag != 1));
    }
    if (d.bs == 6)
        hipLaunchKernelGGL((k_fix_slots<double>), dim3(nblk(d.nslot)), dim3(TPB), 0, d.s, d.K,
                           d.rowptr, d.colidx, d.eq_of_slot, d.nj, d.bs);
    else
        hipLaunchKernelGGL((k_fix_slots<double>), dim3(nblk(d.nslot)), dim3(TPB), 0, d.s, d.K,
                           d.rowptr, d.colidx, d.eq_of_slot, d.nj, d.bs);
    CB_CHECK(hipGetLastError());
    if (want_fp32) {
        hipLaunchKernelGGL(k_cast32, dim3(nblk(nb)), dim3(TPB), 0, d.s, d.K32, d.K, nb);
        CB_CHECK(hipGetLastError());
    }
}

// The tangent in d.K32 for the FP32 precision mode: shell-only models take
// the element stiffness in FP32 (k_shell_tangent<float>) straight into the
// FP32 joint blocks; models with bars or frame members form the FP64 tangent
// and round it once.
void launch_tangent_fp32(Dev& d, int gx, int g) {
    if (d.ntr || d.nfr) {
        launch_tangent(d, gx, g, true);
        return;
    }
    const long nb = d.nnzb * d.bs * d.bs;
    CB_CHECK(hipMemsetAsync(d.K32, 0, sizeof(float) * nb, d.s));
    const long to = d.ntr + d.nfr * 3, io = d.ntr * 2 + d.nfr * 2, po = d.ntr + d.nfr;
    for (size_t c = 0; c + 1 < d.col_ptr_sh.size(); ++c) {
        int n = d.col_ptr_sh[c + 1] - d.col_ptr_sh[c];
        if (!n) continue;
        hipLaunchKernelGGL(k_shell_tangent<float>, dim3(nblk(n, 128)), dim3(128), 0, d.s,
                           d.col_el_sh + d.col_ptr_sh[c], n, d.minc, d.blk_sh, d.x[gx], d.c1[g],
                           d.c2[g], d.c3[g], d.emod, d.nu, d.thick, d.xlocal, d.farea, d.slength,
                           d.deffarea[g], d.K32, io, to, po, int(d.anaflag != 1));
    }
    hipLaunchKernelGGL((k_fix_slots<float>), dim3(nblk(d.nslot)), dim3(TPB), 0, d.s, d.K32, d.rowptr, d.colidx,
                       d.eq_of_slot, d.nj, d.bs);
    CB_CHECK(hipGetLastError());
}

void launch_forces(Dev& d, int gx, double* f, const double* dd) {
    const long to = d.ntr + d.nfr * 3, io = d.ntr * 2 + d.nfr * 2, po = d.ntr + d.nfr;
    const long mo = d.ntr * 6 + d.nfr * 14, eo = d.ntr * 2 + d.nfr * 14;
    if (d.anaflag == 1) {
        // linear elastic: dd is the total displacement; frames take their
        // elastic matrix alone (the geometric one is off for ANAFLAG 1)
        for (size_t c = 0; c + 1 < d.col_ptr_tr.size(); ++c) {
            int n = d.col_ptr_tr[c + 1] - d.col_ptr_tr[c];
            if (!n) continue;
            hipLaunchKernelGGL(k_truss_forces_lin, dim3(nblk(n)), dim3(TPB), 0, d.s,
                               d.col_el_tr + d.col_ptr_tr[c], n, d.mcode, dd, d.c1[G_I], d.c2[G_I],
                               d.c3[G_I], d.emod, d.carea, d.llen, d.ef[G_I], f);
        }
        if (d.nfr) launch_frame_forces(d, f, dd);
        for (size_t c = 0; c + 1 < d.col_ptr_sh.size(); ++c) {
            int n = d.col_ptr_sh[c + 1] - d.col_ptr_sh[c];
            if (!n) continue;
            hipLaunchKernelGGL(k_shell_forces_lin, dim3(nblk(n, 128)), dim3(128), 0, d.s,
                               d.col_el_sh + d.col_ptr_sh[c], n, d.mcode, dd, d.c1[G_IP], d.c2[G_IP],
                               d.c3[G_IP], d.c1[G_I], d.c2[G_I], d.c3[G_I], d.emod, d.nu, d.thick,
                               d.xlocal, d.farea, d.slength, d.ef[G_I], f, to, po, mo, eo);
        }
        CB_CHECK(hipGetLastError());
        return;
    }
    for (size_t c = 0; c + 1 < d.col_ptr_tr.size(); ++c) {
        int n = d.col_ptr_tr[c + 1] - d.col_ptr_tr[c];
        if (!n) continue;
        hipLaunchKernelGGL(k_truss_forces, dim3(nblk(n)), dim3(TPB), 0, d.s,
                           d.col_el_tr + d.col_ptr_tr[c], n, d.mcode, d.c1[G_I], d.c2[G_I], d.c3[G_I],
                           d.emod, d.carea, d.llen, d.defllen[G_I], d.anaflag == 3 ? d.yield : nullptr,
                           d.ef[G_I], f);
    }
    if (d.nfr) launch_frame_forces(d, f, dd);
    for (size_t c = 0; c + 1 < d.col_ptr_sh.size(); ++c) {
        int n = d.col_ptr_sh[c + 1] - d.col_ptr_sh[c];
        if (!n) continue;
        hipLaunchKernelGGL(k_shell_forces, dim3(nblk(n, 128)), dim3(128), 0, d.s,
                           d.col_el_sh + d.col_ptr_sh[c], n, d.minc, d.mcode, d.x[gx], dd,
                           d.c1[G_IP], d.c2[G_IP], d.c3[G_IP], d.c1[G_I], d.c2[G_I], d.c3[G_I],
                           d.emod, d.nu, d.thick, d.xlocal, d.farea, d.slength, d.ef[G_IP],
                           d.ef[G_I], f, io, to, po, mo, eo);
    }
    CB_CHECK(hipGetLastError());
}

void launch_updatc(Dev& d, int gx, const double* dd) {
    hipLaunchKernelGGL(k_update_joints, dim3(nblk(d.nj * 3)), dim3(TPB), 0, d.s, d.jcode, dd,
                       d.x[gx], d.x[G_IP], d.nj);
    if (d.ntr)
        hipLaunchKernelGGL(k_update_truss, dim3(nblk(d.ntr)), dim3(TPB), 0, d.s, d.minc, d.x[gx],
                           d.defllen[G_I], d.c1[G_I], d.c2[G_I], d.c3[G_I], d.ntr);
    launch_frame_update(d, gx);
    if (d.nsh)
        hipLaunchKernelGGL(k_update_shell, dim3(nblk(d.nsh)), dim3(TPB), 0, d.s, d.minc, d.x[gx],
                           d.defslen[G_I], d.deffarea[G_I], d.c1[G_I], d.c2[G_I], d.c3[G_I], d.nsh,
                           d.ntr * 2 + d.nfr * 2, d.ntr + d.nfr * 3);
    CB_CHECK(hipGetLastError());
}

void copy_gen(Dev& d, int dst, int src, bool with_ef, bool with_x) {
    const long ntri = d.ntr + d.nfr * 3 + d.nsh * 3, nef = d.ntr * 2 + d.nfr * 14 + d.nsh * 18;
    d2d(d.c1[dst], d.c1[src], ntri, d.s);
    d2d(d.c2[dst], d.c2[src], ntri, d.s);
    d2d(d.c3[dst], d.c3[src], ntri, d.s);
    if (with_ef) d2d(d.ef[dst], d.ef[src], nef, d.s);
    d2d(d.defllen[dst], d.defllen[src], d.ntr + d.nfr, d.s);
    d2d(d.deffarea[dst], d.deffarea[src], d.nsh, d.s);
    d2d(d.defslen[dst], d.defslen[src], d.nsh * 3, d.s);
    if (with_x) d2d(d.x[dst], d.x[src], d.nj * 3, d.s);
}

void launch_dots(hipStream_t s, int n, const double* const* a, const double* const* b, long len, double* part,
                 double* out) {
    DotArgs args{};
    for (int k = 0; k < n; ++k) { args.a[k] = a[k]; args.b[k] = b[k]; }
    int nbk = int(std::min<long>(DOT_BLOCKS, std::max<long>(1, (len + TPB - 1) / TPB)));
    hipLaunchKernelGGL(k_dot_partial, dim3(nbk), dim3(TPB), 0, s, args, n, len, part);
    hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(TPB), 0, s, part, nbk, n, out);
    CB_CHECK(hipGetLastError());
}

void multi_dot(Dev& d, int n, const double* const* a, const double* const* b, long len, double* out) {
    launch_dots(d.s, n, a, b, len, d.red, d.red + DOTS_PARTS);
    CB_CHECK(hipMemcpyAsync(d.red_host, d.red + DOTS_PARTS, sizeof(double) * n, hipMemcpyDeviceToHost, d.s));
    CB_CHECK(hipStreamSynchronize(d.s));
    for (int k = 0; k < n; ++k) out[k] = d.red_host[k];
}

void vfill(Dev& d, double* y, double v, long n) {
    if (n) hipLaunchKernelGGL(k_fill, dim3(nblk(n)), dim3(TPB), 0, d.s, y, v, n);
}
void vcopy(Dev& d, double* y, const double* x, long n) { d2d(y, x, n, d.s); }
void vaxpby(Dev& d, double* y, double a, const double* x, double b, long n) {
    if (n) hipLaunchKernelGGL(k_axpby, dim3(nblk(n)), dim3(TPB), 0, d.s, y, a, x, b, n);
}
void vresidual(Dev& d, double* r, double lam, const double* q, const double* f, long n) {
    if (n) hipLaunchKernelGGL(k_residual, dim3(nblk(n)), dim3(TPB), 0, d.s, r, lam, q, f, n);
}
void to_slots(Dev& d, double* ext, const double* v) {
    hipLaunchKernelGGL(k_to_slots, dim3(nblk(d.nslot)), dim3(TPB), 0, d.s, ext, v, d.eq_of_slot, d.nslot);
}
void from_slots(Dev& d, double* v, const double* ext) {
    hipLaunchKernelGGL(k_from_slots, dim3(nblk(d.nslot)), dim3(TPB), 0, d.s, v, ext, d.eq_of_slot, d.nslot);
}

void bsr_spmv(Dev& d, const double* x, double* y) {
    if (d.bs == 7)
        hipLaunchKernelGGL((k_bsr_spmv<double, 7>), dim3(nblk(d.nslot)), dim3(TPB), 0, d.s, d.K,
                           d.rowptr, d.colidx, x, y, d.nj);
    else if (d.bs == 6)
        hipLaunchKernelGGL((k_bsr_spmv<double, 6>), dim3(nblk(d.nslot)), dim3(TPB), 0, d.s, d.K,
                           d.rowptr, d.colidx, x, y, d.nj);
    else
        hipLaunchKernelGGL((k_bsr_spmv<double, 3>), dim3(nblk(d.nslot)), dim3(TPB), 0, d.s, d.K,
                           d.rowptr, d.colidx, x, y, d.nj);
}
void bsr_spmv32(Dev& d, const double* x, double* y) {
    if (d.bs == 7)
        hipLaunchKernelGGL((k_bsr_spmv<float, 7>), dim3(nblk(d.nslot)), dim3(TPB), 0, d.s, d.K32,
                           d.rowptr, d.colidx, x, y, d.nj);
    else if (d.bs == 6)
        hipLaunchKernelGGL((k_bsr_spmv<float, 6>), dim3(nblk(d.nslot)), dim3(TPB), 0, d.s, d.K32,
                           d.rowptr, d.colidx, x, y, d.nj);
    else
        hipLaunchKernelGGL((k_bsr_spmv<float, 3>), dim3(nblk(d.nslot)), dim3(TPB), 0, d.s, d.K32,
                           d.rowptr, d.colidx, x, y, d.nj);
}

// ============================================================== setup
void dev_create(Dev& d, const Model& m) {
    static std::atomic<uint64_t> next_uid{1};
    d.uid = next_uid.fetch_add(1);
    CB_CHECK(hipStreamCreateWithFlags(&d.s, hipStreamNonBlocking));
    d.nj = m.nj; d.ntr = m.ntr; d.nfr = m.nfr; d.nsh = m.nsh; d.neq = m.neq;
    d.bs = m.bs; d.nslot = m.nj * m.bs; d.nnzb = long(m.colidx.size());
    d.anaflag = m.anaflag;
    auto up32 = [&](const std::vector<long>& v) {
        std::vector<int32_t> t(v.begin(), v.end());
        int32_t* p = dalloc<int32_t>(t.size());
        h2d(p, t.data(), t.size(), d.s);
        return p;
    };
    auto upi = [&](const std::vector<int32_t>& v) {
        int32_t* p = dalloc<int32_t>(v.size());
        h2d(p, v.data(), v.size(), d.s);
        return p;
    };
    auto upd = [&](const std::vector<double>& v) {
        double* p = dalloc<double>(v.size());
        h2d(p, v.data(), v.size(), d.s);
        return p;
    };
    d.minc = up32(m.minc);
    d.mcode = up32(m.mcode);
    d.jcode = up32(m.jcode);
    d.eq_of_slot = upi(m.eq_of_slot);
    d.rowptr = upi(m.rowptr);
    d.colidx = upi(m.colidx);
    d.diagblk = upi(m.diagblk);
    d.blk_tr = upi(m.blk_tr);
    d.blk_sh = upi(m.blk_sh);
    d.blk_fr = upi(m.blk_fr);
    d.col_el_tr = upi(m.col_el_tr);
    d.col_el_sh = upi(m.col_el_sh);
    d.col_el_fr = upi(m.col_el_fr);
    d.col_ptr_tr = m.col_ptr_tr;
    d.col_ptr_sh = m.col_ptr_sh;
    d.col_ptr_fr = m.col_ptr_fr;
    d.blkrow = upi(m.blkrow);
    d.gat_ptr_fr = upi(m.gat_ptr_fr);
    d.gat_item_fr = upi(m.gat_item_fr);
    d.jnt_ptr_fr = upi(m.jnt_ptr_fr);
    d.jnt_item_fr = upi(m.jnt_item_fr);
    d.fh = dalloc<double>(size_t(m.nfr * 14));   // 2 bs end forces per member
    d.emod = upd(m.emod);
    d.carea = upd(m.carea);
    d.llen = upd(m.llength);
    d.dens = upd(m.dens);
    d.gmod = upd(m.gmod);
    d.istrong = upd(m.istrong);
    d.iweak = upd(m.iweak);
    d.ipolar = upd(m.ipolar);
    d.iwarp = upd(m.iwarp);
    d.auxpt = upd(m.auxpt);
    if (std::any_of(m.osflag.begin(), m.osflag.end(), [](int v) { return v != 0; })) d.ofs = upd(m.offset);
    if (std::any_of(m.relmask.begin(), m.relmask.end(), [](int32_t v) { return v != 0; }))
        d.relmask = upi(m.relmask);
    d.nu = upd(m.nu);
    d.thick = upd(m.thick);
    d.xlocal = upd(m.xlocal);
    d.farea = upd(m.farea);
    d.slength = upd(m.slength);
    d.yield = upd(m.yield);
    d.zstrong = upd(m.zstrong);
    d.zweak = upd(m.zweak);
    d.yld = dalloc<int32_t>(size_t(m.nfr * 2));
    d.pl_efc = dalloc<double>(size_t(m.nfr * 14));
    d.pl_tau = dalloc<double>(size_t(m.nfr));
    d.pl_code = dalloc<int32_t>(size_t(m.nfr * 4));
    d.pl_st = dalloc<PlasticStatus>(1);
    CB_CHECK(hipHostMalloc(&d.pl_host, sizeof(PlasticStatus), hipHostMallocDefault));
    const long ntri = m.ntr + m.nfr * 3 + m.nsh * 3, nef = m.ntr * 2 + m.nfr * 14 + m.nsh * 18;
    for (int g = 0; g < 3; ++g) {
        d.x[g] = dalloc<double>(size_t(m.nj * 3));
        d.c1[g] = dalloc<double>(size_t(ntri));
        d.c2[g] = dalloc<double>(size_t(ntri));
        d.c3[g] = dalloc<double>(size_t(ntri));
        d.ef[g] = dalloc<double>(size_t(nef));
        d.defllen[g] = dalloc<double>(size_t(m.ntr + m.nfr));
        d.deffarea[g] = dalloc<double>(size_t(m.nsh));
        d.defslen[g] = dalloc<double>(size_t(m.nsh * 3));
    }
    d.K = dalloc<double>(size_t(d.nnzb * d.bs * d.bs));
    d.K32 = dalloc<float>(size_t(d.nnzb * d.bs * d.bs));
    d.red = dalloc<double>(size_t(DOTS_PARTS + 64));
    CB_CHECK(hipHostMalloc(&d.red_host, sizeof(double) * 64, hipHostMallocDefault));
    dev_reset_state(d, m);
}

void dev_reset_state(Dev& d, const Model& m) {
    const long ntri = m.ntr + m.nfr * 3 + m.nsh * 3, nef = m.ntr * 2 + m.nfr * 14 + m.nsh * 18;
    for (int g = 0; g < 3; ++g) {
        h2d(d.x[g], m.x.data(), size_t(m.nj * 3), d.s);
        h2d(d.c1[g], m.c1.data(), size_t(ntri), d.s);
        h2d(d.c2[g], m.c2.data(), size_t(ntri), d.s);
        h2d(d.c3[g], m.c3.data(), size_t(ntri), d.s);
        CB_CHECK(hipMemsetAsync(d.ef[g], 0, sizeof(double) * (nef ? nef : 1), d.s));
        h2d(d.defllen[g], m.llength.data(), size_t(m.ntr + m.nfr), d.s);
        h2d(d.deffarea[g], m.farea.data(), size_t(m.nsh), d.s);
        h2d(d.defslen[g], m.slength.data(), size_t(m.nsh * 3), d.s);
    }
    h2d(d.llen, m.llength.data(), m.llength.size(), d.s);
    h2d(d.farea, m.farea.data(), m.farea.size(), d.s);
    h2d(d.slength, m.slength.data(), m.slength.size(), d.s);
    CB_CHECK(hipMemsetAsync(d.yld, 0, sizeof(int32_t) * size_t(m.nfr * 2 > 0 ? m.nfr * 2 : 1), d.s));
    CB_CHECK(hipStreamSynchronize(d.s));
}

void dev_destroy(Dev& d) {
    void* ps[] = {d.minc, d.mcode, d.jcode, d.eq_of_slot, d.rowptr, d.colidx, d.diagblk, d.blk_tr,
                  d.blk_sh, d.blk_fr, d.col_el_tr, d.col_el_sh, d.col_el_fr, d.blkrow, d.gat_ptr_fr,
                  d.gat_item_fr, d.jnt_ptr_fr, d.jnt_item_fr, d.fh, d.emod, d.carea,
                  d.llen, d.dens, d.gmod, d.istrong, d.iweak, d.ipolar, d.iwarp, d.auxpt, d.ofs, d.relmask, d.nu,
                  d.thick, d.xlocal, d.farea, d.slength, d.K, d.K32, d.red, d.yield, d.zstrong,
                  d.zweak, d.yld, d.pl_efc, d.pl_tau, d.pl_code, d.pl_st};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    for (int g = 0; g < 3; ++g) {
        void* qs[] = {d.x[g], d.c1[g], d.c2[g], d.c3[g], d.ef[g], d.defllen[g], d.deffarea[g], d.defslen[g]};
        for (void* p : qs)
            if (p) (void)hipFree(p);
    }
    if (d.red_host) (void)hipHostFree(d.red_host);
    if (d.pl_host) (void)hipHostFree(d.pl_host);
    if (d.s) (void)hipStreamDestroy(d.s);
    d = Dev();
}

}  // namespace cb
