// Corotational 3-D frame element (elastic + geometric, ANAFLAG 2; plastic
// hinges at the member ends, ANAFLAG 3) on the device.
//
// One thread per member.  The 14x14 local matrix is never stored: every entry
// is a closed-form expression of ~40 per-member scalars (stiffe_fr /
// stiffg_fr, frame.c:354-579), evaluated where the fully unrolled transform
// loops ask for it, so the compiler folds the zero pattern away.  Joints carry
// BS = 6 slots (warping fixed everywhere: 12 active DOF per member) or BS = 7
// (the warping DOF in slot 6, shared by the members at the joint as codes()
// numbers it, model.c:951-957; the transformation is 1 on it).  Members of one
// colour share no joint, so the block adds and force scatters of one launch
// never collide.
#include <hip/hip_runtime.h>

#include <cmath>
#include <type_traits>

#include "dev.h"

namespace cb {
namespace {

constexpr int TPB = 128;
inline int nblk(long n, int tpb = TPB) { return int((n + tpb - 1) / tpb); }

// Scalars of the local tangent of one member: elastic terms from E, G, A, L0,
// I_strong, I_weak, J, C_w and geometric terms from the previous iterate's end
// forces (P = f[7], My1 = f[4], Mz1 = f[5], T = f[10], My2 = f[11],
// Mz2 = f[12]) over its deformed length L.
// T: the arithmetic of the local tangent (double; float for the FP32 element
// tangent of frame batches)
template <class T>
struct FrameCoefT {
    T ax, s12, s6, s4, s2, w12, w6, w4, w2, t33, t63, t66, t136;
    T gP, pv, pt, pr, pw, p10, pq, pl, pw30, tl, th;
    T a, b, my1, my2, c, e, mz1, mz2;
    T x34, x35, x311, x312, x410, x510, x1011, x1012;
    T w46, w56, w413, w513, w611, w612, w1113, w1213;
};
using FrameCoef = FrameCoefT<double>;

template <class T>
__device__ inline void frame_coef(FrameCoefT<T>& k, double E_, double G_, double A_, double L0_,
                                  double Is_, double Iw_, double J_, double Cw_, const double* f,
                                  double L_, bool geometric) {
    const T E = T(E_), G = T(G_), A = T(A_), L0 = T(L0_), Is = T(Is_), Iw = T(Iw_), J = T(J_), Cw = T(Cw_);
    const T L = T(L_);
    const T L2 = L0 * L0, L3 = L0 * L0 * L0;
    k.ax = E * A / L0;
    k.s12 = 12 * E * Is / L3; k.s6 = 6 * E * Is / L2; k.s4 = 4 * E * Is / L0; k.s2 = 2 * E * Is / L0;
    k.w12 = 12 * E * Iw / L3; k.w6 = 6 * E * Iw / L2; k.w4 = 4 * E * Iw / L0; k.w2 = 2 * E * Iw / L0;
    k.t33 = 6 * G * J / (5 * L0) + 12 * E * Cw / L3;
    k.t63 = G * J / 10 + 6 * E * Cw / L2;
    k.t66 = 2 * G * J * L0 / 15 + 4 * E * Cw / L0;
    k.t136 = -(G * J * L0 / 30 - 2 * E * Cw / L0);
    const T P = T(geometric ? f[7] : 0.0), My1 = T(geometric ? f[4] : 0.0);
    const T Mz1 = T(geometric ? f[5] : 0.0), Tq = T(geometric ? f[10] : 0.0);
    const T My2 = T(geometric ? f[11] : 0.0), Mz2 = T(geometric ? f[12] : 0.0);
    k.gP = P / L; k.pv = 6 * P / (5 * L); k.pt = 6 * P * J / (5 * A * L);
    k.pr = 2 * P * L / 15; k.pw = 2 * P * J / (15 * A); k.p10 = P / 10;
    k.pq = P * J / (10 * A); k.pl = P * L / 30; k.pw30 = P * J / (30 * A);
    k.tl = Tq / L; k.th = Tq / 2;
    k.a = (11 * My1 - My2) / (10 * L); k.b = (My1 - 11 * My2) / (10 * L);
    k.my1 = My1 / 10; k.my2 = My2 / 10;
    k.c = (11 * Mz1 - Mz2) / (10 * L); k.e = (Mz1 - 11 * Mz2) / (10 * L);
    k.mz1 = Mz1 / 10; k.mz2 = Mz2 / 10;
    k.x34 = -(2 * Mz1 - Mz2) / 5; k.x35 = (2 * My1 - My2) / 5;
    k.x311 = -(2 * Mz1 + Mz2) / 10; k.x312 = (2 * My1 + My2) / 10;
    k.x410 = -(Mz1 + 2 * Mz2) / 10; k.x510 = (My1 + 2 * My2) / 10;
    k.x1011 = (Mz1 - 2 * Mz2) / 5; k.x1012 = -(My1 - 2 * My2) / 5;
    k.w46 = -(3 * Mz1 - Mz2) * L / 30; k.w56 = (3 * My1 - My2) * L / 30;
    k.w413 = Mz1 * L / 30; k.w513 = -My1 * L / 30;
    k.w611 = -Mz2 * L / 30; k.w612 = My2 * L / 30;
    k.w1113 = -(Mz1 - 3 * Mz2) * L / 30; k.w1213 = (My1 - 3 * My2) * L / 30;
}

// Entry (i, j) of the symmetric local tangent, DOF order per end u, v, w,
// theta_x, theta_y, theta_z, warping (end 2 at +7).
template <class T>
__device__ __forceinline__ T kf(const FrameCoefT<T>& k, int i, int j) {
    if (i > j) { int t = i; i = j; j = t; }
    switch (i * 14 + j) {
    case 0 * 14 + 0: case 7 * 14 + 7: return k.ax + k.gP;
    case 0 * 14 + 7: return -k.ax + -k.gP;
    case 1 * 14 + 1: case 8 * 14 + 8: return k.s12 + k.pv;
    case 1 * 14 + 8: return -k.s12 + -k.pv;
    case 2 * 14 + 2: case 9 * 14 + 9: return k.w12 + k.pv;
    case 2 * 14 + 9: return -k.w12 + -k.pv;
    case 3 * 14 + 3: case 10 * 14 + 10: return k.t33 + k.pt;
    case 3 * 14 + 10: return -k.t33 + -k.pt;
    case 4 * 14 + 4: case 11 * 14 + 11: return k.w4 + k.pr;
    case 5 * 14 + 5: case 12 * 14 + 12: return k.s4 + k.pr;
    case 6 * 14 + 6: case 13 * 14 + 13: return k.t66 + k.pw;
    case 1 * 14 + 5: case 1 * 14 + 12: return k.s6 + k.p10;
    case 5 * 14 + 8: case 8 * 14 + 12: return -k.s6 + -k.p10;
    case 4 * 14 + 9: case 9 * 14 + 11: return k.w6 + k.p10;
    case 2 * 14 + 4: case 2 * 14 + 11: return -k.w6 + -k.p10;
    case 3 * 14 + 6: case 3 * 14 + 13: return k.t63 + k.pq;
    case 6 * 14 + 10: case 10 * 14 + 13: return -k.t63 + -k.pq;
    case 5 * 14 + 12: return k.s2 + -k.pl;
    case 4 * 14 + 11: return k.w2 + -k.pl;
    case 6 * 14 + 13: return k.t136 + -k.pw30;
    case 1 * 14 + 4: case 2 * 14 + 5: case 8 * 14 + 11: case 9 * 14 + 12: return k.tl;
    case 1 * 14 + 11: case 2 * 14 + 12: case 4 * 14 + 8: case 5 * 14 + 9: return -k.tl;
    case 4 * 14 + 12: return k.th;
    case 5 * 14 + 11: return -k.th;
    case 1 * 14 + 3: return k.a;
    case 3 * 14 + 8: return -k.a;
    case 8 * 14 + 10: return k.b;
    case 1 * 14 + 10: return -k.b;
    case 1 * 14 + 6: return k.my1;
    case 6 * 14 + 8: return -k.my1;
    case 8 * 14 + 13: return k.my2;
    case 1 * 14 + 13: return -k.my2;
    case 2 * 14 + 3: return k.c;
    case 3 * 14 + 9: return -k.c;
    case 9 * 14 + 10: return k.e;
    case 2 * 14 + 10: return -k.e;
    case 2 * 14 + 6: return k.mz1;
    case 6 * 14 + 9: return -k.mz1;
    case 9 * 14 + 13: return k.mz2;
    case 2 * 14 + 13: return -k.mz2;
    case 3 * 14 + 4: return k.x34;
    case 3 * 14 + 5: return k.x35;
    case 3 * 14 + 11: return k.x311;
    case 3 * 14 + 12: return k.x312;
    case 4 * 14 + 10: return k.x410;
    case 5 * 14 + 10: return k.x510;
    case 10 * 14 + 11: return k.x1011;
    case 10 * 14 + 12: return k.x1012;
    case 4 * 14 + 6: return k.w46;
    case 5 * 14 + 6: return k.w56;
    case 4 * 14 + 13: return k.w413;
    case 5 * 14 + 13: return k.w513;
    case 6 * 14 + 11: return k.w611;
    case 6 * 14 + 12: return k.w612;
    case 11 * 14 + 13: return k.w1113;
    case 12 * 14 + 13: return k.w1213;
    default: return T(0);
    }
}

// local index of slot s (0..5) of end a
__device__ __forceinline__ constexpr int lidx(int a, int s) { return 7 * a + s; }

// ---------------------------------------------------------- plastic hinges
// Orbison's yield function of the end forces over the squash load and the
// plastic moments (frame.c:613-620) and its gradient in the end forces
// (P, My, Mz) at local DOFs 7a, 7a+4, 7a+5 (frame.c:626-637).
constexpr double PHITOL = 1e-4;

__device__ inline double orbison(double p, double my, double mz) {
    const double p2 = p * p, y2 = my * my, z2 = mz * mz;
    return p2 + z2 + y2 * y2 + 3.5 * p2 * z2 + 3 * (p2 * p2 * p2) * y2 + 4.5 * (z2 * z2) * y2;
}

struct YieldRef {
    double Py, Mpy, Mpz;
};

__device__ inline double end_phi(const double* f, int a, const YieldRef& y) {
    return orbison(f[7 * a] / y.Py, f[7 * a + 4] / y.Mpy, f[7 * a + 5] / y.Mpz);
}

__device__ inline void end_grad(const double* f, int a, const YieldRef& y, double g[3]) {
    const double p = f[7 * a] / y.Py, my = f[7 * a + 4] / y.Mpy, mz = f[7 * a + 5] / y.Mpz;
    const double p2 = p * p, y2 = my * my, z2 = mz * mz;
    g[0] = (2 * p + 7 * p * z2 + 18 * (p2 * p2 * p) * y2) / y.Py;
    g[1] = (4 * y2 * my + 6 * (p2 * p2 * p2) * my + 9 * (z2 * z2) * my) / y.Mpy;
    g[2] = (2 * mz + 7 * p2 * mz + 18 * (z2 * mz) * y2) / y.Mpz;
}

// The plastic reduction k_p = k - U H U^T of a member's local tangent, U = k G
// for up to two yield-gradient columns G (one per end, zero where unused),
// H = (G^T k G)^-1 by the adjugate over the determinant as frame.c:677-683
// forms it (a zero column with the other in use leaves det = 0).
struct FramePl {
    double u[2][14];
    double h[2][2];
};

template <int A>
__device__ inline void pl_column(const FrameCoef& k, const double g[3], double u[14]) {
#pragma unroll
    for (int i = 0; i < 14; ++i)
        u[i] = kf(k, i, 7 * A) * g[0] + kf(k, i, 7 * A + 4) * g[1] + kf(k, i, 7 * A + 5) * g[2];
}

// columns: end of column c (-1: zero column), gradients g[c]; nc columns in use
__device__ inline void pl_setup(FramePl& pl, const FrameCoef& k, const int end[2], const double g[2][3], int nc) {
    double H[2][2] = {{0, 0}, {0, 0}};
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        if (end[c] == 0) pl_column<0>(k, g[c], pl.u[c]);
        else if (end[c] == 1) pl_column<1>(k, g[c], pl.u[c]);
        else
#pragma unroll
            for (int i = 0; i < 14; ++i) pl.u[c][i] = 0.0;
    }
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int b = 0; b < 2; ++b)
            if (end[b] >= 0) {
                const int o = 7 * end[b];
                H[c][b] = pl.u[c][o] * g[b][0] + pl.u[c][o + 4] * g[b][1] + pl.u[c][o + 5] * g[b][2];
            }
    pl.h[0][1] = pl.h[1][0] = pl.h[1][1] = 0.0;
    if (nc == 1) {
        pl.h[0][0] = 1 / H[0][0];
    } else if (nc == 2) {
        const double det = H[0][0] * H[1][1] - H[0][1] * H[1][0];
        pl.h[0][0] = H[1][1] / det;
        pl.h[1][1] = H[0][0] / det;
        pl.h[0][1] = H[0][1] * (-1 / det);
        pl.h[1][0] = H[1][0] * (-1 / det);
    } else {
        pl.h[0][0] = 0.0;
    }
}

// stiffm_fr (frame.c:580-797) selecting the ends from the total end forces f
// of the previous iterate: both ends on the surface (the column of an end
// flagged 2 zero), else the one end on it not flagged 2; none when both ends
// are flagged 2 (stiff_fr skips the call, frame.c:270-275).
__device__ inline void frame_plastic(FramePl& pl, const FrameCoef& k, const double* f, int y0, int y1,
                                     const YieldRef& yr) {
    int end[2] = {-1, -1}, nc = 0;
    double g[2][3] = {{0, 0, 0}, {0, 0, 0}};
    if (y0 != 2 || y1 != 2) {
        const double phi0 = end_phi(f, 0, yr), phi1 = end_phi(f, 1, yr);
        if (phi0 >= 1 - PHITOL && phi1 >= 1 - PHITOL) {
            nc = 2;
            if (y0 != 2) { end[0] = 0; end_grad(f, 0, yr, g[0]); }
            if (y1 != 2) { end[1] = 1; end_grad(f, 1, yr, g[1]); }
        } else if (phi0 >= 1 - PHITOL && y0 != 2) {
            nc = 1; end[0] = 0; end_grad(f, 0, yr, g[0]);
        } else if (phi1 >= 1 - PHITOL && y1 != 2) {
            nc = 1; end[0] = 1; end_grad(f, 1, yr, g[0]);
        }
    }
    pl_setup(pl, k, end, g, nc);
}

// entry (i, j) of the local tangent, less its plastic part when PL
template <bool PL, class T = double>
__device__ __forceinline__ T kent(const FrameCoefT<T>& k, const FramePl& pl, int i, int j) {
    T v = kf(k, i, j);
    if constexpr (PL)
        v -= pl.u[0][i] * (pl.h[0][0] * pl.u[0][j] + pl.h[0][1] * pl.u[1][j]) +
             pl.u[1][i] * (pl.h[1][0] * pl.u[0][j] + pl.h[1][1] * pl.u[1][j]);
    return v;
}

// Member yield properties: Py = A fy, Mpy = Z_weak fy, Mpz = Z_strong fy.
__device__ inline YieldRef yield_ref(const double* carea, const double* yield, const double* zstrong,
                                     const double* zweak, long e, long ntr) {
    const double fy = yield[ntr + e];
    return YieldRef{carea[ntr + e] * fy, zweak[e] * fy, zstrong[e] * fy};
}

__device__ inline void load_triad(const double* c1, const double* c2, const double* c3, long t,
                                  double R[3][3]) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        R[0][k] = c1[t + k];
        R[1][k] = c2[t + k];
        R[2][k] = c3[t + k];
    }
}

// local index of element of triple t (0: u1, 1: theta1, 2: u2, 3: theta2)
__device__ __forceinline__ constexpr int tidx(int t, int m) { return lidx(t / 2, (t % 2) * 3 + m); }

// Member-end offsets (rigid links, frame.c:297-324): the member end sits at
// joint + o (o fixed, in global axes), so u_end = u + S theta with
// S = -[o]x, and the joint takes the end forces back as (F_u, F_t + S^T F_u).
__device__ inline void link_S(const double* o, double S[3][3]) {
    S[0][0] = 0.0;   S[0][1] = o[2];  S[0][2] = -o[1];
    S[1][0] = -o[2]; S[1][1] = 0.0;   S[1][2] = o[0];
    S[2][0] = o[1];  S[2][1] = -o[0]; S[2][2] = 0.0;
}

// The members' offsets and end releases, each null when no member has any:
// ofs [NE_FR * 6] (global axes, zero for a member without), rel [NE_FR] the
// release flags as a mask (bit a: flag a of the input line).
struct MemberXtr {
    const double* ofs;
    const int32_t* rel;
};

// S of both ends of member e, zero without offsets (G = I then, and every
// term it adds is an exact zero).
__device__ inline void member_links(const MemberXtr& x, long e, double S[2][3][3]) {
    if (x.ofs) {
        link_S(x.ofs + e * 6, S[0]);
        link_S(x.ofs + e * 6 + 3, S[1]);
    } else {
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c) S[a][r][c] = 0.0;
    }
}

// Member end bending releases (release, frame.c:798-896): the released local
// rotations are condensed out of the member's tangent, k <- k - V W V^T with
// V = k G the columns of k at the released DOFs and W = (G^T k G)^-1.  The
// four candidates in the input's flag order: strong axis end 1 (local 5),
// weak axis end 1 (4), strong axis end 2 (12), weak axis end 2 (11).  V and W
// are kept over all four, zero where a DOF is not released, so a member
// without releases subtracts exact zeros.
constexpr int REL_DOF[4] = {5, 4, 12, 11};

struct FrameRel {
    double v[4][14];
    double w[4][4];
};

// inverse (misc.c:284-342) of a 4 x 4 matrix in place: Gauss-Jordan on
// [H | I] without normalising the pivot rows, divided out at the end; the
// pivot search scans row i from the diagonal for its first non-zero entry and
// swaps rows i and that column's index.  Rows and columns of identity among
// the others are exact no-ops, so the inverse of the released block embedded
// in the identity equals the reference's inverse of the compact block.
__device__ inline void gj_inverse4(double H[4][4]) {
    double aug[4][8];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j) aug[i][j] = j < 4 ? H[i][j] : (j - 4 == i ? 1.0 : 0.0);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        int kk = i;
#pragma unroll
        for (int j = 3; j >= i; --j)
            if (aug[i][j] != 0) kk = j;
#pragma unroll
        for (int r = i + 1; r < 4; ++r)
            if (r == kk)
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    const double t = aug[i][q];
                    aug[i][q] = aug[r][q];
                    aug[r][q] = t;
                }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (j == i) continue;
            const double m = aug[j][i] / aug[i][i];
#pragma unroll
            for (int q = 0; q < 8; ++q) aug[j][q] -= m * aug[i][q];
        }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) H[i][j] = aug[i][4 + j] / aug[i][i];
}

// V and W of the release mask on the local tangent kent<PL>; one release
// inverts by 1/x, two by the adjugate over the determinant, three or four by
// gj_inverse4, as release does.
template <bool PL>
__device__ inline void rel_setup(FrameRel& rl, const FrameCoef& k, const FramePl& pl, int mask) {
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int i = 0; i < 14; ++i) rl.v[a][i] = (mask >> a & 1) ? kent<PL>(k, pl, i, REL_DOF[a]) : 0.0;
    // (G^T k G)[a][b] = k[R_b][R_a] = v[a][R_b] on the released DOFs
    double H[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            H[a][b] = ((mask >> a & 1) && (mask >> b & 1)) ? rl.v[a][REL_DOF[b]] : (a == b ? 1.0 : 0.0);
            rl.w[a][b] = 0.0;
        }
    const int n = __popc(mask);
    if (n == 1) {
#pragma unroll
        for (int a = 0; a < 4; ++a)
            if (mask == (1 << a)) rl.w[a][a] = 1 / H[a][a];
    } else if (n == 2) {
#pragma unroll
        for (int p = 0; p < 4; ++p)
#pragma unroll
            for (int q = p + 1; q < 4; ++q)
                if (mask == ((1 << p) | (1 << q))) {
                    const double det = H[p][p] * H[q][q] - H[p][q] * H[q][p];
                    rl.w[p][p] = H[q][q] / det;
                    rl.w[q][q] = H[p][p] / det;
                    rl.w[p][q] = H[p][q] * (-1 / det);
                    rl.w[q][p] = H[q][p] * (-1 / det);
                }
    } else if (n > 2) {
        gj_inverse4(H);
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b)
                if ((mask >> a & 1) && (mask >> b & 1)) rl.w[a][b] = H[a][b];
    }
}

// entry (i, j) of the local tangent, less its plastic part when PL, with the
// member's releases condensed out when REL
template <bool PL, bool REL>
__device__ __forceinline__ double kx(const FrameCoef& k, const FramePl& pl, const FrameRel& rl, int i, int j) {
    double v = kent<PL>(k, pl, i, j);
    if constexpr (REL) {
        double s = 0.0;
#pragma unroll
        for (int a = 0; a < 4; ++a)
            s += rl.v[a][i] * (rl.w[a][0] * rl.v[0][j] + rl.w[a][1] * rl.v[1][j] + rl.w[a][2] * rl.v[2][j] +
                               rl.w[a][3] * rl.v[3][j]);
        v -= s;
    }
    return v;
}

// Yield data of the members under ANAFLAG 3 (null otherwise).
struct PlasticIn {
    const int32_t* yld;      // yldflag, 2 per member
    const double* yield;     // [NE_TR + NE_FR + NE_SH]
    const double* zstrong;   // [NE_FR]
    const double* zweak;
};

// stiff_fr (frame.c:226-349) for the members of one colour: T^T k T with
// T = blockdiag(R, R, 1, R, R, 1), one 3x3 sub-block (pair of triples) at a
// time, added into the four joint blocks of the member.  PL: less the plastic
// part of k (ANAFLAG 3).  XTR: members with end offsets, K_joint =
// G^T K_end G with G = [I S; 0 I] per end (frame.c:297-324), the four 3x3
// blocks of each end pair combined before they are added, and end releases
// condensed out of k (frame.c:280-283).
template <bool PL, int BS, bool XTR>
__global__ void __launch_bounds__(TPB) k_frame_tangent(
    const int32_t* __restrict__ els, int n, const int32_t* __restrict__ blk,
    const double* __restrict__ c1, const double* __restrict__ c2, const double* __restrict__ c3,
    const double* __restrict__ emod, const double* __restrict__ gmod, const double* __restrict__ carea,
    const double* __restrict__ llen, const double* __restrict__ istrong, const double* __restrict__ iweak,
    const double* __restrict__ ipolar, const double* __restrict__ iwarp,
    const double* __restrict__ defllen, const double* __restrict__ ef, double* __restrict__ K,
    long ntr, int geometric, PlasticIn pin, MemberXtr xtr) {
    int it = blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= n) return;
    const long e = els[it], p = ntr + e;
    const double* fe = ef + ntr * 2 + e * 14;
    FrameCoef k;
    frame_coef(k, emod[p], gmod[e], carea[p], llen[p], istrong[e], iweak[e], ipolar[e], iwarp[e], fe,
               defllen[p], geometric != 0);
    FramePl pl;
    if constexpr (PL)
        frame_plastic(pl, k, fe, pin.yld[e * 2], pin.yld[e * 2 + 1],
                      yield_ref(carea, pin.yield, pin.zstrong, pin.zweak, e, ntr));
    double R[3][3];
    load_triad(c1, c2, c3, ntr + e * 3, R);
    int32_t bk[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) bk[q] = blk[e * 4 + q];
    if constexpr (XTR) {
        double S[2][3][3];
        member_links(xtr, e, S);
        FrameRel rl;
        rel_setup<PL>(rl, k, pl, xtr.rel ? xtr.rel[e] : 0);
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                // the end blocks E[h1][h2] = R^T k R of (triple h1 of end a, triple h2 of end b)
                double E[2][2][3][3];
#pragma unroll
                for (int h1 = 0; h1 < 2; ++h1)
#pragma unroll
                    for (int h2 = 0; h2 < 2; ++h2) {
                        const int t1 = 2 * a + h1, t2 = 2 * b + h2;
                        double tmp[3][3];
#pragma unroll
                        for (int r = 0; r < 3; ++r)
#pragma unroll
                            for (int l = 0; l < 3; ++l)
                                tmp[r][l] = R[0][r] * kx<PL, true>(k, pl, rl, tidx(t1, 0), tidx(t2, l)) +
                                            R[1][r] * kx<PL, true>(k, pl, rl, tidx(t1, 1), tidx(t2, l)) +
                                            R[2][r] * kx<PL, true>(k, pl, rl, tidx(t1, 2), tidx(t2, l));
#pragma unroll
                        for (int r = 0; r < 3; ++r)
#pragma unroll
                            for (int s = 0; s < 3; ++s)
                                E[h1][h2][r][s] = tmp[r][0] * R[0][s] + tmp[r][1] * R[1][s] + tmp[r][2] * R[2][s];
                    }
                // joint blocks: J_uu = E00, J_ut = E00 S_b + E01, J_tu = S_a^T E00 + E10,
                // J_tt = S_a^T (E00 S_b + E01) + E10 S_b + E11
                double J[2][2][3][3];
#pragma unroll
                for (int r = 0; r < 3; ++r)
#pragma unroll
                    for (int s = 0; s < 3; ++s) {
                        double x = E[0][1][r][s], y = E[1][1][r][s];
#pragma unroll
                        for (int m = 0; m < 3; ++m) {
                            x += E[0][0][r][m] * S[b][m][s];
                            y += E[1][0][r][m] * S[b][m][s];
                        }
                        J[0][0][r][s] = E[0][0][r][s];
                        J[0][1][r][s] = x;
                        J[1][1][r][s] = y;   // E10 S_b + E11, S_a^T J01 added below
                    }
#pragma unroll
                for (int r = 0; r < 3; ++r)
#pragma unroll
                    for (int s = 0; s < 3; ++s) {
                        double u = E[1][0][r][s], v = J[1][1][r][s];
#pragma unroll
                        for (int m = 0; m < 3; ++m) {
                            u += S[a][m][r] * E[0][0][m][s];
                            v += S[a][m][r] * J[0][1][m][s];
                        }
                        J[1][0][r][s] = u;
                        J[1][1][r][s] = v;
                    }
                double* Kb = K + (long)bk[a * 2 + b] * (BS * BS);
#pragma unroll
                for (int h1 = 0; h1 < 2; ++h1)
#pragma unroll
                    for (int h2 = 0; h2 < 2; ++h2)
#pragma unroll
                        for (int r = 0; r < 3; ++r)
#pragma unroll
                            for (int s = 0; s < 3; ++s) Kb[(h1 * 3 + r) * BS + h2 * 3 + s] += J[h1][h2][r][s];
                if constexpr (BS == 7) {
                    // warping column of end b and row of end a through G
                    double cu[3], ct[3], ru[3], rt[3];
#pragma unroll
                    for (int r = 0; r < 3; ++r) {
                        cu[r] = R[0][r] * kx<PL, true>(k, pl, rl, tidx(2 * a, 0), 7 * b + 6) +
                                R[1][r] * kx<PL, true>(k, pl, rl, tidx(2 * a, 1), 7 * b + 6) +
                                R[2][r] * kx<PL, true>(k, pl, rl, tidx(2 * a, 2), 7 * b + 6);
                        ct[r] = R[0][r] * kx<PL, true>(k, pl, rl, tidx(2 * a + 1, 0), 7 * b + 6) +
                                R[1][r] * kx<PL, true>(k, pl, rl, tidx(2 * a + 1, 1), 7 * b + 6) +
                                R[2][r] * kx<PL, true>(k, pl, rl, tidx(2 * a + 1, 2), 7 * b + 6);
                        ru[r] = kx<PL, true>(k, pl, rl, 7 * a + 6, tidx(2 * b, 0)) * R[0][r] +
                                kx<PL, true>(k, pl, rl, 7 * a + 6, tidx(2 * b, 1)) * R[1][r] +
                                kx<PL, true>(k, pl, rl, 7 * a + 6, tidx(2 * b, 2)) * R[2][r];
                        rt[r] = kx<PL, true>(k, pl, rl, 7 * a + 6, tidx(2 * b + 1, 0)) * R[0][r] +
                                kx<PL, true>(k, pl, rl, 7 * a + 6, tidx(2 * b + 1, 1)) * R[1][r] +
                                kx<PL, true>(k, pl, rl, 7 * a + 6, tidx(2 * b + 1, 2)) * R[2][r];
                    }
#pragma unroll
                    for (int r = 0; r < 3; ++r) {
                        double x = ct[r], y = rt[r];
#pragma unroll
                        for (int m = 0; m < 3; ++m) {
                            x += S[a][m][r] * cu[m];
                            y += ru[m] * S[b][m][r];
                        }
                        Kb[r * 7 + 6] += cu[r];
                        Kb[(3 + r) * 7 + 6] += x;
                        Kb[6 * 7 + r] += ru[r];
                        Kb[6 * 7 + 3 + r] += y;
                    }
                    Kb[48] += kx<PL, true>(k, pl, rl, 7 * a + 6, 7 * b + 6);
                }
            }
        return;
    } else {
#pragma unroll
    for (int t1 = 0; t1 < 4; ++t1)
#pragma unroll
        for (int t2 = 0; t2 < 4; ++t2) {
            double tmp[3][3];
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int l = 0; l < 3; ++l)
                    tmp[r][l] = R[0][r] * kent<PL>(k, pl, tidx(t1, 0), tidx(t2, l)) +
                                R[1][r] * kent<PL>(k, pl, tidx(t1, 1), tidx(t2, l)) +
                                R[2][r] * kent<PL>(k, pl, tidx(t1, 2), tidx(t2, l));
            double* Kb = K + (long)bk[(t1 / 2) * 2 + t2 / 2] * (BS * BS) + ((t1 % 2) * 3) * BS + (t2 % 2) * 3;
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int s = 0; s < 3; ++s)
                    Kb[r * BS + s] += tmp[r][0] * R[0][s] + tmp[r][1] * R[1][s] + tmp[r][2] * R[2][s];
        }
    if constexpr (BS == 7) {
        // warping rows and columns: T is 1 on them
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                double* Kb = K + (long)bk[a * 2 + b] * 49;
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int t = 2 * a + h, tb = 2 * b + h;
#pragma unroll
                    for (int r = 0; r < 3; ++r) {
                        Kb[(h * 3 + r) * 7 + 6] += R[0][r] * kent<PL>(k, pl, tidx(t, 0), 7 * b + 6) +
                                                   R[1][r] * kent<PL>(k, pl, tidx(t, 1), 7 * b + 6) +
                                                   R[2][r] * kent<PL>(k, pl, tidx(t, 2), 7 * b + 6);
                        Kb[6 * 7 + h * 3 + r] += kent<PL>(k, pl, 7 * a + 6, tidx(tb, 0)) * R[0][r] +
                                                 kent<PL>(k, pl, 7 * a + 6, tidx(tb, 1)) * R[1][r] +
                                                 kent<PL>(k, pl, 7 * a + 6, tidx(tb, 2)) * R[2][r];
                    }
                }
                Kb[48] += kent<PL>(k, pl, 7 * a + 6, 7 * b + 6);
            }
    }
    }
}

// The same tangent as k_frame_tangent, gathered: two threads per joint block,
// each summing rows 3u..3u+2 of the block over the block's members (each
// member adds the 3 x 3 sub-blocks of its end pair (RA, CB)) in colour order,
// which is the order the colour launches add them in, and writing them once
// with the rows and columns of fixed slots as identity (k_fix_slots).  No
// zeroing and no read-modify-write of K; a member's scalars are evaluated by
// each thread whose block it touches.
// U 0 / 1: rows 3U..3U+2 of the block (a triple of end RA) over the columns
// of end CB; U 2: the warping row of end RA.  BS 7 adds the warping column.
template <int RA, int CB, int U, bool PL, int BS, class T>
__device__ __forceinline__ void frame_half_add(T acc[3][7], const FrameCoefT<T>& k, const FramePl& pl,
                                               const T R[3][3]) {
    if constexpr (U == 2) {
#pragma unroll
        for (int v = 0; v < 2; ++v) {
            const int t2 = 2 * CB + v;
#pragma unroll
            for (int s = 0; s < 3; ++s)
                acc[0][v * 3 + s] += kent<PL>(k, pl, 7 * RA + 6, tidx(t2, 0)) * R[0][s] +
                                     kent<PL>(k, pl, 7 * RA + 6, tidx(t2, 1)) * R[1][s] +
                                     kent<PL>(k, pl, 7 * RA + 6, tidx(t2, 2)) * R[2][s];
        }
        acc[0][6] += kent<PL>(k, pl, 7 * RA + 6, 7 * CB + 6);
        return;
    } else {
    if constexpr (BS == 7) {
        const int t1 = 2 * RA + U;
#pragma unroll
        for (int r = 0; r < 3; ++r)
            acc[r][6] += R[0][r] * kent<PL>(k, pl, tidx(t1, 0), 7 * CB + 6) +
                         R[1][r] * kent<PL>(k, pl, tidx(t1, 1), 7 * CB + 6) +
                         R[2][r] * kent<PL>(k, pl, tidx(t1, 2), 7 * CB + 6);
    }
#pragma unroll
    for (int v = 0; v < 2; ++v) {
        const int t1 = 2 * RA + U, t2 = 2 * CB + v;
        T tmp[3][3];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int l = 0; l < 3; ++l)
                tmp[r][l] = R[0][r] * kent<PL>(k, pl, tidx(t1, 0), tidx(t2, l)) +
                            R[1][r] * kent<PL>(k, pl, tidx(t1, 1), tidx(t2, l)) +
                            R[2][r] * kent<PL>(k, pl, tidx(t1, 2), tidx(t2, l));
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int s = 0; s < 3; ++s)
                acc[r][v * 3 + s] += tmp[r][0] * R[0][s] + tmp[r][1] * R[1][s] + tmp[r][2] * R[2][s];
    }
    }
}

// TO_BAND: instead of the joint block, the entries of free rows and columns on
// or below the diagonal go straight into the band tiles (the copy
// BandLDLT / BandBatch make from K, with the same values): tile (I, J) of the
// equation pair (i, j) is tiles[(I (w + 1) + J - I + w) T^2], entry (j % T, i % T)
// column-major; an entry more than w tiles below the diagonal sets *bad = 2.
// T: the arithmetic of the member tangents (float: the FP32 element tangent of
// frame batches, summed in FP32 and rounded once into K).
template <bool TO_BAND, bool PL, int BS, class T = double>
__global__ void __launch_bounds__(TPB) k_frame_tangent_gather(
    const int32_t* __restrict__ gptr, const int32_t* __restrict__ gitem, const int32_t* __restrict__ blkrow,
    const int32_t* __restrict__ colidx, const int32_t* __restrict__ eq, long nnzb,
    const double* __restrict__ c1, const double* __restrict__ c2, const double* __restrict__ c3,
    const double* __restrict__ emod, const double* __restrict__ gmod, const double* __restrict__ carea,
    const double* __restrict__ llen, const double* __restrict__ istrong, const double* __restrict__ iweak,
    const double* __restrict__ ipolar, const double* __restrict__ iwarp,
    const double* __restrict__ defllen, const double* __restrict__ ef, double* __restrict__ K, int geometric,
    int w, int* __restrict__ bad, PlasticIn pin) {
    // NU threads per block: rows 0-2 and 3-5, and with BS 7 the warping row
    constexpr int NU = BS == 7 ? 3 : 2;
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const long kb = t / NU;
    const int u = int(t % NU);
    if (kb >= nnzb) return;
    T acc[3][7];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 7; ++j) acc[i][j] = T(0);
    for (int32_t it = gptr[kb]; it < gptr[kb + 1]; ++it) {
        const int32_t item = gitem[it];
        const long e = item >> 2;
        FrameCoefT<T> k;
        frame_coef<T>(k, emod[e], gmod[e], carea[e], llen[e], istrong[e], iweak[e], ipolar[e], iwarp[e],
                      ef + e * 14, defllen[e], geometric != 0);
        FramePl pl;
        if constexpr (PL)
            frame_plastic(pl, k, ef + e * 14, pin.yld[e * 2], pin.yld[e * 2 + 1],
                          yield_ref(carea, pin.yield, pin.zstrong, pin.zweak, e, 0));
        double Rd[3][3];
        load_triad(c1, c2, c3, e * 3, Rd);
        T R[3][3];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) R[r][c] = T(Rd[r][c]);
        switch ((item & 3) * NU + u) {
        case 0 * NU + 0: frame_half_add<0, 0, 0, PL, BS, T>(acc, k, pl, R); break;
        case 0 * NU + 1: frame_half_add<0, 0, 1, PL, BS, T>(acc, k, pl, R); break;
        case 1 * NU + 0: frame_half_add<0, 1, 0, PL, BS, T>(acc, k, pl, R); break;
        case 1 * NU + 1: frame_half_add<0, 1, 1, PL, BS, T>(acc, k, pl, R); break;
        case 2 * NU + 0: frame_half_add<1, 0, 0, PL, BS, T>(acc, k, pl, R); break;
        case 2 * NU + 1: frame_half_add<1, 0, 1, PL, BS, T>(acc, k, pl, R); break;
        case 3 * NU + 0: frame_half_add<1, 1, 0, PL, BS, T>(acc, k, pl, R); break;
        case 3 * NU + 1: frame_half_add<1, 1, 1, PL, BS, T>(acc, k, pl, R); break;
        default:
            if constexpr (BS == 7) {
                switch (item & 3) {
                case 0: frame_half_add<0, 0, 2, PL, BS, T>(acc, k, pl, R); break;
                case 1: frame_half_add<0, 1, 2, PL, BS, T>(acc, k, pl, R); break;
                case 2: frame_half_add<1, 0, 2, PL, BS, T>(acc, k, pl, R); break;
                default: frame_half_add<1, 1, 2, PL, BS, T>(acc, k, pl, R); break;
                }
            }
            break;
        }
    }
    const long A = blkrow[kb], B = colidx[kb];
    const int rows = u < 2 ? 3 : 1;
    if constexpr (TO_BAND) {
        constexpr long BT = 64;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            if (i >= rows) break;
            const int32_t qi = eq[A * BS + u * 3 + i];
            if (qi < 0) continue;
#pragma unroll
            for (int j = 0; j < BS; ++j) {
                const int32_t qj = eq[B * BS + j];
                if (qj < 0 || qj > qi) continue;
                const long I = qi / BT, J = qj / BT;
                if (I - J > w) {
                    *bad = 2;
                    continue;
                }
                K[(I * (w + 1) + (J - I + w)) * BT * BT + (qj % BT) * BT + (qi % BT)] = double(acc[i][j]);
            }
        }
    } else {
        double* Kb = K + kb * (BS * BS) + u * 3 * BS;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            if (i >= rows) break;
            const bool rowfix = eq[A * BS + u * 3 + i] < 0;
#pragma unroll
            for (int j = 0; j < BS; ++j) {
                const bool colfix = eq[B * BS + j] < 0;
                Kb[i * BS + j] = (rowfix || colfix) ? ((A == B && u * 3 + i == j) ? 1.0 : 0.0) : double(acc[i][j]);
            }
        }
    }
}

// forces_fr, elastic branch (frame.c:902-1096, 1273-1309):
//   dl = T_ip DD;  def = k(ef_ip) dl;  ef_i = T_i T_ip^T (ef_ip + def);  f += T_i^T ef_i,
// one output triple at a time.
// One member of forces_fr: ef_i, and the member's end forces in global axes
// (T_i^T ef_i) handed to out(local DOF, value); with BS 7 the warping
// increments enter and the bimoments are handed out too; XTR: with end
// offsets the joint increments reach the ends through u + S theta and the
// joints take (F_u, F_t + S^T F_u), and end releases are condensed out of k.
template <int BS, bool XTR, class Out>
__device__ __forceinline__ void frame_member_forces(
    long e, const int32_t* __restrict__ mc, const double* __restrict__ dd, const double* __restrict__ c1p,
    const double* __restrict__ c2p, const double* __restrict__ c3p, const double* __restrict__ c1c,
    const double* __restrict__ c2c, const double* __restrict__ c3c, const double* __restrict__ emod,
    const double* __restrict__ gmod, const double* __restrict__ carea, const double* __restrict__ llen,
    const double* __restrict__ istrong, const double* __restrict__ iweak, const double* __restrict__ ipolar,
    const double* __restrict__ iwarp, const double* __restrict__ defllen_ip, const double* __restrict__ ef_ip,
    double* __restrict__ ef_i, long ntr, int geometric, MemberXtr xtr, Out&& out) {
    const long p = ntr + e;
    const double* fe = ef_ip + ntr * 2 + e * 14;
    FrameCoef k;
    frame_coef(k, emod[p], gmod[e], carea[p], llen[p], istrong[e], iweak[e], ipolar[e], iwarp[e], fe,
               defllen_ip[p], geometric != 0);
    double Rp[3][3], Ri[3][3];
    load_triad(c1p, c2p, c3p, ntr + e * 3, Rp);
    load_triad(c1c, c2c, c3c, ntr + e * 3, Ri);
    // the joint increments of the four triples
    double Dj[4][3];
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            int32_t q = mc[tidx(t, m)];
            Dj[t][m] = q ? dd[q - 1] : 0.0;
        }
    double S[2][3][3];
    FrameRel rl;
    FramePl pl;
    if constexpr (XTR) {
        member_links(xtr, e, S);
        rel_setup<false>(rl, k, pl, xtr.rel ? xtr.rel[e] : 0);
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int m = 0; m < 3; ++m)
                Dj[2 * a][m] += S[a][m][0] * Dj[2 * a + 1][0] + S[a][m][1] * Dj[2 * a + 1][1] +
                                S[a][m][2] * Dj[2 * a + 1][2];
    }
    // local increments of the four triples, and of the warping DOFs (T is 1
    // there; zero with BS 6, where they are fixed)
    double dl[4][3];
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int m = 0; m < 3; ++m) dl[t][m] = Rp[m][0] * Dj[t][0] + Rp[m][1] * Dj[t][1] + Rp[m][2] * Dj[t][2];
    double dw[2] = {0.0, 0.0};
    if constexpr (BS == 7)
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            const int32_t q = mc[7 * a + 6];
            dw[a] = q ? dd[q - 1] : 0.0;
        }
    double Q[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) Q[r][c] = Ri[r][0] * Rp[c][0] + Ri[r][1] * Rp[c][1] + Ri[r][2] * Rp[c][2];
    double* ei = ef_i + ntr * 2 + e * 14;
    double Fu[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int to = 0; to < 4; ++to) {
        double v[3];
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            double s = 0;
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int m = 0; m < 3; ++m) s += kx<false, XTR>(k, pl, rl, tidx(to, r), tidx(t, m)) * dl[t][m];
            if constexpr (BS == 7)
                s += kx<false, XTR>(k, pl, rl, tidx(to, r), 6) * dw[0] +
                     kx<false, XTR>(k, pl, rl, tidx(to, r), 13) * dw[1];
            v[r] = fe[tidx(to, r)] + s;
        }
        double g[3];
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            g[r] = Q[r][0] * v[0] + Q[r][1] * v[1] + Q[r][2] * v[2];
            ei[tidx(to, r)] = g[r];
        }
        // T_i^T ef_i on this triple (G^T with offsets)
        double F[3];
#pragma unroll
        for (int m = 0; m < 3; ++m) F[m] = Ri[0][m] * g[0] + Ri[1][m] * g[1] + Ri[2][m] * g[2];
        if constexpr (XTR) {
            if (to % 2 == 0) {
#pragma unroll
                for (int m = 0; m < 3; ++m) Fu[m] = F[m];
            } else {
#pragma unroll
                for (int m = 0; m < 3; ++m)
                    F[m] += S[to / 2][0][m] * Fu[0] + S[to / 2][1][m] * Fu[1] + S[to / 2][2][m] * Fu[2];
            }
        }
#pragma unroll
        for (int m = 0; m < 3; ++m) out(tidx(to, m), F[m]);
    }
    // warping rows (bimoments): T_i T_ip^T is 1 there
#pragma unroll
    for (int w = 0; w < 2; ++w) {
        const int i = 7 * w + 6;
        double s = 0;
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int m = 0; m < 3; ++m) s += kx<false, XTR>(k, pl, rl, i, tidx(t, m)) * dl[t][m];
        if constexpr (BS == 7)
            s += kx<false, XTR>(k, pl, rl, i, 6) * dw[0] + kx<false, XTR>(k, pl, rl, i, 13) * dw[1];
        ei[i] = fe[i] + s;
        if constexpr (BS == 7) out(i, ei[i]);
    }
}

// forces_fr, elastic branch (frame.c:902-1096, 1273-1309), for the members of
// one colour, scattered by mcode into f.
template <int BS, bool XTR>
__global__ void __launch_bounds__(TPB) k_frame_forces(
    const int32_t* __restrict__ els, int n, const int32_t* __restrict__ mcode,
    const double* __restrict__ dd, const double* __restrict__ c1p, const double* __restrict__ c2p,
    const double* __restrict__ c3p, const double* __restrict__ c1c, const double* __restrict__ c2c,
    const double* __restrict__ c3c, const double* __restrict__ emod, const double* __restrict__ gmod,
    const double* __restrict__ carea, const double* __restrict__ llen,
    const double* __restrict__ istrong, const double* __restrict__ iweak,
    const double* __restrict__ ipolar, const double* __restrict__ iwarp,
    const double* __restrict__ defllen_ip, const double* __restrict__ ef_ip,
    double* __restrict__ ef_i, double* __restrict__ f, long ntr, int geometric, MemberXtr xtr) {
    int it = blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= n) return;
    const long e = els[it];
    const int32_t* mc = mcode + ntr * 6 + e * 14;
    frame_member_forces<BS, XTR>(e, mc, dd, c1p, c2p, c3p, c1c, c2c, c3c, emod, gmod, carea, llen, istrong, iweak,
                            ipolar, iwarp, defllen_ip, ef_ip, ef_i, ntr, geometric, xtr, [&](int i, double val) {
                                const int32_t q = mc[i];
                                if (q) f[q - 1] += val;
                            });
}

// Frame-only models: every member in one launch, its global end forces into
// h[e * 2 BS + BS * end + slot] ...
template <int BS, bool XTR>
__global__ void __launch_bounds__(TPB) k_frame_forces_members(
    long n, const int32_t* __restrict__ mcode, const double* __restrict__ dd, const double* __restrict__ c1p,
    const double* __restrict__ c2p, const double* __restrict__ c3p, const double* __restrict__ c1c,
    const double* __restrict__ c2c, const double* __restrict__ c3c, const double* __restrict__ emod,
    const double* __restrict__ gmod, const double* __restrict__ carea, const double* __restrict__ llen,
    const double* __restrict__ istrong, const double* __restrict__ iweak, const double* __restrict__ ipolar,
    const double* __restrict__ iwarp, const double* __restrict__ defllen_ip, const double* __restrict__ ef_ip,
    double* __restrict__ ef_i, double* __restrict__ h, int geometric, MemberXtr xtr) {
    const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (e >= n) return;
    double* he = h + e * (2 * BS);
    frame_member_forces<BS, XTR>(e, mcode + e * 14, dd, c1p, c2p, c3p, c1c, c2c, c3c, emod, gmod, carea, llen,
                            istrong, iweak, ipolar, iwarp, defllen_ip, ef_ip, ef_i, 0, geometric, xtr,
                            [&](int i, double val) { he[BS * (i / 7) + i % 7] = val; });
}

// ... then one thread per free slot adds the end forces of the members at its
// joint in colour order, the order in which the colour launches add them;
// with st, only the members before st->first (the members forces_fr went
// through before it returned).
template <int BS>
__global__ void k_frame_forces_gather(const int32_t* __restrict__ jptr, const int32_t* __restrict__ jitem,
                                      const int32_t* __restrict__ eq, long nslot, const double* __restrict__ h,
                                      double* __restrict__ f, const PlasticStatus* __restrict__ st) {
    const long r = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (r >= nslot) return;
    const int32_t q = eq[r];
    if (q < 0) return;
    const long j = r / BS;
    const int s = int(r % BS);
    const int32_t lim = st ? st->first : INT32_MAX;
    double acc = f[q];
    for (int32_t it = jptr[j]; it < jptr[j + 1]; ++it) {
        const int32_t item = jitem[it];
        if ((item >> 1) >= lim) continue;
        acc += h[long(item >> 1) * (2 * BS) + BS * (item & 1) + s];
    }
    f[q] = acc;
}

// regula_falsi (frame.c:1397-1455): the fraction tau of the end-force
// increment (dp, dmy, dmz) from (p, my, mz) at which phi = 1, by false
// position on [0, 1], the estimate from the end values and one further step
// (the reference's loop test can never hold, so its body runs once).
__device__ inline double regula_falsi(double p, double dp, double my, double dmy, double mz, double dmz) {
    double tu = 1, tl = 0;
    double fu = orbison(p + tu * dp, my + tu * dmy, mz + tu * dmz);
    double fl = orbison(p + tl * dp, my + tl * dmy, mz + tl * dmz);
    const double tr = tu - (fu - 1) * (tl - tu) / (fl - fu);
    const double fr = orbison(p + tr * dp, my + tr * dmy, mz + tr * dmz);
    if ((fl - 1 > 0 && fr - 1 > 0) || (fl - 1 < 0 && fr - 1 < 0)) {
        tl = tr;
        fl = fr;
    } else {
        tu = tr;
        fu = fr;
    }
    return tu - (fu - 1) * (tl - tu) / (fl - fu);
}

// forces_fr under ANAFLAG 3 (frame.c:902-1309) for every member at once, each
// as if the reference's member loop reached it: the plastic tangent of the
// previous iterate's end forces (stiffm_fr) gives ef_i, kept in efc with the
// member's global end forces in h; then the yield checks at the new end
// forces: an end through the surface (phi > 1 + tol, the larger of the two,
// not flagged 2) gives code 1 with tau from regula_falsi and that end flagged
// 1; otherwise the ends on the surface and not flagged 2 are flagged 1 and,
// when an end is flagged 1, the unloading test of unload (frame.c:1457-1658:
// lambda = (G^T k G)^-1 G^T k dl with the elastic + geometric k and the
// gradients at the new end forces, an end unloading when lambda < -1e-8)
// flags the unloading ends 2 with code 2.  code[4e] = code, code[4e+1..2] =
// the member's flags after its pass; st->first <- the lowest member with a
// code.  With BS 7 the warping increments and bimoments enter as in the
// elastic pass, and end offsets and releases (XTR) as in frame_member_forces;
// the fixed-end forces are zero (no distributed loads).
template <int BS, bool XTR>
__global__ void __launch_bounds__(64) k_frame_forces_plastic(
    long n, const int32_t* __restrict__ mcode, const double* __restrict__ dd, const double* __restrict__ c1p,
    const double* __restrict__ c2p, const double* __restrict__ c3p, const double* __restrict__ c1c,
    const double* __restrict__ c2c, const double* __restrict__ c3c, const double* __restrict__ emod,
    const double* __restrict__ gmod, const double* __restrict__ carea, const double* __restrict__ llen,
    const double* __restrict__ istrong, const double* __restrict__ iweak, const double* __restrict__ ipolar,
    const double* __restrict__ iwarp, const double* __restrict__ defllen_ip, const double* __restrict__ ef_ip,
    PlasticIn pin, double* __restrict__ efc, double* __restrict__ h, int32_t* __restrict__ code,
    double* __restrict__ tau, PlasticStatus* __restrict__ st, long ntr, MemberXtr xtr) {
    const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (e >= n) return;
    const long p = ntr + e;
    const int32_t* mc = mcode + ntr * 6 + e * 14;
    const double* fe = ef_ip + ntr * 2 + e * 14;
    FrameCoef k;
    frame_coef(k, emod[p], gmod[e], carea[p], llen[p], istrong[e], iweak[e], ipolar[e], iwarp[e], fe,
               defllen_ip[p], true);
    const YieldRef yr = yield_ref(carea, pin.yield, pin.zstrong, pin.zweak, e, ntr);
    int y[2] = {pin.yld[e * 2], pin.yld[e * 2 + 1]};
    FramePl pl;
    frame_plastic(pl, k, fe, y[0], y[1], yr);
    double Rp[3][3], Ri[3][3];
    load_triad(c1p, c2p, c3p, ntr + e * 3, Rp);
    load_triad(c1c, c2c, c3c, ntr + e * 3, Ri);
    double Dj[4][3];
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            const int32_t q = mc[tidx(t, m)];
            Dj[t][m] = q ? dd[q - 1] : 0.0;
        }
    double S[2][3][3];
    FrameRel rl;
    if constexpr (XTR) {
        member_links(xtr, e, S);
        rel_setup<true>(rl, k, pl, xtr.rel ? xtr.rel[e] : 0);
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int m = 0; m < 3; ++m)
                Dj[2 * a][m] += S[a][m][0] * Dj[2 * a + 1][0] + S[a][m][1] * Dj[2 * a + 1][1] +
                                S[a][m][2] * Dj[2 * a + 1][2];
    }
    double dl[4][3];
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int m = 0; m < 3; ++m) dl[t][m] = Rp[m][0] * Dj[t][0] + Rp[m][1] * Dj[t][1] + Rp[m][2] * Dj[t][2];
    double dw[2] = {0.0, 0.0};
    if constexpr (BS == 7)
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            const int32_t q = mc[7 * a + 6];
            dw[a] = q ? dd[q - 1] : 0.0;
        }
    double Q[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) Q[r][c] = Ri[r][0] * Rp[c][0] + Ri[r][1] * Rp[c][1] + Ri[r][2] * Rp[c][2];
    double fi[14];
    double* he = h + e * (2 * BS);
#pragma unroll
    for (int to = 0; to < 4; ++to) {
        double v[3];
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            double s = 0;
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int m = 0; m < 3; ++m) s += kx<true, XTR>(k, pl, rl, tidx(to, r), tidx(t, m)) * dl[t][m];
            if constexpr (BS == 7)
                s += kx<true, XTR>(k, pl, rl, tidx(to, r), 6) * dw[0] + kx<true, XTR>(k, pl, rl, tidx(to, r), 13) * dw[1];
            v[r] = fe[tidx(to, r)] + s;
        }
#pragma unroll
        for (int r = 0; r < 3; ++r) fi[tidx(to, r)] = Q[r][0] * v[0] + Q[r][1] * v[1] + Q[r][2] * v[2];
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            const int i = tidx(to, m);
            he[BS * (i / 7) + i % 7] =
                Ri[0][m] * fi[tidx(to, 0)] + Ri[1][m] * fi[tidx(to, 1)] + Ri[2][m] * fi[tidx(to, 2)];
        }
        if (XTR && to % 2 == 1) {
            // joint moments: F_t + S^T F_u of this end
            const int a = to / 2;
            double Fu[3];
#pragma unroll
            for (int m = 0; m < 3; ++m) Fu[m] = he[BS * a + m];
#pragma unroll
            for (int m = 0; m < 3; ++m)
                he[BS * a + 3 + m] += S[a][0][m] * Fu[0] + S[a][1][m] * Fu[1] + S[a][2][m] * Fu[2];
        }
    }
#pragma unroll
    for (int w = 0; w < 2; ++w) {
        const int i = 7 * w + 6;
        double s = 0;
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int m = 0; m < 3; ++m) s += kx<true, XTR>(k, pl, rl, i, tidx(t, m)) * dl[t][m];
        if constexpr (BS == 7) s += kx<true, XTR>(k, pl, rl, i, 6) * dw[0] + kx<true, XTR>(k, pl, rl, i, 13) * dw[1];
        fi[i] = fe[i] + s;
        if constexpr (BS == 7) he[BS * w + 6] = fi[i];
    }
#pragma unroll
    for (int i = 0; i < 14; ++i) efc[e * 14 + i] = fi[i];

    // yield checks at the new end forces (frame.c:1157-1228)
    const double phi[2] = {end_phi(fi, 0, yr), end_phi(fi, 1, yr)};
    int c = 0;
    double t = 1.0;
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        const int o = 1 - a, b = 7 * a;
        if (!c && phi[a] > phi[o] && phi[a] > 1 + PHITOL && y[a] != 2) {
            t = regula_falsi(fe[b] / yr.Py, (fi[b] - fe[b]) / yr.Py, fe[b + 4] / yr.Mpy,
                             (fi[b + 4] - fe[b + 4]) / yr.Mpy, fe[b + 5] / yr.Mpz, (fi[b + 5] - fe[b + 5]) / yr.Mpz);
            y[a] = 1;
            c = 1;
        }
    }
    if (!c) {
        const bool on0 = phi[0] >= 1 - PHITOL && y[0] != 2, on1 = phi[1] >= 1 - PHITOL && y[1] != 2;
        if (on0 && on1) y[0] = y[1] = 1;
        else if (on0) y[0] = 1;
        else if (on1) y[1] = 1;
        if (y[0] == 1 || y[1] == 1) {
            // unload: columns at the new end forces, k without its plastic part
            int end[2] = {-1, -1}, nc = 0;
            double g[2][3] = {{0, 0, 0}, {0, 0, 0}};
            if (phi[0] >= 1 - PHITOL && phi[1] >= 1 - PHITOL) {
                nc = 2; end[0] = 0; end[1] = 1;
                end_grad(fi, 0, yr, g[0]);
                end_grad(fi, 1, yr, g[1]);
            } else if (phi[0] >= 1 - PHITOL) {
                nc = 1; end[0] = 0; end_grad(fi, 0, yr, g[0]);
            } else if (phi[1] >= 1 - PHITOL) {
                nc = 1; end[0] = 1; end_grad(fi, 1, yr, g[0]);
            }
            if (nc) {
                FramePl ul;
                pl_setup(ul, k, end, g, nc);
                double w[2] = {0, 0};
#pragma unroll
                for (int cc = 0; cc < 2; ++cc)
#pragma unroll
                    for (int tt = 0; tt < 4; ++tt)
#pragma unroll
                        for (int m = 0; m < 3; ++m) w[cc] += ul.u[cc][tidx(tt, m)] * dl[tt][m];
                if constexpr (BS == 7)
#pragma unroll
                    for (int cc = 0; cc < 2; ++cc) w[cc] += ul.u[cc][6] * dw[0] + ul.u[cc][13] * dw[1];
                const double l0 = ul.h[0][0] * w[0] + ul.h[0][1] * w[1];
                const double l1 = ul.h[1][0] * w[0] + ul.h[1][1] * w[1];
                int un = 0;
                if (nc == 2) un = (l0 < -1e-8 && l1 < -1e-8) ? 1 : (l0 < -1e-8) ? 2 : (l1 < -1e-8) ? 3 : 0;
                else if (l0 < -1e-8) un = 2 + end[0];
                if (un) {
                    if (un == 1 || un == 2) y[0] = 2;
                    if (un == 1 || un == 3) y[1] = 2;
                    c = 2;
                }
            }
        }
    }
    code[e * 4] = c;
    code[e * 4 + 1] = y[0];
    code[e * 4 + 2] = y[1];
    tau[e] = t;
    if (c) atomicMin(&st->first, int32_t(e));
}

__global__ void k_plastic_init(PlasticStatus* st, long n) {
    st->first = int32_t(n);
    st->code = 0;
    st->tau = 1.0;
}

// The members up to the first with a code take their pass: ef_i <- efc and
// their flags; that member's code and tau go to st.  The members after it
// keep ef_i and their flags, as the reference's loop never reaches them.
__global__ void k_frame_plastic_commit(long n, const double* __restrict__ efc, const int32_t* __restrict__ code,
                                       const double* __restrict__ tau, PlasticStatus* __restrict__ st,
                                       int32_t* __restrict__ yld, double* __restrict__ ef_i, long ntr) {
    const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (e >= n) return;
    const long first = st->first;
    if (e > first) return;
#pragma unroll
    for (int i = 0; i < 14; ++i) ef_i[ntr * 2 + e * 14 + i] = efc[e * 14 + i];
    yld[e * 2] = code[e * 4 + 1];
    yld[e * 2 + 1] = code[e * 4 + 2];
    if (e == first) {
        st->code = code[e * 4];
        st->tau = tau[e];
    }
}

// yldflag 2 -> 0 at a commit (main.c:2107-2112, 3086-3091, 3887-3892)
__global__ void k_unloaded_elastic(int32_t* __restrict__ yld, long n2) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n2 && yld[i] == 2) yld[i] = 0;
}

// misc.c:122-144: member length and triad from the current chord and the
// fixed input auxiliary point; the member ends sit at joint + offset (fixed
// in global axes; ofs null when no member has offsets).
__global__ void k_update_frame(const int32_t* __restrict__ minc, const double* __restrict__ x,
                               const double* __restrict__ auxpt, double* __restrict__ defllen,
                               double* __restrict__ c1, double* __restrict__ c2,
                               double* __restrict__ c3, long n, long ntr, const double* __restrict__ ofs) {
    long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (e >= n) return;
    const double* xa = x + (minc[ntr * 2 + e * 2] - 1) * 3;
    const double* xb = x + (minc[ntr * 2 + e * 2 + 1] - 1) * 3;
    double a[3] = {xa[0], xa[1], xa[2]}, b[3] = {xb[0], xb[1], xb[2]};
    if (ofs)
        for (int k = 0; k < 3; ++k) {
            a[k] += ofs[e * 6 + k];
            b[k] += ofs[e * 6 + 3 + k];
        }
    double el[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    double L = sqrt(el[0] * el[0] + el[1] * el[1] + el[2] * el[2]);
    defllen[ntr + e] = L;
    double lx[3] = {el[0] / L, el[1] / L, el[2] / L};
    double tp[3] = {auxpt[e * 3] - a[0], auxpt[e * 3 + 1] - a[1], auxpt[e * 3 + 2] - a[2]};
    double lz[3] = {lx[1] * tp[2] - lx[2] * tp[1], lx[2] * tp[0] - lx[0] * tp[2], lx[0] * tp[1] - lx[1] * tp[0]};
    double nz = sqrt(lz[0] * lz[0] + lz[1] * lz[1] + lz[2] * lz[2]);
    lz[0] /= nz; lz[1] /= nz; lz[2] /= nz;
    double ly[3] = {lz[1] * lx[2] - lz[2] * lx[1], lz[2] * lx[0] - lz[0] * lx[2], lz[0] * lx[1] - lz[1] * lx[0]};
    double ny = sqrt(ly[0] * ly[0] + ly[1] * ly[1] + ly[2] * ly[2]);
    const long t = ntr + e * 3;
    for (int k = 0; k < 3; ++k) {
        c1[t + k] = lx[k];
        c2[t + k] = ly[k] / ny;
        c3[t + k] = lz[k];
    }
}

}  // namespace

namespace {
PlasticIn plastic_in(const Dev& d) { return PlasticIn{d.yld, d.yield, d.zstrong, d.zweak}; }
MemberXtr member_xtr(const Dev& d) { return MemberXtr{d.ofs, d.relmask}; }
bool has_xtr(const Dev& d) { return d.ofs || d.relmask; }

// f(std::integral_constant<int, BS>) for the joint slots of the model: 7 when
// a warping DOF is free, else 6 (frames never run with 3)
template <class F>
void with_bs(const Dev& d, F&& f) {
    if (d.bs == 7) f(std::integral_constant<int, 7>{});
    else if (d.bs == 6) f(std::integral_constant<int, 6>{});
    else throw DeviceError("frame members need 6 or 7 slots per joint");
}
}  // namespace

// The geometric terms are on for ANAFLAG 2 and 3 (frame.c:266-277).
void launch_frame_tangent(Dev& d, int g) {
    const int geo = d.anaflag == 2 || d.anaflag == 3;
    with_bs(d, [&](auto B) {
        constexpr int BS = decltype(B)::value;
        for (size_t c = 0; c + 1 < d.col_ptr_fr.size(); ++c) {
            int n = d.col_ptr_fr[c + 1] - d.col_ptr_fr[c];
            if (!n) continue;
            auto go = [&](auto kern, PlasticIn pin) {
                hipLaunchKernelGGL(kern, dim3(nblk(n)), dim3(TPB), 0, d.s, d.col_el_fr + d.col_ptr_fr[c], n,
                                   d.blk_fr, d.c1[g], d.c2[g], d.c3[g], d.emod, d.gmod, d.carea, d.llen,
                                   d.istrong, d.iweak, d.ipolar, d.iwarp, d.defllen[g], d.ef[g], d.K, d.ntr, geo,
                                   pin, member_xtr(d));
            };
            if (d.anaflag == 3 && has_xtr(d)) go(k_frame_tangent<true, BS, true>, plastic_in(d));
            else if (d.anaflag == 3) go(k_frame_tangent<true, BS, false>, plastic_in(d));
            else if (has_xtr(d)) go(k_frame_tangent<false, BS, true>, PlasticIn{});
            else go(k_frame_tangent<false, BS, false>, PlasticIn{});
        }
    });
    CB_CHECK(hipGetLastError());
}

void launch_frame_tangent_gather(Dev& d, int g) {
    if (has_xtr(d))
        throw DeviceError("gathered frame tangent: members with end offsets or releases take the colour launches");
    const int geo = d.anaflag == 2 || d.anaflag == 3;
    with_bs(d, [&](auto B) {
        constexpr int BS = decltype(B)::value;
        const long nth = (BS == 7 ? 3 : 2) * d.nnzb;
        if (d.anaflag == 3)
            hipLaunchKernelGGL((k_frame_tangent_gather<false, true, BS>), dim3(nblk(nth)), dim3(TPB), 0, d.s,
                               d.gat_ptr_fr, d.gat_item_fr, d.blkrow, d.colidx, d.eq_of_slot, d.nnzb, d.c1[g],
                               d.c2[g], d.c3[g], d.emod, d.gmod, d.carea, d.llen, d.istrong, d.iweak, d.ipolar,
                               d.iwarp, d.defllen[g], d.ef[g], d.K, geo, 0, nullptr, plastic_in(d));
        else
            hipLaunchKernelGGL((k_frame_tangent_gather<false, false, BS>), dim3(nblk(nth)), dim3(TPB), 0, d.s,
                               d.gat_ptr_fr, d.gat_item_fr, d.blkrow, d.colidx, d.eq_of_slot, d.nnzb, d.c1[g],
                               d.c2[g], d.c3[g], d.emod, d.gmod, d.carea, d.llen, d.istrong, d.iweak, d.ipolar,
                               d.iwarp, d.defllen[g], d.ef[g], d.K, geo, 0, nullptr, PlasticIn{});
    });
    CB_CHECK(hipGetLastError());
}

void launch_frame_tangent_band(Dev& d, int g, double* tiles, int w, int* bad, bool fp32) {
    if (d.bs != 6 || d.ntr || d.nsh) throw DeviceError("frame tangent into the band: frame-only models, warping fixed");
    if (d.anaflag == 3) throw DeviceError("frame tangent into the band: elastic (ANAFLAG 2) members only");
    if (has_xtr(d)) throw DeviceError("frame tangent into the band: members without end offsets or releases only");
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(nblk(2 * d.nnzb)), dim3(TPB), 0, d.s, d.gat_ptr_fr, d.gat_item_fr, d.blkrow,
                           d.colidx, d.eq_of_slot, d.nnzb, d.c1[g], d.c2[g], d.c3[g], d.emod, d.gmod, d.carea,
                           d.llen, d.istrong, d.iweak, d.ipolar, d.iwarp, d.defllen[g], d.ef[g], tiles,
                           int(d.anaflag == 2), w, bad, PlasticIn{});
    };
    if (fp32) go(k_frame_tangent_gather<true, false, 6, float>);
    else go(k_frame_tangent_gather<true, false, 6, double>);
    CB_CHECK(hipGetLastError());
}

// forces_fr under ANAFLAG 3: every member's pass at once, then the members up
// to the first that returns take it and their end forces before that member
// are gathered into f (the reference's loop stops at a return).
static void launch_frame_forces_plastic(Dev& d, double* f, const double* dd) {
    with_bs(d, [&](auto B) {
        constexpr int BS = decltype(B)::value;
        hipLaunchKernelGGL(k_plastic_init, dim3(1), dim3(1), 0, d.s, d.pl_st, d.nfr);
        auto go = [&](auto kern) {
            hipLaunchKernelGGL(kern, dim3(nblk(d.nfr, 64)), dim3(64), 0, d.s, d.nfr, d.mcode, dd, d.c1[G_IP],
                               d.c2[G_IP], d.c3[G_IP], d.c1[G_I], d.c2[G_I], d.c3[G_I], d.emod, d.gmod, d.carea,
                               d.llen, d.istrong, d.iweak, d.ipolar, d.iwarp, d.defllen[G_IP], d.ef[G_IP],
                               plastic_in(d), d.pl_efc, d.fh, d.pl_code, d.pl_tau, d.pl_st, d.ntr, member_xtr(d));
        };
        if (has_xtr(d)) go(k_frame_forces_plastic<BS, true>);
        else go(k_frame_forces_plastic<BS, false>);
        hipLaunchKernelGGL(k_frame_plastic_commit, dim3(nblk(d.nfr, 256)), dim3(256), 0, d.s, d.nfr, d.pl_efc,
                           d.pl_code, d.pl_tau, d.pl_st, d.yld, d.ef[G_I], d.ntr);
        hipLaunchKernelGGL((k_frame_forces_gather<BS>), dim3(nblk(d.nslot, 256)), dim3(256), 0, d.s,
                           d.jnt_ptr_fr, d.jnt_item_fr, d.eq_of_slot, d.nslot, d.fh, f, d.pl_st);
    });
    CB_CHECK(hipGetLastError());
}

int frame_plastic_result(Dev& d, double* tau) {
    if (d.anaflag != 3 || !d.nfr) {
        *tau = 1.0;
        return 0;
    }
    d2h(d.pl_host, d.pl_st, 1, d.s);
    CB_CHECK(hipStreamSynchronize(d.s));
    *tau = d.pl_host->tau;
    return d.pl_host->code;
}

void launch_unloaded_elastic(Dev& d) {
    if (d.anaflag != 3 || !d.nfr) return;
    hipLaunchKernelGGL(k_unloaded_elastic, dim3(nblk(d.nfr * 2, 256)), dim3(256), 0, d.s, d.yld, d.nfr * 2);
    CB_CHECK(hipGetLastError());
}

void launch_frame_forces(Dev& d, double* f, const double* dd) {
    if (d.anaflag == 3) {
        launch_frame_forces_plastic(d, f, dd);
        return;
    }
    with_bs(d, [&](auto B) {
        constexpr int BS = decltype(B)::value;
        if (!d.ntr && !d.nsh) {
            auto go = [&](auto kern) {
                hipLaunchKernelGGL(kern, dim3(nblk(d.nfr)), dim3(TPB), 0, d.s, d.nfr, d.mcode, dd, d.c1[G_IP],
                                   d.c2[G_IP], d.c3[G_IP], d.c1[G_I], d.c2[G_I], d.c3[G_I], d.emod, d.gmod,
                                   d.carea, d.llen, d.istrong, d.iweak, d.ipolar, d.iwarp, d.defllen[G_IP],
                                   d.ef[G_IP], d.ef[G_I], d.fh, int(d.anaflag == 2), member_xtr(d));
            };
            if (has_xtr(d)) go(k_frame_forces_members<BS, true>);
            else go(k_frame_forces_members<BS, false>);
            hipLaunchKernelGGL((k_frame_forces_gather<BS>), dim3(nblk(d.nslot, 256)), dim3(256), 0, d.s,
                               d.jnt_ptr_fr, d.jnt_item_fr, d.eq_of_slot, d.nslot, d.fh, f, nullptr);
            return;
        }
        for (size_t c = 0; c + 1 < d.col_ptr_fr.size(); ++c) {
            int n = d.col_ptr_fr[c + 1] - d.col_ptr_fr[c];
            if (!n) continue;
            auto go = [&](auto kern) {
                hipLaunchKernelGGL(kern, dim3(nblk(n)), dim3(TPB), 0, d.s, d.col_el_fr + d.col_ptr_fr[c], n,
                                   d.mcode, dd, d.c1[G_IP], d.c2[G_IP], d.c3[G_IP], d.c1[G_I], d.c2[G_I],
                                   d.c3[G_I], d.emod, d.gmod, d.carea, d.llen, d.istrong, d.iweak, d.ipolar,
                                   d.iwarp, d.defllen[G_IP], d.ef[G_IP], d.ef[G_I], f, d.ntr,
                                   int(d.anaflag == 2), member_xtr(d));
            };
            if (has_xtr(d)) go(k_frame_forces<BS, true>);
            else go(k_frame_forces<BS, false>);
        }
    });
    CB_CHECK(hipGetLastError());
}

void launch_frame_update(Dev& d, int gx) {
    if (!d.nfr) return;
    hipLaunchKernelGGL(k_update_frame, dim3(nblk(d.nfr, 256)), dim3(256), 0, d.s, d.minc, d.x[gx],
                       d.auxpt, d.defllen[G_I], d.c1[G_I], d.c2[G_I], d.c3[G_I], d.nfr, d.ntr, d.ofs);
    CB_CHECK(hipGetLastError());
}

}  // namespace cb
