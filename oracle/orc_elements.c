/*
 * orc_elements.c -- oracle: truss and CST+DKT shell element routines and the
 * geometry update, restated from the reference.
 *
 * TEST INFRASTRUCTURE ONLY (see cb_oracle.h).
 *
 * Shell: the DKT bending stiffness is evaluated by exact 3-point (mid-side)
 * quadrature of the Batoz-Bathe-Ho (1980) curvature interpolation instead of
 * the reference's explicit Batoz (1982) product form (shell.c:533-658).  Both
 * integrate the same quadratic integrand exactly, so they agree to rounding;
 * the parity tests pin this against the reference routine compiled from source.
 * The 18x18 transformation T^T k T (misc.c:41-69) is evaluated block by block:
 * T is block-diagonal, so each output entry of the reference's dense product has
 * exactly three non-zero terms, summed here in the same order.
 */
#include "cb_oracle.h"

#include <math.h>
#include <string.h>

static inline double dot3(const double *a, const double *b)
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

static void cross3(const double *a, const double *b, double *c, int unit)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
    if (unit) {
        double l = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
        c[0] /= l; c[1] /= l; c[2] /= l;
    }
}

/* Skyline accumulation of the upper triangle of an element matrix
 * (truss.c:167-189, shell.c:307-329). */
static void skyline_add(double *ss, const long *maxa, const long *mc, int n, const double *K)
{
    for (int b = 0; b < n; ++b) {
        long j = mc[b];
        if (!j) continue;
        for (int a = 0; a <= b; ++a) {
            long i = mc[a];
            if (!i) continue;
            long addr = (i > j) ? maxa[i - 1] + (i - j) : maxa[j - 1] + (j - i);
            ss[addr - 1] += K[a * n + b];
        }
    }
}

/* ---------------------------------------------------------------- truss --- */

/* truss.c:82-204 (ANAFLAG 2): k = EA L^2/L0^3 [1 -1;-1 1] rotated to 6x6, plus
 * the geometric term N/L on the translation diagonal blocks.  ANAFLAG 3: the
 * same k less its plastic part when (N/Py)^2 > 1 + tol (stiffm_tr,
 * truss.c:206-229: k - (k g)(k g)^T / g^T k g with g = 2 (N1, N2) / Py^2),
 * and the geometric term only on a bar at its squash load,
 * (N/Py)^2 >= 1 - tol (truss.c:155-166). */
#define TR_PHITOL 1e-4
void orc_stiff_tr(const orc_dims *d, double *ss, const double *emod, const double *carea,
                  const double *llength, const double *defllen_ip, const double *yield,
                  const double *c1_ip, const double *c2_ip, const double *c3_ip,
                  const double *ef_ip, const long *maxa, const long *mcode)
{
    for (long n = 0; n < d->NE_TR; ++n) {
        double ka;
        if (d->ANAFLAG == 1)
            ka = emod[n] * carea[n] / llength[n];
        else
            ka = emod[n] * carea[n] * pow(defllen_ip[n], 2) / pow(llength[n], 3);
        double k2[2][2] = {{ka, -ka}, {-ka, ka}};
        double Py = 0;
        if (d->ANAFLAG == 3) {
            Py = carea[n] * yield[n];
            double r = ef_ip[n * 2] / Py;
            if (r * r > 1 + TR_PHITOL) {
                double g0 = 2 * ef_ip[n * 2] / (Py * Py), g1 = 2 * ef_ip[n * 2 + 1] / (Py * Py);
                double u0 = k2[0][0] * g0 + k2[0][1] * g1, u1 = k2[1][0] * g0 + k2[1][1] * g1;
                double gkg = u0 * g0 + u1 * g1;
                double r00 = u0 * u0 / gkg, r01 = u0 * u1 / gkg, r11 = u1 * u1 / gkg;
                k2[0][0] -= r00; k2[0][1] -= r01; k2[1][0] -= r01; k2[1][1] -= r11;
            }
        }
        double c[3] = {c1_ip[n], c2_ip[n], c3_ip[n]};
        double K[36];
        /* T^T k T with T = [c 0; 0 c]; the two-term sums keep the order of
           truss.c:133-150 */
        for (int a = 0; a < 6; ++a) {
            double ta0 = a < 3 ? c[a] : 0.0, ta1 = a < 3 ? 0.0 : c[a - 3];
            double tk0 = ta0 * k2[0][0] + ta1 * k2[1][0];
            double tk1 = ta0 * k2[0][1] + ta1 * k2[1][1];
            for (int b = 0; b < 6; ++b) {
                double t0b = (b < 3) ? c[b] : 0.0, t1b = (b < 3) ? 0.0 : c[b - 3];
                K[a * 6 + b] = tk0 * t0b + tk1 * t1b;
            }
        }
        int geo = d->ANAFLAG == 2;
        if (d->ANAFLAG == 3) {
            double r = ef_ip[n * 2] / Py;
            geo = r * r >= 1 - TR_PHITOL;
        }
        if (geo) {
            double g = ef_ip[n * 2] / defllen_ip[n];
            for (int a = 0; a < 6; ++a) K[a * 6 + a] += g;
            for (int a = 0; a < 3; ++a) {
                K[a * 6 + a + 3] -= g;
                K[(a + 3) * 6 + a] -= g;
            }
        }
        skyline_add(ss, maxa, mcode + n * 6, 6, K);
    }
}

/* truss.c:327-376: N = EA (eps + eps^2/2) L/L0, f -= N c at end 1, += at end 2;
 * under ANAFLAG 3, N at or beyond the squash load ((N/Py)^2 >= 1 - tol) is
 * set to +-Py (truss.c:336-347): the strain keeps no plastic part. */
void orc_forces_tr(const orc_dims *d, double *f, double *ef_i, const double *emod,
                   const double *carea, const double *llength, const double *defllen_i,
                   const double *yield, const double *c1_i, const double *c2_i,
                   const double *c3_i, const long *mcode)
{
    for (long n = 0; n < d->NE_TR; ++n) {
        double eps = (defllen_i[n] - llength[n]) / llength[n];
        double N = emod[n] * carea[n] * (eps + 0.5 * pow(eps, 2)) * defllen_i[n] / llength[n];
        if (d->ANAFLAG == 3) {
            double Py = yield[n] * carea[n], r = N / Py;
            if (r * r >= 1 - TR_PHITOL) N = N < 0 ? -Py : Py;
        }
        ef_i[n * 2] = N;
        ef_i[n * 2 + 1] = -N;
        double c[3] = {c1_i[n], c2_i[n], c3_i[n]};
        for (int j = 0; j < 6; ++j) {
            long k = mcode[n * 6 + j];
            if (k) f[k - 1] -= ef_i[n * 2 + j / 3] * c[j % 3];
        }
    }
}

/* ---------------------------------------------------------------- shell --- */

/* Local in-plane coordinates (x2', x3', y3') of the vertices in triad c1/c2
 * (mem_coord, shell.c:2402-2446). */
static void shell_local_xy(const double *p1, const double *p2, const double *p3,
                           const double *c1, const double *c2, double out[3])
{
    double e2[3], e3[3];
    for (int k = 0; k < 3; ++k) { e2[k] = p2[k] - p1[k]; e3[k] = p3[k] - p1[k]; }
    /* the reference forms T*X with a zero first column; the products summed are
       the same three */
    out[0] = c1[0] * e2[0] + c1[1] * e2[1] + c1[2] * e2[2];
    out[1] = c1[0] * e3[0] + c1[1] * e3[1] + c1[2] * e3[2];
    out[2] = c2[0] * e3[0] + c2[1] * e3[1] + c2[2] * e3[2];
}

/* Shape-function gradients of the constant-strain triangle with vertex 1 at the
 * origin, vertex 2 at (x2,0), vertex 3 at (x3,y3), divided by 2A. */
static void cst_gradients(const double xl[3], double twoA, double gx[3], double gy[3])
{
    gx[0] = -(xl[2] / twoA);
    gx[1] = xl[2] / twoA;
    gx[2] = 0.0;
    gy[0] = (xl[1] - xl[0]) / twoA;
    gy[1] = -(xl[1] / twoA);
    gy[2] = xl[0] / twoA;
}

/* Plane-stress membrane stiffness t A Bm^T C Bm over (u1,v1,u2,v2,u3,v3)
 * (stiffe_m_sh, shell.c:487-531). */
static void cst_membrane(double E, double nu, double t, const double xl[3], double A,
                         double km[6][6])
{
    double gx[3], gy[3];
    cst_gradients(xl, 2 * A, gx, gy);
    double c0 = E / (1 - pow(nu, 2));
    double C[3][3] = {{c0, c0 * nu, 0}, {c0 * nu, c0, 0}, {0, 0, c0 * (1 - nu) / 2}};
    double Bm[3][6] = {{gx[0], 0, gx[1], 0, gx[2], 0},
                       {0, gy[0], 0, gy[1], 0, gy[2]},
                       {gy[0], gx[0], gy[1], gx[1], gy[2], gx[2]}};
    double BC[6][3];
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0;
            for (int k = 0; k < 3; ++k) s += Bm[k][i] * C[k][j];
            BC[i][j] = s;
        }
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) {
            double s = 0;
            for (int k = 0; k < 3; ++k) s += BC[i][k] * Bm[k][j];
            km[i][j] = t * A * s;
        }
}

/* DKT bending stiffness over (w, theta_x, theta_y) at each vertex.  Rotations of
 * the normal are beta_x = Hx.U, beta_y = Hy.U with the Batoz-Bathe-Ho quadratic
 * interpolation; curvatures are linear, so the 3 mid-side points integrate
 * B^T D B exactly.  Squared side lengths come from the 3-D side lengths sl, as
 * in shell.c:560-562. */
static void dkt_bending(double E, double nu, double t, const double xl[3], double A,
                        const double sl[3], double kb[9][9])
{
    const double x2 = xl[0], x3 = xl[1], y3 = xl[2];
    /* sides k = 4 (2-3), 5 (3-1), 6 (1-2): x_ij, y_ij = (x_i - x_j, y_i - y_j) */
    const double xs[3] = {x2 - x3, x3, -x2};
    const double ys[3] = {-y3, y3, 0.0};
    const double l2[3] = {pow(sl[1], 2), pow(sl[2], 2), pow(sl[0], 2)};
    double a[3], b[3], c[3], dd[3], e[3];
    for (int s = 0; s < 3; ++s) {
        a[s] = -xs[s] / l2[s];
        b[s] = 0.75 * xs[s] * ys[s] / l2[s];
        c[s] = (0.25 * xs[s] * xs[s] - 0.5 * ys[s] * ys[s]) / l2[s];
        dd[s] = -ys[s] / l2[s];
        e[s] = (0.25 * ys[s] * ys[s] - 0.5 * xs[s] * xs[s]) / l2[s];
    }
    const double x31 = x3, y31 = y3, x12 = -x2, y12 = 0.0;
    double D = E * pow(t, 3) / (12 * (1 - pow(nu, 2)));
    double Db[3][3] = {{D, D * nu, 0}, {D * nu, D, 0}, {0, 0, D * (1 - nu) / 2}};
    static const double gp[3][2] = {{0.5, 0.0}, {0.5, 0.5}, {0.0, 0.5}};
    memset(kb, 0, sizeof(double) * 81);
    for (int g = 0; g < 3; ++g) {
        double xi = gp[g][0], eta = gp[g][1], L = 1 - xi - eta;
        /* derivatives of the six quadratic functions: N1..N3 vertices, N4 on
           side 2-3, N5 on 3-1, N6 on 1-2 */
        double dN[6][2] = {{-(4 * L - 1), -(4 * L - 1)},
                           {4 * xi - 1, 0},
                           {0, 4 * eta - 1},
                           {4 * eta, 4 * xi},
                           {-4 * eta, 4 * L - 4 * eta},
                           {4 * L - 4 * xi, -4 * xi}};
        double Hx[9][2], Hy[9][2];
        for (int r = 0; r < 2; ++r) {
            /* vertex v uses the two sides meeting at it: side "next" (opposite
               order) and side "prev" as in the BBH tables */
            const double *N4 = &dN[3][r], *N5 = &dN[4][r], *N6 = &dN[5][r];
            double n4 = *N4, n5 = *N5, n6 = *N6;
            double n1 = dN[0][r], n2 = dN[1][r], n3 = dN[2][r];
            /* indices into a..e: side 4 -> 0, 5 -> 1, 6 -> 2 */
            Hx[0][r] = 1.5 * (a[2] * n6 - a[1] * n5);
            Hx[1][r] = b[1] * n5 + b[2] * n6;
            Hx[2][r] = n1 - c[1] * n5 - c[2] * n6;
            Hx[3][r] = 1.5 * (a[0] * n4 - a[2] * n6);
            Hx[4][r] = b[2] * n6 + b[0] * n4;
            Hx[5][r] = n2 - c[2] * n6 - c[0] * n4;
            Hx[6][r] = 1.5 * (a[1] * n5 - a[0] * n4);
            Hx[7][r] = b[0] * n4 + b[1] * n5;
            Hx[8][r] = n3 - c[0] * n4 - c[1] * n5;
            Hy[0][r] = 1.5 * (dd[2] * n6 - dd[1] * n5);
            Hy[1][r] = -n1 + e[1] * n5 + e[2] * n6;
            Hy[2][r] = -Hx[1][r];
            Hy[3][r] = 1.5 * (dd[0] * n4 - dd[2] * n6);
            Hy[4][r] = -n2 + e[2] * n6 + e[0] * n4;
            Hy[5][r] = -Hx[4][r];
            Hy[6][r] = 1.5 * (dd[1] * n5 - dd[0] * n4);
            Hy[7][r] = -n3 + e[0] * n4 + e[1] * n5;
            Hy[8][r] = -Hx[7][r];
        }
        double B[3][9];
        for (int i = 0; i < 9; ++i) {
            B[0][i] = y31 * Hx[i][0] + y12 * Hx[i][1];
            B[1][i] = -x31 * Hy[i][0] - x12 * Hy[i][1];
            B[2][i] = -x31 * Hx[i][0] - x12 * Hx[i][1] + y31 * Hy[i][0] + y12 * Hy[i][1];
        }
        double DB[3][9];
        for (int p = 0; p < 3; ++p)
            for (int i = 0; i < 9; ++i)
                DB[p][i] = Db[p][0] * B[0][i] + Db[p][1] * B[1][i] + Db[p][2] * B[2][i];
        for (int i = 0; i < 9; ++i)
            for (int j = 0; j < 9; ++j)
                kb[i][j] += (B[0][i] * DB[0][j] + B[1][i] * DB[1][j] + B[2][i] * DB[2][j]) / 6.0;
    }
    for (int i = 0; i < 9; ++i)
        for (int j = 0; j < 9; ++j) kb[i][j] /= 2 * A;
}

static const int MEM_DOF[6] = {0, 1, 6, 7, 12, 13};
static const int BEND_DOF[9] = {2, 3, 4, 8, 9, 10, 14, 15, 16};

/* Local 18x18 tangent: membrane + bending + drilling (kb/1e4) + membrane
 * geometric stiffness (stiffe_sh shell.c:345-485, stiffg_sh shell.c:660-840). */
static void shell_local_tangent(double E, double nu, double t, const double xl[3], double A0,
                                const double sl[3], double A_ip, const double dm[6],
                                int geometric, double k[18][18])
{
    double km[6][6], kb[9][9];
    memset(k, 0, sizeof(double) * 324);
    cst_membrane(E, nu, t, xl, A0, km);
    dkt_bending(E, nu, t, xl, A0, sl, kb);
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) k[MEM_DOF[i]][MEM_DOF[j]] = km[i][j];
    for (int i = 0; i < 9; ++i)
        for (int j = 0; j < 9; ++j) k[BEND_DOF[i]][BEND_DOF[j]] = kb[i][j];
    k[5][5] = kb[1][1] / 10000;
    k[11][11] = kb[4][4] / 10000;
    k[17][17] = kb[7][7] / 10000;
    if (!geometric) return;
    /* membrane force from the deformed-area strain of dm, then
       kg[(a,i),(b,i)] = A_ip grad_a . N . grad_b for each translation i */
    double gx[3], gy[3];
    cst_gradients(xl, 2 * A_ip, gx, gy);
    double c0 = E / (1 - pow(nu, 2));
    double C[3][3] = {{c0, c0 * nu, 0}, {c0 * nu, c0, 0}, {0, 0, c0 * (1 - nu) / 2}};
    double Bm[3][6] = {{gx[0], 0, gx[1], 0, gx[2], 0},
                       {0, gy[0], 0, gy[1], 0, gy[2]},
                       {gy[0], gx[0], gy[1], gx[1], gy[2], gx[2]}};
    double Nm[3];
    for (int i = 0; i < 3; ++i) {
        double CB[6];
        for (int j = 0; j < 6; ++j) CB[j] = C[i][0] * Bm[0][j] + C[i][1] * Bm[1][j] + C[i][2] * Bm[2][j];
        double s = 0;
        for (int j = 0; j < 6; ++j) s += CB[j] * dm[j];
        Nm[i] = t * s;
    }
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            /* (Bnl^T N) then (.) Bnl in the order of shell.c:735-752 */
            double r0 = gx[a] * Nm[0] + gy[a] * Nm[2];
            double r1 = gx[a] * Nm[2] + gy[a] * Nm[1];
            double g = A_ip * (r0 * gx[b] + r1 * gy[b]);
            for (int i = 0; i < 3; ++i) k[6 * a + i][6 * b + i] += g;
        }
}

/* K = T^T k T with T = blockdiag of six triads (rows c1, c2, c3). */
static void rotate_18(const double k[18][18], const double *c1, const double *c2,
                      const double *c3, double *K)
{
    const double *R[3] = {c1, c2, c3};
    double tmp[18][18];
    for (int i = 0; i < 18; ++i) {
        int bi = i / 3 * 3, ci = i % 3;
        for (int j = 0; j < 18; ++j)
            tmp[i][j] = R[0][ci] * k[bi][j] + R[1][ci] * k[bi + 1][j] + R[2][ci] * k[bi + 2][j];
    }
    for (int i = 0; i < 18; ++i)
        for (int j = 0; j < 18; ++j) {
            int bj = j / 3 * 3, cj = j % 3;
            K[i * 18 + j] = tmp[i][bj] * R[0][cj] + tmp[i][bj + 1] * R[1][cj] + tmp[i][bj + 2] * R[2][cj];
        }
}

void orc_shell_elem_tangent(double E, double nu, double t, const double xl[3], double A0,
                            const double sl[3], double A_ip, const double dm[6],
                            const double c1[3], const double c2[3], const double c3[3],
                            int geometric, double K[18 * 18])
{
    double k[18][18];
    shell_local_tangent(E, nu, t, xl, A0, sl, A_ip, dm, geometric, k);
    rotate_18(k, c1, c2, c3, K);
}

/* stiff_sh (shell.c:110-344), elastic branches. */
void orc_stiff_sh(const orc_dims *d, double *ss, const double *emod, const double *nu,
                  const double *x_temp, const double *xlocal, const double *thick,
                  const double *farea, const double *deffarea_ip, const double *slength,
                  const double *c1_ip, const double *c2_ip, const double *c3_ip,
                  const long *maxa, const long *minc, const long *mcode)
{
    long poff = d->NE_TR + d->NE_FR, ioff = d->NE_TR * 2 + d->NE_FR * 2;
    long toff = d->NE_TR + d->NE_FR * 3, moff = d->NE_TR * 6 + d->NE_FR * 14;
    double K[324];
    for (long n = 0; n < d->NE_SH; ++n) {
        const double *ci1 = c1_ip + toff + n * 3, *ci2 = c2_ip + toff + n * 3, *ci3 = c3_ip + toff + n * 3;
        double dm[6] = {0, 0, 0, 0, 0, 0};
        int geo = (d->ANAFLAG == 2);
        if (geo) {
            const long *v = minc + ioff + n * 3;
            double xl[3];
            shell_local_xy(x_temp + (v[0] - 1) * 3, x_temp + (v[1] - 1) * 3, x_temp + (v[2] - 1) * 3,
                           ci1, ci2, xl);
            dm[2] = xl[0] - xlocal[n * 3];
            dm[4] = xl[1] - xlocal[n * 3 + 1];
            dm[5] = xl[2] - xlocal[n * 3 + 2];
        }
        orc_shell_elem_tangent(emod[poff + n], nu[n], thick[n], xlocal + n * 3, farea[n],
                               slength + n * 3, deffarea_ip[n], dm, ci1, ci2, ci3, geo, K);
        skyline_add(ss, maxa, mcode + moff + n * 18, 18, K);
    }
}

/* forces_sh (shell.c:1593-1786, 2305-2400), ANAFLAG 2: membrane force in total
 * form from the current triad, bending in incremental form from the previous
 * triad, the sum rotated by T_i T_ip^T; the membrane entries of ef_ip are
 * cleared in place (shell.c:1773). */
void orc_forces_sh(const orc_dims *d, double *f, double *ef_ip, double *ef_i,
                   const double *dd, const double *x_temp, const double *emod,
                   const double *nu, const double *xlocal, const double *thick,
                   const double *farea, const double *slength,
                   const double *c1_ip, const double *c2_ip, const double *c3_ip,
                   const double *c1_i, const double *c2_i, const double *c3_i,
                   const long *minc, const long *mcode)
{
    long poff = d->NE_TR + d->NE_FR, ioff = d->NE_TR * 2 + d->NE_FR * 2;
    long toff = d->NE_TR + d->NE_FR * 3, moff = d->NE_TR * 6 + d->NE_FR * 14;
    long eoff = d->NE_TR * 2 + d->NE_FR * 14;
    for (long n = 0; n < d->NE_SH; ++n) {
        const double *Rp[3] = {c1_ip + toff + n * 3, c2_ip + toff + n * 3, c3_ip + toff + n * 3};
        const double *Rc[3] = {c1_i + toff + n * 3, c2_i + toff + n * 3, c3_i + toff + n * 3};
        const long *mc = mcode + moff + n * 18;
        double *efp = ef_ip + eoff + n * 18, *efc = ef_i + eoff + n * 18;
        double km[6][6], kb[9][9];
        cst_membrane(emod[poff + n], nu[n], thick[n], xlocal + n * 3, farea[n], km);
        dkt_bending(emod[poff + n], nu[n], thick[n], xlocal + n * 3, farea[n], slength + n * 3, kb);

        const long *v = minc + ioff + n * 3;
        double xl[3], dm[6] = {0, 0, 0, 0, 0, 0};
        shell_local_xy(x_temp + (v[0] - 1) * 3, x_temp + (v[1] - 1) * 3, x_temp + (v[2] - 1) * 3,
                       Rc[0], Rc[1], xl);
        dm[2] = xl[0] - xlocal[n * 3];
        dm[4] = xl[1] - xlocal[n * 3 + 1];
        dm[5] = xl[2] - xlocal[n * 3 + 2];

        double DD[18];
        for (int i = 0; i < 18; ++i) DD[i] = mc[i] ? dd[mc[i] - 1] : 0.0;
        /* bending increment in the previous triad: (w, tx, ty) per vertex */
        double ddb[9];
        for (int i = 0; i < 9; ++i) {
            int r = BEND_DOF[i], blk = r / 3 * 3, row = r % 3;
            ddb[i] = Rp[row][0] * DD[blk] + Rp[row][1] * DD[blk + 1] + Rp[row][2] * DD[blk + 2];
        }
        double eft[18], def[18];
        memset(eft, 0, sizeof eft);
        memset(def, 0, sizeof def);
        for (int i = 0; i < 6; ++i) {
            double s = 0;
            for (int j = 0; j < 6; ++j) s += km[i][j] * dm[j];
            efp[MEM_DOF[i]] = 0;
            eft[MEM_DOF[i]] = s;
        }
        for (int i = 0; i < 9; ++i) {
            double s = 0;
            for (int j = 0; j < 9; ++j) s += kb[i][j] * ddb[j];
            def[BEND_DOF[i]] = s;
        }
        /* rotation previous -> current triad: Q[a][b] = Rc_a . Rp_b */
        double Q[3][3];
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                Q[a][b] = Rc[a][0] * Rp[b][0] + Rc[a][1] * Rp[b][1] + Rc[a][2] * Rp[b][2];
        for (int i = 0; i < 18; ++i) {
            int blk = i / 3 * 3, row = i % 3;
            double s = Q[row][0] * (def[blk] + efp[blk]) + Q[row][1] * (def[blk + 1] + efp[blk + 1]) +
                       Q[row][2] * (def[blk + 2] + efp[blk + 2]);
            efc[i] = eft[i] + s;
        }
        for (int i = 0; i < 18; ++i) {
            int blk = i / 3 * 3, col = i % 3;
            double s = Rc[0][col] * efc[blk] + Rc[1][col] * efc[blk + 1] + Rc[2][col] * efc[blk + 2];
            if (mc[i]) f[mc[i] - 1] += s;
        }
    }
}

/* updatc (misc.c:71-185): coordinates, then per-element lengths, areas and
 * triads from the updated coordinates. */
void orc_updatc(const orc_dims *d, double *x_temp, double *x_ip, double *xfr_temp,
                const double *dd, double *defllen_i, double *deffarea_i, double *defslen_i,
                const double *offset, const int *osflag, const double *auxpt,
                double *c1_i, double *c2_i, double *c3_i, const long *minc,
                const long *jcode)
{
    for (long j = 0; j < d->NJ; ++j)
        for (int k = 0; k < 3; ++k) {
            long e = jcode[j * 7 + k];
            x_ip[j * 3 + k] = x_temp[j * 3 + k];
            if (e) x_temp[j * 3 + k] += dd[e - 1];
        }
    for (long n = 0; n < d->NE_TR; ++n) {
        const double *a = x_temp + (minc[n * 2] - 1) * 3, *b = x_temp + (minc[n * 2 + 1] - 1) * 3;
        double el[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        defllen_i[n] = sqrt(dot3(el, el));
        c1_i[n] = el[0] / defllen_i[n];
        c2_i[n] = el[1] / defllen_i[n];
        c3_i[n] = el[2] / defllen_i[n];
    }
    long ioff = d->NE_TR * 2;
    for (long n = 0; n < d->NE_FR; ++n) {
        const double *a = x_temp + (minc[ioff + n * 2] - 1) * 3;
        const double *b = x_temp + (minc[ioff + n * 2 + 1] - 1) * 3;
        double *xe = xfr_temp + n * 6;
        for (int k = 0; k < 3; ++k) {
            xe[k] = a[k] + (osflag[n] ? offset[n * 6 + k] : 0.0);
            xe[3 + k] = b[k] + (osflag[n] ? offset[n * 6 + 3 + k] : 0.0);
        }
        double el[3] = {xe[3] - xe[0], xe[4] - xe[1], xe[5] - xe[2]};
        double L = sqrt(dot3(el, el));
        defllen_i[d->NE_TR + n] = L;
        double lx[3] = {el[0] / L, el[1] / L, el[2] / L};
        double tmp[3] = {auxpt[n * 3] - xe[0], auxpt[n * 3 + 1] - xe[1], auxpt[n * 3 + 2] - xe[2]};
        double lz[3], ly[3];
        cross3(lx, tmp, lz, 1);
        cross3(lz, lx, ly, 1);
        for (int k = 0; k < 3; ++k) {
            c1_i[d->NE_TR + n * 3 + k] = lx[k];
            c2_i[d->NE_TR + n * 3 + k] = ly[k];
            c3_i[d->NE_TR + n * 3 + k] = lz[k];
        }
    }
    ioff += d->NE_FR * 2;
    long toff = d->NE_TR + d->NE_FR * 3;
    for (long n = 0; n < d->NE_SH; ++n) {
        const double *p1 = x_temp + (minc[ioff + n * 3] - 1) * 3;
        const double *p2 = x_temp + (minc[ioff + n * 3 + 1] - 1) * 3;
        const double *p3 = x_temp + (minc[ioff + n * 3 + 2] - 1) * 3;
        double e12[3], e23[3], e31[3], nrm[3];
        for (int k = 0; k < 3; ++k) {
            e23[k] = p3[k] - p2[k];
            e31[k] = p3[k] - p1[k];
            e12[k] = p2[k] - p1[k];
        }
        defslen_i[n * 3 + 1] = sqrt(dot3(e23, e23));
        defslen_i[n * 3 + 2] = sqrt(dot3(e31, e31));
        defslen_i[n * 3] = sqrt(dot3(e12, e12));
        cross3(e12, e31, nrm, 0);
        deffarea_i[n] = 0.5 * sqrt(dot3(nrm, nrm));
        double lx[3], lz[3], ly[3];
        for (int k = 0; k < 3; ++k) {
            lx[k] = e12[k] / defslen_i[n * 3];
            lz[k] = nrm[k] / (2 * deffarea_i[n]);
        }
        cross3(lz, lx, ly, 1);
        for (int k = 0; k < 3; ++k) {
            c1_i[toff + n * 3 + k] = lx[k];
            c2_i[toff + n * 3 + k] = ly[k];
            c3_i[toff + n * 3 + k] = lz[k];
        }
    }
}

/* ---- linear elastic forces (ANAFLAG 1) -------------------------------------- */

/* truss.c:243-326: d = T D in the element's triad, ef = EA/L0 [1 -1; -1 1] d,
 * f -= ef c at each end. */
void orc_forces_tr_lin(const orc_dims *d, double *f, double *ef_i, const double *D,
                       const double *emod, const double *carea, const double *llength,
                       const double *c1, const double *c2, const double *c3, const long *mcode)
{
    for (long n = 0; n < d->NE_TR; ++n) {
        double De[6];
        for (int i = 0; i < 6; ++i) {
            long j = mcode[n * 6 + i];
            De[i] = j ? D[j - 1] : 0.0;
        }
        double c[3] = {c1[n], c2[n], c3[n]};
        double dl[2];
        for (int i = 0; i < 2; ++i) {
            double s = 0;
            for (int j = 0; j < 6; ++j) s += (j / 3 == i ? c[j % 3] : 0.0) * De[j];
            dl[i] = s;
        }
        double ka = carea[n] * emod[n] / llength[n];
        ef_i[n * 2] = ka * dl[0] + (-ka) * dl[1];
        ef_i[n * 2 + 1] = (-ka) * dl[0] + ka * dl[1];
        for (int j = 0; j < 6; ++j) {
            long k = mcode[n * 6 + j];
            if (k) f[k - 1] -= ef_i[n * 2 + j / 3] * c[j % 3];
        }
    }
}

/* shell.c:1695-1727: ef = k_sh (T D), the elastic matrix of stiffe_sh with its
 * drilling term, then f += T^T ef (shell.c:2385-2397). */
void orc_forces_sh_lin(const orc_dims *d, double *f, double *ef_i, const double *D,
                       const double *emod, const double *nu, const double *xlocal,
                       const double *thick, const double *farea, const double *slength,
                       const double *c1, const double *c2, const double *c3, const long *mcode)
{
    long poff = d->NE_TR + d->NE_FR, toff = d->NE_TR + d->NE_FR * 3;
    long moff = d->NE_TR * 6 + d->NE_FR * 14, eoff = d->NE_TR * 2 + d->NE_FR * 14;
    static const double dm0[6] = {0, 0, 0, 0, 0, 0};
    for (long n = 0; n < d->NE_SH; ++n) {
        double k[18][18];
        shell_local_tangent(emod[poff + n], nu[n], thick[n], xlocal + n * 3, farea[n],
                            slength + n * 3, farea[n], dm0, 0, k);
        const double *R[3] = {c1 + toff + n * 3, c2 + toff + n * 3, c3 + toff + n * 3};
        double De[18], dl[18];
        for (int i = 0; i < 18; ++i) {
            long j = mcode[moff + n * 18 + i];
            De[i] = j ? D[j - 1] : 0.0;
        }
        for (int i = 0; i < 18; ++i) {          /* T rows: c1, c2, c3 per node block */
            int b = i / 3 * 3, r = i % 3;
            dl[i] = R[r][0] * De[b] + R[r][1] * De[b + 1] + R[r][2] * De[b + 2];
        }
        double *ef = ef_i + eoff + n * 18;
        for (int i = 0; i < 18; ++i) {
            double s = 0;
            for (int j = 0; j < 18; ++j) s += k[i][j] * dl[j];
            ef[i] = s;
        }
        for (int i = 0; i < 18; ++i) {
            int b = i / 3 * 3, c = i % 3;
            double s = R[0][c] * ef[b] + R[1][c] * ef[b + 1] + R[2][c] * ef[b + 2];
            long j = mcode[moff + n * 18 + i];
            if (j) f[j - 1] += s;
        }
    }
}
