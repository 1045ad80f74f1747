/*
 * orc_frame.c -- oracle: 3-D corotational frame element (elastic, ANAFLAG 2;
 * plastic hinges at the member ends, ANAFLAG 3) and the lumped mass routines
 * of truss, frame and shell.
 *
 * TEST INFRASTRUCTURE ONLY (see cb_oracle.h).
 *
 * Frame element DOF order (per end, end 2 at +7): u, v, w, theta_x, theta_y,
 * theta_z, warping.  The local frame is (c1, c2, c3) of the element; the
 * transformation is block diagonal with the 3x3 triad on the translation and
 * rotation triples and 1 on the warping DOF (frame.c:286-295).
 */
#include "cb_oracle.h"

#include <math.h>
#include <string.h>

static void skyline_add_n(double *ss, const long *maxa, const long *mc, int n, const double *K)
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

/* Symmetric set: k[i][j] = k[j][i] = v. */
static inline void sset(double k[14][14], int i, int j, double v) { k[i][j] = v; k[j][i] = v; }
/* Symmetric add: k[i][j] += v, k[j][i] += v (once on the diagonal). */
static inline void sadd(double k[14][14], int i, int j, double v)
{
    k[i][j] += v;
    if (i != j) k[j][i] += v;
}

/* Elastic 14x14 local stiffness (stiffe_fr, frame.c:354-408): axial EA/L,
 * strong-axis (v, theta_z) and weak-axis (w, theta_y) Euler-Bernoulli bending,
 * and Vlasov torsion coupling theta_x with the warping DOF. */
void orc_frame_ke(double E, double G, double A, double L, double Is, double Iw, double J,
                  double Cw, double k[14][14])
{
    memset(k, 0, sizeof(double) * 14 * 14);
    double L2 = pow(L, 2), L3 = pow(L, 3);
    double ax = E * A / L;
    sset(k, 0, 0, ax); sset(k, 7, 7, ax); sset(k, 0, 7, -ax);
    /* strong axis: v (1, 8) with theta_z (5, 12) */
    double s12 = 12 * E * Is / L3, s6 = 6 * E * Is / L2, s4 = 4 * E * Is / L, s2 = 2 * E * Is / L;
    sset(k, 1, 1, s12); sset(k, 8, 8, s12); sset(k, 1, 8, -s12);
    sset(k, 1, 5, s6); sset(k, 1, 12, s6); sset(k, 5, 8, -s6); sset(k, 8, 12, -s6);
    sset(k, 5, 5, s4); sset(k, 12, 12, s4); sset(k, 5, 12, s2);
    /* weak axis: w (2, 9) with theta_y (4, 11) */
    double w12 = 12 * E * Iw / L3, w6 = 6 * E * Iw / L2, w4 = 4 * E * Iw / L, w2 = 2 * E * Iw / L;
    sset(k, 2, 2, w12); sset(k, 9, 9, w12); sset(k, 2, 9, -w12);
    sset(k, 4, 9, w6); sset(k, 9, 11, w6); sset(k, 2, 4, -w6); sset(k, 2, 11, -w6);
    sset(k, 4, 4, w4); sset(k, 11, 11, w4); sset(k, 4, 11, w2);
    /* torsion with warping: theta_x (3, 10), warping (6, 13) */
    double t33 = 6 * G * J / (5 * L) + 12 * E * Cw / L3;
    double t63 = G * J / 10 + 6 * E * Cw / L2;
    double t66 = 2 * G * J * L / 15 + 4 * E * Cw / L;
    double t136 = -(G * J * L / 30 - 2 * E * Cw / L);
    sset(k, 3, 3, t33); sset(k, 10, 10, t33); sset(k, 3, 10, -t33);
    sset(k, 3, 6, t63); sset(k, 3, 13, t63); sset(k, 6, 10, -t63); sset(k, 10, 13, -t63);
    sset(k, 6, 6, t66); sset(k, 13, 13, t66); sset(k, 6, 13, t136);
}

/* Geometric stiffness (stiffg_fr, frame.c:410-579) from the total element end
 * forces of the previous iterate: P = f[7] (axial at end 2), end moments
 * My1 = f[4], Mz1 = f[5], My2 = f[11], Mz2 = f[12], torque T = f[10], with
 * r2 = Ip/A.  Grouped by the force each term is proportional to. */
void orc_frame_kg(const double f[14], double L, double A, double Ip, double k[14][14])
{
    double P = f[7], My1 = f[4], Mz1 = f[5], My2 = f[11], Mz2 = f[12], T = f[10];
    double r2 = Ip / A;
    /* axial force P */
    sadd(k, 0, 0, P / L); sadd(k, 7, 7, P / L); sadd(k, 0, 7, -P / L);
    double pv = 6 * P / (5 * L);
    sadd(k, 1, 1, pv); sadd(k, 8, 8, pv); sadd(k, 2, 2, pv); sadd(k, 9, 9, pv);
    sadd(k, 1, 8, -pv); sadd(k, 2, 9, -pv);
    double pt = 6 * P * Ip / (5 * A * L);
    sadd(k, 3, 3, pt); sadd(k, 10, 10, pt); sadd(k, 3, 10, -pt);
    double pr = 2 * P * L / 15;
    sadd(k, 4, 4, pr); sadd(k, 11, 11, pr); sadd(k, 5, 5, pr); sadd(k, 12, 12, pr);
    double pw = 2 * P * Ip / (15 * A);
    sadd(k, 6, 6, pw); sadd(k, 13, 13, pw);
    double p10 = P / 10;
    sadd(k, 1, 5, p10); sadd(k, 1, 12, p10); sadd(k, 4, 9, p10); sadd(k, 9, 11, p10);
    sadd(k, 2, 4, -p10); sadd(k, 2, 11, -p10); sadd(k, 5, 8, -p10); sadd(k, 8, 12, -p10);
    double pq = P * Ip / (10 * A);
    sadd(k, 3, 6, pq); sadd(k, 3, 13, pq); sadd(k, 6, 10, -pq); sadd(k, 10, 13, -pq);
    double pl = P * L / 30;
    sadd(k, 4, 11, -pl); sadd(k, 5, 12, -pl);
    sadd(k, 6, 13, -P * Ip / (30 * A));
    (void)r2;
    /* torque T */
    double tl = T / L;
    sadd(k, 1, 4, tl); sadd(k, 2, 5, tl); sadd(k, 8, 11, tl); sadd(k, 9, 12, tl);
    sadd(k, 1, 11, -tl); sadd(k, 2, 12, -tl); sadd(k, 4, 8, -tl); sadd(k, 5, 9, -tl);
    sadd(k, 4, 12, T / 2); sadd(k, 5, 11, -T / 2);
    /* end moments about the weak axis (My1, My2) with theta_x / warping */
    double a = (11 * My1 - My2) / (10 * L);
    sadd(k, 1, 3, a); sadd(k, 3, 8, -a);
    double b = (My1 - 11 * My2) / (10 * L);
    sadd(k, 8, 10, b); sadd(k, 1, 10, -b);
    sadd(k, 1, 6, My1 / 10); sadd(k, 6, 8, -My1 / 10);
    sadd(k, 8, 13, My2 / 10); sadd(k, 1, 13, -My2 / 10);
    /* end moments about the strong axis (Mz1, Mz2) */
    double c = (11 * Mz1 - Mz2) / (10 * L);
    sadd(k, 2, 3, c); sadd(k, 3, 9, -c);
    double e = (Mz1 - 11 * Mz2) / (10 * L);
    sadd(k, 9, 10, e); sadd(k, 2, 10, -e);
    sadd(k, 2, 6, Mz1 / 10); sadd(k, 6, 9, -Mz1 / 10);
    sadd(k, 9, 13, Mz2 / 10); sadd(k, 2, 13, -Mz2 / 10);
    /* theta_x coupling with the bending rotations */
    sadd(k, 3, 4, -(2 * Mz1 - Mz2) / 5);
    sadd(k, 3, 5, (2 * My1 - My2) / 5);
    sadd(k, 3, 11, -(2 * Mz1 + Mz2) / 10);
    sadd(k, 3, 12, (2 * My1 + My2) / 10);
    sadd(k, 4, 10, -(Mz1 + 2 * Mz2) / 10);
    sadd(k, 5, 10, (My1 + 2 * My2) / 10);
    sadd(k, 10, 11, (Mz1 - 2 * Mz2) / 5);
    sadd(k, 10, 12, -(My1 - 2 * My2) / 5);
    /* warping rows */
    sadd(k, 4, 6, -(3 * Mz1 - Mz2) * L / 30);
    sadd(k, 5, 6, (3 * My1 - My2) * L / 30);
    sadd(k, 4, 13, Mz1 * L / 30);
    sadd(k, 5, 13, -My1 * L / 30);
    sadd(k, 6, 11, -Mz2 * L / 30);
    sadd(k, 6, 12, My2 * L / 30);
    sadd(k, 11, 13, -(Mz1 - 3 * Mz2) * L / 30);
    sadd(k, 12, 13, (My1 - 3 * My2) * L / 30);
}

/* 14x14 transformation T (rows of the local triad on every 3-block). */
static void frame_T(const double *c1, const double *c2, const double *c3, double T[14][14])
{
    memset(T, 0, sizeof(double) * 14 * 14);
    static const int base[4] = {0, 3, 7, 10};
    for (int b = 0; b < 4; ++b)
        for (int j = 0; j < 3; ++j) {
            T[base[b]][base[b] + j] = c1[j];
            T[base[b] + 1][base[b] + j] = c2[j];
            T[base[b] + 2][base[b] + j] = c3[j];
        }
    T[6][6] = T[13][13] = 1;
}

/* Rigid-link matrix G (element end DOFs = G * joint DOFs): the end point sits
 * at joint + offset, so u_end = u_joint + theta x offset (frame.c:297-324). */
static void frame_G(const double *o, double G[14][14])
{
    memset(G, 0, sizeof(double) * 14 * 14);
    for (int i = 0; i < 14; ++i) G[i][i] = 1;
    for (int e = 0; e < 2; ++e) {
        int u = 7 * e, t = 7 * e + 3;
        const double *oe = o + 3 * e;
        G[u + 0][t + 1] = oe[2];  G[u + 0][t + 2] = -oe[1];
        G[u + 1][t + 0] = -oe[2]; G[u + 1][t + 2] = oe[0];
        G[u + 2][t + 0] = oe[1];  G[u + 2][t + 1] = -oe[0];
    }
}

/* C = A^T B A for 14x14 matrices. */
static void congruence(const double A[14][14], const double B[14][14], double C[14][14])
{
    double tmp[14][14];
    for (int i = 0; i < 14; ++i)
        for (int j = 0; j < 14; ++j) {
            double s = 0;
            for (int k = 0; k < 14; ++k) s += A[k][i] * B[k][j];
            tmp[i][j] = s;
        }
    for (int i = 0; i < 14; ++i)
        for (int j = 0; j < 14; ++j) {
            double s = 0;
            for (int k = 0; k < 14; ++k) s += tmp[i][k] * A[k][j];
            C[i][j] = s;
        }
}

static void matvec14(const double A[14][14], const double *x, double *y, int transpose)
{
    for (int i = 0; i < 14; ++i) {
        double s = 0;
        for (int j = 0; j < 14; ++j) s += (transpose ? A[j][i] : A[i][j]) * x[j];
        y[i] = s;
    }
}

/* ------------------------------------------------------ plastic hinges --- */
#define PHITOL 1e-4   /* allowed deviation of phi from 1 (frame.c:37) */

double orc_orbison(double p, double my, double mz)
{
    double p2 = p * p, y2 = my * my, z2 = mz * mz;
    return p2 + z2 + y2 * y2 + 3.5 * p2 * z2 + 3 * (p2 * p2 * p2) * y2 + 4.5 * (z2 * z2) * y2;
}

/* d phi / d(P, My, Mz): the partial derivatives of orc_orbison in p, my, mz
 * over the normalising forces */
void orc_orbison_grad(double p, double my, double mz, double Py, double Mpy, double Mpz,
                      double g[3])
{
    double p2 = p * p, y2 = my * my, z2 = mz * mz;
    g[0] = (2 * p + 7 * p * z2 + 18 * (p2 * p2 * p) * y2) / Py;
    g[1] = (4 * y2 * my + 6 * (p2 * p2 * p2) * my + 9 * (z2 * z2) * my) / Mpy;
    g[2] = (2 * mz + 7 * p2 * mz + 18 * (z2 * mz) * y2) / Mpz;
}

/* phi at both ends of total end forces f (P at 7a, My at 7a+4, Mz at 7a+5) */
static void end_phis(const double f[14], double Py, double Mpy, double Mpz, double phi[2])
{
    for (int a = 0; a < 2; ++a)
        phi[a] = orc_orbison(f[7 * a] / Py, f[7 * a + 4] / Mpy, f[7 * a + 5] / Mpz);
}

/* the yield-gradient column of end a: nonzero at P, My, Mz of that end */
static void end_column(const double f[14], int a, double Py, double Mpy, double Mpz, double col[14])
{
    double g[3];
    memset(col, 0, sizeof(double) * 14);
    orc_orbison_grad(f[7 * a] / Py, f[7 * a + 4] / Mpy, f[7 * a + 5] / Mpz, Py, Mpy, Mpz, g);
    col[7 * a] = g[0];
    col[7 * a + 4] = g[1];
    col[7 * a + 5] = g[2];
}

/* H^-1 of the 1x1 or 2x2 system H = G^T k G by the adjugate over the
 * determinant, as frame.c:677-683 forms it (a zero column leaves det = 0) */
static void small_inverse(double H[2][2], int nc)
{
    if (nc == 1) {
        H[0][0] = 1 / H[0][0];
        return;
    }
    double det = H[0][0] * H[1][1] - H[0][1] * H[1][0];
    double h00 = H[0][0];
    H[0][0] = H[1][1] / det;
    H[1][1] = h00 / det;
    H[0][1] *= -1 / det;
    H[1][0] *= -1 / det;
}

/* k G (14 x nc) and G^T k G for the columns G */
static void kg_products(const double k[14][14], const double G[2][14], int nc, double kG[2][14],
                        double H[2][2])
{
    for (int a = 0; a < nc; ++a)
        for (int i = 0; i < 14; ++i) {
            double s = 0;
            for (int j = 0; j < 14; ++j) s += k[i][j] * G[a][j];
            kG[a][i] = s;
        }
    for (int a = 0; a < nc; ++a)
        for (int b = 0; b < nc; ++b) {
            double s = 0;
            for (int j = 0; j < 14; ++j) s += kG[a][j] * G[b][j];
            H[a][b] = s;
        }
}

/* stiffm_fr (frame.c:580-797).  Both ends on the surface: both columns, the
 * column of an end flagged 2 left zero; otherwise the one end on the surface
 * that is not flagged 2; otherwise no change. */
void orc_frame_plastic(double k[14][14], const double eftot[14], const int yld[2], double Py,
                       double Mpy, double Mpz)
{
    double phi[2], G[2][14], kG[2][14], H[2][2];
    end_phis(eftot, Py, Mpy, Mpz, phi);
    int nc = 0;
    if (phi[0] >= 1 - PHITOL && phi[1] >= 1 - PHITOL) {
        for (int a = 0; a < 2; ++a) {
            if (yld[a] != 2) end_column(eftot, a, Py, Mpy, Mpz, G[a]);
            else memset(G[a], 0, sizeof G[a]);
        }
        nc = 2;
    } else if (phi[0] >= 1 - PHITOL && yld[0] != 2) {
        end_column(eftot, 0, Py, Mpy, Mpz, G[0]);
        nc = 1;
    } else if (phi[1] >= 1 - PHITOL && yld[1] != 2) {
        end_column(eftot, 1, Py, Mpy, Mpz, G[0]);
        nc = 1;
    }
    if (!nc) return;
    kg_products(k, G, nc, kG, H);
    small_inverse(H, nc);
    for (int i = 0; i < 14; ++i)
        for (int j = 0; j < 14; ++j) {
            double s = 0;
            for (int a = 0; a < nc; ++a)
                for (int b = 0; b < nc; ++b) s += kG[a][i] * H[a][b] * kG[b][j];
            k[i][j] -= s;
        }
}

/* False position for phi(tau) = 1 on [0, 1] of the end forces (p, my, mz) +
 * tau (dp, dmy, dmz): the estimate from the two end values, then one further
 * step on the sub-interval whose ends straddle 1 (frame.c:1397-1455; the
 * reference's loop test phi <= 1 - tol && phi >= 1 + tol never holds, so its
 * body runs once). */
double orc_regula_falsi(double p, double dp, double my, double dmy, double mz, double dmz)
{
    double tu = 1, tl = 0;
    double fu = orc_orbison(p + tu * dp, my + tu * dmy, mz + tu * dmz);
    double fl = orc_orbison(p + tl * dp, my + tl * dmy, mz + tl * dmz);
    double tr = tu - (fu - 1) * (tl - tu) / (fl - fu);
    double fr = orc_orbison(p + tr * dp, my + tr * dmy, mz + tr * dmz);
    if ((fl - 1 > 0 && fr - 1 > 0) || (fl - 1 < 0 && fr - 1 < 0)) {
        tl = tr;
        fl = fr;
    } else {
        tu = tr;
        fu = fr;
    }
    return tu - (fu - 1) * (tl - tu) / (fl - fu);
}

/* unload (frame.c:1457-1658): the plastic multipliers lambda = (G^T k G)^-1
 * G^T k dl of the ends on the surface (gradients at the new end forces,
 * k = elastic + geometric); an end unloads when its lambda < -1e-8. */
int orc_frame_unload(const double phi[2], const double eftot_i[14], double Py, double Mpy,
                     double Mpz, const double k[14][14], const double dl[14])
{
    double G[2][14], kG[2][14], H[2][2], lam[2];
    int nc = 0, which = 0;
    if (phi[0] >= 1 - PHITOL && phi[1] >= 1 - PHITOL) {
        end_column(eftot_i, 0, Py, Mpy, Mpz, G[0]);
        end_column(eftot_i, 1, Py, Mpy, Mpz, G[1]);
        nc = 2;
    } else if (phi[0] >= 1 - PHITOL) {
        end_column(eftot_i, 0, Py, Mpy, Mpz, G[0]);
        nc = 1;
        which = 0;
    } else if (phi[1] >= 1 - PHITOL) {
        end_column(eftot_i, 1, Py, Mpy, Mpz, G[0]);
        nc = 1;
        which = 1;
    }
    if (!nc) return 0;
    kg_products(k, G, nc, kG, H);     /* k symmetric: row a of G^T k is kG[a] */
    small_inverse(H, nc);
    double w[2];
    for (int a = 0; a < nc; ++a) {
        double s = 0;
        for (int j = 0; j < 14; ++j) s += kG[a][j] * dl[j];
        w[a] = s;
    }
    for (int a = 0; a < nc; ++a) {
        double s = 0;
        for (int b = 0; b < nc; ++b) s += H[a][b] * w[b];
        lam[a] = s;
    }
    if (nc == 2) {
        if (lam[0] < -1e-8 && lam[1] < -1e-8) return 1;
        if (lam[0] < -1e-8) return 2;
        if (lam[1] < -1e-8) return 3;
        return 0;
    }
    return lam[0] < -1e-8 ? 2 + which : 0;
}

/* inverse (misc.c:284-342): Gauss-Jordan on [A | I] without normalising the
 * pivot rows, which are divided out at the end.  Its pivot search scans row i
 * (not column i) from the diagonal for the first non-zero entry and swaps
 * rows i and that column's index; with a non-zero diagonal, as the released
 * blocks of the frame tangent have, no swap happens. */
void orc_gj_inverse(double *a, int n)
{
    double aug[4][8];
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < 2 * n; ++j)
            aug[i][j] = j < n ? a[i * n + j] : (j - n == i ? 1.0 : 0.0);
    for (int i = 0; i < n; ++i) {
        int k = i;
        for (int j = i; j < n; ++j)
            if (aug[i][j] != 0) {
                k = j;
                break;
            }
        if (k != i)
            for (int j = 0; j < 2 * n; ++j) {
                double t = aug[i][j];
                aug[i][j] = aug[k][j];
                aug[k][j] = t;
            }
        for (int j = 0; j < n; ++j) {
            if (j == i) continue;
            double f = aug[j][i] / aug[i][i];
            for (int q = 0; q < 2 * n; ++q) aug[j][q] -= f * aug[i][q];
        }
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) a[i * n + j] = aug[i][n + j] / aug[i][i];
}

/* release (frame.c:798-896): member end bending releases by static
 * condensation, k <- k - k G (G^T k G)^-1 G^T k, where G selects the released
 * local DOFs in the order strong axis end 1 (5), weak axis end 1 (4), strong
 * axis end 2 (12), weak axis end 2 (11); flags[0..3] say which are released.
 * G^T k is taken as (k G)^T, the reference's order.  One release inverts by
 * 1/x, two by the adjugate over the determinant, three or four by
 * orc_gj_inverse. */
void orc_frame_release(double k[14][14], const int flags[4])
{
    static const int dof[4] = {5, 4, 12, 11};
    int r[4], n = 0;
    for (int a = 0; a < 4; ++a)
        if (flags[a]) r[n++] = dof[a];
    if (!n) return;
    double kG[14][4], H[16];
    for (int i = 0; i < 14; ++i)
        for (int a = 0; a < n; ++a) kG[i][a] = k[i][r[a]];
    for (int a = 0; a < n; ++a)            /* (G^T k G)[a][b] = k[r_b][r_a] */
        for (int b = 0; b < n; ++b) H[a * n + b] = kG[r[b]][a];
    if (n == 1) {
        H[0] = 1 / H[0];
    } else if (n == 2) {
        double det = H[0] * H[3] - H[1] * H[2], t = H[0];
        H[0] = H[3] / det;
        H[3] = t / det;
        H[1] *= -1 / det;
        H[2] *= -1 / det;
    } else {
        orc_gj_inverse(H, n);
    }
    double kGH[14][4];
    for (int i = 0; i < 14; ++i)
        for (int b = 0; b < n; ++b) {
            double s = 0;
            for (int a = 0; a < n; ++a) s += kG[i][a] * H[a * n + b];
            kGH[i][b] = s;
        }
    for (int i = 0; i < 14; ++i)
        for (int j = 0; j < 14; ++j) {
            double s = 0;
            for (int b = 0; b < n; ++b) s += kGH[i][b] * kG[j][b];
            k[i][j] -= s;
        }
}

static void frame_ke_of(const orc_model *m, long n, const double *llength, double k[14][14])
{
    long ne = m->dm.NE_TR + n;
    orc_frame_ke(m->emod[ne], m->gmod[n], m->carea[ne], llength[ne], m->istrong[n], m->iweak[n],
                 m->ipolar[n], m->iwarp[n], k);
}

/* Element tangent of frame n in the local frame: elastic + geometric, under
 * ANAFLAG 3 the plastic reduction unless both ends are flagged 2
 * (frame.c:266-277), then the member's end releases (frame.c:280-283).
 * eftot: the member's total end forces (ef + efFE). */
static void frame_local_tangent(const orc_model *m, long n, const double *llength,
                                const double *defllen_ip, const double *eftot, const int *yld,
                                double k[14][14])
{
    const orc_dims *d = &m->dm;
    long ne = d->NE_TR + n;
    frame_ke_of(m, n, llength, k);
    if (d->ANAFLAG == 2 || d->ANAFLAG == 3)
        orc_frame_kg(eftot, defllen_ip[ne], m->carea[ne], m->ipolar[n], k);
    if (d->ANAFLAG == 3 && (yld[n * 2] != 2 || yld[n * 2 + 1] != 2))
        orc_frame_plastic(k, eftot, yld + n * 2, m->carea[ne] * m->yield[ne],
                          m->zweak[n] * m->yield[ne], m->zstrong[n] * m->yield[ne]);
    if (m->mendrel[n * 5]) orc_frame_release(k, m->mendrel + n * 5 + 1);
}

static void total_forces(const orc_dims *d, long n, const double *ef, const double *efFE,
                         double ft[14])
{
    for (int i = 0; i < 14; ++i) ft[i] = ef[d->NE_TR * 2 + n * 14 + i] + efFE[n * 14 + i];
}

/* stiff_fr (frame.c:226-349): K = G^T T^T k T G accumulated into the skyline. */
void orc_stiff_fr(const orc_dims *d, double *ss, const orc_model *m, const double *llength,
                  const double *defllen_ip, const double *c1_ip, const double *c2_ip,
                  const double *c3_ip, const double *ef_ip, const double *efFE_ip,
                  const int *yldflag)
{
    double k[14][14], T[14][14], K[14][14], G[14][14], K2[14][14], ft[14];
    for (long n = 0; n < d->NE_FR; ++n) {
        total_forces(d, n, ef_ip, efFE_ip, ft);
        frame_local_tangent(m, n, llength, defllen_ip, ft, yldflag, k);
        long t = d->NE_TR + n * 3;
        frame_T(c1_ip + t, c2_ip + t, c3_ip + t, T);
        congruence(T, k, K);
        if (m->osflag[n]) {
            frame_G(m->offset + n * 6, G);
            congruence(G, K, K2);
            memcpy(K, K2, sizeof K);
        }
        skyline_add_n(ss, m->maxa, m->mcode + d->NE_TR * 6 + n * 14, 14, &K[0][0]);
    }
}

/* forces_fr (frame.c:902-1096, 1153-1309):
 *   dd_loc = T_ip G DD12;  def = k(ef_ip + efFE_ip) dd_loc
 *   ef_i   = T_i T_ip^T (ef_ip + def);  efFE_i = T_i T_ip^T efFE_ip
 *   f     += G^T T_i^T ef_i
 * The fixed-end forces carry no distributed load on this path, so the
 * first-iteration dlpf * efFE_ref term of every yldflag branch
 * (frame.c:1098-1151) is zero.
 * ANAFLAG 3, between ef_i and the scatter, with phi at the new total end
 * forces: an end beyond the surface (phi > 1 + tol, the larger of the two,
 * not flagged 2) scales *dlpf by the fraction of this iteration's end-force
 * increment that reaches phi = 1, is flagged 1 and returns 1; otherwise the
 * ends on the surface and not flagged 2 are flagged 1, and when an end is
 * flagged 1 the unloading test of the elastic + geometric k against dd_loc
 * flags the unloading ends 2 and returns 2.  A return leaves the member's ef_i
 * written, its forces out of f, and the members after it untouched. */
int orc_forces_fr(const orc_dims *d, double *f, const double *ef_ip, double *ef_i,
                  const double *efFE_ip, double *efFE_i, const double *dd,
                  const orc_model *m, const double *llength, const double *defllen_ip,
                  const double *c1_ip, const double *c2_ip, const double *c3_ip,
                  const double *c1_i, const double *c2_i, const double *c3_i,
                  int *yldflag, double *dlpf)
{
    double k[14][14], Tp[14][14], Ti[14][14], G[14][14], Q[14][14];
    double D[14], Dj[14], dl[14], def[14], tmp[14], eg[14], eg2[14], fp[14], fi[14];
    for (long n = 0; n < d->NE_FR; ++n) {
        const long *mc = m->mcode + d->NE_TR * 6 + n * 14;
        total_forces(d, n, ef_ip, efFE_ip, fp);
        frame_local_tangent(m, n, llength, defllen_ip, fp, yldflag, k);
        for (int i = 0; i < 14; ++i) D[i] = mc[i] ? dd[mc[i] - 1] : 0.0;
        long t = d->NE_TR + n * 3;
        frame_T(c1_ip + t, c2_ip + t, c3_ip + t, Tp);
        frame_T(c1_i + t, c2_i + t, c3_i + t, Ti);
        if (m->osflag[n]) {
            frame_G(m->offset + n * 6, G);
            matvec14(G, D, Dj, 0);
        } else {
            memcpy(Dj, D, sizeof D);
        }
        matvec14(Tp, Dj, dl, 0);
        matvec14(k, dl, def, 0);
        for (int i = 0; i < 14; ++i)
            for (int j = 0; j < 14; ++j) {
                double s = 0;
                for (int q = 0; q < 14; ++q) s += Ti[i][q] * Tp[j][q];
                Q[i][j] = s;
            }
        const double *ep = ef_ip + d->NE_TR * 2 + n * 14;
        double *ei = ef_i + d->NE_TR * 2 + n * 14;
        for (int i = 0; i < 14; ++i) tmp[i] = ep[i] + def[i];
        matvec14(Q, tmp, ei, 0);
        matvec14(Q, efFE_ip + n * 14, efFE_i + n * 14, 0);
        if (d->ANAFLAG == 3) {
            long ne = d->NE_TR + n;
            double Py = m->carea[ne] * m->yield[ne], Mpy = m->zweak[n] * m->yield[ne];
            double Mpz = m->zstrong[n] * m->yield[ne], phi[2];
            int *y = yldflag + n * 2;
            total_forces(d, n, ef_i, efFE_i, fi);
            end_phis(fi, Py, Mpy, Mpz, phi);
            for (int a = 0; a < 2; ++a) {
                int o = 1 - a;
                if (phi[a] > phi[o] && phi[a] > 1 + PHITOL && y[a] != 2) {
                    *dlpf *= orc_regula_falsi(fp[7 * a] / Py, (fi[7 * a] - fp[7 * a]) / Py,
                                              fp[7 * a + 4] / Mpy,
                                              (fi[7 * a + 4] - fp[7 * a + 4]) / Mpy,
                                              fp[7 * a + 5] / Mpz,
                                              (fi[7 * a + 5] - fp[7 * a + 5]) / Mpz);
                    y[a] = 1;
                    return 1;
                }
            }
            int on0 = phi[0] >= 1 - PHITOL && y[0] != 2, on1 = phi[1] >= 1 - PHITOL && y[1] != 2;
            if (on0 && on1) y[0] = y[1] = 1;
            else if (on0) y[0] = 1;
            else if (on1) y[1] = 1;
            if (y[0] == 1 || y[1] == 1) {
                double ke[14][14];
                frame_ke_of(m, n, llength, ke);
                orc_frame_kg(fp, defllen_ip[ne], m->carea[ne], m->ipolar[n], ke);
                int u = orc_frame_unload(phi, fi, Py, Mpy, Mpz, ke, dl);
                if (u) {
                    if (u == 1 || u == 2) y[0] = 2;
                    if (u == 1 || u == 3) y[1] = 2;
                    return 2;
                }
            }
        }
        matvec14(Ti, ei, eg, 1);
        if (m->osflag[n]) {
            matvec14(G, eg, eg2, 1);
            memcpy(eg, eg2, sizeof eg);
        }
        for (int i = 0; i < 14; ++i)
            if (mc[i]) f[mc[i] - 1] += eg[i];
    }
    return 0;
}

/* Lumped masses, called in main.c's order (main.c:3596-3620) after the
 * stiffness of the same element type: mass_tr (truss.c:381-441), mass_fr
 * (frame.c:1314-1393), mass_sh (shell.c:1505-1591).  Like the reference, each routine recomputes its element geometry from the
 * converged coordinates x and OVERWRITES the length (truss, frame), end point
 * (xfr), face area and side length arrays it is given; all three index their
 * per-element arrays (dens, carea, llength, mcode) from 0 within their own
 * element type, so dens is the one array the reference shares among types. */
/* truss: consistent mass rho A L/6 [2 1; 1 2] per direction, row sums over
   free DOF pairs (truss.c:403-426) */
void orc_mass_tr(const orc_dims *d, double *sm, const orc_model *m, const double *x,
                 double *llength, const double *dens)
{
    for (long i = 0; i < d->NE_TR; ++i) {
        long a = m->minc[i * 2] - 1, b = m->minc[i * 2 + 1] - 1;
        double el[3];
        for (int l = 0; l < 3; ++l) el[l] = x[b * 3 + l] - x[a * 3 + l];
        llength[i] = sqrt(el[0] * el[0] + el[1] * el[1] + el[2] * el[2]);
        double mii = dens[i] * m->carea[i] * llength[i] / 3;
        double mij = dens[i] * m->carea[i] * llength[i] / 6;
        const long *mc = m->mcode + i * 6;
        for (int r = 0; r < 6; ++r) {
            if (!mc[r]) continue;
            for (int c = 0; c < 6; ++c) {
                if (!mc[c]) continue;
                double v = (r == c) ? mii : ((r % 3 == c % 3) ? mij : 0.0);
                sm[mc[r] - 1] += v;
            }
        }
    }
}

/* frame: rho A L/2 on translations, rho A L^3/24 on rotations, nothing on
   warping (frame.c:1357-1363) */
void orc_mass_fr(const orc_dims *d, double *sm, const orc_model *m, const double *x,
                 double *llength, double *xfr, const double *dens)
{
    for (long i = 0; i < d->NE_FR; ++i) {
        long a = m->minc[d->NE_TR * 2 + i * 2] - 1, b = m->minc[d->NE_TR * 2 + i * 2 + 1] - 1;
        double *xe = xfr + i * 6;
        for (int l = 0; l < 3; ++l) {
            double oa = m->osflag[i] ? m->offset[i * 6 + l] : 0.0;
            double ob = m->osflag[i] ? m->offset[i * 6 + 3 + l] : 0.0;
            xe[l] = x[a * 3 + l] + oa;
            xe[3 + l] = x[b * 3 + l] + ob;
        }
        double el[3] = {xe[3] - xe[0], xe[4] - xe[1], xe[5] - xe[2]};
        llength[d->NE_TR + i] = sqrt(el[0] * el[0] + el[1] * el[1] + el[2] * el[2]);
        double mt = dens[i] * m->carea[i] * llength[i] / 24 * 12;
        double mr = dens[i] * m->carea[i] * llength[i] / 24 * pow(llength[i], 2);
        const long *mc = m->mcode + i * 14;
        for (int r = 0; r < 14; ++r) {
            if (!mc[r] || r == 6 || r == 13) continue;
            int rot = (r % 7) >= 3;
            sm[mc[r] - 1] += rot ? mr : mt;
        }
    }
}

/* shell: rho A t/3 on translations, (rho A t/3) t^2/12 on rotations
   (shell.c:1544-1555); side lengths and area from the converged x */
void orc_mass_sh(const orc_dims *d, double *sm, const orc_model *m, const double *x,
                 double *farea, double *slength, const double *dens)
{
    for (long i = 0; i < d->NE_SH; ++i) {
        long off = d->NE_TR * 2 + d->NE_FR * 2 + i * 3;
        const double *p1 = x + (m->minc[off] - 1) * 3;
        const double *p2 = x + (m->minc[off + 1] - 1) * 3;
        const double *p3 = x + (m->minc[off + 2] - 1) * 3;
        double e12[3], e23[3], e31[3], nrm[3];
        for (int l = 0; l < 3; ++l) {
            e23[l] = p3[l] - p2[l];
            e31[l] = p3[l] - p1[l];
            e12[l] = p2[l] - p1[l];
        }
        slength[i * 3 + 1] = sqrt(e23[0] * e23[0] + e23[1] * e23[1] + e23[2] * e23[2]);
        slength[i * 3 + 2] = sqrt(e31[0] * e31[0] + e31[1] * e31[1] + e31[2] * e31[2]);
        slength[i * 3] = sqrt(e12[0] * e12[0] + e12[1] * e12[1] + e12[2] * e12[2]);
        nrm[0] = e12[1] * e31[2] - e12[2] * e31[1];
        nrm[1] = e12[2] * e31[0] - e12[0] * e31[2];
        nrm[2] = e12[0] * e31[1] - e12[1] * e31[0];
        farea[i] = 0.5 * sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
        double mtot = dens[i] * farea[i] * m->thick[i];
        double mt = mtot / 3, mr = mtot / 3 * pow(m->thick[i], 2) / 12;
        const long *mc = m->mcode + i * 18;
        for (int r = 0; r < 18; ++r) {
            if (!mc[r]) continue;
            sm[mc[r] - 1] += ((r % 6) >= 3) ? mr : mt;
        }
    }
}
