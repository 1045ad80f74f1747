/*
 * orc_frame.c -- oracle: 3-D corotational frame element (elastic, ANAFLAG 2)
 * and the lumped mass routines of truss, frame and shell.
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

/* Element tangent of frame n in the local frame: elastic + geometric. */
static void frame_local_tangent(const orc_model *m, long n, const double *llength,
                                const double *defllen_ip, const double *ef_ip,
                                const double *efFE_ip, double k[14][14])
{
    const orc_dims *d = &m->dm;
    long ne = d->NE_TR + n;
    orc_frame_ke(m->emod[ne], m->gmod[n], m->carea[ne], llength[ne], m->istrong[n], m->iweak[n],
                 m->ipolar[n], m->iwarp[n], k);
    if (d->ANAFLAG == 2) {
        double ft[14];
        for (int i = 0; i < 14; ++i)
            ft[i] = ef_ip[d->NE_TR * 2 + n * 14 + i] + efFE_ip[n * 14 + i];
        orc_frame_kg(ft, defllen_ip[ne], m->carea[ne], m->ipolar[n], k);
    }
}

/* stiff_fr (frame.c:226-349): K = G^T T^T k T G accumulated into the skyline. */
void orc_stiff_fr(const orc_dims *d, double *ss, const orc_model *m, const double *llength,
                  const double *defllen_ip, const double *c1_ip, const double *c2_ip,
                  const double *c3_ip, const double *ef_ip, const double *efFE_ip)
{
    double k[14][14], T[14][14], K[14][14], G[14][14], K2[14][14];
    for (long n = 0; n < d->NE_FR; ++n) {
        frame_local_tangent(m, n, llength, defllen_ip, ef_ip, efFE_ip, k);
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

/* forces_fr, elastic branch (frame.c:902-1096, 1273-1309):
 *   dd_loc = T_ip G DD12;  def = k(ef_ip + efFE_ip) dd_loc
 *   ef_i   = T_i T_ip^T (ef_ip + def);  efFE_i = T_i T_ip^T efFE_ip
 *   f     += G^T T_i^T ef_i
 * The fixed-end forces carry no distributed load on this path, so the
 * first-iteration dlpf * efFE_ref term is zero. */
int orc_forces_fr(const orc_dims *d, double *f, const double *ef_ip, double *ef_i,
                  const double *efFE_ip, double *efFE_i, const double *dd,
                  const orc_model *m, const double *llength, const double *defllen_ip,
                  const double *c1_ip, const double *c2_ip, const double *c3_ip,
                  const double *c1_i, const double *c2_i, const double *c3_i)
{
    double k[14][14], Tp[14][14], Ti[14][14], G[14][14], Q[14][14];
    double D[14], Dj[14], dl[14], def[14], tmp[14], eg[14], eg2[14];
    for (long n = 0; n < d->NE_FR; ++n) {
        const long *mc = m->mcode + d->NE_TR * 6 + n * 14;
        frame_local_tangent(m, n, llength, defllen_ip, ef_ip, efFE_ip, k);
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
