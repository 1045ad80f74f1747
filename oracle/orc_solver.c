/*
 * orc_solver.c -- oracle: skyline LDL^T (Bathe's active-column method), the
 * convergence test and the arc-length root selection.
 *
 * TEST INFRASTRUCTURE ONLY (see cb_oracle.h).
 *
 * Follows solve.c:539-653 (skyfact), solve.c:655-698 (skysolve),
 * misc.c:187-250 (test), arc.c:70-157 (quad).  The floating-point operations
 * are issued in the same order as the reference so the pivots that drive the
 * arc-length sign and psi decisions agree to the last bit on the same input.
 */
#include "cb_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>

/* In-place LDL^T of the skyline matrix ss (1-based diagonal addresses maxa).
 * Column n holds, below its diagonal address, rows n-1, n-2, ... up to its
 * height.  After the call those entries hold l = g/d and the diagonal holds d.
 * msal != 0: record pivots in ssd and set *det = 1 if any pivot is negative
 * (zero pivot -> error); otherwise any non-positive pivot is an error. */
int orc_skyfact(long neq, const long *maxa, double *ss, double *ssd, int msal, int *det)
{
    *det = 0;
    for (long n = 1; n <= neq; ++n) {
        long diag = maxa[n - 1];           /* address of d_n           */
        long top = maxa[n] - 1;            /* address of the top row   */
        long height = top - diag;          /* off-diagonal entries     */
        if (height >= 2) {
            /* g_kn = a_kn - sum_r l_rk g_rn for rows k from second-from-top
               down to n-1; the top entry needs no update */
            long k = n - (height - 1);
            for (long j = 1; j <= height - 1; ++j, ++k) {
                long at = top - j;                       /* address of row k   */
                long kd = maxa[k - 1];
                long kh = maxa[k] - kd - 1;              /* height of column k */
                if (kh > 0) {
                    long cnt = kh < j ? kh : j;
                    double c = 0;
                    for (long l = 1; l <= cnt; ++l) c += ss[kd + l - 1] * ss[at + l - 1];
                    ss[at - 1] -= c;
                }
            }
        }
        if (height >= 1) {
            double b = 0;
            long k = n;
            for (long a = diag + 1; a <= top; ++a) {
                k--;
                double c = ss[a - 1] / ss[maxa[k - 1] - 1];
                b += c * ss[a - 1];
                ss[a - 1] = c;
            }
            ss[diag - 1] -= b;
        }
        double piv = ss[diag - 1];
        if (!msal) {
            if (piv <= 0) return 1;
        } else {
            ssd[n - 1] = piv;
            if (piv == 0) return 1;
            if (!(piv > 0 && *det != 1)) *det = 1;
        }
    }
    return 0;
}

#include <time.h>
static double wall_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

int orc_skyfact_partial(long neq, const long *maxa, double *ss, double budget_s, long *cols)
{
    double t0 = wall_s();
    long n;
    for (n = 1; n <= neq; ++n) {
        long diag = maxa[n - 1], top = maxa[n] - 1, height = top - diag;
        if (height >= 2) {
            long k = n - (height - 1);
            for (long j = 1; j <= height - 1; ++j, ++k) {
                long at = top - j, kd = maxa[k - 1], kh = maxa[k] - kd - 1;
                if (kh > 0) {
                    long cnt = kh < j ? kh : j;
                    double c = 0;
                    for (long l = 1; l <= cnt; ++l) c += ss[kd + l - 1] * ss[at + l - 1];
                    ss[at - 1] -= c;
                }
            }
        }
        if (height >= 1) {
            double b = 0;
            long k = n;
            for (long a = diag + 1; a <= top; ++a) {
                k--;
                double c = ss[a - 1] / ss[maxa[k - 1] - 1];
                b += c * ss[a - 1];
                ss[a - 1] = c;
            }
            ss[diag - 1] -= b;
        }
        if ((n & 255) == 0 && wall_s() - t0 > budget_s) break;
    }
    *cols = n > neq ? neq : n;
    return 0;
}

/* Forward reduction, diagonal scaling and back substitution on a factor from
 * orc_skyfact; rhs is overwritten by the solution. */
void orc_skysolve(long neq, const long *maxa, const double *ss, double *rhs)
{
    for (long n = 1; n <= neq; ++n) {
        long lo = maxa[n - 1] + 1, hi = maxa[n] - 1;
        if (hi >= lo) {
            double c = 0;
            long k = n;
            for (long a = lo; a <= hi; ++a) {
                k--;
                c += ss[a - 1] * rhs[k - 1];
            }
            rhs[n - 1] -= c;
        }
    }
    for (long n = 0; n < neq; ++n) rhs[n] /= ss[maxa[n] - 1];
    for (long n = neq; n >= 2; --n) {
        long lo = maxa[n - 1] + 1, hi = maxa[n] - 1;
        long k = n;
        for (long a = lo; a <= hi; ++a) {
            k--;
            rhs[k - 1] -= ss[a - 1] * rhs[n - 1];
        }
    }
}

static double vdot(long n, const double *a, const double *b)
{
    double s = 0;
    for (long i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

/* misc.c:187-250.  convchk += 10 / 100 / 1000 for the displacement, force and
 * energy criteria that fail; returns 1 on a zero-norm error. */
int orc_test(long neq, const double *d_temp, const double *dd, const double *f_temp,
             const double *fp, const double *qtot, const double *f_ip, double intener1,
             int *convchk, double toldisp, double tolforc, double tolener)
{
    *convchk = 0;
    if (getenv("ORC_VERBOSE")) {
        double dn = vdot(neq, dd, dd), tn = vdot(neq, d_temp, d_temp), ui = 0, up = 0, e = 0;
        for (long i = 0; i < neq; ++i) {
            ui += (qtot[i] - f_temp[i]) * (qtot[i] - f_temp[i]);
            up += (qtot[i] - fp[i]) * (qtot[i] - fp[i]);
            e += dd[i] * (qtot[i] - f_ip[i]);
        }
        fprintf(stderr, "test: disp %.12e force %.12e energy %.12e\n", sqrt(dn) / sqrt(tn),
                sqrt(ui) / sqrt(up), fabs(e / intener1));
    }
    if (toldisp < 1) {
        double dn = vdot(neq, dd, dd), tn = vdot(neq, d_temp, d_temp);
        if (tn == 0) return 1;
        if (sqrt(dn) / sqrt(tn) > toldisp) *convchk += 10;
    }
    if (tolforc < 1) {
        double ui = 0, up = 0;
        for (long i = 0; i < neq; ++i) {
            ui += (qtot[i] - f_temp[i]) * (qtot[i] - f_temp[i]);
            up += (qtot[i] - fp[i]) * (qtot[i] - fp[i]);
        }
        if (up == 0) return 1;
        if (sqrt(ui) / sqrt(up) > tolforc) *convchk += 100;
    }
    if (tolener < 1) {
        double e = 0;
        for (long i = 0; i < neq; ++i) e += dd[i] * (qtot[i] - f_ip[i]);
        if (intener1 == 0) return 1;
        if (fabs(e / intener1) > tolener) *convchk += 1000;
    }
    return 0;
}

/* arc.c:70-157.  Roots of a x^2 + b x + c; keeps the root whose displacement
 * increment makes the larger positive projection on the current increment.
 * Returns 2 for complex roots, 3 when neither projection is positive. */
int orc_quad(long neq, double a, double b, double c, double *dt, const double *dp,
             const double *ddr, const double *ddq, double *dd, double *dlpf, double *lpft)
{
    double disc = b * b - 4 * a * c;
    if (!(disc > 0)) return 2;
    double r[2] = {(-b + sqrt(b * b - 4 * a * c)) / (2 * a), (-b - sqrt(b * b - 4 * a * c)) / (2 * a)};
    double g[2] = {0, 0};
    for (int s = 0; s < 2; ++s)
        for (long i = 0; i < neq; ++i) {
            double di = dt[i] - dp[i] + ddr[i] + r[s] * ddq[i];
            g[s] += (dt[i] - dp[i]) * di;
        }
    int pick;
    if (g[0] > g[1] && g[0] > 0) pick = 0;
    else if (g[1] > g[0] && g[1] > 0) pick = 1;
    else return 3;
    for (long i = 0; i < neq; ++i) {
        double di = dt[i] - dp[i] + ddr[i] + r[pick] * ddq[i];
        dt[i] = di + dp[i];
        dd[i] = ddr[i] + r[pick] * ddq[i];
    }
    *dlpf = r[pick];
    *lpft += *dlpf;
    return 0;
}

/* skymult (solve.c:700-760): the upper triangle, the diagonal and the lower
 * triangle of the profile applied to v in turn, column by column. */
void orc_skymult(long neq, const long *maxa, const double *ss, const double *v, double *out)
{
    for (long i = 0; i < neq; ++i) out[i] = 0;
    for (long n = neq; n >= 2; --n) {            /* rows above the diagonal of column n */
        long k = n;
        for (long a = maxa[n - 1] + 1; a <= maxa[n] - 1; ++a) {
            k--;
            out[k - 1] += ss[a - 1] * v[n - 1];
        }
    }
    for (long n = 0; n < neq; ++n) out[n] += v[n] * ss[maxa[n] - 1];
    for (long n = neq; n >= 1; --n) {            /* the same entries as rows of the lower part */
        long k = n;
        double c = 0;
        for (long a = maxa[n - 1] + 1; a <= maxa[n] - 1; ++a) {
            k--;
            c += ss[a - 1] * v[k - 1];
        }
        out[n - 1] += c;
    }
}
