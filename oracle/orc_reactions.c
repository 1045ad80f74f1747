/*
 * Support reactions (TEST INFRASTRUCTURE ONLY; see cb_oracle.h).
 *
 * The reference computes none.  A reaction is the internal force its element
 * routines accumulate into f -- T_i^T ef_i, truss.c:349-372, frame.c:1273-1309
 * (elastic, no member loads), shell.c:1728-1785 -- at a constrained joint slot
 * (jcode 0), where those routines skip the scatter because mcode is 0.  The
 * restatement walks the elements in input order and adds every end force of
 * every element at its joint slot, then keeps the constrained slots.
 */
#include <string.h>

#include "cb_oracle.h"

/* global end force of one element end: rows of T (the triad c1, c2, c3 of the
 * element) applied to a local force/moment triple f3 */
static void rotate_back(const double *c1, const double *c2, const double *c3, const double *f3,
                        double *out)
{
    for (int m = 0; m < 3; ++m) out[m] += c1[m] * f3[0] + c2[m] * f3[1] + c3[m] * f3[2];
}

void orc_reactions(const orc_model *m, const double *ef, const double *c1, const double *c2,
                   const double *c3, double *r)
{
    const orc_dims *d = &m->dm;
    const long ntr = d->NE_TR, nfr = d->NE_FR, nsh = d->NE_SH;
    memset(r, 0, sizeof(double) * (size_t)(d->NJ * 7));
    /* truss: f -= ef[end] * c (truss.c:349-372).  The linear branch
     * (ANAFLAG 1, truss.c:300-305) gives end forces of the opposite sign,
     * EA/L0 (d1 - d2) at end 1, so its f is -K d: the reactions take its end
     * forces with the sign flipped back. */
    const double s_tr = d->ANAFLAG == 1 ? -1.0 : 1.0;
    for (long e = 0; e < ntr; ++e) {
        const double c[3] = {c1[e], c2[e], c3[e]};
        for (int a = 0; a < 2; ++a) {
            long j = m->minc[e * 2 + a] - 1;
            for (int k = 0; k < 3; ++k) r[j * 7 + k] -= s_tr * ef[e * 2 + a] * c[k];
        }
    }
    /* frame: per end a force triple, a moment triple and the warping
     * bimoment; T_i is block diagonal in the triad, 1 on warping
     * (frame.c:286-295); an end offset o moves the end force F to the joint
     * with the moment o x F added (G^T, frame.c:297-324) */
    for (long e = 0; e < nfr; ++e) {
        const long t = ntr + e * 3;
        const double *fe = ef + ntr * 2 + e * 14;
        for (int a = 0; a < 2; ++a) {
            long j = m->minc[ntr * 2 + e * 2 + a] - 1;
            double F[3] = {0, 0, 0};
            rotate_back(c1 + t, c2 + t, c3 + t, fe + 7 * a, F);
            for (int k = 0; k < 3; ++k) r[j * 7 + k] += F[k];
            rotate_back(c1 + t, c2 + t, c3 + t, fe + 7 * a + 3, r + j * 7 + 3);
            if (m->osflag[e]) {
                const double *o = m->offset + e * 6 + 3 * a;
                r[j * 7 + 3] += o[1] * F[2] - o[2] * F[1];
                r[j * 7 + 4] += o[2] * F[0] - o[0] * F[2];
                r[j * 7 + 5] += o[0] * F[1] - o[1] * F[0];
            }
            r[j * 7 + 6] += fe[7 * a + 6];
        }
    }
    /* shell: per vertex a force and a moment triple in the element triad
     * (shell.c:1775-1785) */
    for (long e = 0; e < nsh; ++e) {
        const long t = ntr + nfr * 3 + e * 3;
        const double *fe = ef + ntr * 2 + nfr * 14 + e * 18;
        for (int a = 0; a < 3; ++a) {
            long j = m->minc[ntr * 2 + nfr * 2 + e * 3 + a] - 1;
            rotate_back(c1 + t, c2 + t, c3 + t, fe + 6 * a, r + j * 7);
            rotate_back(c1 + t, c2 + t, c3 + t, fe + 6 * a + 3, r + j * 7 + 3);
        }
    }
    for (long i = 0; i < d->NJ * 7; ++i)
        if (m->jcode[i] != 0) r[i] = 0;
}
