/*
 * orc_drivers.c -- oracle: the nonlinear drivers of main.c restated over the
 * skyline LDL^T, with the element routines taken from a backend table so the
 * same driver runs on the oracle's restatement or on the reference's own
 * element routines (oracle/_ref).
 *
 * TEST INFRASTRUCTURE ONLY (see cb_oracle.h).
 *
 * NR / modified NR: main.c:1800-2157.  MSAL arc-length: main.c:2158-3141.
 * Nonlinear Newmark: main.c:3305-3960 with solve.c:139-245, 469-533.
 */
#include "cb_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* ---- default backend: the oracle's own element routines ---------------- */
static void be_stiff_tr(void *ctx, const orc_dims *d, double *ss, const double *emod,
                        const double *carea, const double *llength, const double *defllen_ip,
                        const double *yield, const double *c1_ip, const double *c2_ip,
                        const double *c3_ip, const double *ef_ip, const long *maxa,
                        const long *mcode)
{
    (void)ctx;
    orc_stiff_tr(d, ss, emod, carea, llength, defllen_ip, yield, c1_ip, c2_ip, c3_ip, ef_ip, maxa,
                 mcode);
}
static void be_forces_tr(void *ctx, const orc_dims *d, double *f, double *ef_i,
                         const double *emod, const double *carea, const double *llength,
                         const double *defllen_i, const double *yield, const double *c1_i,
                         const double *c2_i, const double *c3_i, const long *mcode,
                         const double *d_tot)
{
    (void)ctx;
    if (d->ANAFLAG == 1)
        orc_forces_tr_lin(d, f, ef_i, d_tot, emod, carea, llength, c1_i, c2_i, c3_i, mcode);
    else
        orc_forces_tr(d, f, ef_i, emod, carea, llength, defllen_i, yield, c1_i, c2_i, c3_i, mcode);
}
static void be_stiff_sh(void *ctx, const orc_dims *d, double *ss, const double *emod,
                        const double *nu, const double *x_temp, const double *xlocal,
                        const double *thick, const double *farea, const double *deffarea_ip,
                        const double *slength, const double *defslen_ip, const double *yield,
                        const double *c1_ip, const double *c2_ip, const double *c3_ip,
                        const double *ef_ip, const long *maxa, const long *minc,
                        const long *mcode)
{
    (void)ctx; (void)defslen_ip; (void)yield; (void)ef_ip;
    orc_stiff_sh(d, ss, emod, nu, x_temp, xlocal, thick, farea, deffarea_ip, slength, c1_ip,
                 c2_ip, c3_ip, maxa, minc, mcode);
}
static int be_forces_sh(void *ctx, const orc_dims *d, double *f, double *ef_ip, double *ef_i,
                        const double *dd, const double *x_temp, const double *x_ip,
                        const double *emod, const double *nu, const double *xlocal,
                        const double *thick, const double *farea, const double *deffarea_ip,
                        const double *slength, const double *defslen_ip, const double *yield,
                        const double *c1_ip, const double *c2_ip, const double *c3_ip,
                        const double *c1_i, const double *c2_i, const double *c3_i,
                        const long *minc, const long *mcode, const long *jcode,
                        const double *d_tot)
{
    (void)ctx; (void)x_ip; (void)deffarea_ip; (void)defslen_ip; (void)yield; (void)jcode;
    if (d->ANAFLAG == 1)
        orc_forces_sh_lin(d, f, ef_i, d_tot, emod, nu, xlocal, thick, farea, slength, c1_ip,
                          c2_ip, c3_ip, mcode);
    else
        orc_forces_sh(d, f, ef_ip, ef_i, dd, x_temp, emod, nu, xlocal, thick, farea, slength,
                      c1_ip, c2_ip, c3_ip, c1_i, c2_i, c3_i, minc, mcode);
    return 0;
}
static void be_updatc(void *ctx, const orc_dims *d, double *x_temp, double *x_ip,
                      double *xfr_temp, const double *dd, double *defllen_i,
                      double *deffarea_i, double *defslen_i, const double *offset,
                      const int *osflag, const double *auxpt, double *c1_i, double *c2_i,
                      double *c3_i, const long *minc, const long *jcode)
{
    (void)ctx;
    orc_updatc(d, x_temp, x_ip, xfr_temp, dd, defllen_i, deffarea_i, defslen_i, offset, osflag,
               auxpt, c1_i, c2_i, c3_i, minc, jcode);
}

static void be_stiff_fr(void *ctx, const orc_dims *d, double *ss, const orc_model *m,
                        const double *llength, const double *defllen_ip, const double *c1_ip,
                        const double *c2_ip, const double *c3_ip, const double *ef_ip,
                        const double *efFE_ip, int *yldflag)
{
    (void)ctx;
    orc_stiff_fr(d, ss, m, llength, defllen_ip, c1_ip, c2_ip, c3_ip, ef_ip, efFE_ip, yldflag);
}
static int be_forces_fr(void *ctx, const orc_dims *d, double *f, double *ef_ip, double *ef_i,
                        const double *efFE_ip, double *efFE_i, const double *dd,
                        const orc_model *m, const double *llength, const double *defllen_ip,
                        const double *c1_ip, const double *c2_ip, const double *c3_ip,
                        const double *c1_i, const double *c2_i, const double *c3_i,
                        int *yldflag, double *dlpf, int itecnt)
{
    (void)ctx; (void)itecnt;
    return orc_forces_fr(d, f, ef_ip, ef_i, efFE_ip, efFE_i, dd, m, llength, defllen_ip, c1_ip,
                         c2_ip, c3_ip, c1_i, c2_i, c3_i, yldflag, dlpf);
}
static void be_mass_tr(void *ctx, const orc_dims *d, double *sm, const orc_model *m,
                       const double *x, double *llength, const double *dens)
{
    (void)ctx;
    orc_mass_tr(d, sm, m, x, llength, dens);
}
static void be_mass_fr(void *ctx, const orc_dims *d, double *sm, const orc_model *m,
                       const double *x, double *llength, double *xfr, const double *dens)
{
    (void)ctx;
    orc_mass_fr(d, sm, m, x, llength, xfr, dens);
}
static void be_mass_sh(void *ctx, const orc_dims *d, double *sm, const orc_model *m,
                       const double *x, double *farea, double *slength, const double *dens)
{
    (void)ctx;
    orc_mass_sh(d, sm, m, x, farea, slength, dens);
}

void orc_default_backend(orc_backend *b)
{
    b->ctx = NULL;
    b->stiff_tr = be_stiff_tr;
    b->forces_tr = be_forces_tr;
    b->stiff_sh = be_stiff_sh;
    b->forces_sh = be_forces_sh;
    b->updatc = be_updatc;
    b->stiff_fr = be_stiff_fr;
    b->forces_fr = be_forces_fr;
    b->mass_tr = be_mass_tr;
    b->mass_fr = be_mass_fr;
    b->mass_sh = be_mass_sh;
}

/* ---- state of one analysis -------------------------------------------------
 * Three generations of every per-element quantity, as in the reference:
 * base (last converged), _i (current iterate), _ip (previous iterate). */
typedef struct {
    orc_model *m;
    const orc_backend *be;
    long neq, nj, ntri, nef, nlen, nsh, nfr;
    double *ss, *sm, *ssd, *ssd_o;
    double *d, *dp, *dpp, *d_temp, *dd, *ddq, *ddr, *f, *fp, *f_temp, *f_ip, *qtot, *r;
    double *x, *x_temp, *x_ip, *xfr, *xfr_temp;
    double *ef, *ef_i, *ef_ip;
    double *c1, *c2, *c3, *c1_i, *c2_i, *c3_i, *c1_ip, *c2_ip, *c3_ip;
    double *defllen, *defllen_i, *defllen_ip, *llen_ref;
    double *deffarea, *deffarea_i, *deffarea_ip, *defslen, *defslen_i, *defslen_ip;
    double *efFE, *efFE_i, *efFE_ip;
    double *farea_w, *slen_w;          /* face areas / side lengths the elements read */
    double *keff, *uc, *vc, *ac, *uc_i, *vc_i, *ac_i, *um, *vm, *am, *dyn;
    int *yld;                          /* yldflag: 0, 1 yielded, 2 unloading (per member end) */
    orc_path *out;
    int ev[ORC_NEV];                   /* events since the last recorded step */
    double ev_psi;                     /* psi of the step being taken */
} run_t;

static double *dz(long n) { return calloc((size_t)(n > 0 ? n : 1), sizeof(double)); }
#define CP(dst, src, n) memcpy((dst), (src), sizeof(double) * (size_t)(n))

static void run_init(run_t *R, orc_model *m, const orc_backend *be, orc_path *out)
{
    memset(R, 0, sizeof *R);
    R->ev_psi = 1;
    const orc_dims *d = &m->dm;
    R->m = m; R->be = be; R->out = out;
    R->neq = d->NEQ; R->nj = d->NJ; R->nsh = d->NE_SH; R->nfr = d->NE_FR;
    R->ntri = d->NE_TR + d->NE_FR * 3 + d->NE_SH * 3;
    R->nef = d->NE_TR * 2 + d->NE_FR * 14 + d->NE_SH * 18;
    R->nlen = d->NE_TR + d->NE_FR;
    long n = R->neq;
    R->ss = dz(m->lss); R->sm = dz(n); R->ssd = dz(n); R->ssd_o = dz(n);
    R->d = dz(n); R->dp = dz(n); R->dpp = dz(n); R->d_temp = dz(n); R->dd = dz(n);
    R->ddq = dz(n); R->ddr = dz(n); R->f = dz(n); R->fp = dz(n); R->f_temp = dz(n);
    R->f_ip = dz(n); R->qtot = dz(n); R->r = dz(n);
    R->x = dz(R->nj * 3); R->x_temp = dz(R->nj * 3); R->x_ip = dz(R->nj * 3);
    R->xfr = dz(R->nfr * 6); R->xfr_temp = dz(R->nfr * 6);
    R->ef = dz(R->nef); R->ef_i = dz(R->nef); R->ef_ip = dz(R->nef);
    R->c1 = dz(R->ntri); R->c2 = dz(R->ntri); R->c3 = dz(R->ntri);
    R->c1_i = dz(R->ntri); R->c2_i = dz(R->ntri); R->c3_i = dz(R->ntri);
    R->c1_ip = dz(R->ntri); R->c2_ip = dz(R->ntri); R->c3_ip = dz(R->ntri);
    R->defllen = dz(R->nlen); R->defllen_i = dz(R->nlen); R->defllen_ip = dz(R->nlen);
    R->llen_ref = dz(R->nlen);
    R->deffarea = dz(R->nsh); R->deffarea_i = dz(R->nsh); R->deffarea_ip = dz(R->nsh);
    R->defslen = dz(R->nsh * 3); R->defslen_i = dz(R->nsh * 3); R->defslen_ip = dz(R->nsh * 3);
    R->efFE = dz(R->nfr * 14); R->efFE_i = dz(R->nfr * 14); R->efFE_ip = dz(R->nfr * 14);
    R->farea_w = dz(R->nsh); R->slen_w = dz(R->nsh * 3);
    R->keff = dz(m->dm.ALGFLAG >= 4 ? m->lss : 1);
    R->uc = dz(n); R->vc = dz(n); R->ac = dz(n); R->uc_i = dz(n); R->vc_i = dz(n);
    R->ac_i = dz(n); R->um = dz(n); R->vm = dz(n); R->am = dz(n); R->dyn = dz(n);
    CP(R->farea_w, m->farea, R->nsh);
    CP(R->slen_w, m->slength, R->nsh * 3);
    R->yld = calloc((size_t)(R->nfr * 2 + 1), sizeof(int));

    CP(R->x, m->x, R->nj * 3);
    CP(R->x_temp, m->x, R->nj * 3);
    CP(R->xfr, m->xfr, R->nfr * 6);
    CP(R->xfr_temp, m->xfr, R->nfr * 6);
    CP(R->c1, m->c1, R->ntri); CP(R->c2, m->c2, R->ntri); CP(R->c3, m->c3, R->ntri);
    CP(R->llen_ref, m->llength, R->nlen);
    CP(R->defllen, m->llength, R->nlen);
    CP(R->deffarea, m->farea, R->nsh);
    CP(R->defslen, m->slength, R->nsh * 3);
}

static void run_free(run_t *R)
{
    double **p = &R->ss, **e = &R->dyn;
    for (; p <= e; ++p) free(*p);
    free(R->yld);
}

/* _ip <- _i for the per-element generations (main.c:2390-2411). */
static void shift_ip(run_t *R, int with_ef)
{
    if (with_ef) CP(R->ef_ip, R->ef_i, R->nef);
    CP(R->c1_ip, R->c1_i, R->ntri); CP(R->c2_ip, R->c2_i, R->ntri); CP(R->c3_ip, R->c3_i, R->ntri);
    CP(R->defllen_ip, R->defllen_i, R->nlen);
    CP(R->efFE_ip, R->efFE_i, R->nfr * 14);
    CP(R->deffarea_ip, R->deffarea_i, R->nsh);
    CP(R->defslen_ip, R->defslen_i, R->nsh * 3);
}

/* _i, _ip <- base (main.c:2583-2622). */
static void reset_from_base(run_t *R)
{
    CP(R->x_temp, R->x, R->nj * 3);
    CP(R->ef_i, R->ef, R->nef); CP(R->ef_ip, R->ef, R->nef);
    CP(R->c1_i, R->c1, R->ntri); CP(R->c2_i, R->c2, R->ntri); CP(R->c3_i, R->c3, R->ntri);
    CP(R->c1_ip, R->c1, R->ntri); CP(R->c2_ip, R->c2, R->ntri); CP(R->c3_ip, R->c3, R->ntri);
    CP(R->defllen_i, R->defllen, R->nlen); CP(R->defllen_ip, R->defllen, R->nlen);
    CP(R->xfr_temp, R->xfr, R->nfr * 6);
    CP(R->efFE_i, R->efFE, R->nfr * 14); CP(R->efFE_ip, R->efFE, R->nfr * 14);
    CP(R->deffarea_i, R->deffarea, R->nsh); CP(R->deffarea_ip, R->deffarea, R->nsh);
    CP(R->defslen_i, R->defslen, R->nsh * 3); CP(R->defslen_ip, R->defslen, R->nsh * 3);
}

/* base <- _i (main.c:3055-3110); an unloading end (2) is elastic again
   (main.c:3086-3091). */
static void commit(run_t *R)
{
    for (long i = 0; i < R->nfr * 2; ++i)
        if (R->yld[i] == 2) R->yld[i] = 0;
    CP(R->x, R->x_temp, R->nj * 3);
    CP(R->ef, R->ef_i, R->nef);
    CP(R->c1, R->c1_i, R->ntri); CP(R->c2, R->c2_i, R->ntri); CP(R->c3, R->c3_i, R->ntri);
    CP(R->defllen, R->defllen_i, R->nlen);
    CP(R->xfr, R->xfr_temp, R->nfr * 6);
    CP(R->efFE, R->efFE_i, R->nfr * 14);
    CP(R->deffarea, R->deffarea_i, R->nsh);
    CP(R->defslen, R->defslen_i, R->nsh * 3);
}

/* ss = 0; stiff_tr/fr/sh over the chosen state. */
static void assemble(run_t *R, const double *x, const double *c1, const double *c2,
                     const double *c3, const double *ef, const double *defllen,
                     const double *deffarea, const double *defslen, const double *efFE)
{
    orc_model *m = R->m;
    const orc_dims *d = &m->dm;
    double t0 = now_s();
    memset(R->ss, 0, sizeof(double) * (size_t)m->lss);
    if (d->NE_TR > 0)
        R->be->stiff_tr(R->be->ctx, d, R->ss, m->emod, m->carea, R->llen_ref, defllen, m->yield,
                        c1, c2, c3, ef, m->maxa, m->mcode);
    if (d->NE_FR > 0)
        R->be->stiff_fr(R->be->ctx, d, R->ss, m, R->llen_ref, defllen, c1, c2, c3, ef, efFE,
                        R->yld);
    if (d->NE_SH > 0)
        R->be->stiff_sh(R->be->ctx, d, R->ss, m->emod, m->nu, x, m->xlocal, m->thick, R->farea_w,
                        deffarea, R->slen_w, defslen, m->yield, c1, c2, c3, ef, m->maxa,
                        m->minc, m->mcode);
    if (R->out) { R->out->t_tangent += now_s() - t0; }
}

/* updatc + forces_* into f_out (zeroed by the caller as in the reference);
   returns forces_fr's code, with *dlpf the load factor (increment) that a
   code 1 scales. */
static int update_and_forces(run_t *R, double *x_temp, double *xfr_temp, double *f_out,
                             const double *dd, double *dlpf, int itecnt)
{
    orc_model *m = R->m;
    const orc_dims *d = &m->dm;
    double t0 = now_s();
    R->be->updatc(R->be->ctx, d, x_temp, R->x_ip, xfr_temp, dd, R->defllen_i, R->deffarea_i,
                  R->defslen_i, m->offset, m->osflag, m->auxpt, R->c1_i, R->c2_i, R->c3_i,
                  m->minc, m->jcode);
    if (d->NE_TR > 0)
        R->be->forces_tr(R->be->ctx, d, f_out, R->ef_i, m->emod, m->carea, R->llen_ref,
                         R->defllen_i, m->yield, R->c1_i, R->c2_i, R->c3_i, m->mcode, R->d_temp);
    int chk = 0;
    if (d->NE_FR > 0)
        chk = R->be->forces_fr(R->be->ctx, d, f_out, R->ef_ip, R->ef_i, R->efFE_ip, R->efFE_i, dd,
                               m, R->llen_ref, R->defllen_ip, R->c1_ip, R->c2_ip, R->c3_ip,
                               R->c1_i, R->c2_i, R->c3_i, R->yld, dlpf, itecnt);
    if (d->NE_SH > 0)
        (void)R->be->forces_sh(R->be->ctx, d, f_out, R->ef_ip, R->ef_i, dd, x_temp, R->x_ip,
                               m->emod, m->nu, m->xlocal, m->thick, R->farea_w, R->deffarea_ip,
                               R->slen_w, R->defslen_ip, m->yield, R->c1_ip, R->c2_ip,
                               R->c3_ip, R->c1_i, R->c2_i, R->c3_i, m->minc, m->mcode,
                               m->jcode, R->d_temp);
    if (R->out) R->out->t_forces += now_s() - t0;
    return chk;
}

static int factor(run_t *R, int *det)
{
    double t0 = now_s();
    int e = orc_skyfact(R->neq, R->m->maxa, R->ss, R->ssd, R->m->dm.ALGFLAG == 3, det);
    R->ev[ORC_EV_DET] = *det;
    if (R->out) { R->out->t_factor += now_s() - t0; R->out->n_factor++; }
    return e;
}

static void backsolve(run_t *R, const double *rhs, double *x)
{
    double t0 = now_s();
    CP(x, rhs, R->neq);
    orc_skysolve(R->neq, R->m->maxa, R->ss, x);
    if (R->out) { R->out->t_solve += now_s() - t0; R->out->n_solve++; }
}

static void record(run_t *R, double lpf, int itecnt, const double *d, const double *f)
{
    orc_path *o = R->out;
    if (!o || o->nsteps >= o->max_steps) return;
    if (o->lpf) o->lpf[o->nsteps] = lpf;
    if (o->itecnt) o->itecnt[o->nsteps] = itecnt;
    if (o->d) CP(o->d + (long)o->nsteps * R->neq, d, R->neq);
    if (o->f && f) CP(o->f + (long)o->nsteps * R->neq, f, R->neq);
    if (o->events) memcpy(o->events + (long)o->nsteps * ORC_NEV, R->ev, sizeof R->ev);
    if (o->psi) o->psi[o->nsteps] = R->ev_psi;
    if (o->react)
        orc_reactions(R->m, R->ef, R->c1, R->c2, R->c3, o->react + (long)o->nsteps * R->nj * 7);
    if (o->ef) CP(o->ef + (long)o->nsteps * R->nef, R->ef, R->nef);
    if (o->resid) {
        double rr = 0, qq = 0;
        if (R->m->dm.ALGFLAG <= 3 && f)
            for (long i = 0; i < R->neq; ++i) {
                double t = lpf * R->m->q[i] - f[i];
                rr += t * t;
                qq += R->m->q[i] * R->m->q[i];
            }
        o->resid[2 * o->nsteps] = sqrt(rr);
        o->resid[2 * o->nsteps + 1] = fabs(lpf) * sqrt(qq);
    }
    o->nsteps++;
    memset(R->ev, 0, sizeof R->ev);
}

static double vdot(long n, const double *a, const double *b)
{
    double s = 0;
    for (long i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

#define FAIL(code, txt) do { if (R->out) { R->out->status = (code); snprintf(R->out->msg, sizeof R->out->msg, "%s", (txt)); } return (code); } while (0)

/* ---- NR / modified NR with load control (main.c:1800-2157) --------------- */
static int run_nr(run_t *R)
{
    orc_model *m = R->m;
    long n = R->neq;
    double lpf = m->lpf0, dlpf = m->dlpf0, dlpfp;
    int inccnt = 0, solcnt = 0, subcnt = 0, itecnt = 0, convchk = 0, det = 0;
    double intener1 = 0;
    CP(R->defllen_i, R->defllen, R->nlen); CP(R->defllen_ip, R->defllen, R->nlen);
    do {
        if (lpf > m->lpfmax) lpf = m->lpfmax;
        for (long i = 0; i < n; ++i) {
            R->qtot[i] = m->q[i] * lpf;
            R->fp[i] = R->f[i];
            R->d_temp[i] = R->d[i];
            R->f_temp[i] = R->f[i];
        }
        dlpfp = dlpf;
        reset_from_base(R);
        itecnt = 0;
        int frc = 0;
        do {
            for (long i = 0; i < n; ++i) R->r[i] = R->qtot[i] - R->f_temp[i];
            int full = (m->dm.ALGFLAG == 1 || itecnt == 0);
            if (full)
                assemble(R, R->x_temp, R->c1_ip, R->c2_ip, R->c3_ip, R->ef_ip, R->defllen_ip,
                         R->deffarea_ip, R->defslen_ip, R->efFE_ip);
            if (m->lss == 1) {
                R->dd[0] = R->r[0] / R->ss[0];
            } else {
                /* solve() discards skyfact's status (solve.c:86-88, 535): a
                   non-positive pivot leaves a partial factor that is used as is */
                if (full) factor(R, &det);
                backsolve(R, R->r, R->dd);
            }
            for (long i = 0; i < n; ++i) {
                R->d_temp[i] += R->dd[i];
                R->f_ip[i] = R->f_temp[i];
                R->f_temp[i] = 0;
            }
            frc = update_and_forces(R, R->x_temp, R->xfr_temp, R->f_temp, R->dd, &dlpf, itecnt);
            CP(R->ef_ip, R->ef_i, R->nef);
            if (itecnt == 0) {
                intener1 = 0;
                for (long i = 0; i < n; ++i) intener1 += R->dd[i] * (R->qtot[i] - R->fp[i]);
            }
            if (orc_test(n, R->d_temp, R->dd, R->f_temp, R->fp, R->qtot, R->f_ip, intener1,
                         &convchk, m->toldisp, m->tolforc, m->tolener))
                FAIL(1, "zero norm in convergence test");
            if (getenv("ORC_VERBOSE"))
                fprintf(stderr, "NR lpf=%g it=%d convchk=%d |dd|=%g\n", lpf, itecnt, convchk,
                        sqrt(vdot(n, R->dd, R->dd)));
            shift_ip(R, 0);
            itecnt++;
            if (R->out) R->out->n_newton++;
        } while (convchk != 0 && frc == 0 && itecnt <= m->itemax);

        if (frc == 2) {
            /* an unloading end: the same increment again (main.c:2033-2034) */
            dlpf = dlpfp;
            R->ev[ORC_EV_UNLOAD]++;
        } else if ((convchk != 0 || frc != 0) && subcnt <= m->submax) {
            if (lpf == m->lpfmax) FAIL(1, "maximum load factor attempted without convergence");
            if (dlpfp == m->dlpfmin) FAIL(1, "minimum load factor increment reached");
            /* code 1 keeps the increment forces_fr scaled (main.c:2050-2053) */
            if (frc != 1) dlpf = dlpfp / 2;
            if (dlpf < m->dlpfmin) dlpf = m->dlpfmin;
            lpf = lpf - dlpfp + dlpf;
            subcnt++;
            solcnt = 0;
            R->ev[frc == 1 ? ORC_EV_YIELD : ORC_EV_HALVE]++;
        } else if (subcnt > m->submax) {
            FAIL(1, "maximum number of subdivisions exceeded");
        } else {
            inccnt++;
            CP(R->d, R->d_temp, n);
            CP(R->f, R->f_temp, n);
            commit(R);
            record(R, lpf, itecnt, R->d, R->f);
            if (R->out && R->out->nsteps >= R->out->max_steps) return 0;
            solcnt++;
            subcnt = 0;
            if (solcnt >= m->solmin) {
                dlpf *= 2;
                if (dlpf > m->dlpfmax) dlpf = m->dlpfmax;
                solcnt = 0;
            }
            lpf += dlpf;
        }
    } while (lpf <= m->lpfmax);
    (void)inccnt;
    return 0;
}

/* ---- modified spherical arc length (main.c:2158-3141) -------------------- */
static int run_msal(run_t *R)
{
    orc_model *m = R->m;
    long n = R->neq, k = m->dkdof;
    int det = 0, convchk = 0, itecnt = 0, frc = 0;
    double lpf, dlpf, intener1, lpfp = 0, lpfpp = 0;

    CP(R->defllen_i, R->defllen, R->nlen); CP(R->defllen_ip, R->defllen, R->nlen);
    CP(R->deffarea_i, R->deffarea, R->nsh); CP(R->deffarea_ip, R->deffarea, R->nsh);
    CP(R->defslen_i, R->defslen, R->nsh * 3); CP(R->defslen_ip, R->defslen, R->nsh * 3);
    CP(R->c1_i, R->c1, R->ntri); CP(R->c2_i, R->c2, R->ntri); CP(R->c3_i, R->c3, R->ntri);
    CP(R->c1_ip, R->c1, R->ntri); CP(R->c2_ip, R->c2, R->ntri); CP(R->c3_ip, R->c3, R->ntri);

    /* first increment: displacement control at the MSAL DOF (Bathe-Dvorkin) */
    assemble(R, R->x, R->c1, R->c2, R->c3, R->ef, R->defllen, R->deffarea, R->defslen, R->efFE);
    if (m->lss == 1) {
        R->ddq[0] = m->q[0] / R->ss[0];
        R->ssd[0] = R->ss[0];
    } else {
        factor(R, &det);
        backsolve(R, m->q, R->ddq);
    }
    lpf = m->dk / R->ddq[k];
    intener1 = 0;
    for (long i = 0; i < n; ++i) {
        R->d[i] = R->dd[i] = lpf * R->ddq[i];
        intener1 += R->dd[i] * (lpf * m->q[i]);
        R->ssd_o[i] = R->ssd[i];
    }
    memset(R->f, 0, sizeof(double) * (size_t)n);
    /* the first call hands forces_fr lpf itself as the increment a code 1
       scales, and its code is dropped (main.c:2365-2369, 2411) */
    frc = update_and_forces(R, R->x, R->xfr, R->f, R->dd, &lpf, 0);
    itecnt = 1;
    shift_ip(R, 1);
    frc = 0;
    do {
        for (long i = 0; i < n; ++i) {
            R->qtot[i] = m->q[i] * lpf;
            R->r[i] = R->qtot[i] - R->f[i];
        }
        if (m->lss == 1) {
            R->ddr[0] = R->r[0] / R->ss[0];
            R->ssd[0] = R->ss[0];
        } else {
            backsolve(R, R->r, R->ddr);
        }
        dlpf = -R->ddr[k] / R->ddq[k];
        lpf += dlpf;
        for (long i = 0; i < n; ++i) {
            R->dd[i] = R->ddr[i] + dlpf * R->ddq[i];
            R->d[i] += R->dd[i];
            R->f_ip[i] = R->f[i];
            R->f[i] = 0;
        }
        frc = update_and_forces(R, R->x, R->xfr, R->f, R->dd, &dlpf, itecnt);
        if (orc_test(n, R->d, R->dd, R->f, R->fp, R->qtot, R->f_ip, intener1, &convchk,
                     m->toldisp, m->tolforc, m->tolener))
            FAIL(1, "zero norm in convergence test");
        itecnt++;
        if (R->out) R->out->n_newton++;
        shift_ip(R, 1);
    } while (convchk != 0 && frc == 0 && itecnt <= m->itemax);
    if (convchk != 0 || frc != 0) FAIL(1, "initially prescribed displacement too large");
    CP(R->ef, R->ef_i, R->nef);
    CP(R->c1, R->c1_i, R->ntri); CP(R->c2, R->c2_i, R->ntri); CP(R->c3, R->c3_i, R->ntri);
    CP(R->defllen, R->defllen_i, R->nlen);
    CP(R->efFE, R->efFE_i, R->nfr * 14);
    CP(R->deffarea, R->deffarea_i, R->nsh);
    CP(R->defslen, R->defslen_i, R->nsh * 3);
    record(R, lpf, itecnt, R->d, R->f);
    if (R->out && R->out->nsteps >= R->out->max_steps) return 0;

    double dkc = fabs(R->d[k]), lpfc = fabs(lpf);
    for (long i = 0; i < n; ++i) {
        R->dpp[i] = R->dp[i];
        R->dp[i] = R->d[i];
        R->fp[i] = R->f[i];
    }
    lpfpp = lpfp;
    lpfp = lpf;
    reset_from_base(R);

    double dnorm = sqrt(vdot(n, R->dp, R->dp));
    double dnormallow = m->alpha * sqrt(vdot(n, R->dp, R->dp));
    double beta = sqrt(((double)m->iteopt) / (double)itecnt) * (dnormallow / dnorm);
    double psi = 1;
    for (long i = 0; i < n; ++i) {
        double tq = fabs(R->ssd[i] / R->ssd_o[i]);
        if (tq < psi) psi = tq;
    }

    int errchk2 = 0, subcnt = 0, imagcnt = 0, negcnt = 0;
    double arc = 0, a = 0, b, c, lpf_temp;
    R->ev_psi = psi;
    while (lpfc <= m->lpfmax && dkc <= m->dkimax) {
        CP(R->d_temp, R->d, n);
        lpf_temp = lpf;
        if (errchk2 == 0) {
            double s = 0;
            for (long i = 0; i < n; ++i) s += (R->dp[i] - R->dpp[i]) * (R->dp[i] - R->dpp[i]);
            if (psi >= m->psi_thresh)
                arc = beta * sqrt(s + (lpfp - lpfpp) * (lpfp - lpfpp));
            else
                arc = beta * sqrt(s);
        } else {
            reset_from_base(R);
            errchk2 = 0;
        }
        /* tangent at the last converged state, factorised once per increment */
        assemble(R, R->x, R->c1, R->c2, R->c3, R->ef, R->defllen, R->deffarea, R->defslen,
                 R->efFE);
        if (m->lss == 1) {
            R->ddq[0] = m->q[0] / R->ss[0];
            R->ssd[0] = R->ss[0];
            det = R->ss[0] > 0 ? 0 : 1;
            R->ev[ORC_EV_DET] = det;
        } else {
            factor(R, &det);
            backsolve(R, m->q, R->ddq);
        }
        if (psi >= m->psi_thresh)
            a = vdot(n, m->q, m->q) + vdot(n, R->ddq, R->ddq);
        else
            a = vdot(n, R->ddq, R->ddq);
        dlpf = (det == 0) ? arc * sqrt(1 / a) : -arc * sqrt(1 / a);
        lpf_temp += dlpf;
        intener1 = 0;
        for (long i = 0; i < n; ++i) {
            R->dd[i] = dlpf * R->ddq[i];
            R->d_temp[i] += R->dd[i];
            intener1 += R->dd[i] * (dlpf * m->q[i]);
            R->f_temp[i] = 0;
        }
        /* the predictor's code is not looked at; its flags stay, and a return
           leaves the later members out of f_temp (main.c:2794-2800) */
        update_and_forces(R, R->x_temp, R->xfr_temp, R->f_temp, R->dd, &dlpf, itecnt);
        shift_ip(R, 1);
        itecnt = 1;
        frc = 0;
        if (R->out) R->out->n_newton++;
        do {
            for (long i = 0; i < n; ++i) {
                R->qtot[i] = m->q[i] * lpf_temp;
                R->r[i] = R->qtot[i] - R->f_temp[i];
                R->f_ip[i] = R->f_temp[i];
                R->f_temp[i] = 0;
            }
            if (m->lss == 1)
                R->ddr[0] = R->r[0] / R->ss[0];
            else
                backsolve(R, R->r, R->ddr);
            double qq = vdot(n, m->q, m->q);
            if (psi >= m->psi_thresh) {
                b = 2 * (vdot(n, R->d_temp, R->ddq) - vdot(n, R->dp, R->ddq) +
                         vdot(n, R->ddr, R->ddq) + (lpf_temp - lpfp) * qq);
                c = 2 * (vdot(n, R->d_temp, R->ddr) - vdot(n, R->dp, R->ddr) -
                         vdot(n, R->d_temp, R->dp)) +
                    vdot(n, R->d_temp, R->d_temp) + vdot(n, R->dp, R->dp) +
                    vdot(n, R->ddr, R->ddr) + (lpf_temp - lpfp) * (lpf_temp - lpfp) * qq -
                    arc * arc;
            } else {
                b = 2 * (vdot(n, R->d_temp, R->ddq) - vdot(n, R->dp, R->ddq) +
                         vdot(n, R->ddr, R->ddq));
                c = 2 * (vdot(n, R->d_temp, R->ddr) - vdot(n, R->dp, R->ddr) -
                         vdot(n, R->d_temp, R->dp)) +
                    vdot(n, R->d_temp, R->d_temp) + vdot(n, R->dp, R->dp) +
                    vdot(n, R->ddr, R->ddr) - arc * arc;
            }
            errchk2 = orc_quad(n, a, b, c, R->d_temp, R->dp, R->ddr, R->ddq, R->dd, &dlpf,
                               &lpf_temp);
            if (getenv("ORC_VERBOSE"))
                fprintf(stderr, "msal arc=%.4e psi=%.3f a=%.3e b=%.3e c=%.3e quad=%d lpf=%.6e it=%d\n",
                        arc, psi, a, b, c, errchk2, lpf_temp, itecnt);
            if (errchk2 == 0) {
                frc = update_and_forces(R, R->x_temp, R->xfr_temp, R->f_temp, R->dd, &dlpf,
                                        itecnt);
                if (orc_test(n, R->d_temp, R->dd, R->f_temp, R->fp, R->qtot, R->f_ip, intener1,
                             &convchk, m->toldisp, m->tolforc, m->tolener))
                    FAIL(1, "zero norm in convergence test");
                if (R->out) R->out->n_newton++;
                if (convchk == 0) {
                    dnorm = 0;
                    for (long i = 0; i < n; ++i)
                        dnorm += (R->d_temp[i] - R->dp[i]) * (R->d_temp[i] - R->dp[i]);
                    dnorm = sqrt(dnorm);
                    if (getenv("ORC_VERBOSE"))
                        fprintf(stderr, "dnorm ratio %.12e\n", dnorm / (100 * dnormallow));
                    if (dnorm > 100 * dnormallow) {
                        arc /= beta;
                        beta = dnormallow / dnorm;
                        arc *= beta;
                        CP(R->f_temp, R->f, n);
                        errchk2 = 1;
                        R->ev[ORC_EV_DNORM]++;
                    }
                } else if (frc == 2) {
                    CP(R->f_temp, R->f, n);
                    errchk2 = 1;
                    R->ev[ORC_EV_UNLOAD]++;
                } else if (frc != 0 && subcnt <= m->submax) {
                    arc *= 0.5;
                    CP(R->f_temp, R->f, n);
                    subcnt++;
                    errchk2 = 1;
                    R->ev[ORC_EV_YIELD]++;
                } else {
                    itecnt++;
                    if (itecnt > m->itemax) {
                        arc *= 0.5;
                        CP(R->f_temp, R->f, n);
                        subcnt++;
                        errchk2 = 1;
                        R->ev[ORC_EV_HALVE]++;
                    }
                    shift_ip(R, 1);
                }
            } else if (errchk2 == 2) {
                arc *= 0.5;
                CP(R->f_temp, R->f, n);
                imagcnt++;
                R->ev[ORC_EV_IMAG]++;
            } else if (errchk2 == 3) {
                arc *= 0.5;
                CP(R->f_temp, R->f, n);
                negcnt++;
                R->ev[ORC_EV_NEG]++;
            }
        } while (convchk != 0 && errchk2 == 0 && subcnt <= m->submax &&
                 imagcnt <= m->imagmax && negcnt <= m->negmax);

        if (subcnt > m->submax || imagcnt > m->imagmax || negcnt > m->negmax)
            FAIL(1, "maximum allowable number of subdivisions exceeded");
        if (errchk2 == 0) {
            beta = sqrt(((double)m->iteopt) / ((double)itecnt)) * (dnormallow / dnorm);
            for (long i = 0; i < n; ++i) {
                R->dpp[i] = R->dp[i];
                R->d[i] = R->dp[i] = R->d_temp[i];
                R->f[i] = R->fp[i] = R->f_temp[i];
            }
            lpfpp = lpfp;
            lpf = lpfp = lpf_temp;
            commit(R);
            dkc = fabs(R->d[k]);
            lpfc = fabs(lpf);
            psi = 1;
            for (long i = 0; i < n; ++i) {
                double tq = fabs(R->ssd[i] / R->ssd_o[i]);
                if (tq < psi) psi = tq;
            }
            subcnt = imagcnt = negcnt = 0;
            record(R, lpf, itecnt, R->d, R->f);
            R->ev_psi = psi;
            if (R->out && R->out->nsteps >= R->out->max_steps) return 0;
        }
    }
    return 0;
}

/* ---- nonlinear Newmark (main.c:3305-3960) ------------------------------ */
/* ---- nonlinear Newmark (main.c:3305-3960, solve.c:139-245, 469-533) -------
 * Without prescribed support motion (NBC = 0; the input reader rejects it) and
 * with the skyline solver.  The reference discards the initial conditions
 * read from the input here (uc = vc = ac = 0, main.c:3443-3445), and so does
 * this restatement.  One path row per time step: (time, itecnt, d).
 *
 * Checkpoints (out->ckpt_path): after a committed time step k with
 * k % CHKPT == 0 and k != 0 the state goes to the file (main.c:3904-3907); a
 * model with RFLAG 1 starts from the file instead, at step tstep + 1
 * (main.c:3378-3398).  The reference leaves the frame lengths the elements
 * read (llength_temp, malloc'd at main.c:653) unset on that path, since only
 * a fresh start copies them (main.c:3420); here they come from the file's
 * llength as a fresh start would take them, and the exact file's extension
 * restores every working length and area and the time-increment multiplier. */
static int ckpt_view(run_t *R, orc_ckpt *c, double *sm_l, double *ss_l)
{
    memset(c, 0, sizeof *c);
    c->uc = R->uc; c->vc = R->vc; c->ac = R->ac;
    c->ss = ss_l; c->sm = sm_l;
    c->d = R->d; c->f = R->f; c->ef = R->ef; c->x = R->x;
    c->c1 = R->c1; c->c2 = R->c2; c->c3 = R->c3;
    c->defllen = R->defllen; c->llength = R->llen_ref;
    c->efFE = R->efFE; c->xfr = R->xfr; c->yldflag = R->yld;
    c->deffarea = R->deffarea; c->defslen = R->defslen;
    c->llen_w = R->llen_ref; c->farea_w = R->farea_w; c->slen_w = R->slen_w;
    return 0;
}

static int ckpt_write(run_t *R, long k, double ddt)
{
    orc_model *m = R->m;
    long lss = m->lss;
    double *sm_l = dz(lss);
    orc_ckpt c;
    ckpt_view(R, &c, sm_l, R->ss);
    CP(sm_l, R->sm, R->neq);
    c.tstep = k;
    c.ddt = ddt;
    int rc = orc_ckpt_write(R->out->ckpt_path, &m->dm, lss, &c, R->out->ckpt_exact);
    free(sm_l);
    return rc;
}

static int ckpt_restart(run_t *R, long *tstep, double *ddt)
{
    orc_model *m = R->m;
    long lss = m->lss;
    double *sm_l = dz(lss), *ss_l = dz(lss);
    orc_ckpt c;
    ckpt_view(R, &c, sm_l, ss_l);
    int rc = orc_ckpt_read(R->out->ckpt_path, &m->dm, lss, &c);
    if (rc == 0) {
        CP(R->sm, sm_l, R->neq);
        *tstep = c.tstep;
        *ddt = c.has_ext ? c.ddt : 1;
    }
    free(sm_l); free(ss_l);
    return rc;
}

static int run_newmark(run_t *R)
{
    orc_model *m = R->m;
    const orc_dims *d = &m->dm;
    long n = R->neq;
    double am_, af_;
    double rho = m->spectrds;
    switch (m->numopt) {
    case 0:
        if (rho != 1) FAIL(1, "invalid spectral radius for Newmark without dissipation");
        am_ = 0; af_ = 0; break;
    case 1: am_ = (2 * rho - 1) / (rho + 1); af_ = rho / (rho + 1); break;
    case 2: am_ = 0; af_ = (1 - rho) / (1 + rho); break;
    case 3: am_ = (rho - 1) / (rho + 1); af_ = 0; break;
    default: FAIL(1, "unknown Newmark dissipation option");
    }
    double alpha = pow(1 - am_ + af_, 2) / 4, delta = 0.5 - am_ + af_;
    double dt = m->dt;
    long ntst = (long)(m->ttot / dt + 1);
    long ns_in = m->ntstpsinpt;
    if (ntst > ns_in) FAIL(1, "time steps exceed the load history");
    const double *P = m->pinpt;                /* P[i*ns_in + k] */
    double ddt = 1, sub_dt = 1, dt_temp = dt;
    int tsflag = 0;
    double lpfi = m->lpf0, dlpfi = m->dlpf0;
    int convchk = 1, frcchk = 0, frc_fr = 0, itecnt = 0, det = 0;
    double intener1 = 0;
    /* the frame/truss lengths and the shell areas/side lengths the elements
       read are the run's working copies, overwritten by the mass routines */
    double *llen = R->llen_ref;
    long k = 0;
    const int ckpt = R->out && R->out->ckpt_path;
    if (m->rflag == 1) {
        if (!ckpt) FAIL(1, "restart (RFLAG 1) without a checkpoint file");
        long tstep = 0;
        if (ckpt_restart(R, &tstep, &ddt)) FAIL(1, "unable to read the checkpoint file");
        k = tstep + 1;
        if (k >= ntst) {                  /* the checkpoint is the last step (main.c:3387-3392) */
            if (k == ntst) record(R, (double)tstep * dt, 0, R->d, R->f);
            return 0;
        }
    }
    do {
        dt_temp = ddt * dt;
        sub_dt = ddt;
        tsflag = 0;
        do {
            double a0 = 1 / (alpha * pow(dt_temp, 2));
            double a2 = 1 / (alpha * dt_temp);
            double a3 = 1 / (2 * alpha) - 1;
            double a6 = dt_temp * (1 - delta);
            double a7 = delta * dt_temp;
            int solcnt = 0, subcnt = 0;
            double lpf = lpfi, dlpf = dlpfi, dlpfp;
            do {
                if (lpf > m->lpfmax) lpf = m->lpfmax;
                for (long i = 0; i < n; ++i) {
                    R->fp[i] = R->f[i];
                    R->d_temp[i] = R->d[i];
                    R->f_temp[i] = R->f[i];
                }
                dlpfp = dlpf;
                reset_from_base(R);
                CP(R->um, R->uc, n); CP(R->uc_i, R->uc, n);
                CP(R->vm, R->vc, n); CP(R->vc_i, R->vc, n);
                CP(R->am, R->ac, n); CP(R->ac_i, R->ac, n);
                itecnt = 0;
                frcchk = 0;
                do {
                    if (itecnt == 0) {
                        if (k == 0) {
                            for (long i = 0; i < n; ++i) {
                                R->qtot[i] = P[i * ns_in + k] * sub_dt * lpf;
                                R->r[i] = R->qtot[i] - R->f_temp[i];
                            }
                        } else {
                            for (long i = 0; i < n; ++i) {
                                double p0 = P[i * ns_in + k - 1], p1 = P[i * ns_in + k];
                                R->qtot[i] = (p0 + (p1 - p0) * sub_dt) * lpf;
                                R->r[i] = (R->qtot[i] - R->f_temp[i]) +
                                          af_ / (1 - af_) * (p0 - R->f_temp[i]);
                            }
                        }
                    } else {
                        for (long i = 0; i < n; ++i) {
                            double r = R->f_temp[i] - R->qtot[i];
                            R->r[i] = r + R->sm[i] * R->ac_i[i] -
                                      R->sm[i] * ((a2 * R->vc_i[i] + a3 * R->ac_i[i]) * (1 - am_) -
                                                  R->ac_i[i] * am_) / (1 - af_);
                        }
                    }
                    /* tangent and lumped mass, type by type (main.c:3592-3620) */
                    double t0 = now_s();
                    memset(R->ss, 0, sizeof(double) * (size_t)m->lss);
                    memset(R->sm, 0, sizeof(double) * (size_t)n);
                    if (d->NE_TR > 0) {
                        R->be->stiff_tr(R->be->ctx, d, R->ss, m->emod, m->carea, llen, R->defllen_ip,
                                        m->yield, R->c1_ip, R->c2_ip, R->c3_ip, R->ef_ip, m->maxa,
                                        m->mcode);
                        R->be->mass_tr(R->be->ctx, d, R->sm, m, R->x, llen, m->dens);
                    }
                    if (d->NE_FR > 0) {
                        R->be->stiff_fr(R->be->ctx, d, R->ss, m, llen, R->defllen_ip, R->c1_ip,
                                        R->c2_ip, R->c3_ip, R->ef_ip, R->efFE_ip, R->yld);
                        R->be->mass_fr(R->be->ctx, d, R->sm, m, R->x, llen, R->xfr, m->dens);
                    }
                    if (d->NE_SH > 0) {
                        R->be->stiff_sh(R->be->ctx, d, R->ss, m->emod, m->nu, R->x_temp, m->xlocal,
                                        m->thick, R->farea_w, R->deffarea_ip, R->slen_w,
                                        R->defslen_ip, m->yield, R->c1_ip, R->c2_ip, R->c3_ip,
                                        R->ef_ip, m->maxa, m->minc, m->mcode);
                        R->be->mass_sh(R->be->ctx, d, R->sm, m, R->x, R->farea_w, R->slen_w, m->dens);
                    }
                    if (R->out) R->out->t_tangent += now_s() - t0;
                    if (m->lss == 1) {
                        R->dd[0] = R->r[0] / R->ss[0];
                    } else {
                        /* solve.c:189-207, 469-533: Keff = K + a0(1-am)/(1-af) M on the
                           diagonal, Reff = r + M((1-am)(a2 v + a3 a) - am a)/(1-af) */
                        CP(R->keff, R->ss, m->lss);
                        for (long i = 0; i < n; ++i)
                            R->keff[m->maxa[i] - 1] += a0 * (1 - am_) * R->sm[i] / (1 - af_);
                        double t1 = now_s();
                        orc_skyfact(n, m->maxa, R->keff, R->ssd, 0, &det);
                        if (R->out) { R->out->t_factor += now_s() - t1; R->out->n_factor++; }
                        for (long i = 0; i < n; ++i)
                            R->dd[i] = R->r[i] + R->sm[i] * ((1 - am_) * (R->vc_i[i] * a2 +
                                                                        R->ac_i[i] * a3) -
                                                             am_ * R->ac_i[i]) / (1 - af_);
                        t1 = now_s();
                        orc_skysolve(n, m->maxa, R->keff, R->dd);
                        if (R->out) { R->out->t_solve += now_s() - t1; R->out->n_solve++; }
                    }
                    for (long i = 0; i < n; ++i) {
                        if (itecnt > 0) R->dd[i] = -R->dd[i];
                        R->d_temp[i] += R->dd[i];
                        R->f_ip[i] = R->f_temp[i];
                        R->f_temp[i] = 0;
                    }
                    for (long i = 0; i < n; ++i) {
                        R->uc_i[i] = R->d_temp[i];
                        R->ac_i[i] = (R->uc_i[i] - R->um[i]) * a0 - a2 * R->vm[i] - a3 * R->am[i];
                        R->vc_i[i] = R->vm[i] + a6 * R->am[i] + a7 * R->ac_i[i];
                    }
                    /* updatc + forces with the working lengths (main.c:3656-3680) */
                    t0 = now_s();
                    R->be->updatc(R->be->ctx, d, R->x_temp, R->x_ip, R->xfr_temp, R->dd,
                                  R->defllen_i, R->deffarea_i, R->defslen_i, m->offset, m->osflag,
                                  m->auxpt, R->c1_i, R->c2_i, R->c3_i, m->minc, m->jcode);
                    if (d->NE_TR > 0)
                        R->be->forces_tr(R->be->ctx, d, R->f_temp, R->ef_i, m->emod, m->carea, llen,
                                         R->defllen_i, m->yield, R->c1_i, R->c2_i, R->c3_i,
                                         m->mcode, R->d_temp);
                    int frc_sh = 0;
                    frc_fr = 0;
                    if (d->NE_FR > 0)
                        frc_fr = R->be->forces_fr(R->be->ctx, d, R->f_temp, R->ef_ip, R->ef_i,
                                                  R->efFE_ip, R->efFE_i, R->dd, m, llen,
                                                  R->defllen_ip, R->c1_ip, R->c2_ip, R->c3_ip,
                                                  R->c1_i, R->c2_i, R->c3_i, R->yld, &dlpf,
                                                  itecnt);
                    if (d->NE_SH > 0)
                        frc_sh = R->be->forces_sh(R->be->ctx, d, R->f_temp, R->ef_ip, R->ef_i,
                                                  R->dd, R->x_temp, R->x_ip, m->emod, m->nu,
                                                  m->xlocal, m->thick, R->farea_w,
                                                  R->deffarea_ip, R->slen_w, R->defslen_ip,
                                                  m->yield, R->c1_ip, R->c2_ip, R->c3_ip,
                                                  R->c1_i, R->c2_i, R->c3_i, m->minc, m->mcode,
                                                  m->jcode, R->d_temp);
                    if (R->out) R->out->t_forces += now_s() - t0;
                    frcchk = frc_fr || frc_sh;
                    for (long i = 0; i < n; ++i) R->dyn[i] = R->qtot[i] - R->sm[i] * R->ac_i[i];
                    if (itecnt == 0) {
                        intener1 = 0;
                        for (long i = 0; i < n; ++i) intener1 += R->dd[i] * (R->dyn[i] - R->fp[i]);
                    }
                    if (orc_test(n, R->d_temp, R->dd, R->f_temp, R->fp, R->dyn, R->f_ip, intener1,
                                 &convchk, m->toldisp, m->tolforc, m->tolener))
                        FAIL(1, "zero norm in convergence test");
                    if (getenv("ORC_VERBOSE"))
                        fprintf(stderr, "NM k=%ld it=%d convchk=%d |dd|=%g\n", k, itecnt, convchk,
                                sqrt(vdot(n, R->dd, R->dd)));
                    shift_ip(R, 1);
                    if (R->out) R->out->n_newton++;
                    itecnt++;
                } while (convchk != 0 && frcchk == 0 && itecnt <= m->itemax);
                if (frc_fr == 2) {
                    dlpf = dlpfp;                         /* main.c:3773-3774 */
                    R->ev[ORC_EV_UNLOAD]++;
                } else if ((convchk != 0 || frcchk) && subcnt <= m->submax) {
                    if (lpf == m->lpfmax) break;          /* main.c:3783-3788 */
                    if (dlpfp == m->dlpfmin) break;
                    if (frc_fr != 1) dlpf = dlpfp / 2;
                    if (dlpf < m->dlpfmin) dlpf = m->dlpfmin;
                    lpf = lpf - dlpfp + dlpf;
                    subcnt++;
                    R->ev[frc_fr == 1 ? ORC_EV_YIELD : ORC_EV_HALVE]++;
                    solcnt = 0;
                } else if (subcnt > m->submax) {
                    break;
                } else {
                    solcnt++;
                    subcnt = 0;
                    if (solcnt >= m->solmin) { dlpf *= 2; solcnt = 0; }
                    lpf += dlpf;
                }
            } while (lpf <= m->lpfmax);
            int committed = 0;
            if (convchk != 0) {
                /* halve the time increment (main.c:3826-3846) */
                ddt = ddt / 2;
                dt_temp = ddt * dt;
                R->ev[ORC_EV_HALVE]++;
                if (tsflag == 0) sub_dt = ddt;
                else if (tsflag == 1) sub_dt = sub_dt - ddt;
            } else {
                CP(R->d, R->d_temp, n);
                CP(R->f, R->f_temp, n);
                CP(R->uc, R->uc_i, n); CP(R->vc, R->vc_i, n); CP(R->ac, R->ac_i, n);
                commit(R);
                committed = 1;
            }
            if (!frcchk) {
                if (ddt < 1) {
                    if (sub_dt <= 1) { sub_dt = sub_dt + ddt; tsflag = 1; }
                    if (sub_dt > 1) { ddt = ddt * 2; tsflag = 2; }
                } else if (ddt == 1) {
                    tsflag = 2;
                }
            }
            /* written after the multiplier update, which changes no array, so the
               file holds what checkPoint writes at the commit (main.c:3904-3907)
               and the multiplier the next step starts from */
            if (committed && ckpt && k % m->chkpt == 0 && k != 0 && ckpt_write(R, k, ddt))
                FAIL(1, "unable to write the checkpoint file");
        } while (ddt >= 0.0001 && tsflag != 2);
        if (convchk != 0 || frcchk)
            FAIL(3, "minimum time increment interval reached without convergence");
        record(R, (double)k * dt, itecnt, R->d, R->f);
        if (R->out && R->out->nsteps >= R->out->max_steps) break;
        ++k;
    } while (k < ntst);
    return 0;
}

/* ---- linear static analysis: ANAFLAG 1 under ALGFLAG 1/2 (main.c:1714-1798) --
 * One tangent at the input geometry, K d = lpfmax q, and the linear element
 * forces of d; one path row (lpfmax, 0, d).  The reference prints its
 * iteration counter there without setting it on this path (main.c:1670). */
static int run_linear_static(run_t *R)
{
    orc_model *m = R->m;
    const orc_dims *d = &m->dm;
    long n = R->neq;
    for (long i = 0; i < n; ++i) R->qtot[i] = m->q[i] * m->lpfmax;
    assemble(R, R->x, R->c1, R->c2, R->c3, R->ef, R->defllen, R->deffarea, R->defslen, R->efFE);
    if (m->lss == 1) {
        R->d[0] = R->qtot[0] / R->ss[0];
    } else {
        int det = 0;
        (void)factor(R, &det);             /* solve() ignores skyfact's status (solve.c:78-89) */
        backsolve(R, R->qtot, R->d);
    }
    memset(R->f, 0, sizeof(double) * (size_t)n);
    if (d->NE_TR > 0)
        R->be->forces_tr(R->be->ctx, d, R->f, R->ef, m->emod, m->carea, R->llen_ref, R->defllen,
                         m->yield, R->c1, R->c2, R->c3, m->mcode, R->d);
    if (d->NE_FR > 0) {
        double dl = 0;
        R->be->forces_fr(R->be->ctx, d, R->f, R->ef, R->ef, R->efFE, R->efFE, R->d, m, R->llen_ref,
                         R->defllen, R->c1, R->c2, R->c3, R->c1, R->c2, R->c3, R->yld, &dl, 0);
    }
    if (d->NE_SH > 0)
        R->be->forces_sh(R->be->ctx, d, R->f, R->ef, R->ef, R->d, R->x, R->x, m->emod, m->nu,
                         m->xlocal, m->thick, R->farea_w, R->deffarea, R->slen_w, R->defslen,
                         m->yield, R->c1, R->c2, R->c3, R->c1, R->c2, R->c3, m->minc, m->mcode,
                         m->jcode, R->d);
    record(R, m->lpfmax, 0, R->d, R->f);
    return 0;
}

/* ---- linear Newmark: ALGFLAG 4 (main.c:3143-3303, solve.c:139-245, 246-468) --
 * K and the lumped M at the input geometry, Keff = K + a0 (1-am)/(1-af) M on
 * the diagonal factored once, then per time step
 *   Reff = P_k [+ af/(1-af) P_{k-1}] + M ((1-am)(a0 u + a2 v + a3 a) - am a)/(1-af)
 *          [- af/(1-af) K u  when af != 0]
 *   Keff u_new = Reff,  a_new = a0 (u_new - u) - a2 v - a3 a,  v_new = v + a6 a + a7 a_new
 * starting from the input's initial conditions.  Each step's row is
 * (k dt, 0, u_new); after the loop the reference computes the forces of d,
 * which this driver never sets, and writes (lpfmax, counter, 0) as the last
 * row.  The checkpoint (CHKPT) holds u, v, a; RFLAG 1 starts from it. */
static int run_linear_newmark(run_t *R)
{
    orc_model *m = R->m;
    const orc_dims *d = &m->dm;
    long n = R->neq;
    double am_ = 0, af_ = 0, rho = m->spectrds;
    switch (m->numopt) {
    case 0:
        if (rho != 1) FAIL(1, "invalid spectral radius for Newmark without dissipation");
        break;
    case 1: am_ = (2 * rho - 1) / (rho + 1); af_ = rho / (rho + 1); break;
    case 2: af_ = (1 - rho) / (1 + rho); break;
    case 3: am_ = (rho - 1) / (rho + 1); break;
    default: FAIL(1, "unknown Newmark dissipation option");
    }
    double alpha = pow(1 - am_ + af_, 2) / 4, delta = 0.5 - am_ + af_;
    double dt = m->dt;
    long ntst = (long)(m->ttot / dt + 1);
    long ns_in = m->ntstpsinpt;
    if (ntst > ns_in) FAIL(1, "time steps exceed the load history");
    const double *P = m->pinpt;
    assemble(R, R->x, R->c1, R->c2, R->c3, R->ef, R->defllen, R->deffarea, R->defslen, R->efFE);
    memset(R->sm, 0, sizeof(double) * (size_t)n);
    if (d->NE_TR > 0) R->be->mass_tr(R->be->ctx, d, R->sm, m, R->x, R->llen_ref, m->dens);
    if (d->NE_FR > 0) R->be->mass_fr(R->be->ctx, d, R->sm, m, R->x, R->llen_ref, R->xfr, m->dens);
    if (d->NE_SH > 0) R->be->mass_sh(R->be->ctx, d, R->sm, m, R->x, R->farea_w, R->slen_w, m->dens);
    double a0 = 1 / (alpha * pow(dt, 2)), a2 = 1 / (alpha * dt), a3 = 1 / (2 * alpha) - 1;
    double a6 = dt * (1 - delta), a7 = delta * dt;
    CP(R->keff, R->ss, m->lss);
    for (long i = 0; i < n; ++i) R->keff[m->maxa[i] - 1] += a0 * (1 - am_) * R->sm[i] / (1 - af_);
    int det = 0;
    double t1 = now_s();
    orc_skyfact(n, m->maxa, R->keff, R->ssd, 0, &det);
    if (R->out) { R->out->t_factor += now_s() - t1; R->out->n_factor++; }
    CP(R->um, m->um, n); CP(R->vm, m->vm, n); CP(R->am, m->am, n);
    const int ckpt = R->out && R->out->ckpt_path;
    long k = 0;
    if (m->rflag == 1) {
        if (!ckpt) FAIL(1, "restart (RFLAG 1) without a checkpoint file");
        long tstp = 0;
        if (orc_ckpt_read_lin(R->out->ckpt_path, n, &tstp, R->um, R->vm, R->am))
            FAIL(1, "unable to read the checkpoint file");
        if (tstp + 1 == ntst) record(R, (double)tstp * dt, 0, R->um, NULL);
        k = tstp + 1;
    }
    double *Kum = R->dyn;
    for (; k < ntst; ++k) {
        for (long i = 0; i < n; ++i) {
            double meff = R->sm[i] * ((1 - am_) * (R->um[i] * a0 + R->vm[i] * a2 + R->am[i] * a3) -
                                      am_ * R->am[i]) / (1 - af_);
            R->r[i] = (k == 0) ? P[i * ns_in + k] + meff
                               : P[i * ns_in + k] + af_ / (1 - af_) * P[i * ns_in + k - 1] + meff;
        }
        if (af_ != 0) {
            orc_skymult(n, m->maxa, R->ss, R->um, Kum);
            for (long i = 0; i < n; ++i) R->r[i] -= af_ / (1 - af_) * Kum[i];
        }
        t1 = now_s();
        orc_skysolve(n, m->maxa, R->keff, R->r);
        if (R->out) { R->out->t_solve += now_s() - t1; R->out->n_solve++; }
        for (long i = 0; i < n; ++i) {
            R->uc[i] = R->r[i];
            R->ac[i] = a0 * (R->uc[i] - R->um[i]) - a2 * R->vm[i] - a3 * R->am[i];
            R->vc[i] = R->vm[i] + a6 * R->am[i] + a7 * R->ac[i];
        }
        record(R, (double)k * dt, 0, R->uc, NULL);
        CP(R->um, R->uc, n); CP(R->vm, R->vc, n); CP(R->am, R->ac, n);
        if (ckpt && k % m->chkpt == 0 && k != 0 &&
            orc_ckpt_write_lin(R->out->ckpt_path, n, k, R->um, R->vm, R->am, R->out->ckpt_exact))
            FAIL(1, "unable to write the checkpoint file");
        if (R->out && R->out->nsteps >= R->out->max_steps) return 0;
    }
    record(R, m->lpfmax, 0, R->d, R->f);
    return 0;
}

int orc_run(orc_model *m, const orc_backend *b, orc_path *out)
{
    orc_backend own;
    if (!b) { orc_default_backend(&own); b = &own; }
    run_t R;
    run_init(&R, m, b, out);
    if (out) { out->nsteps = 0; out->status = 0; out->msg[0] = 0; }
    int rc;
    switch (m->dm.ALGFLAG) {
    case 1: case 2: rc = m->dm.ANAFLAG == 1 ? run_linear_static(&R) : run_nr(&R); break;
    case 3: rc = run_msal(&R); break;
    case 4: rc = run_linear_newmark(&R); break;
    case 5: rc = run_newmark(&R); break;
    default: rc = 2; if (out) { out->status = 2; snprintf(out->msg, sizeof out->msg, "unsupported ALGFLAG"); }
    }
    run_free(&R);
    return rc;
}

/* Bounded timing sample of the serial CPU path (see cb_oracle.h). */
int orc_time_sample(orc_model *m, double factor_budget_s, orc_sample *out)
{
    const orc_dims *d = &m->dm;
    long neq = d->NEQ, nj = d->NJ, nsh = d->NE_SH, ntri = d->NE_TR + d->NE_FR * 3 + d->NE_SH * 3;
    memset(out, 0, sizeof *out);
    double *ss = calloc((size_t)m->lss, sizeof(double));
    double *f = calloc((size_t)neq, sizeof(double));
    double *dd = calloc((size_t)neq, sizeof(double));
    double *x = malloc(sizeof(double) * (size_t)nj * 3), *xip = malloc(sizeof(double) * (size_t)nj * 3);
    double *efp = calloc((size_t)nsh * 18 + 1, sizeof(double)), *efi = calloc((size_t)nsh * 18 + 1, sizeof(double));
    double *c1 = malloc(sizeof(double) * (size_t)ntri), *c2 = malloc(sizeof(double) * (size_t)ntri);
    double *c3 = malloc(sizeof(double) * (size_t)ntri);
    double *da = calloc((size_t)nsh + 1, sizeof(double)), *ds = calloc((size_t)nsh * 3 + 1, sizeof(double));
    double *dl = calloc((size_t)(d->NE_TR + d->NE_FR) + 1, sizeof(double));
    double *xfr = calloc((size_t)d->NE_FR * 6 + 1, sizeof(double));
    if (!ss || !f || !dd || !x || !xip || !efp || !efi || !c1 || !c2 || !c3 || !da || !ds || !dl || !xfr) return 1;
    memcpy(x, m->x, sizeof(double) * (size_t)nj * 3);
    double t0 = now_s();
    if (d->NE_TR > 0)
        orc_stiff_tr(d, ss, m->emod, m->carea, m->llength, m->llength, m->yield, m->c1, m->c2, m->c3,
                     efp, m->maxa, m->mcode);
    if (nsh > 0)
        orc_stiff_sh(d, ss, m->emod, m->nu, m->x, m->xlocal, m->thick, m->farea, m->farea, m->slength,
                     m->c1, m->c2, m->c3, m->maxa, m->minc, m->mcode);
    out->t_tangent = now_s() - t0;
    t0 = now_s();
    orc_updatc(d, x, xip, xfr, dd, dl, da, ds, m->offset, m->osflag, m->auxpt, c1, c2, c3, m->minc, m->jcode);
    if (d->NE_TR > 0)
        orc_forces_tr(d, f, efi, m->emod, m->carea, m->llength, dl, m->yield, c1, c2, c3, m->mcode);
    if (nsh > 0)
        orc_forces_sh(d, f, efp, efi, dd, x, m->emod, m->nu, m->xlocal, m->thick, m->farea, m->slength,
                      m->c1, m->c2, m->c3, c1, c2, c3, m->minc, m->mcode);
    out->t_forces = now_s() - t0;
    /* a back-substitution sweeps the same skyline array whether or not it is
       factorized; time it on the assembled array */
    memcpy(f, m->q, sizeof(double) * (size_t)neq);
    t0 = now_s();
    orc_skysolve(neq, m->maxa, ss, f);
    out->t_solve = now_s() - t0;
    t0 = now_s();
    orc_skyfact_partial(neq, m->maxa, ss, factor_budget_s, &out->cols_done);
    out->t_factor_part = now_s() - t0;
    double wall = 0, wdone = 0;
    for (long i = 0; i < neq; ++i) {
        double h = (double)m->kht[i];
        wall += h * h + 1;
        if (i < out->cols_done) wdone += h * h + 1;
    }
    out->factor_fraction = wall > 0 ? wdone / wall : 1.0;
    out->t_factor_est = out->factor_fraction > 0 ? out->t_factor_part / out->factor_fraction : 0;
    free(ss); free(f); free(dd); free(x); free(xip); free(efp); free(efi); free(c1); free(c2); free(c3);
    free(da); free(ds); free(dl); free(xfr);
    return 0;
}
