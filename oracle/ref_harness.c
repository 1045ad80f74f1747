/*
 * ref_harness.c -- test harness around the reference's own element, model and
 * utility routines, compiled from their sources in place (see oracle/Makefile).
 *
 * TEST INFRASTRUCTURE ONLY: built into oracle/_ref/libcubens_ref.so, loaded
 * only by tests and the fixture generator under tests/golden/.
 *
 * solve.c is not compiled (it includes umfpack.h, which this image lacks) and
 * main.c is not compiled (it is the program driver; its skyline calls go through
 * solve.c).  This file therefore defines the globals main.c would define
 * (main.c:323-328) and exposes:
 *   - ref_open(): the reference's input sequence (main.c:340-427 header reads,
 *     then struc, codes, skylin, coordinates, prop_tr/fr/sh and load, in the
 *     order of main.c:1063-1653) with array getters;
 *   - thin wrappers with the reference signatures, and an orc_backend table so
 *     the oracle's driver can run on the reference's element routines.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "prototypes.h"
#include "cb_oracle.h"

/* globals of main.c:323-328 */
long NJ, NE_TR, NE_FR, NE_SH, NE_SBR, NE_FBR, NE_BR, NEQ, NBC, SNDOF, FNDOF, NTSTPS, ntstpsinpt;
double dt, ttot;
int ANAFLAG = 666, ALGFLAG, OPTFLAG, SLVFLAG, FSIFLAG, FSIINCFLAG, brFSI_FLAG, shFSI_FLAG;
FILE *IFP[4], *OFP[8];
int CHKPT, RFLAG;

static void open_sinks(void)
{
    for (int i = 0; i < 8; ++i)
        if (!OFP[i]) OFP[i] = fopen("/dev/null", "w");
}

void ref_set_dims(const orc_dims *d)
{
    open_sinks();
    NJ = d->NJ; NE_TR = d->NE_TR; NE_FR = d->NE_FR; NE_SH = d->NE_SH; NEQ = d->NEQ;
    NE_SBR = NE_FBR = NE_BR = 0;
    ANAFLAG = d->ANAFLAG; ALGFLAG = d->ALGFLAG; SLVFLAG = d->SLVFLAG; OPTFLAG = d->OPTFLAG;
}

/* scratch for the plasticity arguments the elastic branches never read */
static double *zeros(long n)
{
    static double *buf = NULL;
    static long cap = 0;
    if (n > cap) { free(buf); buf = calloc((size_t)n, sizeof(double)); cap = n; }
    return buf;
}

#define U(p) ((double *)(p))
#define UL(p) ((long *)(p))

static void w_stiff_tr(void *ctx, const orc_dims *d, double *ss, const double *emod,
                       const double *carea, const double *llength, const double *defllen_ip,
                       const double *yield, const double *c1_ip, const double *c2_ip,
                       const double *c3_ip, const double *ef_ip, const long *maxa,
                       const long *mcode)
{
    (void)ctx;
    ref_set_dims(d);
    stiff_tr(ss, U(emod), U(carea), U(llength), U(defllen_ip), U(yield), U(c1_ip), U(c2_ip),
             U(c3_ip), U(ef_ip), UL(maxa), UL(mcode));
}

static void w_forces_tr(void *ctx, const orc_dims *d, double *f, double *ef_i,
                        const double *emod, const double *carea, const double *llength,
                        const double *defllen_i, const double *yield, const double *c1_i,
                        const double *c2_i, const double *c3_i, const long *mcode,
                        const double *d_tot)
{
    (void)ctx;
    ref_set_dims(d);
    forces_tr(f, ef_i, d_tot ? U(d_tot) : zeros(d->NEQ + 1), U(emod), U(carea), U(llength),
              U(defllen_i), U(yield), U(c1_i), U(c2_i), U(c3_i), UL(mcode));
}

static void w_stiff_sh(void *ctx, const orc_dims *d, double *ss, const double *emod,
                       const double *nu, const double *x_temp, const double *xlocal,
                       const double *thick, const double *farea, const double *deffarea_ip,
                       const double *slength, const double *defslen_ip, const double *yield,
                       const double *c1_ip, const double *c2_ip, const double *c3_ip,
                       const double *ef_ip, const long *maxa, const long *minc,
                       const long *mcode)
{
    (void)ctx;
    ref_set_dims(d);
    double *z = zeros(d->NE_SH * 9 + d->NEQ + 1);
    stiff_sh(ss, U(emod), U(nu), U(x_temp), U(xlocal), U(thick), U(farea), U(deffarea_ip),
             U(slength), U(defslen_ip), U(yield), U(c1_ip), U(c2_ip), U(c3_ip), U(ef_ip), z, z, z,
             z, UL(maxa), UL(minc), UL(mcode));
}

static int w_forces_sh(void *ctx, const orc_dims *d, double *f, double *ef_ip, double *ef_i,
                       const double *dd, const double *x_temp, const double *x_ip,
                       const double *emod, const double *nu, const double *xlocal,
                       const double *thick, const double *farea, const double *deffarea_ip,
                       const double *slength, const double *defslen_ip, const double *yield,
                       const double *c1_ip, const double *c2_ip, const double *c3_ip,
                       const double *c1_i, const double *c2_i, const double *c3_i,
                       const long *minc, const long *mcode, const long *jcode,
                       const double *d_tot)
{
    (void)ctx;
    ref_set_dims(d);
    double *z = zeros(d->NE_SH * 9 + d->NEQ + 1);
    return forces_sh(f, ef_ip, ef_i, z, z, U(dd), d_tot ? U(d_tot) : z, z, U(x_temp), U(x_ip),
                     U(emod), U(nu),
                     U(xlocal), U(thick), U(farea), U(deffarea_ip), U(slength), U(defslen_ip),
                     U(yield), U(c1_ip), U(c2_ip), U(c3_ip), U(c1_i), U(c2_i), U(c3_i), UL(minc),
                     UL(mcode), UL(jcode));
}

static void w_updatc(void *ctx, const orc_dims *d, double *x_temp, double *x_ip,
                     double *xfr_temp, const double *dd, double *defllen_i, double *deffarea_i,
                     double *defslen_i, const double *offset, const int *osflag,
                     const double *auxpt, double *c1_i, double *c2_i, double *c3_i,
                     const long *minc, const long *jcode)
{
    (void)ctx;
    ref_set_dims(d);
    updatc(x_temp, x_ip, xfr_temp, U(dd), defllen_i, deffarea_i, defslen_i, U(offset),
           (int *)osflag, U(auxpt), c1_i, c2_i, c3_i, UL(minc), UL(jcode));
}

static int *zeros_int(long n)
{
    static int *buf = NULL;
    static long cap = 0;
    if (n > cap) { free(buf); buf = calloc((size_t)n, sizeof(int)); cap = n; }
    return buf;
}

static void w_stiff_fr(void *ctx, const orc_dims *d, double *ss, const orc_model *m,
                       const double *llength, const double *defllen_ip, const double *c1_ip,
                       const double *c2_ip, const double *c3_ip, const double *ef_ip,
                       const double *efFE_ip, int *yldflag)
{
    (void)ctx;
    ref_set_dims(d);
    stiff_fr(ss, m->emod, m->gmod, m->carea, m->offset, m->osflag, U(llength), U(defllen_ip),
             m->istrong, m->iweak, m->ipolar, m->iwarp,
             yldflag ? yldflag : zeros_int(d->NE_FR * 2 + 2), m->yield,
             m->zstrong, m->zweak, U(c1_ip), U(c2_ip), U(c3_ip), U(ef_ip), U(efFE_ip),
             m->mendrel, m->maxa, m->mcode);
}

static int w_forces_fr(void *ctx, const orc_dims *d, double *f, double *ef_ip, double *ef_i,
                       const double *efFE_ip, double *efFE_i, const double *dd,
                       const orc_model *m, const double *llength, const double *defllen_ip,
                       const double *c1_ip, const double *c2_ip, const double *c3_ip,
                       const double *c1_i, const double *c2_i, const double *c3_i,
                       int *yldflag, double *dlpf, int itecnt)
{
    (void)ctx;
    ref_set_dims(d);
    int it = itecnt;
    return forces_fr(f, ef_ip, ef_i, zeros(d->NE_FR * 14 + 1), U(efFE_ip), efFE_i,
                     yldflag ? yldflag : zeros_int(d->NE_FR * 2 + 2), U(dd), m->emod, m->gmod, m->carea, m->offset,
                     m->osflag, U(llength), U(defllen_ip), m->istrong, m->iweak, m->ipolar,
                     m->iwarp, m->yield, m->zstrong, m->zweak, U(c1_ip), U(c2_ip), U(c3_ip),
                     U(c1_i), U(c2_i), U(c3_i), m->mendrel, m->mcode, dlpf, &it);
}

static void w_mass_tr(void *ctx, const orc_dims *d, double *sm, const orc_model *m,
                      const double *x, double *llength, const double *dens)
{
    (void)ctx;
    ref_set_dims(d);
    mass_tr(sm, m->carea, llength, U(dens), U(x), m->minc, m->mcode, zeros(1));
}

static void w_mass_fr(void *ctx, const orc_dims *d, double *sm, const orc_model *m,
                      const double *x, double *llength, double *xfr, const double *dens)
{
    (void)ctx;
    ref_set_dims(d);
    mass_fr(sm, m->carea, llength, m->istrong, m->iweak, m->ipolar, m->iwarp, U(dens),
            m->osflag, m->offset, U(x), xfr, m->minc, m->mcode, zeros(1));
}

static void w_mass_sh(void *ctx, const orc_dims *d, double *sm, const orc_model *m,
                      const double *x, double *farea, double *slength, const double *dens)
{
    (void)ctx;
    ref_set_dims(d);
    mass_sh(sm, m->carea, U(dens), m->thick, farea, slength, U(x), m->minc, m->mcode, zeros(1));
}

void ref_fill_backend(orc_backend *b)
{
    b->ctx = NULL;
    b->stiff_tr = w_stiff_tr;
    b->forces_tr = w_forces_tr;
    b->stiff_sh = w_stiff_sh;
    b->forces_sh = w_forces_sh;
    b->updatc = w_updatc;
    b->stiff_fr = w_stiff_fr;
    b->forces_fr = w_forces_fr;
    b->mass_tr = w_mass_tr;
    b->mass_fr = w_mass_fr;
    b->mass_sh = w_mass_sh;
}

/* Flat entry points with the reference signatures, sizes passed explicitly. */
void ref_stiff_sh(const orc_dims *d, double *ss, const double *emod, const double *nu,
                  const double *x_temp, const double *xlocal, const double *thick,
                  const double *farea, const double *deffarea_ip, const double *slength,
                  const double *c1_ip, const double *c2_ip, const double *c3_ip,
                  const long *maxa, const long *minc, const long *mcode)
{
    double *y = calloc((size_t)(d->NE_TR + d->NE_FR + d->NE_SH + 1), sizeof(double));
    double *s3 = calloc((size_t)(d->NE_SH * 3 + 1), sizeof(double));
    double *ef = calloc((size_t)(d->NE_TR * 2 + d->NE_FR * 14 + d->NE_SH * 18 + 1), sizeof(double));
    w_stiff_sh(NULL, d, ss, emod, nu, x_temp, xlocal, thick, farea, deffarea_ip, slength, s3, y,
               c1_ip, c2_ip, c3_ip, ef, maxa, minc, mcode);
    free(y); free(s3); free(ef);
}

int ref_forces_sh(const orc_dims *d, double *f, double *ef_ip, double *ef_i, const double *dd,
                  const double *x_temp, const double *emod, const double *nu,
                  const double *xlocal, const double *thick, const double *farea,
                  const double *slength, const double *c1_ip, const double *c2_ip,
                  const double *c3_ip, const double *c1_i, const double *c2_i,
                  const double *c3_i, const long *minc, const long *mcode, const long *jcode)
{
    double *y = calloc((size_t)(d->NE_TR + d->NE_FR + d->NE_SH + 1), sizeof(double));
    double *s3 = calloc((size_t)(d->NE_SH * 3 + 1), sizeof(double));
    double *a = calloc((size_t)(d->NE_SH + 1), sizeof(double));
    double *xip = calloc((size_t)(d->NJ * 3 + 1), sizeof(double));
    int r = w_forces_sh(NULL, d, f, ef_ip, ef_i, dd, x_temp, xip, emod, nu, xlocal, thick, farea,
                        a, slength, s3, y, c1_ip, c2_ip, c3_ip, c1_i, c2_i, c3_i, minc, mcode,
                        jcode, NULL);
    free(y); free(s3); free(a); free(xip);
    return r;
}

void ref_stiff_tr(const orc_dims *d, double *ss, const double *emod, const double *carea,
                  const double *llength, const double *defllen_ip, const double *c1_ip,
                  const double *c2_ip, const double *c3_ip, const double *ef_ip,
                  const long *maxa, const long *mcode)
{
    double *y = calloc((size_t)(d->NE_TR + d->NE_FR + d->NE_SH + 1), sizeof(double));
    w_stiff_tr(NULL, d, ss, emod, carea, llength, defllen_ip, y, c1_ip, c2_ip, c3_ip, ef_ip,
               maxa, mcode);
    free(y);
}

void ref_forces_tr(const orc_dims *d, double *f, double *ef_i, const double *emod,
                   const double *carea, const double *llength, const double *defllen_i,
                   const double *c1_i, const double *c2_i, const double *c3_i,
                   const long *mcode)
{
    double *y = calloc((size_t)(d->NE_TR + d->NE_FR + d->NE_SH + 1), sizeof(double));
    w_forces_tr(NULL, d, f, ef_i, emod, carea, llength, defllen_i, y, c1_i, c2_i, c3_i, mcode, NULL);
    free(y);
}

void ref_updatc(const orc_dims *d, double *x_temp, double *x_ip, double *xfr_temp,
                const double *dd, double *defllen_i, double *deffarea_i, double *defslen_i,
                const double *offset, const int *osflag, const double *auxpt, double *c1_i,
                double *c2_i, double *c3_i, const long *minc, const long *jcode)
{
    w_updatc(NULL, d, x_temp, x_ip, xfr_temp, dd, defllen_i, deffarea_i, defslen_i, offset,
             osflag, auxpt, c1_i, c2_i, c3_i, minc, jcode);
}

int ref_test(const orc_dims *d, double *d_temp, double *dd, double *f_temp, double *fp,
             double *qtot, double *f_ip, double intener1, int *convchk, double toldisp,
             double tolforc, double tolener)
{
    ref_set_dims(d);
    return test(d_temp, dd, f_temp, fp, qtot, f_ip, &intener1, convchk, &toldisp, &tolforc,
                &tolener);
}

int ref_quad(const orc_dims *d, double a, double b, double c, double *dt_, double *dp,
             double *ddr, double *ddq, double *dd, double *dlpf, double *lpft)
{
    ref_set_dims(d);
    return quad(&a, &b, &c, dt_, dp, ddr, ddq, dd, dlpf, lpft);
}

/* ---- the reference's own input sequence ------------------------------------ */
static struct {
    int open;
    long lss;
    long *jcode, *minc, *mcode, *jnt, *maxa, *kht, *pmot;
    int *wrpres, *osflag, *mendrel;
    double *x, *xfr, *emod, *gmod, *dens, *offset, *auxpt, *carea, *llength, *istrong, *iweak,
        *ipolar, *iwarp, *yield, *zstrong, *zweak, *c1, *c2, *c3, *nu, *xlocal, *thick, *farea,
        *slength, *q, *efFE_ref, *tinpt, *pinpt, *dinpt, *pdisp, *um, *vm, *am;
} RM;

static void *za(long n, size_t sz) { return calloc((size_t)(n > 0 ? n : 1) + 8, sz); }

/* Reads model_def.txt with the reference's routines.  Returns 0 on success.
 * dims_out receives the globals the reference computed. */
int ref_open(const char *path, orc_dims *dims_out, long *lss_out)
{
    open_sinks();
    memset(&RM, 0, sizeof RM);
    IFP[0] = fopen(path, "r");
    if (!IFP[0]) return 1;
    /* header (main.c:340-427) */
    if (fscanf(IFP[0], "%d,", &ANAFLAG) != 1) return 1;
    if (ANAFLAG == 4) return 1;
    if (fscanf(IFP[0], "%d\n", &ALGFLAG) != 1) return 1;
    if (ALGFLAG > 3 && fscanf(IFP[0], "%d,%d\n", &CHKPT, &RFLAG) != 2) return 1;
    if (fscanf(IFP[0], "%d\n", &SLVFLAG) != 1) return 1;
    if (fscanf(IFP[0], "%d\n", &OPTFLAG) != 1) return 1;
    if (fscanf(IFP[0], "%ld\n", &NJ) != 1) return 1;
    if (fscanf(IFP[0], "%ld,%ld,%ld,%ld,%ld\n", &NE_TR, &NE_FR, &NE_SH, &NE_SBR, &NE_FBR) != 5)
        return 1;
    NE_BR = NE_SBR + NE_FBR;
    if (NE_BR) return 1;
    /* with OPTFLAG 2 the reference also writes the renumbered input to
     * model_def_OPT.txt through IFP[1] (main.c:446-455) */
    if (OPTFLAG == 2 && !IFP[1]) IFP[1] = fopen("/dev/null", "w");
    long ntri = NE_TR + NE_FR * 3 + NE_SH * 3, nprop = NE_TR + NE_FR + NE_SH;
    RM.jcode = za(NJ * 7, sizeof(long));
    RM.minc = za(NE_TR * 2 + NE_FR * 2 + NE_SH * 3, sizeof(long));
    RM.wrpres = za(NJ * 3, sizeof(int));
    RM.jnt = za(NJ, sizeof(long));
    if (struc(RM.jcode, RM.minc, RM.wrpres, RM.jnt)) return 1;
    RM.mcode = za(NE_TR * 6 + NE_FR * 14 + NE_SH * 18, sizeof(long));
    codes(RM.mcode, RM.jcode, RM.minc, RM.wrpres);
    RM.maxa = za(NEQ + 1, sizeof(long));
    RM.kht = za(NEQ, sizeof(long));       /* zeroed: skylin only raises entries */
    RM.pmot = za(NEQ, sizeof(long));
    if (skylin(RM.maxa, RM.mcode, &RM.lss, RM.jcode, RM.kht, RM.pmot)) return 1;
    RM.x = za(NJ * 3, sizeof(double));
    for (long i = 0; i < NJ; ++i)
        if (fscanf(IFP[0], "%lf,%lf,%lf\n", &RM.x[RM.jnt[i] * 3], &RM.x[RM.jnt[i] * 3 + 1],
                   &RM.x[RM.jnt[i] * 3 + 2]) != 3)
            return 1;
    RM.emod = za(nprop, sizeof(double));
    RM.yield = za(nprop, sizeof(double));
    RM.dens = za(nprop, sizeof(double));
    RM.carea = za(NE_TR + NE_FR, sizeof(double));
    RM.llength = za(NE_TR + NE_FR, sizeof(double));
    RM.c1 = za(ntri, sizeof(double)); RM.c2 = za(ntri, sizeof(double)); RM.c3 = za(ntri, sizeof(double));
    RM.xfr = za(NE_FR * 6, sizeof(double));
    RM.gmod = za(NE_FR, sizeof(double));
    RM.offset = za(NE_FR * 6, sizeof(double));
    RM.osflag = za(NE_FR, sizeof(int));
    RM.auxpt = za(NE_FR * 3, sizeof(double));
    RM.istrong = za(NE_FR, sizeof(double)); RM.iweak = za(NE_FR, sizeof(double));
    RM.ipolar = za(NE_FR, sizeof(double)); RM.iwarp = za(NE_FR, sizeof(double));
    RM.zstrong = za(NE_FR, sizeof(double)); RM.zweak = za(NE_FR, sizeof(double));
    RM.mendrel = za(NE_FR * 5, sizeof(int));
    RM.nu = za(NE_SH, sizeof(double)); RM.thick = za(NE_SH, sizeof(double));
    RM.xlocal = za(NE_SH * 3, sizeof(double)); RM.farea = za(NE_SH, sizeof(double));
    RM.slength = za(NE_SH * 3, sizeof(double));
    if (NE_TR > 0) prop_tr(RM.x, RM.emod, RM.carea, RM.dens, RM.llength, RM.yield, RM.c1, RM.c2, RM.c3, RM.minc);
    if (NE_FR > 0)
        prop_fr(RM.x, RM.xfr, RM.emod, RM.gmod, RM.dens, RM.offset, RM.osflag, RM.auxpt, RM.carea,
                RM.llength, RM.istrong, RM.iweak, RM.ipolar, RM.iwarp, RM.yield, RM.zstrong,
                RM.zweak, RM.c1, RM.c2, RM.c3, RM.mendrel, RM.minc);
    if (NE_SH > 0)
        prop_sh(RM.x, RM.emod, RM.nu, RM.xlocal, RM.thick, RM.dens, RM.farea, RM.slength, RM.yield,
                RM.c1, RM.c2, RM.c3, RM.minc);
    if (ALGFLAG > 3) {
        if (fscanf(IFP[0], "%ld,%lf\n", &ntstpsinpt, &ttot) != 2) return 1;
        dt = ttot / ntstpsinpt;
        ntstpsinpt += 1;
    } else {
        ntstpsinpt = 0;
    }
    RM.q = za(NEQ, sizeof(double));
    RM.efFE_ref = za(NE_FR * 14, sizeof(double));
    RM.tinpt = za(ntstpsinpt, sizeof(double));
    RM.pinpt = za(NEQ * ntstpsinpt, sizeof(double));
    RM.dinpt = za(NEQ * ntstpsinpt, sizeof(double));
    RM.pdisp = za(NEQ * ntstpsinpt, sizeof(double));
    RM.um = za(NEQ, sizeof(double)); RM.vm = za(NEQ, sizeof(double)); RM.am = za(NEQ, sizeof(double));
    if (load(RM.q, RM.efFE_ref, RM.x, RM.llength, RM.offset, RM.osflag, RM.c1, RM.c2, RM.c3,
             RM.jnt, RM.mcode, RM.jcode, RM.minc, RM.tinpt, RM.pinpt, RM.dinpt, RM.pdisp, RM.um,
             RM.vm, RM.am))
        return 1;
    fclose(IFP[0]);
    IFP[0] = NULL;
    RM.open = 1;
    if (dims_out) {
        dims_out->NJ = NJ; dims_out->NE_TR = NE_TR; dims_out->NE_FR = NE_FR; dims_out->NE_SH = NE_SH;
        dims_out->NEQ = NEQ; dims_out->ANAFLAG = ANAFLAG; dims_out->ALGFLAG = ALGFLAG;
        dims_out->SLVFLAG = SLVFLAG; dims_out->OPTFLAG = OPTFLAG;
    }
    if (lss_out) *lss_out = RM.lss;
    return 0;
}

/* Copies a named array of the last ref_open into out (n entries). */
int ref_get_long(const char *name, long *out, long n)
{
    long *src = NULL;
    if (!strcmp(name, "jcode")) src = RM.jcode;
    else if (!strcmp(name, "minc")) src = RM.minc;
    else if (!strcmp(name, "mcode")) src = RM.mcode;
    else if (!strcmp(name, "maxa")) src = RM.maxa;
    else if (!strcmp(name, "kht")) src = RM.kht;
    else if (!strcmp(name, "jnt")) src = RM.jnt;
    if (!src || !RM.open) return 1;
    memcpy(out, src, sizeof(long) * (size_t)n);
    return 0;
}

int ref_get_double(const char *name, double *out, long n)
{
    double *src = NULL;
    if (!strcmp(name, "x")) src = RM.x;
    else if (!strcmp(name, "emod")) src = RM.emod;
    else if (!strcmp(name, "yield")) src = RM.yield;
    else if (!strcmp(name, "carea")) src = RM.carea;
    else if (!strcmp(name, "llength")) src = RM.llength;
    else if (!strcmp(name, "c1")) src = RM.c1;
    else if (!strcmp(name, "c2")) src = RM.c2;
    else if (!strcmp(name, "c3")) src = RM.c3;
    else if (!strcmp(name, "nu")) src = RM.nu;
    else if (!strcmp(name, "thick")) src = RM.thick;
    else if (!strcmp(name, "xlocal")) src = RM.xlocal;
    else if (!strcmp(name, "farea")) src = RM.farea;
    else if (!strcmp(name, "slength")) src = RM.slength;
    else if (!strcmp(name, "q")) src = RM.q;
    else if (!strcmp(name, "pinpt")) src = RM.pinpt;
    else if (!strcmp(name, "auxpt")) src = RM.auxpt;
    else if (!strcmp(name, "gmod")) src = RM.gmod;
    else if (!strcmp(name, "dens")) src = RM.dens;
    else if (!strcmp(name, "xfr")) src = RM.xfr;
    else if (!strcmp(name, "offset")) src = RM.offset;
    else if (!strcmp(name, "istrong")) src = RM.istrong;
    else if (!strcmp(name, "iweak")) src = RM.iweak;
    else if (!strcmp(name, "ipolar")) src = RM.ipolar;
    else if (!strcmp(name, "iwarp")) src = RM.iwarp;
    if (!src || !RM.open) return 1;
    memcpy(out, src, sizeof(double) * (size_t)n);
    return 0;
}

/* The reference's own results writer (misc.c:345-451) over a recorded path:
 * results1..7.txt are opened in dir as main.c:373-379 opens them, output()
 * writes the headers (flag 0) and then one row per step (flag 1), as the
 * drivers call it (main.c:2556-2559, 3132).  lpf [nsteps], itecnt [nsteps],
 * d [nsteps*NEQ], ef [nsteps*(NE_TR*2 + NE_FR*14 + NE_SH*18)].  The model's
 * sizes and flags must have been set by ref_open. */
int ref_output_files(const char *dir, int nsteps, const double *lpf, const int *itecnt,
                     const double *d, const double *ef)
{
    FILE *keep[8];
    char path[512];
    for (int i = 0; i < 8; ++i) keep[i] = OFP[i];
    for (int i = 0; i < 7; ++i) {
        snprintf(path, sizeof path, "%s/results%d.txt", dir, i + 1);
        OFP[i] = fopen(path, "w");
        if (!OFP[i]) {
            for (int k = 0; k < i; ++k) fclose(OFP[k]);
            for (int k = 0; k < 8; ++k) OFP[k] = keep[k];
            return 1;
        }
    }
    const long nef = NE_TR * 2 + NE_FR * 14 + NE_SH * 18;
    double l = lpf[0];
    int it = itecnt[0];
    output(&l, &it, (double *)d, (double *)ef, 0);
    for (int s = 0; s < nsteps; ++s) {
        l = lpf[s];
        it = itecnt[s];
        output(&l, &it, (double *)(d + (long)s * NEQ), (double *)(ef + (long)s * nef), 1);
    }
    for (int i = 0; i < 7; ++i) fclose(OFP[i]);
    for (int k = 0; k < 8; ++k) OFP[k] = keep[k];
    return 0;
}

/* The reference's checkpoint writer and reader (misc.c:494-716), for the
 * golden checkpoint files: checkPoint writes results8.txt into dir (it opens
 * that name in the working directory); restartStep reads a checkpoint file
 * after its first line, which main.c reads itself (main.c:3381-3386).  ANAFLAG 2
 * only: the plasticity arrays chi, efN, efM are passed but not touched. */
int ref_checkpoint(const char *dir, long tstep, long lss, double *uc, double *vc, double *ac,
                   double *ss, double *sm, double *d, double *f, double *ef, double *x, double *c1,
                   double *c2, double *c3, double *defllen, double *llength, double *efFE,
                   double *xfr, int *yldflag, double *deffarea, double *defslen)
{
    char cwd[4096];
    if (ALGFLAG != 5 || !getcwd(cwd, sizeof cwd) || chdir(dir) != 0) return 1;
    double one = 0;
    checkPoint(tstep, lss, uc, vc, ac, ss, sm, d, f, ef, x, c1, c2, c3, defllen, llength, efFE, xfr,
               yldflag, deffarea, defslen, &one, &one, &one);
    return chdir(cwd) != 0;
}

int ref_restart(const char *path, long *tstep, long lss, double *uc, double *vc, double *ac,
                double *ss, double *sm, double *d, double *f, double *ef, double *x, double *c1,
                double *c2, double *c3, double *defllen, double *llength, double *efFE,
                double *xfr, int *yldflag, double *deffarea, double *defslen)
{
    if (ALGFLAG != 5) return 1;
    FILE *keep = IFP[3];
    IFP[3] = fopen(path, "r");
    if (!IFP[3]) { IFP[3] = keep; return 1; }
    int rc = fscanf(IFP[3], "%ld\n", tstep) == 1 ? 0 : 2;
    double one = 0;
    if (rc == 0)
        restartStep(lss, uc, vc, ac, ss, sm, d, f, ef, x, c1, c2, c3, defllen, llength, efFE, xfr,
                    yldflag, deffarea, defslen, &one, &one, &one);
    fclose(IFP[3]);
    IFP[3] = keep;
    return rc;
}
