/*
 * orc_checkpoint.c -- oracle: the checkpoint file of the nonlinear Newmark
 * driver, checkPoint and restartStep (misc.c:494-716), restated.
 *
 * TEST INFRASTRUCTURE ONLY (see cb_oracle.h).
 *
 * Layout (one value group per line, misc.c:508-598):
 *   tstep
 *   NEQ x "uc,vc,ac"            then "0,0,0"
 *   lss x "ss,sm"               then "0,0"
 *   NEQ x "d,f"
 *   ef, x                       one value a line
 *   NE_TR+NE_FR*3+NE_SH*3 x "c1,c2,c3"
 *   NE_TR x defllen;  NE_FR x "defllen,llength"
 *   efFE (14 a member), xfr (6 a member), yldflag (2 a member, integers)
 *   ANAFLAG 2: deffarea, defslen                      then "0,0,0"
 * The reference prints every value with %e.  Exact mode prints %.17g, which
 * the reference's %le reads back to the same double, and appends after the
 * closing "0,0,0" a block the reference never reads: the marker line
 * "cubens-exact", the time-increment multiplier of the next step and the
 * working lengths, face areas and side lengths the mass routines overwrite
 * (truss.c:381-441, frame.c:1314-1393, shell.c:1505-1591), which the
 * reference's file does not hold.
 */
#include "cb_oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static void put(FILE *f, int exact, int n, const double *v)
{
    for (int i = 0; i < n; ++i) {
        if (i) fputc(',', f);
        fprintf(f, exact ? "%.17g" : "%e", v[i]);
    }
    fputc('\n', f);
}

int orc_ckpt_write(const char *path, const orc_dims *d, long lss, const orc_ckpt *c, int exact)
{
    FILE *f = fopen(path, "w");
    if (!f) return 1;
    const long neq = d->NEQ, ntr = d->NE_TR, nfr = d->NE_FR, nsh = d->NE_SH;
    fprintf(f, "%ld\n", c->tstep);
    for (long i = 0; i < neq; ++i) {
        double v[3] = {c->uc[i], c->vc[i], c->ac[i]};
        put(f, exact, 3, v);
    }
    fprintf(f, "%d,%d,%d\n", 0, 0, 0);
    for (long i = 0; i < lss; ++i) {
        double v[2] = {c->ss[i], c->sm[i]};
        put(f, exact, 2, v);
    }
    fprintf(f, "%d,%d\n", 0, 0);
    for (long i = 0; i < neq; ++i) {
        double v[2] = {c->d[i], c->f[i]};
        put(f, exact, 2, v);
    }
    for (long i = 0; i < ntr * 2 + nfr * 14 + nsh * 18; ++i) put(f, exact, 1, c->ef + i);
    for (long i = 0; i < d->NJ * 3; ++i) put(f, exact, 1, c->x + i);
    for (long i = 0; i < ntr + nfr * 3 + nsh * 3; ++i) {
        double v[3] = {c->c1[i], c->c2[i], c->c3[i]};
        put(f, exact, 3, v);
    }
    for (long i = 0; i < ntr; ++i) put(f, exact, 1, c->defllen + i);
    for (long i = 0; i < nfr; ++i) {
        double v[2] = {c->defllen[ntr + i], c->llength[ntr + i]};
        put(f, exact, 2, v);
    }
    for (long i = 0; i < nfr * 14; ++i) put(f, exact, 1, c->efFE + i);
    for (long i = 0; i < nfr * 6; ++i) put(f, exact, 1, c->xfr + i);
    for (long i = 0; i < nfr * 2; ++i) fprintf(f, "%d\n", c->yldflag[i]);
    for (long i = 0; i < nsh; ++i) put(f, exact, 1, c->deffarea + i);
    for (long i = 0; i < nsh * 3; ++i) put(f, exact, 1, c->defslen + i);
    fprintf(f, "%d,%d,%d\n", 0, 0, 0);
    if (exact) {
        fprintf(f, "cubens-exact\n");
        put(f, 1, 1, &c->ddt);
        for (long i = 0; i < ntr + nfr; ++i) put(f, 1, 1, c->llen_w + i);
        for (long i = 0; i < nsh; ++i) put(f, 1, 1, c->farea_w + i);
        for (long i = 0; i < nsh * 3; ++i) put(f, 1, 1, c->slen_w + i);
    }
    return fclose(f) ? 1 : 0;
}

/* ---- reading: the values in file order, separated by commas and line ends,
 * as restartStep's fscanf calls take them ----------------------------------- */
typedef struct {
    char *buf, *p;
    int bad;
} tok_t;

static double tk(tok_t *t)
{
    while (*t->p == ',' || *t->p == '\n' || *t->p == '\r' || *t->p == ' ' || *t->p == '\t') t->p++;
    char *e;
    double v = strtod(t->p, &e);
    if (e == t->p) { t->bad = 1; return 0; }
    t->p = e;
    return v;
}

static void tk_zeros(tok_t *t, int n)
{
    for (int i = 0; i < n; ++i)
        if (tk(t) != 0) t->bad = 1;
}

int orc_ckpt_read(const char *path, const orc_dims *d, long lss, orc_ckpt *c)
{
    FILE *f = fopen(path, "rb");
    if (!f) return 1;
    fseek(f, 0, SEEK_END);
    long len = ftell(f);
    fseek(f, 0, SEEK_SET);
    tok_t t = {malloc((size_t)len + 1), NULL, 0};
    if (!t.buf) { fclose(f); return 1; }
    size_t got = fread(t.buf, 1, (size_t)len, f);
    fclose(f);
    t.buf[got] = 0;
    t.p = t.buf;
    const long neq = d->NEQ, ntr = d->NE_TR, nfr = d->NE_FR, nsh = d->NE_SH;
    c->tstep = (long)tk(&t);
    for (long i = 0; i < neq; ++i) { c->uc[i] = tk(&t); c->vc[i] = tk(&t); c->ac[i] = tk(&t); }
    tk_zeros(&t, 3);
    for (long i = 0; i < lss; ++i) { c->ss[i] = tk(&t); c->sm[i] = tk(&t); }
    tk_zeros(&t, 2);
    for (long i = 0; i < neq; ++i) { c->d[i] = tk(&t); c->f[i] = tk(&t); }
    for (long i = 0; i < ntr * 2 + nfr * 14 + nsh * 18; ++i) c->ef[i] = tk(&t);
    for (long i = 0; i < d->NJ * 3; ++i) c->x[i] = tk(&t);
    for (long i = 0; i < ntr + nfr * 3 + nsh * 3; ++i) {
        c->c1[i] = tk(&t); c->c2[i] = tk(&t); c->c3[i] = tk(&t);
    }
    for (long i = 0; i < ntr; ++i) c->defllen[i] = tk(&t);
    for (long i = 0; i < nfr; ++i) { c->defllen[ntr + i] = tk(&t); c->llength[ntr + i] = tk(&t); }
    for (long i = 0; i < nfr * 14; ++i) c->efFE[i] = tk(&t);
    for (long i = 0; i < nfr * 6; ++i) c->xfr[i] = tk(&t);
    for (long i = 0; i < nfr * 2; ++i) c->yldflag[i] = (int)tk(&t);
    for (long i = 0; i < nsh; ++i) c->deffarea[i] = tk(&t);
    for (long i = 0; i < nsh * 3; ++i) c->defslen[i] = tk(&t);
    tk_zeros(&t, 3);
    c->has_ext = 0;
    while (*t.p == '\n' || *t.p == '\r' || *t.p == ' ') t.p++;
    if (!t.bad && strncmp(t.p, "cubens-exact", 12) == 0) {
        t.p += 12;
        c->has_ext = 1;
        c->ddt = tk(&t);
        for (long i = 0; i < ntr + nfr; ++i) c->llen_w[i] = tk(&t);
        for (long i = 0; i < nsh; ++i) c->farea_w[i] = tk(&t);
        for (long i = 0; i < nsh * 3; ++i) c->slen_w[i] = tk(&t);
    }
    free(t.buf);
    return t.bad ? 2 : 0;
}

int orc_ckpt_write_lin(const char *path, long neq, long k, const double *u, const double *v,
                       const double *a, int exact)
{
    FILE *f = fopen(path, "w");
    if (!f) return 1;
    fprintf(f, "%ld\n", k);
    for (long i = 0; i < neq; ++i) {
        double w[3] = {u[i], v[i], a[i]};
        put(f, exact, 3, w);
    }
    fprintf(f, "%d,%d,%d\n", 0, 0, 0);
    return fclose(f) ? 1 : 0;
}

int orc_ckpt_read_lin(const char *path, long neq, long *k, double *u, double *v, double *a)
{
    FILE *f = fopen(path, "rb");
    if (!f) return 1;
    fseek(f, 0, SEEK_END);
    long len = ftell(f);
    fseek(f, 0, SEEK_SET);
    tok_t t = {malloc((size_t)len + 1), NULL, 0};
    if (!t.buf) { fclose(f); return 1; }
    size_t got = fread(t.buf, 1, (size_t)len, f);
    fclose(f);
    t.buf[got] = 0;
    t.p = t.buf;
    *k = (long)tk(&t);
    for (long i = 0; i < neq; ++i) { u[i] = tk(&t); v[i] = tk(&t); a[i] = tk(&t); }
    tk_zeros(&t, 3);
    free(t.buf);
    return t.bad ? 2 : 0;
}
