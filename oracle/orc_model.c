/*
 * orc_model.c -- oracle: model_def.txt input, DOF numbering, skyline profile,
 * initial element geometry and the reference load vector.
 *
 * TEST INFRASTRUCTURE ONLY (see cb_oracle.h).
 *
 * Follows: main.c:340-427 (header), model.c:42-285 (struc, without OPTFLAG=2
 * renumbering), model.c:937-1050 (codes), model.c:1145-1283 (skylin),
 * main.c:1394-1432 (coordinates, prop_*), truss.c:41-80, frame.c:43-224,
 * shell.c:43-108 + 2402-2446, main.c:1599-1653 and model.c:1285-1606 (load),
 * arc.c:41-68 (msal) and the driver parameter reads main.c:1809-1812,
 * 2240-2245, 3320-3326.
 */
#include "cb_oracle.h"

#include <ctype.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ---- a comma/whitespace separated number stream --------------------------- */
typedef struct {
    char *buf;
    size_t pos, len;
    int bad;
} tokrd;

static void tk_skip(tokrd *t)
{
    while (t->pos < t->len && (isspace((unsigned char)t->buf[t->pos]) || t->buf[t->pos] == ','))
        t->pos++;
}

static double tk_dbl(tokrd *t)
{
    tk_skip(t);
    if (t->pos >= t->len) { t->bad = 1; return 0.0; }
    char *end;
    double v = strtod(t->buf + t->pos, &end);
    if (end == t->buf + t->pos) { t->bad = 1; return 0.0; }
    t->pos = (size_t)(end - t->buf);
    return v;
}

static long tk_long(tokrd *t)
{
    tk_skip(t);
    if (t->pos >= t->len) { t->bad = 1; return 0; }
    char *end;
    long v = strtol(t->buf + t->pos, &end, 10);
    if (end == t->buf + t->pos) { t->bad = 1; return 0; }
    t->pos = (size_t)(end - t->buf);
    return v;
}

/* whether a comma follows the last value, past white space (the literal ","
   of a scanf format continuing the line) */
static int tk_comma_next(const tokrd *t)
{
    size_t p = t->pos;
    while (p < t->len && isspace((unsigned char)t->buf[p])) p++;
    return p < t->len && t->buf[p] == ',';
}

static void *zalloc(size_t n, size_t sz) { return calloc(n ? n : 1, sz); }

void orc_free(orc_model *m)
{
    if (!m) return;
    void *ptrs[] = {m->minc, m->jcode, m->mcode, m->maxa, m->kht, m->jnt, m->wrpres, m->x,
                    m->emod, m->yield, m->carea, m->llength, m->dens_tr, m->gmod, m->dens_fr,
                    m->istrong, m->iweak, m->ipolar, m->iwarp, m->zstrong, m->zweak,
                    m->offset, m->auxpt, m->xfr, m->osflag, m->mendrel, m->nu, m->thick,
                    m->dens_sh, m->xlocal, m->farea, m->slength, m->c1, m->c2, m->c3, m->q,
                    m->pinpt, m->um, m->vm, m->am, m->dens};
    for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); ++i) free(ptrs[i]);
    free(m);
}

static double norm3(const double *a) { return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }

static void cross3(const double *a, const double *b, double *c, int unit)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
    if (unit) {
        double l = norm3(c);
        c[0] /= l; c[1] /= l; c[2] /= l;
    }
}

/* Joint numbering rules of struc (model.c:222-285) and codes (model.c:941-1050). */
/* struc's last steps (model.c:146-149, 267-285): the members framing into
 * each joint (warping restraint counts) and the slots a joint loses when no
 * element of the types that use them meets it. */
static void autofix(orc_model *m)
{
    orc_dims *d = &m->dm;
    long nj = d->NJ;
    /* element-type incidence counts per joint: [truss, frame, shell] */
    long *cnt = zalloc((size_t)nj * 3, sizeof(long));
    for (long e = 0; e < d->NE_TR; ++e)
        for (int k = 0; k < 2; ++k) cnt[(m->minc[e * 2 + k] - 1) * 3 + 0]++;
    long off = d->NE_TR * 2;
    for (long e = 0; e < d->NE_FR; ++e)
        for (int k = 0; k < 2; ++k) {
            long j = m->minc[off + e * 2 + k] - 1;
            cnt[j * 3 + 1]++;
            m->wrpres[j * 3 + 1]++;
        }
    off += d->NE_FR * 2;
    for (long e = 0; e < d->NE_SH; ++e)
        for (int k = 0; k < 3; ++k) cnt[(m->minc[off + e * 3 + k] - 1) * 3 + 2]++;

    /* joints that carry no element of a type lose the DOFs only that type uses */
    for (long j = 0; j < nj; ++j) {
        long *jc = m->jcode + j * 7;
        if (cnt[j * 3 + 1] != 0) continue;                    /* frame joint: all 7 */
        if (cnt[j * 3 + 2] != 0) { jc[6] = 0; continue; }      /* shell: no warping  */
        if (cnt[j * 3 + 0] != 0) { for (int k = 3; k < 7; ++k) jc[k] = 0; continue; }
        for (int k = 0; k < 7; ++k) jc[k] = 0;                 /* floating joint     */
    }
    free(cnt);
}

/* graph, optnum and updatenum (model.c:649-935), OPTFLAG 2.  adjacency of
 * joint i as graph() lists it: its trusses, frames and shells in element
 * order (a truss or frame adds its other end, a shell its other two vertices
 * in vertex order).  Each joint in turn seeds a level numbering; an attempt
 * stops extending a joint's neighbours at the first one that lands farther
 * than the best bandwidth so far, and a completed attempt replaces the
 * numbering when its bandwidth is smaller, or equal with a smaller profile.
 * jcode rows move to the new numbers and incidences are renamed; wrpres is
 * left in the input numbering (updatenum copies it onto itself,
 * model.c:893-898). */
static int renumber(orc_model *m)
{
    orc_dims *d = &m->dm;
    const long nj = d->NJ, ntr = d->NE_TR, nfr = d->NE_FR, nsh = d->NE_SH;
    const long fo = ntr * 2, so = ntr * 2 + nfr * 2;
    long *deg = zalloc((size_t)nj + 1, sizeof(long));
    for (long e = 0; e < ntr; ++e) { deg[m->minc[e * 2] - 1]++; deg[m->minc[e * 2 + 1] - 1]++; }
    for (long e = 0; e < nfr; ++e) { deg[m->minc[fo + e * 2] - 1]++; deg[m->minc[fo + e * 2 + 1] - 1]++; }
    for (long e = 0; e < nsh; ++e)
        for (int a = 0; a < 3; ++a) deg[m->minc[so + e * 3 + a] - 1] += 2;
    long *xadj = zalloc((size_t)nj + 1, sizeof(long));
    for (long i = 0; i < nj; ++i) xadj[i + 1] = xadj[i] + deg[i];
    long *adj = zalloc((size_t)xadj[nj] + 1, sizeof(long));
    long *fill = zalloc((size_t)nj + 1, sizeof(long));
    for (long i = 0; i < nj; ++i) fill[i] = xadj[i];
    /* three passes keep each joint's list in type order, elements ascending */
    for (long e = 0; e < ntr; ++e) {
        long l = m->minc[e * 2] - 1, r = m->minc[e * 2 + 1] - 1;
        adj[fill[l]++] = r;
        adj[fill[r]++] = l;
    }
    for (long e = 0; e < nfr; ++e) {
        long l = m->minc[fo + e * 2] - 1, r = m->minc[fo + e * 2 + 1] - 1;
        adj[fill[l]++] = r;
        adj[fill[r]++] = l;
    }
    for (long e = 0; e < nsh; ++e) {
        long v[3] = {m->minc[so + e * 3] - 1, m->minc[so + e * 3 + 1] - 1, m->minc[so + e * 3 + 2] - 1};
        for (int a = 0; a < 3; ++a) {
            long j = v[a];
            for (int b = 0; b < 3; ++b)
                if (b != a) adj[fill[j]++] = v[b];
        }
    }
    long band = nj * 7, prof = band * band;
    int changed = 0;
    long *newjt = zalloc((size_t)nj + 1, sizeof(long)), *joint = zalloc((size_t)nj + 1, sizeof(long));
    for (long ik = 0; ik < nj; ++ik) {
        for (long i = 0; i < nj; ++i) joint[i] = newjt[i] = -1;
        long mx = 0, nprof = 0, i = 0, k = 0, kmax = 0;
        newjt[0] = ik;
        joint[ik] = 0;
        do {
            if (newjt[i] != -1) {
                long node = newjt[i];
                for (long jj = xadj[node]; jj < xadj[node + 1]; ++jj) {
                    long k5 = adj[jj];
                    if (joint[k5] == -1) {
                        k++;
                        kmax++;
                        newjt[k] = k5;
                        joint[k5] = k;
                        long nb = labs(i - k);
                        if (nb <= band) {
                            if (nb > mx) mx = nb;
                        } else {
                            break;
                        }
                    }
                }
            } else if (k != 0) {
                k++;
                newjt[k] = newjt[k - 1] - 1;
            } else {
                break;
            }
            nprof += mx;
            i++;
        } while (k < nj - 1);
        if (k == nj - 1 && (mx < band || (mx == band && nprof < prof))) {
            changed = 1;
            band = mx;
            prof = nprof;
            for (long j = 0; j < nj; ++j) m->jnt[j] = joint[j] != -1 ? joint[j] : ++kmax;
        }
    }
    if (changed) {
        long *jc = zalloc((size_t)nj * 7, sizeof(long));
        memcpy(jc, m->jcode, sizeof(long) * (size_t)nj * 7);
        for (long i = 0; i < nj; ++i)
            for (int k = 0; k < 7; ++k) m->jcode[m->jnt[i] * 7 + k] = jc[i * 7 + k];
        free(jc);
        for (long i = 0; i < ntr * 2 + nfr * 2 + nsh * 3; ++i) m->minc[i] = m->jnt[m->minc[i] - 1] + 1;
    }
    free(deg); free(xadj); free(adj); free(fill); free(newjt); free(joint);
    return 0;
}

/* codes (model.c:937-1050): sequential numbering, joint-major, DOFs 1-6 then
 * warping, and the element DOF maps */
static int number_dofs(orc_model *m, char *err, int errlen)
{
    orc_dims *d = &m->dm;
    long nj = d->NJ;
    long off;
    long neq = 0;
    for (long j = 0; j < nj; ++j) {
        long *jc = m->jcode + j * 7;
        int *wr = m->wrpres + j * 3;
        wr[2] = wr[1];
        for (int k = 0; k < 6; ++k)
            if (jc[k] != 0) jc[k] = ++neq;
        if (jc[6] != 0) {
            jc[6] = ++neq;
        } else if (wr[0] == 1) {
            /* warping released at a joint: one warping DOF per framing member */
            jc[6] = ++neq;
            neq += wr[1] - 1;
        }
    }
    d->NEQ = neq;

    /* element DOF maps */
    for (long e = 0; e < d->NE_TR; ++e) {
        long a = m->minc[e * 2] - 1, b = m->minc[e * 2 + 1] - 1;
        for (int k = 0; k < 3; ++k) {
            m->mcode[e * 6 + k] = m->jcode[a * 7 + k];
            m->mcode[e * 6 + 3 + k] = m->jcode[b * 7 + k];
        }
    }
    long moff = d->NE_TR * 6;
    off = d->NE_TR * 2;
    for (long e = 0; e < d->NE_FR; ++e) {
        long *mc = m->mcode + moff + e * 14;
        long ends[2] = {m->minc[off + e * 2] - 1, m->minc[off + e * 2 + 1] - 1};
        for (int s = 0; s < 2; ++s) {
            long j = ends[s];
            for (int k = 0; k < 6; ++k) mc[s * 7 + k] = m->jcode[j * 7 + k];
            int *wr = m->wrpres + j * 3;
            if (wr[0] == 1) {
                mc[s * 7 + 6] = m->jcode[j * 7 + 6] + wr[1] - wr[2];
                wr[2]--;
            } else {
                mc[s * 7 + 6] = m->jcode[j * 7 + 6];
            }
        }
    }
    moff += d->NE_FR * 14;
    off += d->NE_FR * 2;
    for (long e = 0; e < d->NE_SH; ++e)
        for (int v = 0; v < 3; ++v) {
            long j = m->minc[off + e * 3 + v] - 1;
            for (int k = 0; k < 6; ++k) m->mcode[moff + e * 18 + v * 6 + k] = m->jcode[j * 7 + k];
        }
    (void)err; (void)errlen;
    return 0;
}

/* Column heights and diagonal addresses of the skyline (model.c:1204-1281). */
static void skyline_profile(orc_model *m)
{
    orc_dims *d = &m->dm;
    long neq = d->NEQ;
    memset(m->kht, 0, sizeof(long) * (size_t)(neq ? neq : 1));
    struct { long n; int w; long off; } groups[3] = {
        {d->NE_TR, 6, 0}, {d->NE_FR, 14, d->NE_TR * 6}, {d->NE_SH, 18, d->NE_TR * 6 + d->NE_FR * 14}};
    for (int g = 0; g < 3; ++g)
        for (long e = 0; e < groups[g].n; ++e) {
            const long *mc = m->mcode + groups[g].off + e * groups[g].w;
            long lo = neq;
            for (int k = 0; k < groups[g].w; ++k)
                if (mc[k] > 0 && mc[k] < lo) lo = mc[k];
            for (int k = 0; k < groups[g].w; ++k) {
                long c = mc[k];
                if (c != 0 && c - lo > m->kht[c - 1]) m->kht[c - 1] = c - lo;
            }
        }
    m->maxa[0] = 1;
    for (long i = 0; i < neq; ++i) m->maxa[i + 1] = m->maxa[i] + m->kht[i] + 1;
    m->lss = m->maxa[neq] - 1;
}

/* Local membrane coordinates (x2', x3', y3') of a triangle in triad c
 * (mem_coord, shell.c:2402-2446). */
static void local_coords(const double *xa, const double *xb, const double *xc,
                         const double *c1, const double *c2, double out[3])
{
    double e2[3], e3[3];
    for (int k = 0; k < 3; ++k) { e2[k] = xb[k] - xa[k]; e3[k] = xc[k] - xa[k]; }
    out[0] = c1[0] * e2[0] + c1[1] * e2[1] + c1[2] * e2[2];
    out[1] = c1[0] * e3[0] + c1[1] * e3[1] + c1[2] * e3[2];
    out[2] = c2[0] * e3[0] + c2[1] * e3[1] + c2[2] * e3[2];
}

orc_model *orc_open(const char *path, char *err, int errlen)
{
#define FAIL(...) do { if (err) snprintf(err, (size_t)errlen, __VA_ARGS__); orc_free(m); free(t.buf); return NULL; } while (0)
    tokrd t = {0};
    orc_model *m = NULL;
    FILE *fp = fopen(path, "rb");
    if (!fp) { if (err) snprintf(err, (size_t)errlen, "cannot open %s", path); return NULL; }
    fseek(fp, 0, SEEK_END);
    long len = ftell(fp);
    fseek(fp, 0, SEEK_SET);
    t.buf = malloc((size_t)len + 1);
    t.len = fread(t.buf, 1, (size_t)len, fp);
    t.buf[t.len] = 0;
    fclose(fp);

    m = zalloc(1, sizeof(orc_model));
    orc_dims *d = &m->dm;
    d->ANAFLAG = (int)tk_long(&t);
    if (d->ANAFLAG == 4) FAIL("ANAFLAG 4 (acoustic FSI) is not on this path");
    d->ALGFLAG = (int)tk_long(&t);
    if (d->ALGFLAG > 3) { m->chkpt = (int)tk_long(&t); m->rflag = (int)tk_long(&t); }
    d->SLVFLAG = (int)tk_long(&t);
    d->OPTFLAG = (int)tk_long(&t);
    d->NJ = tk_long(&t);
    d->NE_TR = tk_long(&t); d->NE_FR = tk_long(&t); d->NE_SH = tk_long(&t);
    long nsbr = tk_long(&t), nfbr = tk_long(&t);
    if (t.bad) FAIL("truncated header");
    if (nsbr || nfbr) FAIL("brick elements are not on this path");
    if (d->OPTFLAG != 1 && d->OPTFLAG != 2) FAIL("OPTFLAG must be 1 or 2");
    if (d->ANAFLAG < 1 || d->ANAFLAG > 3) FAIL("ANAFLAG %d not supported", d->ANAFLAG);
    if (d->ALGFLAG < 1 || d->ALGFLAG > 5) FAIL("ALGFLAG %d not supported", d->ALGFLAG);
    if (d->ANAFLAG == 1 && (d->ALGFLAG == 3 || d->ALGFLAG == 5))
        FAIL("linear analysis (ANAFLAG 1) runs under ALGFLAG 1, 2 (static) or 4 (linear Newmark)");
    if (d->ALGFLAG == 4 && d->ANAFLAG != 1)
        FAIL("invalid entry for analysis type: ALGFLAG 4 (linear Newmark) takes ANAFLAG 1");
    if (d->SLVFLAG != 0) FAIL("only SLVFLAG 0 (skyline) is restated");
    if (d->ANAFLAG == 3 && d->NE_SH > 0)
        FAIL("ANAFLAG 3 shells (Ivanov's yield criterion, shell.c:842-1503) are not restated");
    long nj = d->NJ, ntr = d->NE_TR, nfr = d->NE_FR, nsh = d->NE_SH;

    m->minc = zalloc((size_t)(ntr * 2 + nfr * 2 + nsh * 3), sizeof(long));
    m->jcode = zalloc((size_t)nj * 7, sizeof(long));
    m->mcode = zalloc((size_t)(ntr * 6 + nfr * 14 + nsh * 18), sizeof(long));
    m->wrpres = zalloc((size_t)nj * 3, sizeof(int));
    m->jnt = zalloc((size_t)nj, sizeof(long));
    m->x = zalloc((size_t)nj * 3, sizeof(double));
    for (long j = 0; j < nj; ++j) m->jnt[j] = j;

    /* incidences (model.c:95-124) */
    for (long i = 0; i < ntr * 2 + nfr * 2 + nsh * 3; ++i) {
        m->minc[i] = tk_long(&t);
        if (m->minc[i] < 1 || m->minc[i] > nj) FAIL("incidence %ld out of range", i);
    }
    /* constraints (model.c:222-264); free = -1, fixed = 0 */
    for (long i = 0; i < nj * 7; ++i) m->jcode[i] = -1;
    for (long i = 0; i < nj * 7; ++i) {
        long j = tk_long(&t), k = tk_long(&t);
        if (t.bad) FAIL("truncated constraints");
        if (j == 0) break;
        if (k < 1 || k > 7 || j > nj) FAIL("joint constraint input not recognized");
        /* model.c:251-254: a bare "j,7" line leaves the restraint 0 */
        if (k == 7 && tk_comma_next(&t)) m->wrpres[(j - 1) * 3] = (int)tk_long(&t);
        m->jcode[(j - 1) * 7 + (k - 1)] = 0;
    }
    autofix(m);
    if (d->OPTFLAG == 2) renumber(m);
    number_dofs(m, err, errlen);
    long neq = d->NEQ;

    m->kht = zalloc((size_t)neq, sizeof(long));
    m->maxa = zalloc((size_t)neq + 1, sizeof(long));
    /* prescribed displacement DOFs (model.c:1156-1202) -- only the terminator is
     * accepted on this path (configs use NBC = 0) */
    if (d->ALGFLAG > 3) {
        long j = tk_long(&t), k = tk_long(&t);
        if (j != 0 || k != 0) FAIL("prescribed displacement DOFs (NBC > 0) are not on this path");
    }
    skyline_profile(m);

    /* coordinates (main.c:1394-1396) */
    for (long j = 0; j < nj; ++j)
        for (int k = 0; k < 3; ++k) m->x[m->jnt[j] * 3 + k] = tk_dbl(&t);

    long nprop = ntr + nfr + nsh;
    m->emod = zalloc((size_t)nprop, sizeof(double));
    m->yield = zalloc((size_t)nprop, sizeof(double));
    m->carea = zalloc((size_t)(ntr + nfr), sizeof(double));
    m->llength = zalloc((size_t)(ntr + nfr), sizeof(double));
    m->c1 = zalloc((size_t)(ntr + nfr * 3 + nsh * 3), sizeof(double));
    m->c2 = zalloc((size_t)(ntr + nfr * 3 + nsh * 3), sizeof(double));
    m->c3 = zalloc((size_t)(ntr + nfr * 3 + nsh * 3), sizeof(double));
    m->dens_tr = zalloc((size_t)ntr, sizeof(double));

    /* truss properties (truss.c:41-80) */
    for (long e = 0; e < ntr; ++e) {
        long a = m->minc[e * 2] - 1, b = m->minc[e * 2 + 1] - 1;
        double el[3];
        for (int k = 0; k < 3; ++k) el[k] = m->x[b * 3 + k] - m->x[a * 3 + k];
        m->llength[e] = sqrt(el[0] * el[0] + el[1] * el[1] + el[2] * el[2]);
        m->c1[e] = el[0] / m->llength[e];
        m->c2[e] = el[1] / m->llength[e];
        m->c3[e] = el[2] / m->llength[e];
        m->emod[e] = tk_dbl(&t);
        m->carea[e] = tk_dbl(&t);
        m->dens_tr[e] = tk_dbl(&t);
        m->yield[e] = tk_dbl(&t);
    }

    /* frame properties (frame.c:43-224) */
    m->gmod = zalloc((size_t)nfr, sizeof(double));
    m->dens_fr = zalloc((size_t)nfr, sizeof(double));
    m->istrong = zalloc((size_t)nfr, sizeof(double));
    m->iweak = zalloc((size_t)nfr, sizeof(double));
    m->ipolar = zalloc((size_t)nfr, sizeof(double));
    m->iwarp = zalloc((size_t)nfr, sizeof(double));
    m->zstrong = zalloc((size_t)nfr, sizeof(double));
    m->zweak = zalloc((size_t)nfr, sizeof(double));
    m->offset = zalloc((size_t)nfr * 6, sizeof(double));
    m->auxpt = zalloc((size_t)nfr * 3, sizeof(double));
    m->xfr = zalloc((size_t)nfr * 6, sizeof(double));
    m->osflag = zalloc((size_t)nfr, sizeof(int));
    m->mendrel = zalloc((size_t)nfr * 5, sizeof(int));
    if (nfr > 0) {
        for (long e = 0; e < nfr; ++e) {
            m->emod[ntr + e] = tk_dbl(&t);
            m->gmod[e] = tk_dbl(&t);
            m->dens_fr[e] = tk_dbl(&t);
            m->carea[ntr + e] = tk_dbl(&t);
            m->istrong[e] = tk_dbl(&t);
            m->iweak[e] = tk_dbl(&t);
            m->ipolar[e] = tk_dbl(&t);
            m->iwarp[e] = tk_dbl(&t);
        }
        for (;;) {   /* member end offsets, terminated by a zero member */
            long j = tk_long(&t);
            double o[6];
            for (int k = 0; k < 6; ++k) o[k] = tk_dbl(&t);
            if (t.bad) FAIL("truncated frame offsets");
            if (j == 0) break;
            m->osflag[j - 1] = 1;
            memcpy(m->offset + (j - 1) * 6, o, sizeof o);
        }
        for (long e = 0; e < nfr; ++e) {
            long a = m->minc[ntr * 2 + e * 2] - 1, b = m->minc[ntr * 2 + e * 2 + 1] - 1;
            double *xe = m->xfr + e * 6;
            for (int k = 0; k < 3; ++k) {
                xe[k] = m->x[a * 3 + k] + (m->osflag[e] ? m->offset[e * 6 + k] : 0.0);
                xe[3 + k] = m->x[b * 3 + k] + (m->osflag[e] ? m->offset[e * 6 + 3 + k] : 0.0);
            }
            double el[3], tmp[3], lx[3], ly[3], lz[3];
            for (int k = 0; k < 3; ++k) el[k] = xe[3 + k] - xe[k];
            double L = sqrt(el[0] * el[0] + el[1] * el[1] + el[2] * el[2]);
            m->llength[ntr + e] = L;
            for (int k = 0; k < 3; ++k) m->auxpt[e * 3 + k] = tk_dbl(&t);
            for (int k = 0; k < 3; ++k) { lx[k] = el[k] / L; tmp[k] = m->auxpt[e * 3 + k] - xe[k]; }
            cross3(lx, tmp, lz, 1);
            cross3(lz, lx, ly, 1);
            for (int k = 0; k < 3; ++k) {
                m->c1[ntr + e * 3 + k] = lx[k];
                m->c2[ntr + e * 3 + k] = ly[k];
                m->c3[ntr + e * 3 + k] = lz[k];
            }
        }
        for (long e = 0; e < nfr; ++e) {
            m->yield[ntr + e] = tk_dbl(&t);
            m->zstrong[e] = tk_dbl(&t);
            m->zweak[e] = tk_dbl(&t);
        }
        for (;;) {   /* member end bending releases (frame.c:193-220) */
            long j = tk_long(&t);
            int r[4];
            for (int k = 0; k < 4; ++k) r[k] = (int)tk_long(&t);
            if (t.bad) FAIL("truncated frame releases");
            if (j == 0) break;
            if (j < 1 || j > nfr) FAIL("frame release member out of range");
            for (int k = 0; k < 4; ++k)
                if (r[k] != 0 && r[k] != 1) FAIL("frame release flags are 0 or 1");
            m->mendrel[(j - 1) * 5] = 1;
            for (int k = 0; k < 4; ++k) m->mendrel[(j - 1) * 5 + 1 + k] = r[k];
        }
    }

    /* shell properties and initial geometry (shell.c:43-108) */
    m->nu = zalloc((size_t)nsh, sizeof(double));
    m->thick = zalloc((size_t)nsh, sizeof(double));
    m->dens_sh = zalloc((size_t)nsh, sizeof(double));
    m->xlocal = zalloc((size_t)nsh * 3, sizeof(double));
    m->farea = zalloc((size_t)nsh, sizeof(double));
    m->slength = zalloc((size_t)nsh * 3, sizeof(double));
    long shoff = ntr * 2 + nfr * 2, toff = ntr + nfr * 3, poff = ntr + nfr;
    for (long e = 0; e < nsh; ++e) {
        m->emod[poff + e] = tk_dbl(&t);
        m->nu[e] = tk_dbl(&t);
        m->thick[e] = tk_dbl(&t);
        m->dens_sh[e] = tk_dbl(&t);
        m->yield[poff + e] = tk_dbl(&t);
        const double *p1 = m->x + (m->minc[shoff + e * 3] - 1) * 3;
        const double *p2 = m->x + (m->minc[shoff + e * 3 + 1] - 1) * 3;
        const double *p3 = m->x + (m->minc[shoff + e * 3 + 2] - 1) * 3;
        double e12[3], e23[3], e31[3], nrm[3];
        for (int k = 0; k < 3; ++k) {
            e12[k] = p2[k] - p1[k];
            e23[k] = p3[k] - p2[k];
            e31[k] = p3[k] - p1[k];
        }
        m->slength[e * 3] = norm3(e12);
        m->slength[e * 3 + 1] = norm3(e23);
        m->slength[e * 3 + 2] = norm3(e31);
        cross3(e12, e31, nrm, 0);
        m->farea[e] = 0.5 * norm3(nrm);
        double lx[3], lz[3], ly[3];
        for (int k = 0; k < 3; ++k) {
            lx[k] = e12[k] / m->slength[e * 3];
            lz[k] = nrm[k] / (2 * m->farea[e]);
        }
        cross3(lz, lx, ly, 1);
        for (int k = 0; k < 3; ++k) {
            m->c1[toff + e * 3 + k] = lx[k];
            m->c2[toff + e * 3 + k] = ly[k];
            m->c3[toff + e * 3 + k] = lz[k];
        }
        local_coords(p1, p2, p3, lx, ly, m->xlocal + e * 3);
    }
    if (t.bad) FAIL("truncated element properties");
    /* the reference keeps one density array that prop_tr, prop_fr and prop_sh
       each fill from index 0 (truss.c:66, frame.c:61, shell.c:61) */
    m->dens = zalloc((size_t)nprop, sizeof(double));
    for (long e = 0; e < ntr; ++e) m->dens[e] = m->dens_tr[e];
    for (long e = 0; e < nfr; ++e) m->dens[e] = m->dens_fr[e];
    for (long e = 0; e < nsh; ++e) m->dens[e] = m->dens_sh[e];

    /* load vectors (model.c:1285-1606) */
    m->q = zalloc((size_t)neq, sizeof(double));
    m->um = zalloc((size_t)neq, sizeof(double));
    m->vm = zalloc((size_t)neq, sizeof(double));
    m->am = zalloc((size_t)neq, sizeof(double));
    if (d->ALGFLAG > 3) {
        long ns = tk_long(&t);
        m->ttot = tk_dbl(&t);
        m->dt = m->ttot / (double)ns;
        m->ntstpsinpt = ns + 1;
        if (m->chkpt == 0) m->chkpt = (int)m->ntstpsinpt + 1;
        m->pinpt = zalloc((size_t)(neq * m->ntstpsinpt), sizeof(double));
        /* per-DOF load histories: consecutive lines of one DOF are consecutive steps */
        long prev = 0, i = 0;
        for (;;) {
            long jt = tk_long(&t);
            int dir = (int)tk_long(&t);
            double mag = tk_dbl(&t);
            if (t.bad) FAIL("truncated load histories");
            if (jt == 0) break;
            if (d->ALGFLAG == 5 && mag == 0) mag = 1e-30;       /* model.c:1517-1521 */
            long ks = m->jcode[(jt - 1) * 7 + dir - 1];
            if (ks != prev) i = 0;
            if (ks != 0 && i < m->ntstpsinpt) m->pinpt[(ks - 1) * m->ntstpsinpt + i] = mag;
            prev = ks;
            ++i;
        }
        long jt = tk_long(&t); tk_long(&t); tk_dbl(&t);
        if (jt != 0) FAIL("prescribed support displacements are not on this path");
        for (;;) {   /* initial displacement / velocity / acceleration */
            long j = tk_long(&t);
            int dir = (int)tk_long(&t);
            double u0 = tk_dbl(&t), v0 = tk_dbl(&t), a0 = tk_dbl(&t);
            if (t.bad) FAIL("truncated initial conditions");
            if (j == 0) break;
            long k = m->jcode[(j - 1) * 7 + dir - 1];
            if (k != 0) {
                if (u0 != 0) m->um[k - 1] = u0;
                if (v0 != 0) m->vm[k - 1] = v0;
                if (a0 != 0) m->am[k - 1] = a0;
            }
        }
    } else {
        int any = 0;
        for (;;) {
            long jt = tk_long(&t);
            int dir = (int)tk_long(&t);
            double mag = tk_dbl(&t);
            if (t.bad) FAIL("truncated joint loads");
            if (jt == 0) break;
            any = 1;
            long k = m->jcode[(m->jnt[jt - 1]) * 7 + dir - 1];
            if (k != 0) m->q[k - 1] = mag;          /* assignment, model.c:1331 */
        }
        if (nfr > 0) {
            long fr = tk_long(&t); tk_long(&t); tk_dbl(&t);
            if (fr != 0) FAIL("frame distributed loads are not on this path");
        }
        if (!any) FAIL("no forces present in input file");
    }

    /* driver parameters */
    if ((d->ALGFLAG == 1 || d->ALGFLAG == 2) && d->ANAFLAG == 1) {
        m->lpfmax = tk_dbl(&t);                            /* main.c:1717 */
    } else if (d->ALGFLAG == 4) {
        m->numopt = (int)tk_dbl(&t); m->spectrds = tk_dbl(&t);   /* main.c:3184-3197 */
        m->lpfmax = tk_dbl(&t);
    } else if (d->ALGFLAG == 1 || d->ALGFLAG == 2) {
        m->lpfmax = tk_dbl(&t); m->lpf0 = tk_dbl(&t); m->dlpf0 = tk_dbl(&t);
        m->dlpfmax = tk_dbl(&t); m->dlpfmin = tk_dbl(&t);
        m->itemax = (int)tk_long(&t); m->submax = (int)tk_long(&t); m->solmin = (int)tk_long(&t);
        m->toldisp = tk_dbl(&t); m->tolforc = tk_dbl(&t); m->tolener = tk_dbl(&t);
    } else if (d->ALGFLAG == 3) {
        long jt = tk_long(&t);
        int dir = (int)tk_long(&t);
        m->dk = tk_dbl(&t);
        if (jt < 1 || jt > nj || dir < 1 || dir > 7) FAIL("bad MSAL control DOF");
        long k = m->jcode[m->jnt[jt - 1] * 7 + dir - 1];
        if (k == 0) FAIL("initially prescribed displacement at fixed degree of freedom");
        m->dkdof = k - 1;
        m->alpha = tk_dbl(&t);
        m->psi_thresh = tk_dbl(&t);
        m->iteopt = (int)tk_long(&t);
        m->lpfmax = tk_dbl(&t); m->dkimax = tk_dbl(&t);
        m->itemax = (int)tk_long(&t); m->submax = (int)tk_long(&t);
        m->imagmax = (int)tk_long(&t); m->negmax = (int)tk_long(&t);
        m->toldisp = tk_dbl(&t); m->tolforc = tk_dbl(&t); m->tolener = tk_dbl(&t);
    } else if (d->ALGFLAG == 5) {
        m->numopt = (int)tk_dbl(&t); m->spectrds = tk_dbl(&t);
        m->lpfmax = tk_dbl(&t); m->lpf0 = tk_dbl(&t); m->dlpf0 = tk_dbl(&t);
        m->dlpfmax = tk_dbl(&t); m->dlpfmin = tk_dbl(&t);
        m->itemax = (int)tk_long(&t); m->submax = (int)tk_long(&t); m->solmin = (int)tk_long(&t);
        m->toldisp = tk_dbl(&t); m->tolforc = tk_dbl(&t); m->tolener = tk_dbl(&t);
    }
    if (t.bad) FAIL("truncated solver parameters");
    free(t.buf);
    return m;
#undef FAIL
}
