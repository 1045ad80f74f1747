/*
 * cb_oracle.h -- CPU restatement of the CU-BENs nonlinear solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path is
 * compared against; it is never linked into, loaded by or called from the
 * product library.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may use it.
 *
 * The restatement follows the reference's algorithms (file:line cited at each
 * function in the .c files) and keeps the reference's flat array layouts so
 * that the same numpy buffers can be handed to the oracle and to the reference
 * element routines compiled from source into oracle/_ref:
 *   per-element properties   [NE_TR | NE_FR | NE_SH]   (emod, yield)
 *   carea, llength           [NE_TR | NE_FR]
 *   nu, thick, dens(shell)   [NE_SH]           (no type offset, as in shell.c:61)
 *   triads c1/c2/c3          [NE_TR | NE_FR*3 | NE_SH*3]
 *   element forces ef        [NE_TR*2 | NE_FR*14 | NE_SH*18]
 *   minc                     [NE_TR*2 | NE_FR*2 | NE_SH*3]   (1-based joints)
 *   mcode                    [NE_TR*6 | NE_FR*14 | NE_SH*18] (1-based eqs, 0 = fixed)
 *   jcode                    NJ*7                            (1-based eqs, 0 = fixed)
 */
#ifndef CB_ORACLE_H
#define CB_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Sizes and flags the reference keeps in globals (main.c:323-328). */
typedef struct {
    long NJ, NE_TR, NE_FR, NE_SH, NEQ;
    int ANAFLAG, ALGFLAG, SLVFLAG, OPTFLAG;
} orc_dims;

/* Parsed model_def.txt plus everything derived from it before the driver runs. */
typedef struct orc_model {
    orc_dims dm;
    int chkpt, rflag;
    long *minc, *jcode, *mcode, *maxa, *kht, *jnt;
    int *wrpres;
    long lss;
    double *x;                        /* NJ*3 */
    /* element properties (layouts in the header comment) */
    double *emod, *yield, *carea, *llength, *dens_tr;
    double *gmod, *dens_fr, *istrong, *iweak, *ipolar, *iwarp, *zstrong, *zweak;
    double *offset, *auxpt, *xfr;     /* NE_FR*6, NE_FR*3, NE_FR*6 */
    int *osflag, *mendrel;            /* NE_FR, NE_FR*5 */
    double *nu, *thick, *dens_sh, *xlocal, *farea, *slength;
    double *dens;                     /* the reference's one density array [NE_TR+NE_FR+NE_SH],
                                         written by prop_tr, prop_fr, prop_sh in turn, each from 0 */
    double *c1, *c2, *c3;
    double *q;                        /* NEQ reference load */
    /* dynamic input (ALGFLAG > 3) */
    long ntstpsinpt; double ttot, dt;
    double *pinpt;                    /* NEQ*ntstpsinpt load histories */
    double *um, *vm, *am;             /* initial conditions */
    /* driver parameters */
    double lpfmax, lpf0, dlpf0, dlpfmax, dlpfmin;          /* NR / Newmark */
    int itemax, submax, solmin, imagmax, negmax, iteopt;
    double toldisp, tolforc, tolener;
    double alpha, psi_thresh, dkimax, dk;                   /* MSAL */
    long dkdof;                                             /* 0-based */
    int numopt; double spectrds;                            /* Newmark */
} orc_model;

/* ---- model input (model.c, main.c:330-1653) ---------------------------- */
orc_model *orc_open(const char *path, char *err, int errlen);
void orc_free(orc_model *m);

/* ---- element routines; argument lists mirror prototypes.h ------------- */
void orc_stiff_tr(const orc_dims *d, double *ss, const double *emod, const double *carea,
                  const double *llength, const double *defllen_ip, const double *yield,
                  const double *c1_ip, const double *c2_ip, const double *c3_ip,
                  const double *ef_ip, const long *maxa, const long *mcode);
void orc_forces_tr(const orc_dims *d, double *f, double *ef_i, const double *emod,
                   const double *carea, const double *llength, const double *defllen_i,
                   const double *yield, const double *c1_i, const double *c2_i,
                   const double *c3_i, const long *mcode);
/* ANAFLAG 1 (linear elastic) forces from the total displacement D:
 * truss.c:243-326 and shell.c:1695-1727 with the scatter of shell.c:2385-2397. */
void orc_forces_tr_lin(const orc_dims *d, double *f, double *ef_i, const double *D,
                       const double *emod, const double *carea, const double *llength,
                       const double *c1, const double *c2, const double *c3, const long *mcode);
void orc_forces_sh_lin(const orc_dims *d, double *f, double *ef_i, const double *D,
                       const double *emod, const double *nu, const double *xlocal,
                       const double *thick, const double *farea, const double *slength,
                       const double *c1, const double *c2, const double *c3, const long *mcode);
void orc_stiff_sh(const orc_dims *d, double *ss, const double *emod, const double *nu,
                  const double *x_temp, const double *xlocal, const double *thick,
                  const double *farea, const double *deffarea_ip, const double *slength,
                  const double *c1_ip, const double *c2_ip, const double *c3_ip,
                  const long *maxa, const long *minc, const long *mcode);
void orc_forces_sh(const orc_dims *d, double *f, double *ef_ip, double *ef_i,
                   const double *dd, const double *x_temp, const double *emod,
                   const double *nu, const double *xlocal, const double *thick,
                   const double *farea, const double *slength,
                   const double *c1_ip, const double *c2_ip, const double *c3_ip,
                   const double *c1_i, const double *c2_i, const double *c3_i,
                   const long *minc, const long *mcode);
void orc_updatc(const orc_dims *d, double *x_temp, double *x_ip, double *xfr_temp,
                const double *dd, double *defllen_i, double *deffarea_i, double *defslen_i,
                const double *offset, const int *osflag, const double *auxpt,
                double *c1_i, double *c2_i, double *c3_i, const long *minc,
                const long *jcode);
/* One element's 18x18 global tangent (stiff_sh body, shell.c:140-305). */
void orc_shell_elem_tangent(double E, double nu, double t, const double xl[3], double A0,
                            const double sl[3], double A_ip, const double dm[6],
                            const double c1[3], const double c2[3], const double c3[3],
                            int geometric, double K[18 * 18]);

/* ---- frame (frame.c) ---------------------------------------------------- */
/* Elastic and geometric local 14x14 frame matrices (frame.c:354-579). */
void orc_frame_ke(double E, double G, double A, double L, double Is, double Iw, double J,
                  double Cw, double k[14][14]);
void orc_frame_kg(const double f[14], double L, double A, double Ip, double k[14][14]);
void orc_gj_inverse(double *a, int n);
void orc_frame_release(double k[14][14], const int flags[4]);
void orc_stiff_fr(const orc_dims *d, double *ss, const orc_model *m, const double *llength,
                  const double *defllen_ip, const double *c1_ip, const double *c2_ip,
                  const double *c3_ip, const double *ef_ip, const double *efFE_ip,
                  const int *yldflag);
/* Returns forces_fr's code (frame.c:1184-1265): 0, 1 when a member end has
 * gone through the yield surface (*dlpf scaled by the regula-falsi fraction,
 * the end flagged 1, the members after it left as they were), 2 when a
 * yielded end unloads (flagged 2).  yldflag may be NULL below ANAFLAG 3. */
int orc_forces_fr(const orc_dims *d, double *f, const double *ef_ip, double *ef_i,
                  const double *efFE_ip, double *efFE_i, const double *dd,
                  const orc_model *m, const double *llength, const double *defllen_ip,
                  const double *c1_ip, const double *c2_ip, const double *c3_ip,
                  const double *c1_i, const double *c2_i, const double *c3_i,
                  int *yldflag, double *dlpf);

/* ---- plastic hinges of the frame member ends (ANAFLAG 3) ------------------
 * Orbison's yield function of end forces normalised by the squash load and
 * the plastic moments (frame.c:613-620); its gradient with respect to the end
 * forces (P, My, Mz) (frame.c:626-637); the plastic reduction of a local
 * tangent k <- k - (k G)(G^T k G)^-1 (k G)^T for the ends stiffm_fr selects
 * from the total end forces eftot (frame.c:580-797); the fraction of an
 * end-force increment at which phi reaches 1 (regula_falsi,
 * frame.c:1397-1455); and the unloading test (unload, frame.c:1457-1658):
 * 0, or 1 / 2 / 3 for both ends / end 1 / end 2 unloading. */
double orc_orbison(double p, double my, double mz);
void orc_orbison_grad(double p, double my, double mz, double Py, double Mpy, double Mpz,
                      double g[3]);
void orc_frame_plastic(double k[14][14], const double eftot[14], const int yld[2], double Py,
                       double Mpy, double Mpz);
double orc_regula_falsi(double p, double dp, double my, double dmy, double mz, double dmz);
int orc_frame_unload(const double phi[2], const double eftot_i[14], double Py, double Mpy,
                     double Mpz, const double k[14][14], const double dl[14]);
/* Lumped diagonal masses into sm[NEQ]; overwrite the geometry arrays they are
 * given, as the reference's mass_* do. */
void orc_mass_tr(const orc_dims *d, double *sm, const orc_model *m, const double *x,
                 double *llength, const double *dens);
void orc_mass_fr(const orc_dims *d, double *sm, const orc_model *m, const double *x,
                 double *llength, double *xfr, const double *dens);
void orc_mass_sh(const orc_dims *d, double *sm, const orc_model *m, const double *x,
                 double *farea, double *slength, const double *dens);

/* Support reactions r[NJ*7] of an element state (ef, triads): T_i^T ef_i of
 * every element summed at the constrained joint slots (jcode 0), zero at the
 * free ones (the reference computes none; orc_reactions.c). */
void orc_reactions(const orc_model *m, const double *ef, const double *c1, const double *c2,
                   const double *c3, double *r);

/* ---- solver pieces (solve.c, misc.c, arc.c) ------------------------------ */
int orc_skyfact(long neq, const long *maxa, double *ss, double *ssd, int msal, int *det);
void orc_skysolve(long neq, const long *maxa, const double *ss, double *rhs);
/* out = K v for the symmetric skyline K (skymult, solve.c:700-760) */
void orc_skymult(long neq, const long *maxa, const double *ss, const double *v, double *out);
int orc_test(long neq, const double *d_temp, const double *dd, const double *f_temp,
             const double *fp, const double *qtot, const double *f_ip, double intener1,
             int *convchk, double toldisp, double tolforc, double tolener);
int orc_quad(long neq, double a, double b, double c, double *dt, const double *dp,
             const double *ddr, const double *ddq, double *dd, double *dlpf, double *lpft);

/* Skyline LDL^T of the leading columns only, stopping once budget_s seconds
 * have passed; *cols receives the number of columns factorized. */
int orc_skyfact_partial(long neq, const long *maxa, double *ss, double budget_s, long *cols);

/* Bounded CPU timing sample of the serial path on a model: one tangent
 * assembly, one internal-force pass (updatc + forces), one skyline
 * back-substitution and a time-bounded partial factorization whose full cost
 * is extrapolated by the column-height weights sum(kht^2). */
typedef struct {
    double t_tangent, t_forces, t_solve, t_factor_part, factor_fraction, t_factor_est;
    long cols_done;
} orc_sample;
int orc_time_sample(orc_model *m, double factor_budget_s, orc_sample *out);

/* ---- element backend: lets the driver run on the oracle's own element
 *      routines or on the reference's (oracle/_ref) ------------------------ */
typedef struct orc_backend {
    void *ctx;
    void (*stiff_tr)(void *ctx, const orc_dims *d, double *ss, const double *emod,
                     const double *carea, const double *llength, const double *defllen_ip,
                     const double *yield, const double *c1_ip, const double *c2_ip,
                     const double *c3_ip, const double *ef_ip, const long *maxa,
                     const long *mcode);
    /* d_tot: the total displacement (pd of forces_tr, pd_temp of forces_sh),
       which the linear-elastic branches (ANAFLAG 1) read */
    void (*forces_tr)(void *ctx, const orc_dims *d, double *f, double *ef_i,
                      const double *emod, const double *carea, const double *llength,
                      const double *defllen_i, const double *yield, const double *c1_i,
                      const double *c2_i, const double *c3_i, const long *mcode,
                      const double *d_tot);
    void (*stiff_sh)(void *ctx, const orc_dims *d, double *ss, const double *emod,
                     const double *nu, const double *x_temp, const double *xlocal,
                     const double *thick, const double *farea, const double *deffarea_ip,
                     const double *slength, const double *defslen_ip, const double *yield,
                     const double *c1_ip, const double *c2_ip, const double *c3_ip,
                     const double *ef_ip, const long *maxa, const long *minc,
                     const long *mcode);
    int (*forces_sh)(void *ctx, const orc_dims *d, double *f, double *ef_ip, double *ef_i,
                     const double *dd, const double *x_temp, const double *x_ip,
                     const double *emod, const double *nu, const double *xlocal,
                     const double *thick, const double *farea, const double *deffarea_ip,
                     const double *slength, const double *defslen_ip, const double *yield,
                     const double *c1_ip, const double *c2_ip, const double *c3_ip,
                     const double *c1_i, const double *c2_i, const double *c3_i,
                     const long *minc, const long *mcode, const long *jcode,
                     const double *d_tot);
    void (*updatc)(void *ctx, const orc_dims *d, double *x_temp, double *x_ip,
                   double *xfr_temp, const double *dd, double *defllen_i,
                   double *deffarea_i, double *defslen_i, const double *offset,
                   const int *osflag, const double *auxpt, double *c1_i, double *c2_i,
                   double *c3_i, const long *minc, const long *jcode);
    void (*stiff_fr)(void *ctx, const orc_dims *d, double *ss, const orc_model *m,
                     const double *llength, const double *defllen_ip, const double *c1_ip,
                     const double *c2_ip, const double *c3_ip, const double *ef_ip,
                     const double *efFE_ip, int *yldflag);
    int (*forces_fr)(void *ctx, const orc_dims *d, double *f, double *ef_ip, double *ef_i,
                     const double *efFE_ip, double *efFE_i, const double *dd,
                     const orc_model *m, const double *llength, const double *defllen_ip,
                     const double *c1_ip, const double *c2_ip, const double *c3_ip,
                     const double *c1_i, const double *c2_i, const double *c3_i,
                     int *yldflag, double *dlpf, int itecnt);
    void (*mass_tr)(void *ctx, const orc_dims *d, double *sm, const orc_model *m,
                    const double *x, double *llength, const double *dens);
    void (*mass_fr)(void *ctx, const orc_dims *d, double *sm, const orc_model *m,
                    const double *x, double *llength, double *xfr, const double *dens);
    void (*mass_sh)(void *ctx, const orc_dims *d, double *sm, const orc_model *m,
                    const double *x, double *farea, double *slength, const double *dens);
} orc_backend;

void orc_default_backend(orc_backend *b);

/* ---- drivers -------------------------------------------------------------- */
/* One converged step of the load-displacement path (misc.c:414-417 writes the
 * same triple to results2.txt). */
typedef struct {
    int max_steps;       /* capacity of the arrays below                      */
    int nsteps;          /* rows written                                      */
    double *lpf;         /* [max_steps]                                       */
    int *itecnt;         /* [max_steps]                                       */
    double *d;           /* [max_steps*NEQ] full-precision displacements      */
    double *f;           /* [max_steps*NEQ] internal force at convergence (opt) */
    int status;          /* 0 = ran to its end or to max_steps, else error code */
    char msg[160];
    double t_tangent, t_factor, t_solve, t_forces, t_other;  /* seconds */
    long n_factor, n_solve, n_newton;
    /* optional per-step driver events [max_steps*ORC_NEV] (ORC_EV_* slots) and
     * the arc-length psi the step was taken with [max_steps] (NULL: not kept) */
    int *events;
    double *psi;
    double *react;       /* optional [max_steps*NJ*7] reactions of each converged step */
    double *resid;       /* optional [max_steps*2]: |lpf q - f| and |lpf q| at convergence
                            (NR/MNR/MSAL; zero under Newmark) */
    double *ef;          /* optional [max_steps*(NE_TR*2+NE_FR*14+NE_SH*18)]: the
                            converged element forces of each step (what output()
                            averages into results3-5.txt, misc.c:419-449) */
    /* Newmark checkpoint file (checkPoint / restartStep, misc.c:494-716): when
       set, written every CHKPT time steps, and read at the start of a run whose
       model has RFLAG 1; ckpt_exact 1 writes full precision plus the working
       lengths, areas and time-increment multiplier (orc_checkpoint.c) */
    const char *ckpt_path;
    int ckpt_exact;
} orc_path;

/* Driver events of one converged step: the pivot-sign flag of the step's last
 * factorization (det, solve.c:608-621), and how often within the step quad
 * found imaginary roots (arc.c:145) or no root with a positive gamma
 * (arc.c:134), the dnorm test restarted the increment (main.c:2936-2953),
 * the arc or load step was halved after itemax iterations (main.c:3000-3012,
 * 2096-2110), and, under ANAFLAG 3, how often forces_fr's code 1 (a member
 * end through the yield surface: load step scaled by the regula-falsi
 * fraction, main.c:2050-2057, or the arc halved, main.c:2961-2970) and code 2
 * (a yielded end unloading: the same step again, main.c:2033-2034, 2954-2960)
 * restarted the step. */
enum { ORC_EV_DET = 0, ORC_EV_IMAG = 1, ORC_EV_NEG = 2, ORC_EV_DNORM = 3, ORC_EV_HALVE = 4,
       ORC_EV_YIELD = 5, ORC_EV_UNLOAD = 6, ORC_NEV = 7 };

/* Runs the analysis selected by ALGFLAG (1/2 NR/MNR main.c:1800-2157,
 * 3 MSAL main.c:2158-3141, 5 nonlinear Newmark main.c:3305-3960) with the
 * element routines of backend b (NULL = oracle's own). */
int orc_run(orc_model *m, const orc_backend *b, orc_path *out);

/* ---- checkpoint file of the nonlinear Newmark driver (orc_checkpoint.c) ----
 * The state checkPoint writes (misc.c:494-603) in the reference's layouts;
 * ANAFLAG 2 (the shell section without the plasticity arrays). */
typedef struct {
    long tstep;
    double *uc, *vc, *ac;            /* NEQ                                  */
    double *ss, *sm;                 /* lss: the tangent K of the last iteration, the lumped mass */
    double *d, *f;                   /* NEQ                                  */
    double *ef;                      /* NE_TR*2 + NE_FR*14 + NE_SH*18        */
    double *x;                       /* NJ*3                                 */
    double *c1, *c2, *c3;            /* NE_TR + NE_FR*3 + NE_SH*3            */
    double *defllen, *llength;       /* NE_TR + NE_FR (llength: frame part)  */
    double *efFE, *xfr;              /* NE_FR*14, NE_FR*6                    */
    int *yldflag;                    /* NE_FR*2                              */
    double *deffarea, *defslen;      /* NE_SH, NE_SH*3                       */
    /* the full-precision extension: present or not (read), written or not */
    int has_ext;
    double ddt;                      /* time-increment multiplier of the next step */
    double *llen_w;                  /* NE_TR + NE_FR working lengths        */
    double *farea_w, *slen_w;        /* NE_SH, NE_SH*3 working areas, side lengths */
} orc_ckpt;
int orc_ckpt_write(const char *path, const orc_dims *d, long lss, const orc_ckpt *c, int exact);
/* The linear Newmark driver's checkpoint (solve.c:445-457, read at 257-282):
 * the time step, then NEQ lines "u,v,a", then "0,0,0". */
int orc_ckpt_write_lin(const char *path, long neq, long k, const double *u, const double *v,
                       const double *a, int exact);
int orc_ckpt_read_lin(const char *path, long neq, long *k, double *u, double *v, double *a);
/* 0 ok, 1 cannot open, 2 malformed */
int orc_ckpt_read(const char *path, const orc_dims *d, long lss, orc_ckpt *c);

#ifdef __cplusplus
}
#endif
#endif
