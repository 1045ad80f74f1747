/*
 * cubens.h -- C ABI of the MI355X CU-BENs nonlinear solve path.
 *
 * The reference has no library boundary: its drivers in main.c call the element,
 * geometry and solver routines declared in prototypes.h directly, over arrays
 * allocated in main().  Each entry point below replaces one of those calls (or
 * a whole driver) and is marked with the reference interface it stands in for.
 * Plain C types only; host arrays use the reference's flat layouts:
 *   triads c1/c2/c3 [NE_TR | NE_FR*3 | NE_SH*3], element forces ef
 *   [NE_TR*2 | NE_FR*14 | NE_SH*18], coordinates x [NJ*3], vectors [NEQ].
 * All hot-path state lives on the GPU; a handle is not thread-safe.
 * Every function returns 0 on success; on failure a nonzero code and
 * cb_last_error() describes the error.  No entry point falls back to a CPU
 * implementation: without a usable GPU cb_open fails.
 */
#ifndef CUBENS_H
#define CUBENS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cb_model cb_model;

typedef struct {
    long nj, ne_tr, ne_fr, ne_sh, neq;
    long lss;          /* skyline length (model.c:1276-1281)            */
    long bandwidth;    /* max column height of the skyline             */
    long nnz_blocks;   /* joint blocks of the device tangent           */
    int bs;            /* DOF slots per joint in the device layout     */
    int anaflag, algflag;
    int ncolours_tr, ncolours_sh;
} cb_info;

/* Linear solvers replacing solve()/skyfact/skysolve (solve.c:51-89, 539-698). */
enum {
    CB_SOLVER_BAND64 = 0,         /* FP64 band LDL^T on the GPU                  */
    CB_SOLVER_BAND32_FGMRES = 1,  /* FGMRES, FP32 band LDL^T preconditioner      */
    CB_SOLVER_BJ_FGMRES = 2,      /* FGMRES, 6x6 joint-block Jacobi              */
    CB_SOLVER_LEARNED_FGMRES = 3, /* FGMRES, learned block approximate inverse   */
    CB_SOLVER_BAND32 = 4          /* FP32 precision mode: FP32 element tangent and
                                     band LDL^T, one solve per Newton iteration;
                                     forces, residual, arc-length constraint and
                                     convergence tests in FP64 (static drivers) */
};

typedef struct {
    int solver;
    double krylov_tol;   /* relative residual target of FGMRES            */
    int restart;
    int maxit;
    int judge;           /* 1: exact FP64 pivots for det/psi               */
    int max_steps;       /* stop after this many converged steps           */
    int keep_d;          /* keep the displacement path for cb_path         */
    int kernel_timing;   /* bitmask of CB_KT_* kernel families timed with HIP events */
    int factor_schedule; /* CB_FACTOR_* schedule of the band factorization  */
    int kernel_timing_every; /* time every n-th launch of a timed family (<= 1: all) */
    int sweep_schedule;  /* CB_SWEEP_* schedule of the band solves' sweeps    */
    int keep_forces;     /* keep the averaged element forces of every step  */
    /* Newmark checkpoint file (checkPoint / restartStep, misc.c:494-716; the
     * reference's results8.txt): when set, written after every CHKPT-th time
     * step (main.c:3904-3907) and, for a model with RFLAG 1, read before the
     * first step, which is then tstep + 1 (main.c:3378-3398).  NULL: none. */
    const char *checkpoint_path;
    /* 0: the reference's %e file; 1: the same layout at full precision
     * (%.17g, which the reference's reader takes) plus, after its closing
     * line, the working lengths and areas and the time-increment multiplier,
     * so that a restart continues the run bit for bit */
    int checkpoint_exact;
    /* batches: 1 replays each round's Newton iteration as one HIP graph
     * captured from the first round's launches (cb_batch_run without
     * kernel_timing); 0 issues the launches one by one */
    int graph;
    /* CB_SOLVER_BAND32: iterative refinement steps per solve, each an FP64
     * residual of the FP32 tangent (rhs - K32 x, FP64 accumulation) solved
     * again with the FP32 factor and added; 0: the FP32 solve alone */
    int refine;
    /* batches: 1 has the first run time the band factor for a few helper-pair
     * counts on the initial tangent and keep the fastest (every path is the
     * same for each count; the choice is remembered per batch shape in the
     * process, and in the file named by the environment variable
     * CUBENS_TUNE_FILE when set); 0 uses the most pairs that fit */
    int autotune;
    /* batches: 1 forms the frame member tangents in FP32 (rounded once into
     * the FP64 band); internal forces, residual, factor and solve stay FP64 */
    int fp32_elements;
} cb_options;

/* Schedules of the band LDL^T factorization (same tile arithmetic, bitwise
 * the same result): one fused launch per tile column whose workgroups hand
 * the panels to the updates through flags, or a diag, a panel and an update
 * launch per tile column (for timing the tile tasks one by one). */
enum { CB_FACTOR_FUSED = 0, CB_FACTOR_SPLIT = 1 };

/* Schedules of the band solves' forward and backward sweeps (the same
 * solution up to rounding), all one tile row per workgroup: a chain step of
 * one 64x64 product with the matrix Linv_I L_{I,I-1} formed after each
 * factorization (default), of two dependent 64x64 products, or both sweeps on
 * a look-ahead system whose rows depend on the rows two and more back, its
 * band (twice the memory of the factor's) formed after each factorization:
 * faster sweeps for the drivers that solve many times per factorization
 * (arc length, modified NR).  CB_SWEEP_LOOK3 deepens that system once more
 * (rows three and more back, by substituting row I-2's equation into row I;
 * one more band each way). */
enum { CB_SWEEP_CHAIN = 0, CB_SWEEP_ROWS = 1, CB_SWEEP_LOOK2 = 2, CB_SWEEP_LOOK3 = 3 };

/* Generations of the per-element state (base = last converged, _i = current
 * iterate, _ip = previous iterate), as the reference's x/x_temp/x_ip, c1/c1_i/
 * c1_ip, ef/ef_i/ef_ip arrays (main.c:497-510, 760-790). */
enum { CB_GEN_BASE = 0, CB_GEN_I = 1, CB_GEN_IP = 2 };

/* Kernel families of the band solver (timed with HIP events on their stream):
 * the fused per-column factorization launch, the diag/panel/update launches
 * of the split schedule, the forward and backward sweeps, and the launch that
 * forms the sweeps' chain matrices after a factorization. */
enum { CB_KT_COLUMN = 0, CB_KT_DIAG = 1, CB_KT_PANEL = 2, CB_KT_UPDATE = 3, CB_KT_FWD = 4, CB_KT_BWD = 5,
       CB_KT_PREP = 6, CB_KT_N = 7 };

typedef struct {
    double t_total, t_tangent, t_factor, t_solve, t_forces, t_judge, t_other;
    long newton, solves, factors, krylov_iters, krylov_fallbacks;
    double worst_relres;
    int status;          /* 0: the driver ran to its end or to max_steps   */
    int nsteps;
} cb_stats;

const char *cb_last_error(void);

/* Reads model_def.txt and builds device state.  Replaces main.c:330-1653:
 * the header reads, struc (model.c:42), codes (model.c:937), skylin
 * (model.c:1145), coordinates, prop_tr/fr/sh (truss.c:41, frame.c:43,
 * shell.c:43) and load (model.c:1285). */
int cb_open(const char *path, cb_model **out);
void cb_close(cb_model *m);
int cb_get_info(const cb_model *m, cb_info *out);

/* DOF maps as produced by codes()/skylin(): jcode [NJ*7], mcode
 * [NE_TR*6 | NE_FR*14 | NE_SH*18], maxa [NEQ+1] (model.c:937-1283). */
int cb_get_dofmaps(const cb_model *m, long *jcode, long *mcode, long *maxa);
/* Joint renumbering jnt [NJ]: input joint i (0-based) is joint jnt[i] of the
 * model (struc, model.c:143; updatenum, model.c:895-925).  The identity under
 * OPTFLAG 1; under OPTFLAG 2 every per-joint array of this interface (x,
 * jcode, the results files) is in the renumbered order, as in the reference. */
int cb_get_joint_numbers(const cb_model *m, long *jnt);
/* Reads a model_def.txt on the host only (no device needed): the input
 * sequence of cb_open (main.c:330-1653) without the device set-up, for tools
 * such as plotting that need the model's sizes, DOF map jcode [NJ*7],
 * coordinates x [NJ*3] and renumbering jnt [NJ].  Any array may be NULL
 * (call once with NULL arrays to size them). */
int cb_model_info(const char *path, cb_info *info, long *jcode, double *x, long *jnt);
/* Joint-block pattern of the device tangent: rowptr [NJ+1], colidx [nnz_blocks]. */
int cb_get_pattern(const cb_model *m, int32_t *rowptr, int32_t *colidx);
/* Reference load vector q [NEQ] (model.c:1331). */
int cb_get_load(const cb_model *m, double *q);

/* Per-element / per-joint state in the reference layouts.  Names: "x", "x_temp",
 * "x_ip", "c1","c2","c3" (converged), "c1_i".., "c1_ip".., "ef", "ef_i",
 * "ef_ip", "defllen[_i|_ip]", "deffarea[_i|_ip]", "defslen[_i|_ip]". */
int cb_state_set(cb_model *m, const char *name, const double *host, long n);
int cb_state_get(cb_model *m, const char *name, double *host, long n);

/* Tangent assembly: replaces ss = 0; stiff_tr/stiff_fr/stiff_sh (truss.c:82,
 * frame.c:226, shell.c:110).  gen_x / gen_state select the coordinates and element state:
 * 0 converged, 1 current (_i / x_temp), 2 previous (_ip).  The MSAL call of
 * main.c:2717-2734 is (0,0); the NR call of main.c:1903-1921 is (1,2). */
int cb_tangent(cb_model *m, int gen_x, int gen_state);
/* Copies the assembled joint-block values [nnz_blocks*bs*bs] (rows/columns of
 * fixed DOFs carry identity). */
int cb_get_tangent(cb_model *m, double *vals);
/* The same tangent as the FP32 precision mode (CB_SOLVER_BAND32) forms it:
 * stiff_sh (shell.c:110-840) with the element stiffness in FP32 for shell-only
 * models, the FP64 tangent rounded once otherwise; vals [nnz_blocks*bs*bs]. */
int cb_tangent_fp32(cb_model *m, int gen_x, int gen_state, float *vals);

/* Internal force: replaces forces_tr/forces_fr/forces_sh (truss.c:231,
 * frame.c:902, shell.c:1593).
 * f_host [NEQ] receives the force (accumulated from zero); dd_host [NEQ] is the
 * increment; reads ef_ip (clears its membrane entries as shell.c:1773 does) and
 * writes ef_i; coordinates of generation gen_x.  Returns forces_fr's code:
 * 0, or under ANAFLAG 3 1 when a member end went through the yield surface
 * (its regula-falsi fraction from cb_get_yield_flags) and 2 when a yielded end
 * unloads (frame.c:1184-1265); the members after the first with a code are
 * left out of f_host and keep their ef_i and flags, as the reference's loop
 * stops there.  Negative on an error. */
int cb_internal_force(cb_model *m, int gen_x, const double *dd_host, double *f_host);
/* ANAFLAG 3: the yldflag array of the frame member ends [NE_FR*2] (0 elastic,
 * 1 yielded, 2 unloading; main.c:802-806, frame.c:1184-1265), which the
 * tangent and the force pass read and update, and the regula-falsi fraction
 * of the last cb_internal_force that returned 1 (either pointer may be NULL);
 * set: replaces the array (values 0..2). */
int cb_get_yield_flags(cb_model *m, int *yld, double *tau);
int cb_set_yield_flags(cb_model *m, const int *yld);

/* Lumped diagonal masses at the converged coordinates into sm[NEQ]: replaces
 * mass_tr (truss.c:381-441), mass_fr (frame.c:1314-1393) and mass_sh
 * (shell.c:1505-1591).  Like them it rewrites the working member lengths,
 * face areas and side lengths the element kernels read ("llength", "farea",
 * "slength" in cb_state_get).  Models of one element type only. */
int cb_mass(cb_model *m, double *sm);

/* Geometry update: replaces updatc (misc.c:71): x[gen_x] += dd, x_ip <- old,
 * lengths/areas/triads of generation _i. */
int cb_update_geometry(cb_model *m, int gen_x, const double *dd_host);

/* Factorization of the last assembled tangent with the solver of the options
 * (skyfact, solve.c:539): det = 1 when a pivot is not positive; pivots [NEQ]
 * may be NULL. */
int cb_factor(cb_model *m, const cb_options *opt, int *det, double *pivots);
/* Solve with the last factorization / preconditioner (skysolve, solve.c:655). */
int cb_solve(cb_model *m, const double *rhs_host, double *x_host);

/* Runs the driver of the model's ALGFLAG: NR / modified NR (main.c:1800-2157),
 * modified spherical arc length (main.c:2158-3141) or nonlinear Newmark
 * (main.c:3305-3960).  Returns 0 when the driver ran to its end or to
 * max_steps; a positive driver status when it stopped as the reference's
 * driver stops (1: an error the reference prints to results1.txt and exits on,
 * e.g. main.c:3041 "maximum allowable number of subdivisions exceeded";
 * 2: no driver for this ALGFLAG; 3: minimum time increment reached,
 * main.c:3841), with cb_last_error() the message and the path up to there
 * kept; a negative code on an argument or device error.  st may be NULL. */
int cb_run(cb_model *m, const cb_options *opt, cb_stats *st);
/* Load-displacement path of the last run, the triple misc.c:414-417 writes
 * to results2.txt: lpf [nsteps], itecnt [nsteps], d [nsteps*NEQ] (full
 * precision). */
int cb_get_path(const cb_model *m, double *lpf, int *itecnt, double *d);
/* Per converged step of the last run: wall seconds since the run started,
 * seconds spent in the FP64 pivot judge, Newton iterations so far (no
 * reference counterpart: the reference keeps no timers). */
int cb_get_step_stats(const cb_model *m, double *t_wall, double *t_judge, long *newton);
/* Accumulated HIP-event times (ms) and counts of the timed launches of the
 * band kernel families [CB_KT_N] selected by cb_options.kernel_timing (every
 * kernel_timing_every-th launch is timed; ms / launches is the average), for the FP64 (precision
 * 64) or FP32 (32) factor; step >= 0 returns the totals at that converged step
 * of the last run, step < 0 the totals now (no reference counterpart). */
int cb_get_kernel_times(const cb_model *m, int precision, int step, double *ms, long *launches);
/* Fused dot products out[k] = sum_i a[k][i] * b[k][i], 1 <= n <= 8, over
 * device vectors of length len on the current device: replaces dot
 * (misc.c:252-262), which the arc-length driver calls about a dozen times per
 * iteration (main.c:2873-2889) and test() per convergence check
 * (misc.c:187-250).  One pass over the vectors with a deterministic two-level
 * reduction (partial sums in a device buffer this call allocates); runs on
 * the null stream and returns the results to the host. */
int cb_dots(int n, const double *const *a, const double *const *b, long len, double *out);

/* The fused dot products of cb_dots without allocating anything, for callers
 * that cannot call hipMalloc (code under HIP graph capture, tight serving
 * loops, harnesses that forbid allocation).  No counterpart in the reference.
 * a[k], b[k] and out [n] are device pointers on the current device; the work
 * is queued on `stream` (a hipStream_t, NULL for the null stream) and the
 * call returns without waiting.  At most 8 products per call (larger n is
 * rejected): the partial sums live in a static 33 KB device array that each
 * GPU the library runs on holds until the process ends and that every call on
 * that GPU shares, so calls that overlap on one GPU race. */
int cb_dots_noalloc(int n, const double *const *a, const double *const *b, long len, double *out,
                    void *stream);

/* Support reactions r [NJ*7] (joint-major, directions 1..7 as jcode) of the
 * element state of generation gen (CB_GEN_BASE after a run: the last converged
 * step): the element forces T_i^T ef_i that forces_tr/forces_fr/forces_sh
 * accumulate into f (truss.c:349-372, frame.c:1273-1309, shell.c:1728-1785),
 * summed at the constrained slots (jcode 0, where those routines skip the
 * scatter); zero at free slots.  In the reference's sign convention of f, so
 * at equilibrium sum(r) + lambda sum(q) = 0 per direction.  No reference
 * counterpart: the reference computes no reactions. */
int cb_get_reactions(cb_model *m, int gen, double *r);

/* Per converged step of the last run: driver events [nsteps*CB_NEV] in the
 * CB_EV_* slots and the arc-length psi the step was taken with [nsteps]
 * (either pointer may be NULL).  The events are the branches of the
 * reference's drivers taken within the step: the pivot-sign flag det of the
 * step's last factorization (solve.c:608-621, which sets the predictor's sign,
 * main.c:2765-2769), quad's imaginary roots (arc.c:145) and its no-positive-
 * gamma return (arc.c:134), the dnorm restart (main.c:2936-2953) and the
 * halvings after itemax iterations (main.c:3000-3012; the load step of
 * main.c:2096-2110 and the time step of main.c:3826-3846 for the other
 * drivers), and under ANAFLAG 3 the restarts on forces_fr's code 1, a member
 * end through the yield surface (the load step scaled by the regula-falsi
 * fraction, main.c:2050-2057, 3789-3797, or the arc halved, main.c:2961-2970),
 * and code 2, a yielded end unloading (the same step again, main.c:2033-2034,
 * 2954-2960, 3773-3774).  No reference counterpart: the reference keeps no
 * such record. */
enum { CB_EV_DET = 0, CB_EV_IMAG = 1, CB_EV_NEG = 2, CB_EV_DNORM = 3, CB_EV_HALVE = 4,
       CB_EV_YIELD = 5, CB_EV_UNLOAD = 6, CB_NEV = 7 };
int cb_get_step_events(const cb_model *m, int *events, double *psi);
/* Per converged step of the last NR/MNR/MSAL run: the 2-norms of the
 * out-of-balance force lambda q - f and of the applied load lambda q at
 * convergence [nsteps] each (zero for the Newmark driver; either pointer may
 * be NULL).  No reference counterpart: test() prints nothing of them. */
int cb_get_step_residuals(const cb_model *m, double *res, double *load);
/* Writes the results2.txt rows of the last run (format of misc.c:350-417);
 * fails when the run kept no displacement path (keep_d = 0). */
int cb_write_results2(const cb_model *m, const char *path);
/* Averaged element forces of every converged step of the last run (a run
 * with keep_forces), as output() writes them to results3-5.txt
 * (misc.c:419-449): per step NE_TR truss values, 7 per frame member, 6 per
 * shell [nsteps*(NE_TR + 7*NE_FR + 6*NE_SH)]. */
int cb_get_path_forces(const cb_model *m, double *forces);
/* Writes results2.txt .. results5.txt of the last run into directory dir as
 * output() (misc.c:345-451) does over a run: the displacement path and, for
 * each element type of the model, its averaged element forces; a model
 * without trusses gets a results3.txt of the tabs output()'s frame and shell
 * headers write there.  Needs keep_d, and keep_forces for a model with
 * elements. */
int cb_write_results(const cb_model *m, const char *dir);

/* ---- learned preconditioner (no reference counterpart) ------------------
 * A message-passing network on the joint graph of the tangent emits a
 * factorized block approximate inverse M = S G^T G S used by
 * CB_SOLVER_LEARNED_FGMRES in place of block-Jacobi; FGMRES keeps the true
 * FP64 residual and falls back to the FP64 band factor (the judge) when it
 * misses its tolerance.  The network is trained on the GPU by its own
 * forward/backward kernels on (tangent, right-hand side, solution) triples of
 * the FP64 solves of an analysis. */
typedef struct cb_learned cb_learned;
typedef struct {
    long steps;                 /* training steps (one per FP64 solve)      */
    double loss_first, loss_last, loss_recent;   /* recent = mean of the last 10 */
} cb_train_stats;

int cb_learned_create(unsigned seed, cb_learned **out);
void cb_learned_free(cb_learned *L);
long cb_learned_nparams(const cb_learned *L);
int cb_learned_get_params(const cb_learned *L, float *params);
int cb_learned_set_params(cb_learned *L, const float *params);
/* Training objective on the error t = M r - K^{-1} r in the Jacobi-scaled
 * space: 1 energy norm t.K t / r.K^{-1}r (default), 0 |K t|^2/|r|^2,
 * 2 |t|^2/|K^{-1}r|^2. */
int cb_learned_set_objective(cb_learned *L, int kind);
/* Use L for CB_SOLVER_LEARNED_FGMRES runs of model m (NULL detaches). */
int cb_learned_attach(cb_model *m, cb_learned *L);
/* Runs m's analysis with the FP64 band solver for max_steps converged steps and
 * takes one Adam step (learning rate lr) per FP64 solve on the pair
 * (right-hand side, solution) plus `probes` random right-hand sides solved
 * with the same factor. */
int cb_learned_train(cb_learned *L, cb_model *m, int max_steps, int probes, float lr,
                     cb_train_stats *st);
/* Loss and gradient (all parameters) at the tangent currently held by m (after
 * cb_tangent) for nvec NEQ-layout pairs r[b*NEQ..], u[b*NEQ..] with K u = r. */
int cb_learned_loss_grad(cb_learned *L, cb_model *m, const double *r, const double *u, int nvec,
                         double *loss, float *grad);
/* Preconditioner blocks at m's current tangent: G diagonal [nv*36] and edge
 * [ne*36] blocks (either pointer may be NULL; nv, ne receive the sizes). */
int cb_learned_blocks(cb_learned *L, cb_model *m, float *gd, float *go, long *nv, long *ne);

/* ---- batches of independent models ---------------------------------------
 * Many frame models of the nonlinear Newmark driver (main.c:3305-3960, one
 * model_def.txt each, ALGFLAG 5) run side by side on one GPU.  Each round
 * takes one Newton iteration of every running model; each model keeps its own
 * time steps, step halving and convergence tests, and its path is the one
 * cb_run gives for it alone (bitwise with CB_SWEEP_ROWS, whose summation
 * order the batch's solve keeps).  The batch has no reference counterpart (the
 * reference runs one model per process); it replaces running main() once per
 * model. */
typedef struct cb_batch cb_batch;
typedef struct {
    double t_total;      /* seconds of the run (device work and host decisions) */
    long rounds;         /* Newton rounds over the batch                  */
    long newton;         /* Newton iterations summed over the models      */
    long steps;          /* recorded time steps summed over the models    */
    int failed;          /* models whose driver stopped with an error     */
    /* with opt->kernel_timing != 0: HIP-event time (ms) of the batched band
     * factor and solve kernels summed over the rounds, and the algorithmic
     * work of one launch of each (flops; bytes of the factor and Linv tile
     * entries the two sweeps read, structurally zero chunks left out) */
    double ms_factor, ms_solve;
    double factor_flops, solve_bytes;
    /* stream operations (launches, copies, graph launches) issued per round on
     * average, and the operations of one round's Newton iteration (the nodes of
     * its graph) */
    double ops_per_round;
    long body_ops;
    /* helper pairs of the band factor in this run and the most it may use;
     * seconds spent tuning (autotune) */
    int pairs, pairs_max;
    double tune_s;
} cb_batch_stats;
/* Reads every model (as cb_open does) and builds the batch's device state;
 * frame-only models without member-end offsets or free warping DOFs. */
int cb_batch_open(const char *const *paths, int count, cb_batch **out);
void cb_batch_close(cb_batch *b);
int cb_batch_count(const cb_batch *b);
int cb_batch_get_info(const cb_batch *b, int i, cb_info *out);
/* Runs every model for at most opt->max_steps recorded steps (keep_d: keep
 * the displacement rows on the GPU for cb_batch_get_path; kernel_timing:
 * time the band kernels). */
int cb_batch_run(cb_batch *b, const cb_options *opt, cb_batch_stats *st);
/* Per model [count]: driver status (0 ran to its end or to max_steps, 1 zero
 * norm in the convergence test, 3 minimum time increment reached without
 * convergence) and recorded steps. */
int cb_batch_get_status(const cb_batch *b, int *status, int *nsteps);
/* The tuner's trials (no counterpart in the reference): up to cap (helper
 * pairs, band factor ms; ms < 0 when a wait bound was hit) pairs; returns the
 * number of trials, 0 when it has not run or its choice came from a cache. */
int cb_batch_get_tuning(const cb_batch *b, int *pairs, double *ms, int cap);
/* Model i's path as cb_get_path gives it: time [nsteps], itecnt [nsteps],
 * d [nsteps*NEQ_i] (any pointer may be NULL). */
int cb_batch_get_path(const cb_batch *b, int i, double *t, int *itecnt, double *d);
#ifdef __cplusplus
}
#endif
#endif
