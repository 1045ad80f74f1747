// Nonlinear drivers on the device: Newton-Raphson / modified NR load control
// (main.c:1800-2157), the modified spherical arc-length method
// (main.c:2158-3141) and nonlinear Newmark time integration (main.c:3305-3960),
// over the device element kernels and linear solvers.
#pragma once

#include <array>
#include <functional>
#include <string>
#include <vector>

#include "band.h"
#include "dev.h"
#include "model.h"
#include "solver.h"

namespace cb {

enum SolverKind {
    SOLVER_BAND64 = 0,          // FP64 band LDL^T, direct (the reference's algorithm)
    SOLVER_BAND32_FGMRES = 1,   // FGMRES(FP64 K) preconditioned by an FP32 band LDL^T
    SOLVER_BJ_FGMRES = 2,       // FGMRES with joint-block Jacobi
    SOLVER_LEARNED_FGMRES = 3,  // FGMRES with the learned block approximate inverse
    // FP32 precision mode: the element tangent in FP32 (shells; bars and frame
    // members rounded from FP64), its FP32 band LDL^T, one solve per Newton
    // iteration; the internal forces, the residual, the arc-length constraint
    // and the convergence tests stay FP64, so the Newton iteration itself
    // refines each FP32 correction against the FP64 equilibrium
    SOLVER_BAND32 = 4,
};

// Driver events of one converged step: the pivot-sign flag of the step's last
// factorization (det, solve.c:608-621), and how often within the step quad
// found imaginary roots (arc.c:145) or no root with a positive gamma
// (arc.c:134), the dnorm test restarted the increment (main.c:2936-2953),
// the arc, load step or time step was halved after itemax iterations
// (main.c:3000-3012, 2096-2110, 3826-3846), and under ANAFLAG 3 how often
// forces_fr's code 1 (a member end through the yield surface) and code 2 (a
// yielded end unloading) restarted the step (main.c:2033-2057, 2954-2970,
// 3773-3795).
enum { EV_DET = 0, EV_IMAG = 1, EV_NEG = 2, EV_DNORM = 3, EV_HALVE = 4, EV_YIELD = 5, EV_UNLOAD = 6,
       EV_N = 7 };

struct RunOptions {
    int solver = SOLVER_BAND64;
    double krylov_tol = 1e-8;
    int restart = 30;
    int maxit = 20000;
    int judge = 1;          // exact FP64 pivots for det/psi when the solver has none
    int max_steps = 1 << 30;
    int keep_d = 1;         // copy d to the host at every converged step
    int kernel_timing = 0;  // KT_* bitmask of band kernels timed with HIP events
    int factor_schedule = 0;  // 0 fused launch per tile column, 1 split diag/panel/update launches
    int kernel_timing_every = 1;  // time every n-th launch of a timed family
    int sweep_schedule = 0;       // 0 sweeps with the chain matrices, 1 the two-product row step,
                                  // 2 both sweeps on the look-ahead system (rows two back)
    int keep_forces = 0;          // record the averaged element forces of every converged step
    // Newmark checkpoint file (checkPoint / restartStep): written every CHKPT
    // time steps when set, read first when the model has RFLAG 1; exact 1
    // writes full precision and the working lengths, areas and multiplier
    std::string checkpoint_path;
    int checkpoint_exact = 0;
    int refine = 0;   // SOLVER_BAND32: iterative refinement steps per solve
};

struct RunStats {
    double t_total = 0, t_tangent = 0, t_factor = 0, t_solve = 0, t_forces = 0, t_judge = 0;
    double t_other = 0;
    long newton = 0, solves = 0, factors = 0, krylov_iters = 0, krylov_fallbacks = 0;
    double worst_relres = 0;
    int status = 0;
    std::string msg;
};

// Hooks for the learned preconditioner: set-up after each tangent assembly and
// application z = M v on slot-layout vectors (CB_SOLVER_LEARNED_FGMRES); and,
// for online training, a call after every FP64 band factorization and after
// every FP64 band solve with its (right-hand side, solution) in NEQ layout.
struct PrecHook {
    std::function<void(Dev&)> setup;
    std::function<void(Dev&, const double*, double*)> apply;
    std::function<void(Dev&)> after_factor;
    std::function<void(Dev&, const double*, const double*)> observe;
};

class Analysis {
public:
    Analysis(Model& m, Dev& d, const RunOptions& o);
    ~Analysis();
    // Runs the driver of ALGFLAG; records each converged step.
    int run(RunStats& st);
    std::vector<double> lpf;
    std::vector<int> itecnt;
    std::vector<double> dpath;   // nsteps * NEQ
    std::vector<double> fpath;   // nsteps * (NE_TR + 7 NE_FR + 6 NE_SH), with keep_forces
    // per converged step: wall time since the run started, judge time and
    // Newton iterations so far
    std::vector<double> step_t, step_judge;
    std::vector<long> step_newton;
    // cumulative band-kernel clocks at each converged step: [step][family]
    std::vector<std::vector<double>> step_kt_ms;
    std::vector<std::vector<long>> step_kt_n;
    // per converged step: driver events (EV_* slots) and the arc-length psi
    // the step was taken with
    std::vector<std::array<int, EV_N>> step_ev;
    std::vector<double> step_psi;
    // per converged step of a static driver: |lambda q - f| and |lambda q|
    std::vector<double> step_res, step_load;
    PrecHook learned;
    // one-shot operations used by the C ABI
    int tangent_and_factor(int gx, int g, int* det);
    // the tangent the solver factors: d.K, or d.K32 in the FP32 precision
    // mode (and d.K too when the FP64 judge pivots are wanted)
    void tangent(int gx, int g);
    int factor_only(int* det);   // factor the tangent already in d.K
    int mass(double* sm_dev);    // lumped masses at the converged coordinates
    int linear_solve(const double* rhs, double* x);   // device NEQ vectors
    double* dev_vec(int which);
    BandLDLT<double>& band64() { return b64_; }
    BandLDLT<float>& band32() { return b32_; }
    std::vector<double> pivots;   // pivots of the last factorization (host)

private:
    Model& m_;
    Dev& d_;
    RunOptions o_;
    RunStats* st_ = nullptr;
    double t_start_ = 0;
    BandLDLT<double> b64_;
    BandLDLT<float> b32_;
    Fgmres gm_;
    BlockJacobi bj_;
    double *q = nullptr, *dv = nullptr, *dp = nullptr, *dpp = nullptr, *dt = nullptr, *dd = nullptr;
    double *ddq = nullptr, *ddr = nullptr, *f = nullptr, *fp = nullptr, *ft = nullptr, *fip = nullptr;
    double *r = nullptr, *eb = nullptr, *ex = nullptr, *tmp = nullptr;
    double *rf = nullptr;   // refinement residual (SOLVER_BAND32 with refine)
    double* favg = nullptr;   // averaged element forces of one step (keep_forces)
    std::vector<double> piv0_;
    std::array<int, EV_N> ev_{};   // events since the last recorded step
    double ev_psi_ = 1;

    int run_nr();
    int run_msal();
    int run_newmark();
    int run_linear_static();     // ANAFLAG 1 under ALGFLAG 1/2 (main.c:1714-1798)
    int run_linear_newmark();    // ALGFLAG 4 (main.c:3143-3303, solve.c:139-468)
    double* kcopy = nullptr;     // the tangent K kept beside Keff (linear Newmark, K u)
    // nonlinear Newmark state (allocated on first use)
    double *sm = nullptr, *qtot = nullptr, *uc = nullptr, *vc = nullptr, *ac = nullptr;
    double *vci = nullptr, *aci = nullptr, *um = nullptr, *vm = nullptr, *amv = nullptr;
    double *dyn = nullptr, *hist = nullptr;
    void record(double lam, int ite, const double* dvec);
    // the Newmark checkpoint of the committed state at time step k, with the
    // mass term cnum * sm / cden taken off the diagonal of d.K (Keff of the
    // last iteration) to give the tangent the file holds
    bool save_checkpoint(long k, double ddt, double cnum, double cden);
    int load_checkpoint(long* tstep, double* ddt);   // 0 ok, 1 cannot open, 2 malformed
    // updatc + internal forces; returns forces_fr's code (ANAFLAG 3; else 0)
    // and scales *dlpf by the regula-falsi fraction on a code 1
    int update_and_forces(int gx, double* fout, const double* ddv, double* dlpf = nullptr);
    int factor_current(int* det);
    double psi_now() const;
    void shift_ip(bool with_ef) { copy_gen(d_, G_IP, G_I, with_ef, false); }
    void reset_from_base();
    void commit();
};

}  // namespace cb
