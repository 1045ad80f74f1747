// Many independent models in one set of device arrays, advanced by the
// nonlinear Newmark driver (main.c:3305-3960) in lockstep: every round each
// running model takes one Newton iteration of its own time step, with its own
// step-size and convergence decisions on the host.
//
// Layout: the models are concatenated (joints, members, equations); each
// model's equations start on a 64-equation tile boundary, so the band of the
// batch holds every model's band unchanged and one workgroup per model factors
// and solves it.  Element kernels run over all members of all models in one
// launch per colour (the colouring of the batch is the union of the models'
// colourings, so a model's tangent is assembled in its own order).
#pragma once

#include <string>
#include <vector>

#include "band.h"
#include "dev.h"
#include "model.h"

namespace cb {

// Per-model values of one round (device copy, refreshed every round).
struct NmParam {
    double sub_dt = 1, lpf = 1, a0 = 0, a2 = 0, a3 = 0, a6 = 0, a7 = 0, am = 0, af = 0;
    double cnum = 0, cden = 1;   // Keff = K + cnum * M / cden on the diagonal (solve.c:199-207)
    long k = 0;                  // time step of the load history
    int op = NM_CORRECT;         // first vector op of the iteration: NM_PREDICT0, NM_PREDICT or NM_CORRECT
    int flip = 0;                // dd sign flip of corrector iterations
    int commit = 0, start = 0;   // state copies before the iteration
    int rec = -1;                // path row recorded before the iteration, or -1
    int pad_ = 0;
};

// Per-model placement in the batch (device copy, fixed).
struct BatchSeg {
    long qb = 0, neq = 0;   // first equation, equations
    long pbase = 0;         // first entry of the model's recorded path
};

// Batched device operations (batch.hip).
struct CopyJob {
    double* dst;
    const double* src;
    long n;
    int per;     // values per owner unit
    int owner;   // 0 equation, 1 joint, 2 member
};
void batch_masked_copy(hipStream_t s, const CopyJob* jobs_dev, int njobs, long nmax, const int32_t* own_eq,
                       const int32_t* own_joint, const int32_t* own_el, const NmParam* prm, int which);
// op: a NewmarkVec, or NM_MODEL_OP for each model's NmParam::op
constexpr int NM_MODEL_OP = -1;
void batch_nm_vec(hipStream_t s, int op, const NewmarkArgs& a, const NmParam* prm, const int32_t* own_eq, long n);
void batch_add_mass_diag(Dev& d, const double* sm, const NmParam* prm, const int32_t* own_joint);
// the same on the diagonal of band tiles of width 64 (w tiles below the diagonal)
void batch_add_mass_band(Dev& d, double* tiles, int w, const double* sm, const NmParam* prm,
                         const int32_t* own_joint);
// per model: sum dd^2, dt^2, (dyn - f)^2, (dyn - fp)^2, dd.(dyn - fip), dd.(dyn - fp)
void batch_test_sums(hipStream_t s, const BatchSeg* segs, int nmod, const double* dd, const double* dt,
                     const double* dyn, const double* f, const double* fp, const double* fip, double* out);
void batch_record(hipStream_t s, const double* dv, double* path, const BatchSeg* segs, const NmParam* prm,
                  const int32_t* own_eq, long n);

struct BatchStats {
    double t_total = 0;
    long rounds = 0, newton = 0, steps = 0;
    int failed = 0;
    // HIP-event time (ms) of the batched band factor and solve kernels, summed
    // over the rounds (when timing is on), and the tile work of one factor
    // launch (flops, T^3 + 2 T^3 m + T^3 m (m+1) per tile column of every model)
    double ms_factor = 0, ms_solve = 0;
    double factor_flops = 0, solve_bytes = 0;
    // stream operations (kernel launches, copies, graph launches) the run
    // issued per round on average, and the operations of one round's Newton
    // iteration (the nodes of its graph)
    double ops_per_round = 0;
    long body_ops = 0;
    // helper pairs of the band factor (the tuned count with autotune) and
    // the most it may use; seconds spent tuning in this run
    int pairs = 0, pairs_max = 0;
    double tune_s = 0;
};

class Batch {
public:
    // Parses every model file, concatenates the models and builds the device
    // state.  Frame-only models of the nonlinear Newmark driver (ALGFLAG 5).
    Batch(const std::vector<std::string>& paths);
    ~Batch();
    int count() const { return int(models_.size()); }
    const Model& model(int b) const { return models_[b]; }
    const Model& all() const { return all_; }
    // Runs every model to its end (or max_steps recorded steps each).  With
    // graph (and no timing), each round's Newton iteration replays one HIP
    // graph captured from the first round's launches.
    // With autotune, the first run times the band factor for a few helper-pair
    // counts on the initial tangent and keeps the fastest (the factor, and so
    // every path, is the same for each count); the choice is remembered for
    // batches of the same shape in this process, and in the file named by
    // CUBENS_TUNE_FILE when that is set.
    // With fp32_elements, the member tangents are formed in FP32 and rounded
    // into the FP64 band; internal forces, residual, factor and solve stay FP64.
    void run(int max_steps, bool keep_path, BatchStats& st, bool timing = false, bool graph = false,
             bool autotune = false, bool fp32_elements = false);
    // the pair counts the tuner timed and their factor times (ms; < 0 a wait
    // bound was hit), empty when it did not run or the choice came from a cache
    std::vector<std::pair<int, double>> tuning;
    // after run(): per model status (0 ran to its end or to max_steps, 1 zero
    // norm in the convergence test, 3 minimum time increment reached), time and
    // Newton iterations of each recorded step, and the displacement rows
    std::vector<int> status;
    std::vector<std::vector<double>> times;
    std::vector<std::vector<int>> itecnt;
    std::vector<std::string> message;
    void path(int b, double* d) const;   // nsteps(b) * neq(b) values

private:
    std::vector<Model> models_;
    Model all_;
    Dev d_;
    BandBatch band_;
    std::vector<BatchSeg> segs_;
    std::vector<long> jb_, eb_;
    long n_ = 0, ld_ = 1;
    int32_t *own_eq_ = nullptr, *own_joint_ = nullptr, *own_el_ = nullptr;
    BatchSeg* segs_dev_ = nullptr;
    NmParam* prm_dev_ = nullptr;
    NmParam* prm_host_ = nullptr;   // pinned
    double* sums_dev_ = nullptr;
    double* sums_host_ = nullptr;   // pinned
    CopyJob* commit_jobs_ = nullptr;
    CopyJob* start_jobs_ = nullptr;
    int ncommit_ = 0, nstart_ = 0;
    double *dv = nullptr, *dt = nullptr, *dd = nullptr, *f = nullptr, *fp = nullptr, *ft = nullptr;
    double *fip = nullptr, *r = nullptr, *tmp = nullptr, *sm = nullptr, *qtot = nullptr, *uc = nullptr;
    double *vc = nullptr, *ac = nullptr, *vci = nullptr, *aci = nullptr, *um = nullptr, *vm = nullptr;
    double *amv = nullptr, *dyn = nullptr, *hist = nullptr, *pathv = nullptr;
    size_t path_len_ = 0;
    bool kept_path_ = false;
    hipGraphExec_t gexec_ = nullptr;   // the captured Newton iteration
    bool tuned_ = false;
    bool fp32_el_ = false;     // the element tangent of iteration() (and of the captured graph)
    void tune_band();
    long body_ops_ = 0;                // its operations (graph nodes)
    void build_jobs();
    void iteration(NewmarkArgs& A);    // one Newton iteration of every model (launches only)
};

}  // namespace cb
