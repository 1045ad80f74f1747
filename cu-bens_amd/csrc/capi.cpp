// C ABI over the host model, device state and drivers (include/cubens.h).
#include "../../include/cubens.h"

#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>

#include "batch.h"
#include "driver.h"
#include "learned.h"
#include "model.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

}  // namespace

namespace cb {
// the message cb_last_error() reports, for entry points defined in other files
int set_last_error(int code, const std::string& msg) { return fail(code, msg); }
}  // namespace cb

struct cb_learned {
    std::unique_ptr<cb::Learned> net;
    // online training state
    bool training = false;
    int probes = 4;
    float lr = 1e-3f;
    uint64_t rng = 0x243F6A8885A308D3ull;
    std::vector<double> losses;
    uint64_t bound = 0;   // uid of the model the training buffers are sized for
    std::vector<double*> rb, ub;   // slot-layout pairs
    double *zq = nullptr, *uq = nullptr;   // NEQ-layout probe and its solution
    void release_buffers() {
        for (double* p : rb) (void)hipFree(p);
        for (double* p : ub) (void)hipFree(p);
        rb.clear();
        ub.clear();
        if (zq) (void)hipFree(zq);
        if (uq) (void)hipFree(uq);
        zq = uq = nullptr;
        bound = 0;
    }
    ~cb_learned() { release_buffers(); }
};

struct cb_model {
    cb::Model m;
    cb::Dev d;
    cb_learned* learned = nullptr;
    std::unique_ptr<cb::Analysis> an;
    cb_options opt{};
    std::vector<double> path_lpf, path_d, path_f, step_t, step_judge;
    std::vector<int> path_ite;
    std::vector<long> step_newton;
    std::vector<std::vector<double>> step_kt_ms;
    std::vector<std::vector<long>> step_kt_n;
    std::vector<std::array<int, cb::EV_N>> step_ev;
    std::vector<double> step_psi, step_res, step_load;
    double last_tau = 1.0;   // regula-falsi fraction of the last force pass with code 1
};

struct cb_batch {
    std::unique_ptr<cb::Batch> bt;
};

namespace {

cb_options default_options() {
    cb_options o{};
    o.solver = CB_SOLVER_BAND64;
    o.krylov_tol = 1e-8;
    o.restart = 30;
    o.maxit = 20000;
    o.judge = 1;
    o.max_steps = 1 << 30;
    o.keep_d = 1;
    o.kernel_timing = 0;
    o.factor_schedule = CB_FACTOR_FUSED;
    o.kernel_timing_every = 1;
    o.sweep_schedule = CB_SWEEP_CHAIN;
    o.keep_forces = 0;
    o.checkpoint_path = nullptr;
    o.checkpoint_exact = 0;
    o.graph = 0;
    o.refine = 0;
    o.autotune = 0;
    o.fp32_elements = 0;
    return o;
}

cb::RunOptions to_run(const cb_options& o) {
    if (o.solver < CB_SOLVER_BAND64 || o.solver > CB_SOLVER_BAND32)
        throw cb::DeviceError("unknown solver " + std::to_string(o.solver));
    cb::RunOptions r;
    r.solver = o.solver;
    r.krylov_tol = o.krylov_tol;
    r.restart = o.restart;
    r.maxit = o.maxit;
    r.judge = o.judge;
    r.max_steps = o.max_steps;
    r.keep_d = o.keep_d;
    r.kernel_timing = o.kernel_timing;
    r.factor_schedule = o.factor_schedule;
    r.kernel_timing_every = o.kernel_timing_every;
    r.sweep_schedule = o.sweep_schedule;
    r.keep_forces = o.keep_forces;
    r.checkpoint_path = o.checkpoint_path ? o.checkpoint_path : "";
    r.checkpoint_exact = o.checkpoint_exact;
    r.refine = o.refine;
    return r;
}

bool same_opts(const cb_options& a, const cb_options& b) {
    return a.solver == b.solver && a.krylov_tol == b.krylov_tol && a.restart == b.restart &&
           a.maxit == b.maxit && a.judge == b.judge && a.max_steps == b.max_steps && a.keep_d == b.keep_d &&
           a.kernel_timing == b.kernel_timing && a.factor_schedule == b.factor_schedule &&
           a.kernel_timing_every == b.kernel_timing_every && a.sweep_schedule == b.sweep_schedule &&
           a.keep_forces == b.keep_forces && a.checkpoint_exact == b.checkpoint_exact && a.graph == b.graph &&
           a.refine == b.refine && a.autotune == b.autotune && a.fp32_elements == b.fp32_elements &&
           std::string(a.checkpoint_path ? a.checkpoint_path : "") ==
               std::string(b.checkpoint_path ? b.checkpoint_path : "");
}

double normal01(uint64_t& st) {
    auto u01 = [&]() {
        st += 0x9E3779B97F4A7C15ull;
        uint64_t z = st;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double(z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    };
    const double a = u01(), b = u01();
    return std::sqrt(-2.0 * std::log(a)) * std::cos(6.283185307179586 * b);
}

// slot-layout buffers of the training batch for model h
void training_buffers(cb_learned* L, cb_model* h) {
    if (L->bound == h->d.uid && int(L->rb.size()) == L->probes + 1) return;
    L->release_buffers();
    for (int b = 0; b <= L->probes; ++b) {
        L->rb.push_back(cb::dalloc<double>(size_t(h->d.nslot)));
        L->ub.push_back(cb::dalloc<double>(size_t(h->d.nslot)));
    }
    L->zq = cb::dalloc<double>(size_t(h->m.neq));
    L->uq = cb::dalloc<double>(size_t(h->m.neq));
    L->bound = h->d.uid;
}

// One online training step on a (rhs, solution) pair of the FP64 band solve
// plus random probes solved with the same factor.
void train_on_pair(cb_model* h, cb::Dev& d, const double* rhs, const double* sol) {
    cb_learned* L = h->learned;
    training_buffers(L, h);
    cb::Learned& net = *L->net;
    net.setup(d);
    cb::to_slots(d, L->rb[0], rhs);
    cb::to_slots(d, L->ub[0], sol);
    std::vector<double> z(size_t(h->m.neq));
    for (int b = 1; b <= L->probes; ++b) {
        for (double& v : z) v = normal01(L->rng);
        cb::h2d(L->zq, z.data(), z.size(), d.s);
        h->an->band64().solve(L->zq, L->uq);
        cb::to_slots(d, L->rb[size_t(b)], L->zq);
        cb::to_slots(d, L->ub[size_t(b)], L->uq);
    }
    std::vector<const double*> r(L->rb.begin(), L->rb.end()), u(L->ub.begin(), L->ub.end());
    const double loss = net.loss_grad(d, r.data(), u.data(), L->probes + 1);
    net.adam_step(L->lr);
    L->losses.push_back(loss);
}

void install_hooks(cb_model* h) {
    cb::PrecHook hk;
    if (cb_learned* L = h->learned) {
        hk.setup = [L](cb::Dev& d) { L->net->setup(d); };
        hk.apply = [L](cb::Dev& d, const double* v, double* z) { L->net->apply(d, v, z); };
        if (L->training)
            hk.observe = [h](cb::Dev& d, const double* rhs, const double* sol) { train_on_pair(h, d, rhs, sol); };
    }
    h->an->learned = hk;
}

cb::Analysis& analysis(cb_model* h, const cb_options* o) {
    cb_options want = o ? *o : (h->an ? h->opt : default_options());
    if (!h->an || !same_opts(want, h->opt)) {
        h->an.reset();
        h->opt = want;
        h->an.reset(new cb::Analysis(h->m, h->d, to_run(want)));
        install_hooks(h);
    }
    return *h->an;
}

// name -> (device pointer, element count)
bool state_ref(cb_model* h, const std::string& name, double** p, long* n) {
    cb::Dev& d = h->d;
    const long ntri = d.ntr + d.nfr * 3 + d.nsh * 3, nef = d.ntr * 2 + d.nfr * 14 + d.nsh * 18;
    std::string base = name;
    int g = cb::G_BASE;
    if (name == "x_temp") { *p = d.x[cb::G_I]; *n = d.nj * 3; return true; }
    if (name == "x_ip") { *p = d.x[cb::G_IP]; *n = d.nj * 3; return true; }
    if (name == "x") { *p = d.x[cb::G_BASE]; *n = d.nj * 3; return true; }
    // working element geometry the kernels read (the mass routines rewrite it)
    if (name == "llength") { *p = d.llen; *n = d.ntr + d.nfr; return true; }
    if (name == "farea") { *p = d.farea; *n = d.nsh; return true; }
    if (name == "slength") { *p = d.slength; *n = d.nsh * 3; return true; }
    if (base.size() > 3 && base.compare(base.size() - 3, 3, "_ip") == 0) {
        g = cb::G_IP;
        base = base.substr(0, base.size() - 3);
    } else if (base.size() > 2 && base.compare(base.size() - 2, 2, "_i") == 0) {
        g = cb::G_I;
        base = base.substr(0, base.size() - 2);
    }
    if (base == "c1") { *p = d.c1[g]; *n = ntri; return true; }
    if (base == "c2") { *p = d.c2[g]; *n = ntri; return true; }
    if (base == "c3") { *p = d.c3[g]; *n = ntri; return true; }
    if (base == "ef") { *p = d.ef[g]; *n = nef; return true; }
    if (base == "defllen") { *p = d.defllen[g]; *n = d.ntr + d.nfr; return true; }
    if (base == "deffarea") { *p = d.deffarea[g]; *n = d.nsh; return true; }
    if (base == "defslen") { *p = d.defslen[g]; *n = d.nsh * 3; return true; }
    return false;
}

#define GUARD_BEGIN try {
#define GUARD_END                                                          \
    }                                                                      \
    catch (const cb::DeviceError& e) { return fail(-2, e.msg); }           \
    catch (const std::exception& e) { return fail(-3, e.what()); }

}  // namespace

extern "C" {

const char* cb_last_error(void) { return g_err.c_str(); }

int cb_open(const char* path, cb_model** out) {
    if (!out) return fail(-1, "null output handle");
    *out = nullptr;
    GUARD_BEGIN
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(-5, "no HIP device available: the CU-BENs GPU path needs an MI355X (gfx950)");
    std::unique_ptr<cb_model> h(new cb_model());
    std::string err;
    if (!cb::load_model(path, h->m, err)) return fail(-4, err);
    cb::build_pattern(h->m);
    cb::dev_create(h->d, h->m);
    h->opt = default_options();
    *out = h.release();
    return 0;
    GUARD_END
}

void cb_close(cb_model* h) {
    if (!h) return;
    try {
        h->an.reset();
        cb::dev_destroy(h->d);
    } catch (...) {
    }
    delete h;
}

static void fill_info(const cb::Model& m, cb_info* out) {
    out->nj = m.nj; out->ne_tr = m.ntr; out->ne_fr = m.nfr; out->ne_sh = m.nsh;
    out->neq = m.neq; out->lss = m.lss; out->bandwidth = m.bandwidth;
    out->nnz_blocks = long(m.colidx.size());
    out->bs = m.bs;
    out->anaflag = m.anaflag; out->algflag = m.algflag;
    out->ncolours_tr = int(m.col_ptr_tr.size()) - 1;
    out->ncolours_sh = int(m.col_ptr_sh.size()) - 1;
}

int cb_get_info(const cb_model* h, cb_info* out) {
    if (!h || !out) return fail(-1, "null argument");
    fill_info(h->m, out);
    return 0;
}

int cb_model_info(const char* path, cb_info* info, long* jcode, double* x, long* jnt) {
    if (!path) return fail(-1, "null path");
    GUARD_BEGIN
    cb::Model m;
    std::string err;
    if (!cb::load_model(path, m, err)) return fail(-4, err);
    cb::build_pattern(m);
    if (info) fill_info(m, info);
    if (jcode) std::memcpy(jcode, m.jcode.data(), sizeof(long) * m.jcode.size());
    if (x) std::memcpy(x, m.x.data(), sizeof(double) * m.x.size());
    if (jnt) std::memcpy(jnt, m.jnt.data(), sizeof(long) * m.jnt.size());
    return 0;
    GUARD_END
}

int cb_get_dofmaps(const cb_model* h, long* jcode, long* mcode, long* maxa) {
    if (!h) return fail(-1, "null handle");
    if (jcode) std::memcpy(jcode, h->m.jcode.data(), sizeof(long) * h->m.jcode.size());
    if (mcode) std::memcpy(mcode, h->m.mcode.data(), sizeof(long) * h->m.mcode.size());
    if (maxa) std::memcpy(maxa, h->m.maxa.data(), sizeof(long) * h->m.maxa.size());
    return 0;
}

int cb_get_joint_numbers(const cb_model* h, long* jnt) {
    if (!h || !jnt) return fail(-1, "null argument");
    std::memcpy(jnt, h->m.jnt.data(), sizeof(long) * h->m.jnt.size());
    return 0;
}

int cb_get_pattern(const cb_model* h, int32_t* rowptr, int32_t* colidx) {
    if (!h) return fail(-1, "null handle");
    if (rowptr) std::memcpy(rowptr, h->m.rowptr.data(), sizeof(int32_t) * h->m.rowptr.size());
    if (colidx) std::memcpy(colidx, h->m.colidx.data(), sizeof(int32_t) * h->m.colidx.size());
    return 0;
}

int cb_get_load(const cb_model* h, double* q) {
    if (!h || !q) return fail(-1, "null argument");
    std::memcpy(q, h->m.q.data(), sizeof(double) * size_t(h->m.neq));
    return 0;
}

int cb_state_set(cb_model* h, const char* name, const double* host, long n) {
    if (!h || !name || !host) return fail(-1, "null argument");
    GUARD_BEGIN
    double* p;
    long cnt;
    if (!state_ref(h, name, &p, &cnt)) return fail(-1, std::string("unknown state array ") + name);
    if (n != cnt) return fail(-1, std::string("size mismatch for ") + name);
    cb::h2d(p, host, size_t(n), h->d.s);
    CB_CHECK(hipStreamSynchronize(h->d.s));
    return 0;
    GUARD_END
}

int cb_state_get(cb_model* h, const char* name, double* host, long n) {
    if (!h || !name || !host) return fail(-1, "null argument");
    GUARD_BEGIN
    double* p;
    long cnt;
    if (!state_ref(h, name, &p, &cnt)) return fail(-1, std::string("unknown state array ") + name);
    if (n != cnt) return fail(-1, std::string("size mismatch for ") + name);
    cb::d2h(host, p, size_t(n), h->d.s);
    CB_CHECK(hipStreamSynchronize(h->d.s));
    return 0;
    GUARD_END
}

int cb_tangent(cb_model* h, int gen_x, int gen_state) {
    if (!h) return fail(-1, "null handle");
    if (gen_x < 0 || gen_x > 2 || gen_state < 0 || gen_state > 2) return fail(-1, "bad generation");
    GUARD_BEGIN
    cb::launch_tangent(h->d, gen_x, gen_state, false);
    CB_CHECK(hipStreamSynchronize(h->d.s));
    return 0;
    GUARD_END
}

int cb_get_tangent(cb_model* h, double* vals) {
    if (!h || !vals) return fail(-1, "null argument");
    GUARD_BEGIN
    cb::d2h(vals, h->d.K, size_t(h->d.nnzb * h->d.bs * h->d.bs), h->d.s);
    CB_CHECK(hipStreamSynchronize(h->d.s));
    return 0;
    GUARD_END
}

int cb_tangent_fp32(cb_model* h, int gen_x, int gen_state, float* vals) {
    if (!h || !vals) return fail(-1, "null argument");
    if (gen_x < 0 || gen_x > 2 || gen_state < 0 || gen_state > 2) return fail(-1, "bad generation");
    GUARD_BEGIN
    cb::launch_tangent_fp32(h->d, gen_x, gen_state);
    cb::d2h(vals, h->d.K32, size_t(h->d.nnzb * h->d.bs * h->d.bs), h->d.s);
    CB_CHECK(hipStreamSynchronize(h->d.s));
    return 0;
    GUARD_END
}

int cb_internal_force(cb_model* h, int gen_x, const double* dd_host, double* f_host) {
    if (!h || !dd_host || !f_host) return fail(-1, "null argument");
    if (gen_x < 0 || gen_x > 2) return fail(-1, "bad generation");
    GUARD_BEGIN
    const long n = h->m.neq;
    double* dd = cb::dalloc<double>(size_t(n));
    double* f = cb::dalloc<double>(size_t(n));
    cb::h2d(dd, dd_host, size_t(n), h->d.s);
    CB_CHECK(hipMemsetAsync(f, 0, sizeof(double) * size_t(n), h->d.s));
    cb::launch_forces(h->d, gen_x, f, dd);
    cb::d2h(f_host, f, size_t(n), h->d.s);
    CB_CHECK(hipStreamSynchronize(h->d.s));
    (void)hipFree(dd);
    (void)hipFree(f);
    double tau = 1.0;
    const int code = cb::frame_plastic_result(h->d, &tau);
    h->last_tau = tau;
    return code;
    GUARD_END
}

int cb_get_yield_flags(cb_model* h, int* yld, double* tau) {
    if (!h) return fail(-1, "null handle");
    GUARD_BEGIN
    if (yld && h->m.nfr) {
        cb::d2h(yld, h->d.yld, size_t(h->m.nfr * 2), h->d.s);
        CB_CHECK(hipStreamSynchronize(h->d.s));
    }
    if (tau) *tau = h->last_tau;
    return 0;
    GUARD_END
}

int cb_set_yield_flags(cb_model* h, const int* yld) {
    if (!h || !yld) return fail(-1, "null argument");
    GUARD_BEGIN
    for (long i = 0; i < h->m.nfr * 2; ++i)
        if (yld[i] < 0 || yld[i] > 2) return fail(-1, "yield flags are 0, 1 or 2");
    if (h->m.nfr) {
        cb::h2d(h->d.yld, yld, size_t(h->m.nfr * 2), h->d.s);
        CB_CHECK(hipStreamSynchronize(h->d.s));
    }
    return 0;
    GUARD_END
}

int cb_mass(cb_model* h, double* sm_host) {
    if (!h || !sm_host) return fail(-1, "null argument");
    GUARD_BEGIN
    const long n = h->m.neq;
    double* smv = cb::dalloc<double>(size_t(n));
    try {
        analysis(h, nullptr).mass(smv);
        cb::d2h(sm_host, smv, size_t(n), h->d.s);
        CB_CHECK(hipStreamSynchronize(h->d.s));
    } catch (...) {
        (void)hipFree(smv);
        throw;
    }
    (void)hipFree(smv);
    return 0;
    GUARD_END
}

int cb_update_geometry(cb_model* h, int gen_x, const double* dd_host) {
    if (!h || !dd_host) return fail(-1, "null argument");
    if (gen_x < 0 || gen_x > 2) return fail(-1, "bad generation");
    GUARD_BEGIN
    const long n = h->m.neq;
    double* dd = cb::dalloc<double>(size_t(n));
    cb::h2d(dd, dd_host, size_t(n), h->d.s);
    cb::launch_updatc(h->d, gen_x, dd);
    CB_CHECK(hipStreamSynchronize(h->d.s));
    (void)hipFree(dd);
    return 0;
    GUARD_END
}

int cb_factor(cb_model* h, const cb_options* opt, int* det, double* pivots) {
    if (!h) return fail(-1, "null handle");
    GUARD_BEGIN
    cb::Analysis& an = analysis(h, opt);
    int dt = 0;
    int rc = an.factor_only(&dt);
    if (det) *det = dt;
    if (pivots && an.pivots.size() >= size_t(h->m.neq))
        std::memcpy(pivots, an.pivots.data(), sizeof(double) * size_t(h->m.neq));
    if (rc) return fail(rc, "singular or out-of-band tangent");
    return 0;
    GUARD_END
}

int cb_solve(cb_model* h, const double* rhs_host, double* x_host) {
    if (!h || !rhs_host || !x_host) return fail(-1, "null argument");
    GUARD_BEGIN
    cb::Analysis& an = analysis(h, nullptr);
    const long n = h->m.neq;
    double* b = cb::dalloc<double>(size_t(n));
    double* x = cb::dalloc<double>(size_t(n));
    cb::h2d(b, rhs_host, size_t(n), h->d.s);
    int rc = an.linear_solve(b, x);
    cb::d2h(x_host, x, size_t(n), h->d.s);
    CB_CHECK(hipStreamSynchronize(h->d.s));
    (void)hipFree(b);
    (void)hipFree(x);
    if (rc) return fail(rc, "iterative solve did not reach its tolerance");
    return 0;
    GUARD_END
}

int cb_run(cb_model* h, const cb_options* opt, cb_stats* st) {
    if (!h) return fail(-1, "null handle");
    GUARD_BEGIN
    cb::Analysis& an = analysis(h, opt);
    cb::RunStats rs;
    int rc = an.run(rs);
    h->path_lpf = an.lpf;
    h->path_ite = an.itecnt;
    h->path_d = an.dpath;
    h->path_f = an.fpath;
    h->step_t = an.step_t;
    h->step_judge = an.step_judge;
    h->step_newton = an.step_newton;
    h->step_kt_ms = an.step_kt_ms;
    h->step_kt_n = an.step_kt_n;
    h->step_ev = an.step_ev;
    h->step_psi = an.step_psi;
    h->step_res = an.step_res;
    h->step_load = an.step_load;
    if (st) {
        st->t_total = rs.t_total; st->t_tangent = rs.t_tangent; st->t_factor = rs.t_factor;
        st->t_solve = rs.t_solve; st->t_forces = rs.t_forces; st->t_judge = rs.t_judge;
        st->t_other = rs.t_other;
        st->newton = rs.newton; st->solves = rs.solves; st->factors = rs.factors;
        st->krylov_iters = rs.krylov_iters; st->krylov_fallbacks = rs.krylov_fallbacks;
        st->worst_relres = rs.worst_relres;
        st->status = rc;
        st->nsteps = int(an.lpf.size());
    }
    if (rc == 4) return fail(-2, rs.msg);
    if (rc) return fail(rc, rs.msg);   // the driver stopped; its path so far is kept
    return 0;
    GUARD_END
}

int cb_get_path(const cb_model* h, double* lpf, int* itecnt, double* d) {
    if (!h) return fail(-1, "null handle");
    if (lpf) std::memcpy(lpf, h->path_lpf.data(), sizeof(double) * h->path_lpf.size());
    if (itecnt) std::memcpy(itecnt, h->path_ite.data(), sizeof(int) * h->path_ite.size());
    if (d) std::memcpy(d, h->path_d.data(), sizeof(double) * h->path_d.size());
    return 0;
}

int cb_get_step_stats(const cb_model* h, double* t_wall, double* t_judge, long* newton) {
    if (!h) return fail(-1, "null handle");
    const size_t n = h->step_t.size();
    if (t_wall) std::memcpy(t_wall, h->step_t.data(), sizeof(double) * n);
    if (t_judge) std::memcpy(t_judge, h->step_judge.data(), sizeof(double) * n);
    if (newton) std::memcpy(newton, h->step_newton.data(), sizeof(long) * n);
    return 0;
}

int cb_get_reactions(cb_model* h, int gen, double* r) {
    if (!h || !r) return fail(-1, "null argument");
    if (gen < cb::G_BASE || gen > cb::G_IP) return fail(-1, "generation must be 0, 1 or 2");
    GUARD_BEGIN
    const size_t n = size_t(h->m.nj) * 7;
    double* rx = cb::dalloc<double>(n);
    cb::launch_reactions(h->d, gen, rx);
    cb::d2h(r, rx, n, h->d.s);
    CB_CHECK(hipStreamSynchronize(h->d.s));
    (void)hipFree(rx);
    return 0;
    GUARD_END
}

int cb_dots(int n, const double* const* a, const double* const* b, long len, double* out) {
    if (n < 1 || n > cb::DOTS_MAX) return fail(-1, "cb_dots: 1 to 8 products per call");
    if (!a || !b || !out || len < 0) return fail(-1, "null argument or negative length");
    for (int k = 0; k < n; ++k)
        if (len > 0 && (!a[k] || !b[k])) return fail(-1, "null vector");
    GUARD_BEGIN
    double* buf = cb::dalloc<double>(size_t(cb::DOTS_PARTS + cb::DOTS_MAX));
    cb::launch_dots(nullptr, n, a, b, len, buf, buf + cb::DOTS_PARTS);
    cb::d2h(out, buf + cb::DOTS_PARTS, size_t(n), nullptr);
    CB_CHECK(hipStreamSynchronize(nullptr));
    (void)hipFree(buf);
    return 0;
    GUARD_END
}

int cb_get_step_events(const cb_model* h, int* events, double* psi) {
    if (!h) return fail(-1, "null handle");
    const size_t n = h->step_ev.size();
    if (events)
        for (size_t s = 0; s < n; ++s)
            std::memcpy(events + s * cb::EV_N, h->step_ev[s].data(), sizeof(int) * cb::EV_N);
    if (psi) std::memcpy(psi, h->step_psi.data(), sizeof(double) * n);
    return 0;
}

int cb_get_step_residuals(const cb_model* h, double* res, double* load) {
    if (!h) return fail(-1, "null handle");
    const size_t n = h->step_res.size();
    if (res) std::memcpy(res, h->step_res.data(), sizeof(double) * n);
    if (load) std::memcpy(load, h->step_load.data(), sizeof(double) * n);
    return 0;
}

int cb_get_kernel_times(const cb_model* h, int precision, int step, double* ms, long* launches) {
    if (!h) return fail(-1, "null handle");
    if (!h->an) return fail(-1, "no analysis has run");
    if (step >= 0) {
        if (size_t(step) >= h->step_kt_ms.size()) return fail(-1, "no such step");
        for (int f = 0; f < cb::KT_N; ++f) {
            if (ms) ms[f] = h->step_kt_ms[size_t(step)][size_t(f)];
            if (launches) launches[f] = h->step_kt_n[size_t(step)][size_t(f)];
        }
        return 0;
    }
    const cb::KernelClock& c = precision == 32 ? h->an->band32().clock : h->an->band64().clock;
    for (int f = 0; f < cb::KT_N; ++f) {
        if (ms) ms[f] = c.ms[f];
        if (launches) launches[f] = c.launches[f];
    }
    return 0;
}

namespace {

// The rows output() (misc.c:345-451) writes over a run: headers (flag 0),
// then one row per converged step (flag 1), to results2.txt (f[0]) and
// results3-5.txt (f[1..3], any of them may be null when not wanted).
void write_results(const cb_model* h, FILE* const f[4]) {
    const cb::Model& m = h->m;
    const long n = m.neq, ntr = m.ntr, nfr = m.nfr, nsh = m.nsh;
    const size_t ns = h->path_lpf.size(), nf = size_t(ntr + 7 * nfr + 6 * nsh);
    if (f[0]) {
        std::fprintf(f[0], "Model Displacements:\n\tLambda\t\tIterations");
        for (long i = 0; i < n; ++i) {
            if (i + 1 <= 1000) std::fprintf(f[0], "\tDOF %ld\t", i + 1);
            else std::fprintf(f[0], "DOF %ld\t", i + 1);
        }
    }
    if (ntr > 0 && f[1]) {
        std::fprintf(f[1], "Truss Element Forces (averaged):\n\tLambda\t\tIterations\t");
        for (long i = 0; i < ntr; ++i) std::fprintf(f[1], "Element %ld\t", i + 1);
    }
    if (nfr > 0) {
        if (f[2]) std::fprintf(f[2], "Frame Element Forces (averaged):\n\tLambda\t\tIterations");
        if (f[1]) std::fprintf(f[1], "\t");   // output() writes this tab to results3.txt
        if (f[2]) {
            for (long i = 0; i < nfr; ++i) std::fprintf(f[2], "Element %ld\t\t\t\t\t\t\t\t\t\t\t\t\t", i + 1);
            std::fprintf(f[2], "\n\t\t\t\t\t");
            for (long i = 0; i < nfr; ++i)
                std::fprintf(f[2], "X-Force\t\tY-Force\t\tZ-Force\t\tX-Moment\tY-Moment\tZ-Moment\tBi-Moment\t");
        }
    }
    if (nsh > 0) {
        if (f[3]) std::fprintf(f[3], "Shell Element Forces (averaged):\n\tLambda\t\tIterations");
        if (f[1]) std::fprintf(f[1], "\t");   // likewise
        if (f[3]) {
            for (long i = 0; i < nsh; ++i) std::fprintf(f[3], "Element %ld\t\t\t\t\t\t\t\t\t\t\t", i + 1);
            std::fprintf(f[3], "\n\t\t\t\t\t");
            for (long i = 0; i < nsh; ++i)
                std::fprintf(f[3], "X-Force\t\tY-Force\t\tZ-Force\t\tX-Moment\tY-Moment\tZ-Moment\t");
        }
    }
    for (size_t s = 0; s < ns; ++s) {
        const double lam = h->path_lpf[s];
        const int ite = h->path_ite[s];
        if (f[0]) {
            std::fprintf(f[0], "\n\t%e\t%d\t", lam, ite);
            for (long i = 0; i < n; ++i) std::fprintf(f[0], "\t%e", h->path_d[s * size_t(n) + size_t(i)]);
        }
        const double* fa = h->path_f.size() >= (s + 1) * nf ? h->path_f.data() + s * nf : nullptr;
        if (!fa) continue;
        if (ntr > 0 && f[1]) {
            std::fprintf(f[1], "\n\t%e\t%d\t", lam, ite);
            for (long i = 0; i < ntr; ++i) std::fprintf(f[1], "\t%e", fa[i]);
        }
        if (nfr > 0 && f[2]) {
            std::fprintf(f[2], "\n\t%e\t%d\t", lam, ite);
            for (long i = 0; i < 7 * nfr; ++i) std::fprintf(f[2], "\t%e", fa[ntr + i]);
        }
        if (nsh > 0 && f[3]) {
            std::fprintf(f[3], "\n\t%e\t%d\t", lam, ite);
            for (long i = 0; i < 6 * nsh; ++i) std::fprintf(f[3], "\t%e", fa[ntr + 7 * nfr + i]);
        }
    }
}

}  // namespace

int cb_write_results2(const cb_model* h, const char* path) {
    if (!h || !path) return fail(-1, "null argument");
    const size_t ns = h->path_lpf.size();
    if (h->path_d.size() < ns * size_t(h->m.neq))
        return fail(-1, "the last run kept no displacement path (keep_d = 0)");
    FILE* fp = std::fopen(path, "w");
    if (!fp) return fail(-1, std::string("cannot write ") + path);
    FILE* const f[4] = {fp, nullptr, nullptr, nullptr};
    write_results(h, f);
    std::fclose(fp);
    return 0;
}

int cb_get_path_forces(const cb_model* h, double* forces) {
    if (!h || !forces) return fail(-1, "null argument");
    std::memcpy(forces, h->path_f.data(), sizeof(double) * h->path_f.size());
    return 0;
}

int cb_write_results(const cb_model* h, const char* dir) {
    if (!h || !dir) return fail(-1, "null argument");
    const cb::Model& m = h->m;
    const size_t ns = h->path_lpf.size(), nf = size_t(m.ntr + 7 * m.nfr + 6 * m.nsh);
    if (h->path_d.size() < ns * size_t(m.neq))
        return fail(-1, "the last run kept no displacement path (keep_d = 0)");
    if (nf > 0 && h->path_f.size() < ns * nf)
        return fail(-1, "the last run kept no element forces (keep_forces = 0)");
    FILE* f[4] = {nullptr, nullptr, nullptr, nullptr};
    for (int k = 0; k < 4; ++k) {
        const std::string p = std::string(dir) + "/results" + std::to_string(k + 2) + ".txt";
        f[k] = std::fopen(p.c_str(), "w");
        if (!f[k]) {
            for (int j = 0; j < k; ++j) std::fclose(f[j]);
            return fail(-1, "cannot write " + p);
        }
    }
    write_results(h, f);
    for (FILE* fp : f) std::fclose(fp);
    return 0;
}


/* ---- learned preconditioner ---------------------------------------------- */
int cb_learned_create(unsigned seed, cb_learned** out) {
    if (!out) return fail(-1, "null argument");
    GUARD_BEGIN
    int nd = 0;
    if (hipGetDeviceCount(&nd) != hipSuccess || nd < 1)
        return fail(-2, "no HIP device available (the learned preconditioner runs on the GPU only)");
    auto* L = new cb_learned();
    L->net.reset(new cb::Learned(seed));
    *out = L;
    return 0;
    GUARD_END
}

void cb_learned_free(cb_learned* L) { delete L; }

long cb_learned_nparams(const cb_learned* L) { return L ? L->net->nparams() : -1; }

int cb_learned_get_params(const cb_learned* L, float* p) {
    if (!L || !p) return fail(-1, "null argument");
    GUARD_BEGIN
    L->net->get_params(p);
    return 0;
    GUARD_END
}

int cb_learned_set_params(cb_learned* L, const float* p) {
    if (!L || !p) return fail(-1, "null argument");
    GUARD_BEGIN
    L->net->set_params(p);
    return 0;
    GUARD_END
}

int cb_learned_set_objective(cb_learned* L, int kind) {
    if (!L) return fail(-1, "null argument");
    if (kind < 0 || kind > 2) return fail(-1, "objective must be 0, 1 or 2");
    L->net->set_objective(kind);
    return 0;
}

int cb_learned_attach(cb_model* h, cb_learned* L) {
    if (!h) return fail(-1, "null handle");
    h->learned = L;
    if (h->an) install_hooks(h);
    return 0;
}

int cb_learned_train(cb_learned* L, cb_model* h, int max_steps, int probes, float lr, cb_train_stats* st) {
    if (!L || !h) return fail(-1, "null argument");
    if (probes < 0 || probes > 15) return fail(-1, "probes must be 0..15");
    GUARD_BEGIN
    cb_learned* keep = h->learned;
    h->learned = L;
    L->training = true;
    L->probes = probes;
    L->lr = lr;
    L->losses.clear();
    cb_options o = default_options();
    o.solver = CB_SOLVER_BAND64;
    o.judge = 0;
    o.max_steps = max_steps;
    o.keep_d = 0;
    h->an.reset();
    int rc = 0;
    std::string msg;
    try {
        cb::Analysis& an = analysis(h, &o);
        cb::RunStats rs;
        rc = an.run(rs);
        msg = rs.msg;
    } catch (...) {
        L->training = false;
        h->learned = keep;
        h->an.reset();
        throw;
    }
    L->training = false;
    h->learned = keep;
    h->an.reset();
    if (st) {
        st->steps = long(L->losses.size());
        st->loss_first = L->losses.empty() ? 0 : L->losses.front();
        st->loss_last = L->losses.empty() ? 0 : L->losses.back();
        double m = 0;
        const size_t k = std::min<size_t>(L->losses.size(), 10);
        for (size_t i = L->losses.size() - k; i < L->losses.size(); ++i) m += L->losses[i];
        st->loss_recent = k ? m / double(k) : 0;
    }
    if (rc) return fail(rc, "training analysis stopped: " + msg);
    return 0;
    GUARD_END
}

int cb_learned_loss_grad(cb_learned* L, cb_model* h, const double* r_host, const double* u_host, int nvec,
                         double* loss, float* grad) {
    if (!L || !h || !r_host || !u_host) return fail(-1, "null argument");
    if (nvec < 1 || nvec > 16) return fail(-1, "nvec must be 1..16");
    GUARD_BEGIN
    cb::Dev& d = h->d;
    const long n = h->m.neq;
    cb::Learned& net = *L->net;
    net.setup(d);
    std::vector<double*> rs, us;
    double* tq = cb::dalloc<double>(size_t(n));
    for (int b = 0; b < nvec; ++b) {
        rs.push_back(cb::dalloc<double>(size_t(d.nslot)));
        us.push_back(cb::dalloc<double>(size_t(d.nslot)));
        cb::h2d(tq, r_host + long(b) * n, size_t(n), d.s);
        cb::to_slots(d, rs.back(), tq);
        cb::h2d(tq, u_host + long(b) * n, size_t(n), d.s);
        cb::to_slots(d, us.back(), tq);
        CB_CHECK(hipStreamSynchronize(d.s));
    }
    std::vector<const double*> rc(rs.begin(), rs.end()), uc(us.begin(), us.end());
    const double l = net.loss_grad(d, rc.data(), uc.data(), nvec);
    if (loss) *loss = l;
    if (grad) net.get_grad(grad);
    for (double* p : rs) (void)hipFree(p);
    for (double* p : us) (void)hipFree(p);
    (void)hipFree(tq);
    return 0;
    GUARD_END
}

int cb_learned_blocks(cb_learned* L, cb_model* h, float* gd, float* go, long* nv, long* ne) {
    if (!L || !h) return fail(-1, "null argument");
    GUARD_BEGIN
    L->net->setup(h->d);
    if (nv) *nv = L->net->nv();
    if (ne) *ne = L->net->ne();
    L->net->get_blocks(gd, go);
    return 0;
    GUARD_END
}

int cb_batch_open(const char* const* paths, int count, cb_batch** out) {
    if (!out || !paths || count < 1) return fail(-1, "null argument or empty batch");
    *out = nullptr;
    GUARD_BEGIN
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(-5, "no HIP device available: the CU-BENs GPU path needs an MI355X (gfx950)");
    std::vector<std::string> ps;
    for (int i = 0; i < count; ++i) {
        if (!paths[i]) return fail(-1, "null model path");
        ps.emplace_back(paths[i]);
    }
    std::unique_ptr<cb_batch> h(new cb_batch());
    try {
        h->bt.reset(new cb::Batch(ps));
    } catch (const std::runtime_error& e) {
        return fail(-4, e.what());
    }
    *out = h.release();
    return 0;
    GUARD_END
}

void cb_batch_close(cb_batch* b) {
    if (!b) return;
    try {
        b->bt.reset();
    } catch (...) {
    }
    delete b;
}

int cb_batch_count(const cb_batch* b) { return (b && b->bt) ? b->bt->count() : 0; }

int cb_batch_get_info(const cb_batch* b, int i, cb_info* out) {
    if (!b || !b->bt || !out) return fail(-1, "null argument");
    if (i < 0 || i >= b->bt->count()) return fail(-1, "model index out of range");
    fill_info(b->bt->model(i), out);
    return 0;
}

int cb_batch_run(cb_batch* b, const cb_options* opt, cb_batch_stats* st) {
    if (!b || !b->bt) return fail(-1, "null handle");
    const cb_options o = opt ? *opt : default_options();
    if (o.solver != CB_SOLVER_BAND64) return fail(-1, "a batch runs the FP64 band solver only");
    GUARD_BEGIN
    cb::BatchStats bs;
    b->bt->run(o.max_steps, o.keep_d != 0, bs, o.kernel_timing != 0, o.graph != 0, o.autotune != 0,
               o.fp32_elements != 0);
    if (st) {
        st->pairs = bs.pairs;
        st->pairs_max = bs.pairs_max;
        st->tune_s = bs.tune_s;
        st->ops_per_round = bs.ops_per_round;
        st->body_ops = bs.body_ops;
        st->t_total = bs.t_total;
        st->rounds = bs.rounds;
        st->newton = bs.newton;
        st->steps = bs.steps;
        st->failed = bs.failed;
        st->ms_factor = bs.ms_factor;
        st->ms_solve = bs.ms_solve;
        st->factor_flops = bs.factor_flops;
        st->solve_bytes = bs.solve_bytes;
    }
    return 0;
    GUARD_END
}

int cb_batch_get_tuning(const cb_batch* b, int* pairs, double* ms, int cap) {
    if (!b || !b->bt) return fail(-1, "null handle");
    const auto& t = b->bt->tuning;
    const int n = int(t.size());
    for (int i = 0; i < n && i < cap; ++i) {
        if (pairs) pairs[i] = t[size_t(i)].first;
        if (ms) ms[i] = t[size_t(i)].second;
    }
    return n;
}

int cb_batch_get_status(const cb_batch* b, int* status, int* nsteps) {
    if (!b || !b->bt) return fail(-1, "null handle");
    const cb::Batch& bt = *b->bt;
    if (bt.status.size() != size_t(bt.count())) return fail(-1, "the batch has not run");
    for (int i = 0; i < bt.count(); ++i) {
        if (status) status[i] = bt.status[size_t(i)];
        if (nsteps) nsteps[i] = int(bt.times[size_t(i)].size());
    }
    return 0;
}

int cb_batch_get_path(const cb_batch* b, int i, double* t, int* itecnt, double* d) {
    if (!b || !b->bt) return fail(-1, "null handle");
    const cb::Batch& bt = *b->bt;
    if (bt.status.size() != size_t(bt.count())) return fail(-1, "the batch has not run");
    if (i < 0 || i >= bt.count()) return fail(-1, "model index out of range");
    GUARD_BEGIN
    const auto& tv = bt.times[size_t(i)];
    const auto& iv = bt.itecnt[size_t(i)];
    if (t) std::memcpy(t, tv.data(), sizeof(double) * tv.size());
    if (itecnt) std::memcpy(itecnt, iv.data(), sizeof(int) * iv.size());
    if (d && !tv.empty()) bt.path(i, d);
    return 0;
    GUARD_END
}

}  // extern "C"
