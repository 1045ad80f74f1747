// C ABI over the host model, device state and drivers (include/cubens.h).
#include "../../include/cubens.h"

#include <cstdio>
#include <cstring>
#include <memory>
#include <string>

#include "driver.h"
#include "model.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

}  // namespace

struct cb_model {
    cb::Model m;
    cb::Dev d;
    std::unique_ptr<cb::Analysis> an;
    cb_options opt{};
    std::vector<double> path_lpf, path_d, step_t, step_judge;
    std::vector<int> path_ite;
    std::vector<long> step_newton;
    std::vector<std::vector<double>> step_kt_ms;
    std::vector<std::vector<long>> step_kt_n;
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
    return o;
}

cb::RunOptions to_run(const cb_options& o) {
    cb::RunOptions r;
    r.solver = o.solver;
    r.krylov_tol = o.krylov_tol;
    r.restart = o.restart;
    r.maxit = o.maxit;
    r.judge = o.judge;
    r.max_steps = o.max_steps;
    r.keep_d = o.keep_d;
    r.kernel_timing = o.kernel_timing;
    return r;
}

bool same_opts(const cb_options& a, const cb_options& b) {
    return a.solver == b.solver && a.krylov_tol == b.krylov_tol && a.restart == b.restart &&
           a.maxit == b.maxit && a.judge == b.judge && a.max_steps == b.max_steps && a.keep_d == b.keep_d &&
           a.kernel_timing == b.kernel_timing;
}

cb::Analysis& analysis(cb_model* h, const cb_options* o) {
    cb_options want = o ? *o : (h->an ? h->opt : default_options());
    if (!h->an || !same_opts(want, h->opt)) {
        h->an.reset();
        h->opt = want;
        h->an.reset(new cb::Analysis(h->m, h->d, to_run(want)));
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
    if (h->m.col_ptr_sh.size() > 65 || h->m.col_ptr_tr.size() > 65)
        return fail(-4, "element colouring exceeded 64 classes");
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

int cb_get_info(const cb_model* h, cb_info* out) {
    if (!h || !out) return fail(-1, "null argument");
    out->nj = h->m.nj; out->ne_tr = h->m.ntr; out->ne_fr = h->m.nfr; out->ne_sh = h->m.nsh;
    out->neq = h->m.neq; out->lss = h->m.lss; out->bandwidth = h->m.bandwidth;
    out->nnz_blocks = long(h->m.colidx.size());
    out->bs = h->m.bs;
    out->anaflag = h->m.anaflag; out->algflag = h->m.algflag;
    out->ncolours_tr = int(h->m.col_ptr_tr.size()) - 1;
    out->ncolours_sh = int(h->m.col_ptr_sh.size()) - 1;
    return 0;
}

int cb_get_dofmaps(const cb_model* h, long* jcode, long* mcode, long* maxa) {
    if (!h) return fail(-1, "null handle");
    if (jcode) std::memcpy(jcode, h->m.jcode.data(), sizeof(long) * h->m.jcode.size());
    if (mcode) std::memcpy(mcode, h->m.mcode.data(), sizeof(long) * h->m.mcode.size());
    if (maxa) std::memcpy(maxa, h->m.maxa.data(), sizeof(long) * h->m.maxa.size());
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
    h->step_t = an.step_t;
    h->step_judge = an.step_judge;
    h->step_newton = an.step_newton;
    h->step_kt_ms = an.step_kt_ms;
    h->step_kt_n = an.step_kt_n;
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
    if (rc) g_err = rs.msg;
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

int cb_write_results2(const cb_model* h, const char* path) {
    if (!h || !path) return fail(-1, "null argument");
    FILE* fp = std::fopen(path, "w");
    if (!fp) return fail(-1, std::string("cannot write ") + path);
    const long n = h->m.neq;
    std::fprintf(fp, "Model Displacements:\n\tLambda\t\tIterations");
    for (long i = 0; i < n; ++i) {
        if (i + 1 <= 1000) std::fprintf(fp, "\tDOF %ld\t", i + 1);
        else std::fprintf(fp, "DOF %ld\t", i + 1);
    }
    const size_t ns = h->path_lpf.size();
    for (size_t s = 0; s < ns; ++s) {
        std::fprintf(fp, "\n\t%e\t%d\t", h->path_lpf[s], h->path_ite[s]);
        if (h->path_d.size() >= (s + 1) * size_t(n))
            for (long i = 0; i < n; ++i) std::fprintf(fp, "\t%e", h->path_d[s * size_t(n) + size_t(i)]);
    }
    std::fclose(fp);
    return 0;
}

}  // extern "C"
