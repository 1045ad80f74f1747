// Nonlinear drivers over device state.  The control flow restates the
// reference's drivers (NR/MNR main.c:1800-2157, MSAL main.c:2158-3141,
// arc.c:70-157, misc.c:187-250); all vectors and element state stay on the
// device, the host sees scalars (dots, pivots summaries, flags).
#include "driver.h"

#include <algorithm>
#include <functional>

#include "checkpoint.h"
#include <chrono>
#include <cmath>
#include <cstring>

namespace cb {
namespace {

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct Timer {
    double* acc;
    hipStream_t s;
    double t0;
    Timer(double* a, hipStream_t st) : acc(a), s(st), t0(now_s()) {}
    ~Timer() {
        (void)hipStreamSynchronize(s);
        *acc += now_s() - t0;
    }
};

}  // namespace

Analysis::Analysis(Model& m, Dev& d, const RunOptions& o) : m_(m), d_(d), o_(o) {
    const long n = m.neq;
    double** vs[] = {&q, &dv, &dp, &dpp, &dt, &dd, &ddq, &ddr, &f, &fp, &ft, &fip, &r, &tmp};
    for (double** p : vs) {
        *p = dalloc<double>(size_t(n));
        CB_CHECK(hipMemsetAsync(*p, 0, sizeof(double) * size_t(n), d.s));
    }
    eb = dalloc<double>(size_t(d.nslot));
    ex = dalloc<double>(size_t(d.nslot));
    if (o.keep_forces) favg = dalloc<double>(size_t(m.ntr + 7 * m.nfr + 6 * m.nsh));
    if (o.solver == SOLVER_BAND32 && o.refine > 0) rf = dalloc<double>(size_t(n));
    h2d(q, m.q.data(), size_t(n), d.s);
    const bool b32 = o.solver == SOLVER_BAND32_FGMRES || o.solver == SOLVER_BAND32;
    const bool need64 = o.solver == SOLVER_BAND64 || o.judge || !b32;
    if (need64) b64_.setup(n, m.bandwidth, d.s);
    if (b32) b32_.setup(n, m.bandwidth, d.s);
    b32_.src32 = o.solver == SOLVER_BAND32;
    b64_.clock.mask = o.kernel_timing;
    b32_.clock.mask = o.kernel_timing;
    b64_.clock.every = o.kernel_timing_every;
    b32_.clock.every = o.kernel_timing_every;
    b64_.split = o.factor_schedule == 1;
    b32_.split = o.factor_schedule == 1;
    b64_.chained = o.sweep_schedule != 1;
    b32_.chained = o.sweep_schedule != 1;
    b64_.look = (o.sweep_schedule == 2 || o.sweep_schedule == 3) ? o.sweep_schedule : 0;
    b32_.look = b64_.look;
    if (o.solver != SOLVER_BAND64 && o.solver != SOLVER_BAND32) gm_.setup(d.nslot, o.restart);
    CB_CHECK(hipStreamSynchronize(d.s));
}

Analysis::~Analysis() {
    double* vs[] = {q, dv, dp, dpp, dt, dd, ddq, ddr, f, fp, ft, fip, r, tmp, eb, ex, favg, rf,
                    sm, qtot, uc, vc, ac, vci, aci, um, vm, amv, dyn, hist, kcopy};
    for (double* p : vs)
        if (p) (void)hipFree(p);
    b64_.release();
    b32_.release();
    gm_.release();
    bj_.release();
}

double* Analysis::dev_vec(int which) {
    double* v[] = {q, dv, dp, dpp, dt, dd, ddq, ddr, f, fp, ft, fip, r};
    return (which >= 0 && which < 13) ? v[which] : nullptr;
}

void Analysis::record(double lam, int ite, const double* dvec) {
    lpf.push_back(lam);
    itecnt.push_back(ite);
    CB_CHECK(hipStreamSynchronize(d_.s));
    step_t.push_back(now_s() - t_start_);
    step_judge.push_back(st_->t_judge);
    step_newton.push_back(st_->newton);
    const KernelClock& kc =
        (o_.solver == SOLVER_BAND32_FGMRES || o_.solver == SOLVER_BAND32) ? b32_.clock : b64_.clock;
    step_kt_ms.emplace_back(kc.ms, kc.ms + KT_N);
    step_kt_n.emplace_back(kc.launches, kc.launches + KT_N);
    step_ev.push_back(ev_);
    step_psi.push_back(ev_psi_);
    ev_.fill(0);
    if (m_.algflag <= 3) {
        // equilibrium of the converged step: |lambda q - f| and |lambda q|
        vresidual(d_, tmp, lam, q, f, m_.neq);
        const double* a[2] = {tmp, q};
        double s[2];
        multi_dot(d_, 2, a, a, m_.neq, s);
        step_res.push_back(std::sqrt(s[0]));
        step_load.push_back(std::fabs(lam) * std::sqrt(s[1]));
    } else {
        step_res.push_back(0.0);
        step_load.push_back(0.0);
    }
    if (o_.keep_forces) {
        const size_t nf = size_t(m_.ntr + 7 * m_.nfr + 6 * m_.nsh);
        const size_t off = fpath.size();
        fpath.resize(off + nf);
        launch_avg_forces(d_, G_BASE, favg);
        d2h(fpath.data() + off, favg, nf, d_.s);
        CB_CHECK(hipStreamSynchronize(d_.s));
    }
    if (o_.keep_d) {
        size_t off = dpath.size();
        dpath.resize(off + size_t(m_.neq));
        d2h(dpath.data() + off, dvec, size_t(m_.neq), d_.s);
        CB_CHECK(hipStreamSynchronize(d_.s));
    }
}

int Analysis::update_and_forces(int gx, double* fout, const double* ddv, double* dlpf) {
    Timer t(&st_->t_forces, d_.s);
    launch_updatc(d_, gx, ddv);
    launch_forces(d_, gx, fout, ddv);
    double tau = 1;
    const int code = frame_plastic_result(d_, &tau);
    if (code == 1 && dlpf) *dlpf *= tau;
    return code;
}

void Analysis::reset_from_base() {
    copy_gen(d_, G_I, G_BASE, true, true);
    copy_gen(d_, G_IP, G_BASE, true, false);
}

// base <- _i; an unloading member end (flag 2) is elastic again
// (main.c:2107-2112, 3086-3091, 3887-3892)
void Analysis::commit() {
    copy_gen(d_, G_BASE, G_I, true, true);
    launch_unloaded_elastic(d_);
}

// Factorization (or preconditioner set-up) of the tangent in d.K, pivots for
// the arc-length decisions.  Returns 0 or a nonzero code on a singular factor.
int Analysis::factor_current(int* det) {
    int rc = 0;
    bool have_piv = false;
    {
        Timer t(&st_->t_factor, d_.s);
        if (o_.solver == SOLVER_BAND64) {
            rc = b64_.factor(d_);
            pivots = b64_.piv_host;
            have_piv = true;
        } else if (o_.solver == SOLVER_BAND32_FGMRES || o_.solver == SOLVER_BAND32) {
            rc = b32_.factor(d_);
            pivots = b32_.piv_host;
            have_piv = true;
        } else if (o_.solver == SOLVER_BJ_FGMRES) {
            bj_.setup(d_);
        } else if (learned.setup) {
            learned.setup(d_);
        } else {
            bj_.setup(d_);
        }
        st_->factors++;
    }
    if (o_.solver == SOLVER_BAND64 && learned.after_factor) learned.after_factor(d_);
    if (o_.solver != SOLVER_BAND64 && o_.judge) {
        // exact FP64 pivots for det/psi (and the direct-solve fallback)
        Timer t(&st_->t_judge, d_.s);
        int rj = b64_.factor(d_);
        pivots = b64_.piv_host;
        have_piv = true;
        if (rc == 0) rc = rj;
    }
    // A tangent entry outside the band means the band was sized wrongly; the
    // reference's skyline holds every entry, so this is never its behaviour.
    // A zero or non-finite pivot (1) carries on, as solve() ignores skyfact's
    // status (solve.c:51-89).
    if (rc == 2) throw DeviceError("tangent entry outside the band of the factorization");
    if (det) {
        *det = 0;
        if (have_piv)
            for (long i = 0; i < m_.neq; ++i)
                if (!(pivots[i] > 0)) { *det = 1; break; }
        ev_[EV_DET] = *det;
    }
    return rc;
}

double Analysis::psi_now() const {
    double psi = 1;
    if (pivots.size() < size_t(m_.neq) || piv0_.size() < size_t(m_.neq)) return psi;
    for (long i = 0; i < m_.neq; ++i) {
        double t = std::fabs(pivots[i] / piv0_[i]);
        if (t < psi) psi = t;
    }
    return psi;
}

void Analysis::tangent(int gx, int g) {
    Timer t(&st_->t_tangent, d_.s);
    if (o_.solver == SOLVER_BAND32) {
        launch_tangent_fp32(d_, gx, g);
        // bars or frame members: launch_tangent_fp32 formed d.K on the way
        if (o_.judge && !d_.ntr && !d_.nfr) launch_tangent(d_, gx, g, false);
    } else {
        launch_tangent(d_, gx, g, false);
    }
}

int Analysis::tangent_and_factor(int gx, int g, int* det) {
    RunStats tmpst;
    RunStats* keep = st_;
    if (!st_) st_ = &tmpst;
    tangent(gx, g);
    int rc = factor_current(det);
    st_ = keep;
    return rc;
}

int Analysis::factor_only(int* det) {
    RunStats tmpst;
    RunStats* keep = st_;
    if (!st_) st_ = &tmpst;
    int rc = factor_current(det);
    st_ = keep;
    return rc;
}

int Analysis::linear_solve(const double* rhs, double* x) {
    RunStats tmpst;
    RunStats* keep = st_;
    if (!st_) st_ = &tmpst;
    Timer t(&st_->t_solve, d_.s);
    st_->solves++;
    int rc = 0;
    if (o_.solver == SOLVER_BAND64) {
        b64_.solve(rhs, x);
        if (learned.observe) learned.observe(d_, rhs, x);
    } else if (o_.solver == SOLVER_BAND32) {
        b32_.solve(rhs, x);
        for (int k = 0; k < o_.refine; ++k) {
            // rf = rhs - K32 x in FP64, then x += (LDL^T)^-1 rf
            to_slots(d_, ex, x);
            bsr_spmv32(d_, ex, eb);
            from_slots(d_, rf, eb);
            vaxpby(d_, rf, 1.0, rhs, -1.0, m_.neq);
            b32_.solve(rf, rf);
            vaxpby(d_, x, 1.0, rf, 1.0, m_.neq);
        }
    } else {
        std::function<void(const double*, double*)> prec;
        if (o_.solver == SOLVER_BAND32_FGMRES) {
            prec = [&](const double* v, double* z) {
                from_slots(d_, tmp, v);
                b32_.solve(tmp, tmp);
                to_slots(d_, z, tmp);
            };
        } else if (o_.solver == SOLVER_LEARNED_FGMRES && learned.apply) {
            prec = [&](const double* v, double* z) { learned.apply(d_, v, z); };
        } else {
            prec = [&](const double* v, double* z) { bj_.apply(d_, v, z); };
        }
        to_slots(d_, eb, rhs);
        double rel = 0;
        int it = gm_.solve(d_, prec, eb, ex, o_.krylov_tol, o_.maxit, &rel);
        st_->krylov_iters += it;
        st_->worst_relres = std::max(st_->worst_relres, rel);
        if (!(rel <= 10 * o_.krylov_tol) && b64_.tiles) {
            // fall back to the direct solve (the judge's factor)
            st_->krylov_fallbacks++;
            b64_.solve(rhs, x);
        } else {
            from_slots(d_, x, ex);
            if (!(rel <= 10 * o_.krylov_tol)) rc = 3;
        }
    }
    st_ = keep;
    return rc;
}

int Analysis::run(RunStats& st) {
    st_ = &st;
    lpf.clear();
    itecnt.clear();
    dpath.clear();
    fpath.clear();
    step_t.clear();
    step_judge.clear();
    step_newton.clear();
    step_kt_ms.clear();
    step_kt_n.clear();
    step_ev.clear();
    step_psi.clear();
    step_res.clear();
    step_load.clear();
    ev_.fill(0);
    ev_psi_ = 1;
    dev_reset_state(d_, m_);
    for (double* v : {q, dv, dp, dpp, dt, dd, ddq, ddr, f, fp, ft, fip, r})
        CB_CHECK(hipMemsetAsync(v, 0, sizeof(double) * size_t(m_.neq), d_.s));
    h2d(q, m_.q.data(), size_t(m_.neq), d_.s);
    CB_CHECK(hipStreamSynchronize(d_.s));
    double t0 = now_s();
    t_start_ = t0;
    int rc = 0;
    try {
        if (o_.solver == SOLVER_BAND32 && m_.algflag >= 4)
            throw DeviceError("the FP32 precision mode runs the static drivers (ALGFLAG 1-3)");
        if ((m_.algflag == 1 || m_.algflag == 2) && m_.anaflag == 1) rc = run_linear_static();
        else if (m_.algflag == 1 || m_.algflag == 2) rc = run_nr();
        else if (m_.algflag == 3) rc = run_msal();
        else if (m_.algflag == 4) rc = run_linear_newmark();
        else if (m_.algflag == 5) rc = run_newmark();
        else { rc = 2; st.msg = "driver not available for this ALGFLAG"; }
    } catch (const DeviceError& e) {
        rc = 4;
        st.msg = e.msg;
    }
    CB_CHECK(hipStreamSynchronize(d_.s));
    st.t_total = now_s() - t0;
    st.status = rc;
    st_ = nullptr;
    return rc;
}

#define FAIL(code, text) do { st_->msg = (text); return (code); } while (0)

// misc.c:187-250 on the fused sums; returns 1 on a zero-norm error.
static int test_decide(const double s[5], double intener1, double toldisp, double tolforc,
                       double tolener, int* convchk) {
    *convchk = 0;
    if (toldisp < 1) {
        if (s[1] == 0) return 1;
        if (std::sqrt(s[0]) / std::sqrt(s[1]) > toldisp) *convchk += 10;
    }
    if (tolforc < 1) {
        if (s[3] == 0) return 1;
        if (std::sqrt(s[2]) / std::sqrt(s[3]) > tolforc) *convchk += 100;
    }
    if (tolener < 1) {
        if (intener1 == 0) return 1;
        if (std::fabs(s[4] / intener1) > tolener) *convchk += 1000;
    }
    return 0;
}

// ---------------------------------------------------------------- NR / MNR
int Analysis::run_nr() {
    const long n = m_.neq;
    double lam = m_.lpf0, dlam = m_.dlpf0, dlamp;
    int solcnt = 0, subcnt = 0, itecnt_ = 0, convchk = 0, det = 0;
    double intener1 = 0;
    copy_gen(d_, G_I, G_BASE, true, true);
    copy_gen(d_, G_IP, G_BASE, true, true);
    do {
        if (lam > m_.lpfmax) lam = m_.lpfmax;
        vcopy(d_, fp, f, n);
        vcopy(d_, dt, dv, n);
        vcopy(d_, ft, f, n);
        dlamp = dlam;
        reset_from_base();
        itecnt_ = 0;
        int frc = 0;
        do {
            vresidual(d_, r, lam, q, ft, n);
            const bool full = (m_.algflag == 1 || itecnt_ == 0);
            if (full) {
                tangent(G_I, G_IP);
                // solve() discards the factor status (solve.c:86-88, 535)
                factor_current(&det);
            }
            linear_solve(r, dd);
            vaxpby(d_, dt, 1.0, dd, 1.0, n);
            vcopy(d_, fip, ft, n);
            vfill(d_, ft, 0.0, n);
            frc = update_and_forces(G_I, ft, dd, &dlam);
            d2d(d_.ef[G_IP], d_.ef[G_I], size_t(d_.ntr * 2 + d_.nfr * 14 + d_.nsh * 18), d_.s);
            double s[5];
            test_sums(d_, dd, dt, q, lam, ft, fp, fip, n, s);
            if (itecnt_ == 0) {
                // intener1 = dd . (qtot - fp): the fifth sum with fp in place of f_ip
                double s2[5];
                test_sums(d_, dd, dt, q, lam, ft, fp, fp, n, s2);
                intener1 = s2[4];
            }
            if (test_decide(s, intener1, m_.toldisp, m_.tolforc, m_.tolener, &convchk))
                FAIL(1, "zero norm in convergence test");
            shift_ip(true);
            itecnt_++;
            st_->newton++;
        } while (convchk != 0 && frc == 0 && itecnt_ <= m_.itemax);

        if (frc == 2) {
            // a yielded end unloads: the same increment again (main.c:2033-2034)
            dlam = dlamp;
            ev_[EV_UNLOAD]++;
        } else if ((convchk != 0 || frc != 0) && subcnt <= m_.submax) {
            if (lam == m_.lpfmax) FAIL(1, "maximum load factor attempted without convergence");
            if (dlamp == m_.dlpfmin) FAIL(1, "minimum load factor increment reached");
            // a code 1 keeps the increment the regula falsi scaled (main.c:2050-2053)
            if (frc != 1) dlam = dlamp / 2;
            if (dlam < m_.dlpfmin) dlam = m_.dlpfmin;
            lam = lam - dlamp + dlam;
            subcnt++;
            solcnt = 0;
            ev_[frc == 1 ? EV_YIELD : EV_HALVE]++;
        } else if (subcnt > m_.submax) {
            FAIL(1, "maximum number of subdivisions exceeded");
        } else {
            vcopy(d_, dv, dt, n);
            vcopy(d_, f, ft, n);
            commit();
            record(lam, itecnt_, dv);
            if (long(lpf.size()) >= o_.max_steps) return 0;
            solcnt++;
            subcnt = 0;
            if (solcnt >= m_.solmin) {
                dlam *= 2;
                if (dlam > m_.dlpfmax) dlam = m_.dlpfmax;
                solcnt = 0;
            }
            lam += dlam;
        }
    } while (lam <= m_.lpfmax);
    return 0;
}

// ---------------------------------------------------------------- MSAL
int Analysis::run_msal() {
    const long n = m_.neq, k = m_.dkdof;
    int det = 0, convchk = 0, itecnt_ = 0;
    double lam, dlam, intener1, lamp = 0, lampp = 0;
    copy_gen(d_, G_I, G_BASE, true, false);
    copy_gen(d_, G_IP, G_BASE, true, false);
    auto fetch = [&](const double* v, long i) {
        double h;
        d2h(&h, v + i, 1, d_.s);
        CB_CHECK(hipStreamSynchronize(d_.s));
        return h;
    };
    auto dots = [&](int cnt, std::initializer_list<const double*> a, std::initializer_list<const double*> b,
                    double* out) {
        std::vector<const double*> va(a), vb(b);
        Timer t(&st_->t_other, d_.s);
        multi_dot(d_, cnt, va.data(), vb.data(), n, out);
    };

    // first increment: displacement control at the arc-length DOF
    tangent(G_BASE, G_BASE);
    factor_current(&det);
    piv0_ = pivots;
    linear_solve(q, ddq);
    lam = m_.dk / fetch(ddq, k);
    vaxpby(d_, dv, lam, ddq, 0.0, n);
    vaxpby(d_, dd, lam, ddq, 0.0, n);
    {
        double s;
        dots(1, {dd}, {q}, &s);
        intener1 = lam * s;
    }
    vfill(d_, f, 0.0, n);
    // the first call hands forces_fr the load factor itself as the increment
    // a code 1 scales, and its code is dropped (main.c:2365-2369, 2411)
    update_and_forces(G_BASE, f, dd, &lam);
    itecnt_ = 1;
    shift_ip(true);
    int frc = 0;
    do {
        // qtot is formed before the load factor moves and test() sees that
        // qtot (main.c:2415-2485)
        const double lam_q = lam;
        vresidual(d_, r, lam, q, f, n);
        linear_solve(r, ddr);
        dlam = -fetch(ddr, k) / fetch(ddq, k);
        lam += dlam;
        vcopy(d_, dd, ddr, n);
        vaxpby(d_, dd, dlam, ddq, 1.0, n);
        vaxpby(d_, dv, 1.0, dd, 1.0, n);
        vcopy(d_, fip, f, n);
        vfill(d_, f, 0.0, n);
        frc = update_and_forces(G_BASE, f, dd, &dlam);
        double s[5];
        test_sums(d_, dd, dv, q, lam_q, f, fp, fip, n, s);
        if (test_decide(s, intener1, m_.toldisp, m_.tolforc, m_.tolener, &convchk))
            FAIL(1, "zero norm in convergence test");
        itecnt_++;
        st_->newton++;
        shift_ip(true);
    } while (convchk != 0 && frc == 0 && itecnt_ <= m_.itemax);
    if (convchk != 0 || frc != 0) FAIL(1, "initially prescribed displacement too large");
    copy_gen(d_, G_BASE, G_I, true, false);
    record(lam, itecnt_, dv);
    if (long(lpf.size()) >= o_.max_steps) return 0;

    double dkc = std::fabs(fetch(dv, k)), lamc = std::fabs(lam);
    vcopy(d_, dpp, dp, n);
    vcopy(d_, dp, dv, n);
    vcopy(d_, fp, f, n);
    lampp = lamp;
    lamp = lam;
    reset_from_base();
    double dpdp;
    dots(1, {dp}, {dp}, &dpdp);
    double dnorm = std::sqrt(dpdp), dnormallow = m_.alpha * std::sqrt(dpdp);
    double beta = std::sqrt(double(m_.iteopt) / double(itecnt_)) * (dnormallow / dnorm);
    double psi = psi_now();

    int errchk2 = 0, subcnt = 0, imagcnt = 0, negcnt = 0;
    double arc = 0, a = 0, b, c, lamt;
    double qq;
    dots(1, {q}, {q}, &qq);
    ev_psi_ = psi;
    while (lamc <= m_.lpfmax && dkc <= m_.dkimax) {
        vcopy(d_, dt, dv, n);
        lamt = lam;
        if (errchk2 == 0) {
            // |dp - dpp|^2
            double sq[3];
            quad_sums(d_, dp, dpp, dpp, dpp, n, sq);
            if (psi >= m_.psi_thresh)
                arc = beta * std::sqrt(sq[0] + (lamp - lampp) * (lamp - lampp));
            else
                arc = beta * std::sqrt(sq[0]);
        } else {
            reset_from_base();
            errchk2 = 0;
        }
        tangent(G_BASE, G_BASE);
        factor_current(&det);
        linear_solve(q, ddq);
        double qd;
        dots(1, {ddq}, {ddq}, &qd);
        a = (psi >= m_.psi_thresh) ? qq + qd : qd;
        dlam = (det == 0) ? arc * std::sqrt(1 / a) : -arc * std::sqrt(1 / a);
        lamt += dlam;
        vaxpby(d_, dd, dlam, ddq, 0.0, n);
        vaxpby(d_, dt, 1.0, dd, 1.0, n);
        {
            double s;
            dots(1, {dd}, {q}, &s);
            intener1 = dlam * s;
        }
        vfill(d_, ft, 0.0, n);
        // the predictor's code is not looked at; its flags stay, and a return
        // leaves the later members out of f_temp (main.c:2794-2800)
        update_and_forces(G_I, ft, dd, &dlam);
        shift_ip(true);
        itecnt_ = 1;
        st_->newton++;
        do {
            const double lam_q = lamt;   // qtot of this iteration (main.c:2845-2855)
            vresidual(d_, r, lamt, q, ft, n);
            vcopy(d_, fip, ft, n);
            vfill(d_, ft, 0.0, n);
            linear_solve(r, ddr);
            double s8[8], s2[2];
            dots(8, {dt, dp, ddr, dt, dp, dt, dt, dp}, {ddq, ddq, ddq, ddr, ddr, dp, dt, dp}, s8);
            dots(1, {ddr}, {ddr}, s2);
            const double dtq = s8[0], dpq = s8[1], drq = s8[2], dtr = s8[3], dpr = s8[4];
            const double dtp = s8[5], dtt = s8[6], dpp_ = s8[7], drr = s2[0];
            if (psi >= m_.psi_thresh) {
                b = 2 * (dtq - dpq + drq + (lamt - lamp) * qq);
                c = 2 * (dtr - dpr - dtp) + dtt + dpp_ + drr + (lamt - lamp) * (lamt - lamp) * qq - arc * arc;
            } else {
                b = 2 * (dtq - dpq + drq);
                c = 2 * (dtr - dpr - dtp) + dtt + dpp_ + drr - arc * arc;
            }
            // quad (arc.c:70-157)
            double disc = b * b - 4 * a * c;
            if (!(disc > 0)) {
                errchk2 = 2;
            } else {
                double r1 = (-b + std::sqrt(b * b - 4 * a * c)) / (2 * a);
                double r2 = (-b - std::sqrt(b * b - 4 * a * c)) / (2 * a);
                double sq[3];
                quad_sums(d_, dt, dp, ddr, ddq, n, sq);
                double g1 = sq[0] + sq[1] + r1 * sq[2], g2 = sq[0] + sq[1] + r2 * sq[2];
                double root;
                if (g1 > g2 && g1 > 0) { root = r1; errchk2 = 0; }
                else if (g2 > g1 && g2 > 0) { root = r2; errchk2 = 0; }
                else { errchk2 = 3; root = 0; }
                if (errchk2 == 0) {
                    quad_update(d_, dt, dp, ddr, ddq, dd, root, n);
                    dlam = root;
                    lamt += dlam;
                }
            }
            if (errchk2 == 0) {
                frc = update_and_forces(G_I, ft, dd, &dlam);
                double s[5];
                test_sums(d_, dd, dt, q, lam_q, ft, fp, fip, n, s);
                if (test_decide(s, intener1, m_.toldisp, m_.tolforc, m_.tolener, &convchk))
                    FAIL(1, "zero norm in convergence test");
                st_->newton++;
                if (convchk == 0) {
                    double sq[3];
                    quad_sums(d_, dt, dp, dp, dp, n, sq);
                    dnorm = std::sqrt(sq[0]);
                    if (dnorm > 100 * dnormallow) {
                        arc /= beta;
                        beta = dnormallow / dnorm;
                        arc *= beta;
                        vcopy(d_, ft, f, n);
                        errchk2 = 1;
                        ev_[EV_DNORM]++;
                    }
                } else if (frc == 2) {
                    vcopy(d_, ft, f, n);
                    errchk2 = 1;
                    ev_[EV_UNLOAD]++;
                } else if (frc != 0 && subcnt <= m_.submax) {
                    arc *= 0.5;
                    vcopy(d_, ft, f, n);
                    subcnt++;
                    errchk2 = 1;
                    ev_[EV_YIELD]++;
                } else {
                    itecnt_++;
                    if (itecnt_ > m_.itemax) {
                        arc *= 0.5;
                        vcopy(d_, ft, f, n);
                        subcnt++;
                        errchk2 = 1;
                        ev_[EV_HALVE]++;
                    }
                    shift_ip(true);
                }
            } else if (errchk2 == 2) {
                arc *= 0.5;
                vcopy(d_, ft, f, n);
                imagcnt++;
                ev_[EV_IMAG]++;
            } else if (errchk2 == 3) {
                arc *= 0.5;
                vcopy(d_, ft, f, n);
                negcnt++;
                ev_[EV_NEG]++;
            }
        } while (convchk != 0 && errchk2 == 0 && subcnt <= m_.submax && imagcnt <= m_.imagmax &&
                 negcnt <= m_.negmax);
        if (subcnt > m_.submax || imagcnt > m_.imagmax || negcnt > m_.negmax)
            FAIL(1, "maximum allowable number of subdivisions exceeded");
        if (errchk2 == 0) {
            beta = std::sqrt(double(m_.iteopt) / double(itecnt_)) * (dnormallow / dnorm);
            vcopy(d_, dpp, dp, n);
            vcopy(d_, dv, dt, n);
            vcopy(d_, dp, dt, n);
            vcopy(d_, f, ft, n);
            vcopy(d_, fp, ft, n);
            lampp = lamp;
            lam = lamp = lamt;
            commit();
            dkc = std::fabs(fetch(dv, k));
            lamc = std::fabs(lam);
            psi = psi_now();
            subcnt = imagcnt = negcnt = 0;
            record(lam, itecnt_, dv);
            ev_psi_ = psi;
            if (long(lpf.size()) >= o_.max_steps) return 0;
        }
    }
    return 0;
}

// ---------------------------------------------------------------- Newmark
// Lumped masses of every element type at the converged coordinates.
int Analysis::mass(double* smv) {
    const int types = (m_.ntr > 0) + (m_.nfr > 0) + (m_.nsh > 0);
    if (types > 1)
        throw DeviceError("lumped masses of mixed element types index the per-type arrays from 0 "
                          "in the reference (frame.c:1356-1368); single-type models only");
    vfill(d_, smv, 0.0, m_.neq);
    if (m_.ntr) launch_mass(d_, smv, 0);
    if (m_.nfr) launch_mass(d_, smv, 1);
    if (m_.nsh) launch_mass(d_, smv, 2);
    return 0;
}

// main.c:3305-3960 with solve.c:139-245, 469-533, without prescribed support
// motion; one path row per time step (time, itecnt, d).  The initial
// conditions read from the input are discarded here as in the reference
// (main.c:3443-3445).
int Analysis::run_newmark() {
    const long n = m_.neq;
    if (!sm) {
        double** vs[] = {&sm, &qtot, &uc, &vc, &ac, &vci, &aci, &um, &vm, &amv, &dyn};
        for (double** p : vs) *p = dalloc<double>(size_t(n));
        hist = dalloc<double>(size_t(n * m_.ntstpsinpt));
        h2d(hist, m_.pinpt.data(), size_t(n * m_.ntstpsinpt), d_.s);
    }
    for (double* v : {sm, qtot, uc, vc, ac, vci, aci, um, vm, amv, dyn})
        CB_CHECK(hipMemsetAsync(v, 0, sizeof(double) * size_t(n), d_.s));
    double am_ = 0, af_ = 0;
    const double rho = m_.spectrds;
    switch (m_.numopt) {
    case 0:
        if (rho != 1) FAIL(1, "invalid spectral radius for Newmark without dissipation");
        break;
    case 1: am_ = (2 * rho - 1) / (rho + 1); af_ = rho / (rho + 1); break;
    case 2: af_ = (1 - rho) / (1 + rho); break;
    case 3: am_ = (rho - 1) / (rho + 1); break;
    default: FAIL(1, "unknown Newmark dissipation option");
    }
    const double alpha = std::pow(1 - am_ + af_, 2) / 4, delta = 0.5 - am_ + af_;
    const double dtime = m_.dt;
    const long ntst = long(m_.ttot / dtime + 1);
    if (ntst > m_.ntstpsinpt) FAIL(1, "time steps exceed the load history");
    double ddt = 1, sub_dt = 1, dt_temp = dtime;
    int tsflag = 0, convchk = 1, itecnt_ = 0, det = 0, frc = 0;
    double intener1 = 0;
    long k = 0;
    const bool ckpt = !o_.checkpoint_path.empty();
    // restart (main.c:3378-3398): the state of the checkpoint file, then the
    // time steps after it; a checkpoint of the last step gives that step's
    // row alone (main.c:3387-3392) rather than a step past the load history
    if (m_.rflag == 1) {
        if (!ckpt) FAIL(1, "restart (RFLAG 1) without a checkpoint file");
        long tstep = 0;
        if (load_checkpoint(&tstep, &ddt)) FAIL(1, "unable to read the checkpoint file");
        k = tstep + 1;
        if (k >= ntst) {
            if (k == ntst) record(double(tstep) * dtime, 0, dv);
            return 0;
        }
    }
    copy_gen(d_, G_I, G_BASE, true, true);
    copy_gen(d_, G_IP, G_BASE, true, true);
    NewmarkArgs A;
    A.P = hist; A.ld = m_.ntstpsinpt; A.am = am_; A.af = af_;
    A.qtot = qtot; A.r = r; A.ft = ft; A.fip = fip; A.dt = dt; A.dd = dd; A.sm = sm;
    A.ac = aci; A.vc = vci; A.um = um; A.vm = vm; A.amv = amv;
    do {
        dt_temp = ddt * dtime;
        sub_dt = ddt;
        tsflag = 0;
        do {
            A.a0 = 1 / (alpha * std::pow(dt_temp, 2));
            A.a2 = 1 / (alpha * dt_temp);
            A.a3 = 1 / (2 * alpha) - 1;
            A.a6 = dt_temp * (1 - delta);
            A.a7 = delta * dt_temp;
            int solcnt = 0, subcnt = 0;
            double lam = m_.lpf0, dlam = m_.dlpf0, dlamp;
            do {
                if (lam > m_.lpfmax) lam = m_.lpfmax;
                vcopy(d_, fp, f, n);
                vcopy(d_, dt, dv, n);
                vcopy(d_, ft, f, n);
                dlamp = dlam;
                reset_from_base();
                vcopy(d_, um, uc, n); vcopy(d_, vm, vc, n); vcopy(d_, amv, ac, n);
                vcopy(d_, vci, vc, n); vcopy(d_, aci, ac, n);
                itecnt_ = 0;
                do {
                    A.k = k; A.sub_dt = sub_dt; A.lpf = lam;
                    if (itecnt_ == 0) newmark_vec(d_, k == 0 ? NM_PREDICT0 : NM_PREDICT, A);
                    else newmark_vec(d_, NM_CORRECT, A);
                    {
                        Timer t(&st_->t_tangent, d_.s);
                        launch_tangent(d_, G_I, G_IP, false);
                        mass(sm);
                        add_mass_diag(d_, sm, A.a0 * (1 - am_), 1 - af_);
                    }
                    // solve() discards the factor status (solve.c:208-212, 535)
                    factor_current(&det);
                    A.out = tmp;
                    newmark_vec(d_, NM_REFF, A);
                    linear_solve(tmp, dd);
                    A.flip = itecnt_ > 0;
                    newmark_vec(d_, NM_UPDATE, A);
                    frc = update_and_forces(G_I, ft, dd, &dlam);
                    A.out = dyn;
                    newmark_vec(d_, NM_DYN, A);
                    double s[5];
                    test_sums(d_, dd, dt, dyn, 1.0, ft, fp, fip, n, s);
                    if (itecnt_ == 0) {
                        double s2[5];
                        test_sums(d_, dd, dt, dyn, 1.0, ft, fp, fp, n, s2);
                        intener1 = s2[4];
                    }
                    if (test_decide(s, intener1, m_.toldisp, m_.tolforc, m_.tolener, &convchk))
                        FAIL(1, "zero norm in convergence test");
                    shift_ip(true);
                    itecnt_++;
                    st_->newton++;
                } while (convchk != 0 && frc == 0 && itecnt_ <= m_.itemax);
                if (frc == 2) {
                    dlam = dlamp;                       // main.c:3773-3774
                    ev_[EV_UNLOAD]++;
                } else if ((convchk != 0 || frc != 0) && subcnt <= m_.submax) {
                    if (lam == m_.lpfmax) break;          // main.c:3783-3788
                    if (dlamp == m_.dlpfmin) break;
                    if (frc != 1) dlam = dlamp / 2;
                    if (dlam < m_.dlpfmin) dlam = m_.dlpfmin;
                    lam = lam - dlamp + dlam;
                    subcnt++;
                    solcnt = 0;
                    ev_[frc == 1 ? EV_YIELD : EV_HALVE]++;
                } else if (subcnt > m_.submax) {
                    break;
                } else {
                    solcnt++;
                    subcnt = 0;
                    if (solcnt >= m_.solmin) { dlam *= 2; solcnt = 0; }
                    lam += dlam;
                }
            } while (lam <= m_.lpfmax);
            bool committed = false;
            if (convchk != 0) {
                ddt = ddt / 2;                          // main.c:3826-3846
                dt_temp = ddt * dtime;
                ev_[EV_HALVE]++;
                if (tsflag == 0) sub_dt = ddt;
                else if (tsflag == 1) sub_dt = sub_dt - ddt;
            } else {
                vcopy(d_, dv, dt, n);
                vcopy(d_, f, ft, n);
                vcopy(d_, uc, dt, n); vcopy(d_, vc, vci, n); vcopy(d_, ac, aci, n);
                commit();
                committed = true;
            }
            // the multiplier moves only after a force pass without a code
            // (main.c:3922-3935)
            if (frc == 0) {
                if (ddt < 1) {
                    if (sub_dt <= 1) { sub_dt = sub_dt + ddt; tsflag = 1; }
                    if (sub_dt > 1) { ddt = ddt * 2; tsflag = 2; }
                } else if (ddt == 1) {
                    tsflag = 2;
                }
            }
            // checkPoint at the commit (main.c:3904-3907), written after the
            // multiplier update, which changes no state, so that the file also
            // holds the multiplier the next step starts from
            if (committed && ckpt && k % m_.chkpt == 0 && k != 0 &&
                !save_checkpoint(k, ddt, A.a0 * (1 - am_), 1 - af_))
                FAIL(1, "unable to write the checkpoint file");
        } while (ddt >= 0.0001 && tsflag != 2);
        if (convchk != 0 || frc != 0) FAIL(3, "minimum time increment interval reached without convergence");
        record(double(k) * dtime, itecnt_, dv);
        if (long(lpf.size()) >= o_.max_steps) return 0;
        ++k;
    } while (k < ntst);
    return 0;
}

bool Analysis::save_checkpoint(long k, double ddt, double cnum, double cden) {
    Checkpoint c;
    c.size_for(m_);
    c.tstep = k;
    c.ddt = ddt;
    const size_t n = size_t(m_.neq), bs = size_t(d_.bs);
    std::vector<double> smv(n), K(size_t(d_.nnzb) * bs * bs);
    d2h(c.uc.data(), uc, n, d_.s);
    d2h(c.vc.data(), vc, n, d_.s);
    d2h(c.ac.data(), ac, n, d_.s);
    d2h(smv.data(), sm, n, d_.s);
    d2h(c.d.data(), dv, n, d_.s);
    d2h(c.f.data(), f, n, d_.s);
    d2h(c.ef.data(), d_.ef[G_BASE], c.ef.size(), d_.s);
    d2h(c.x.data(), d_.x[G_BASE], c.x.size(), d_.s);
    d2h(c.c1.data(), d_.c1[G_BASE], c.c1.size(), d_.s);
    d2h(c.c2.data(), d_.c2[G_BASE], c.c2.size(), d_.s);
    d2h(c.c3.data(), d_.c3[G_BASE], c.c3.size(), d_.s);
    d2h(c.defllen.data(), d_.defllen[G_BASE], c.defllen.size(), d_.s);
    d2h(c.deffarea.data(), d_.deffarea[G_BASE], c.deffarea.size(), d_.s);
    d2h(c.defslen.data(), d_.defslen[G_BASE], c.defslen.size(), d_.s);
    d2h(c.llen_w.data(), d_.llen, c.llen_w.size(), d_.s);
    d2h(c.farea_w.data(), d_.farea, c.farea_w.size(), d_.s);
    d2h(c.slen_w.data(), d_.slength, c.slen_w.size(), d_.s);
    d2h(c.yldflag.data(), d_.yld, c.yldflag.size(), d_.s);
    d2h(K.data(), d_.K, K.size(), d_.s);
    CB_CHECK(hipStreamSynchronize(d_.s));
    c.llength = c.llen_w;      // the frame lengths at the commit (main.c:3880)
    std::copy(smv.begin(), smv.end(), c.sm.begin());
    // member end coordinates: joint + offset (misc.c:122-131)
    for (long e = 0; e < m_.nfr; ++e)
        for (int end = 0; end < 2; ++end) {
            const long j = m_.minc[size_t(m_.ntr * 2 + e * 2 + end)] - 1;
            for (int a = 0; a < 3; ++a)
                c.xfr[size_t(e * 6 + end * 3 + a)] =
                    c.x[size_t(j * 3 + a)] + (m_.osflag[size_t(e)] ? m_.offset[size_t(e * 6 + end * 3 + a)] : 0.0);
        }
    // skyline ss from the joint-block tangent: column j holds rows j - kht[j]..j,
    // the diagonal at maxa[j] - 1 and row i at maxa[j] - 1 + (j - i)
    std::vector<long> slot(n, -1);
    for (size_t s = 0; s < m_.eq_of_slot.size(); ++s)
        if (m_.eq_of_slot[s] >= 0) slot[size_t(m_.eq_of_slot[s])] = long(s);
    auto entry = [&](long si, long sj) {
        if (si < 0 || sj < 0) return 0.0;
        const long a = si / long(bs), b = sj / long(bs);
        const int32_t* lo = m_.colidx.data() + m_.rowptr[size_t(a)];
        const int32_t* hi = m_.colidx.data() + m_.rowptr[size_t(a) + 1];
        const int32_t* it = std::lower_bound(lo, hi, int32_t(b));
        if (it == hi || *it != b) return 0.0;
        const size_t blk = size_t(it - m_.colidx.data());
        return K[blk * bs * bs + size_t(si % long(bs)) * bs + size_t(sj % long(bs))];
    };
    for (long j = 0; j < long(n); ++j) {
        for (long h = 0; h <= m_.kht[size_t(j)]; ++h) {
            double v = entry(slot[size_t(j - h)], slot[size_t(j)]);
            if (h == 0) v -= cnum * smv[size_t(j)] / cden;
            c.ss[size_t(m_.maxa[size_t(j)] - 1 + h)] = v;
        }
    }
    return write_checkpoint(o_.checkpoint_path, m_, c, o_.checkpoint_exact != 0);
}

int Analysis::load_checkpoint(long* tstep, double* ddt) {
    Checkpoint c;
    const int rc = read_checkpoint(o_.checkpoint_path, m_, c);
    if (rc) return rc;
    const size_t n = size_t(m_.neq);
    h2d(uc, c.uc.data(), n, d_.s);
    h2d(vc, c.vc.data(), n, d_.s);
    h2d(ac, c.ac.data(), n, d_.s);
    h2d(sm, c.sm.data(), n, d_.s);
    h2d(dv, c.d.data(), n, d_.s);
    h2d(f, c.f.data(), n, d_.s);
    h2d(d_.ef[G_BASE], c.ef.data(), c.ef.size(), d_.s);
    h2d(d_.x[G_BASE], c.x.data(), c.x.size(), d_.s);
    h2d(d_.c1[G_BASE], c.c1.data(), c.c1.size(), d_.s);
    h2d(d_.c2[G_BASE], c.c2.data(), c.c2.size(), d_.s);
    h2d(d_.c3[G_BASE], c.c3.data(), c.c3.size(), d_.s);
    h2d(d_.defllen[G_BASE], c.defllen.data(), c.defllen.size(), d_.s);
    h2d(d_.deffarea[G_BASE], c.deffarea.data(), c.deffarea.size(), d_.s);
    h2d(d_.defslen[G_BASE], c.defslen.data(), c.defslen.size(), d_.s);
    h2d(d_.yld, c.yldflag.data(), c.yldflag.size(), d_.s);
    // working lengths: the frames' from the file's llength, as a fresh start
    // takes them (main.c:3420; the reference's restart leaves them unset), the
    // rest as the model has them, unless the full-precision extension holds all
    std::vector<double> llen(m_.llength);
    for (long e = 0; e < m_.nfr; ++e) llen[size_t(m_.ntr + e)] = c.llength[size_t(m_.ntr + e)];
    if (c.has_ext) {
        llen = c.llen_w;
        h2d(d_.farea, c.farea_w.data(), c.farea_w.size(), d_.s);
        h2d(d_.slength, c.slen_w.data(), c.slen_w.size(), d_.s);
    }
    h2d(d_.llen, llen.data(), llen.size(), d_.s);
    CB_CHECK(hipStreamSynchronize(d_.s));
    *tstep = c.tstep;
    *ddt = c.has_ext ? c.ddt : 1.0;
    return 0;
}

// Linear static analysis, ANAFLAG 1 under ALGFLAG 1/2 (main.c:1714-1798): the
// tangent at the input geometry (no geometric terms), K d = lpfmax q, the
// linear element forces of d and one path row (lpfmax, 0, d).  The reference
// prints its iteration counter there without setting it (main.c:1670); 0 here.
int Analysis::run_linear_static() {
    const long n = m_.neq;
    double* qt = tmp;
    vcopy(d_, qt, q, n);
    vaxpby(d_, qt, 0.0, q, m_.lpfmax, n);
    tangent(G_BASE, G_BASE);
    int det = 0;
    factor_current(&det);                 // solve() ignores skyfact's status (solve.c:78-89)
    linear_solve(qt, dv);
    CB_CHECK(hipMemsetAsync(f, 0, sizeof(double) * size_t(n), d_.s));
    {
        Timer t(&st_->t_forces, d_.s);
        launch_forces(d_, G_BASE, f, dv);
    }
    commit();                             // the element forces become the state's
    st_->newton++;
    record(m_.lpfmax, 0, dv);
    return 0;
}

// Linear Newmark, ALGFLAG 4 (main.c:3143-3303, solve.c:139-245, 246-468): K
// and the lumped M at the input geometry, Keff = K + a0 (1-am)/(1-af) M on the
// diagonal factored once, then one solve per time step from the input's
// initial conditions; row (k dt, 0, u) per step.  After the loop the reference
// computes the forces of d, which it never sets on this path, and writes
// (lpfmax, counter, 0): that row ends the path here too.  The checkpoint file
// (CHKPT) holds u, v, a; a model with RFLAG 1 starts from it.
int Analysis::run_linear_newmark() {
    const long n = m_.neq;
    if (!sm) {
        double** vs[] = {&sm, &qtot, &uc, &vc, &ac, &vci, &aci, &um, &vm, &amv, &dyn};
        for (double** p : vs) *p = dalloc<double>(size_t(n));
        hist = dalloc<double>(size_t(n * m_.ntstpsinpt));
        h2d(hist, m_.pinpt.data(), size_t(n * m_.ntstpsinpt), d_.s);
    }
    if (!kcopy) kcopy = dalloc<double>(size_t(d_.nnzb) * size_t(d_.bs * d_.bs));
    double am_ = 0, af_ = 0;
    const double rho = m_.spectrds;
    switch (m_.numopt) {
    case 0:
        if (rho != 1) FAIL(1, "invalid spectral radius for Newmark without dissipation");
        break;
    case 1: am_ = (2 * rho - 1) / (rho + 1); af_ = rho / (rho + 1); break;
    case 2: af_ = (1 - rho) / (1 + rho); break;
    case 3: am_ = (rho - 1) / (rho + 1); break;
    default: FAIL(1, "unknown Newmark dissipation option");
    }
    const double alpha = std::pow(1 - am_ + af_, 2) / 4, delta = 0.5 - am_ + af_;
    const double dtime = m_.dt;
    const long ntst = long(m_.ttot / dtime + 1);
    if (ntst > m_.ntstpsinpt) FAIL(1, "time steps exceed the load history");
    {
        Timer t(&st_->t_tangent, d_.s);
        launch_tangent(d_, G_BASE, G_BASE, false);
        d2d(kcopy, d_.K, size_t(d_.nnzb) * size_t(d_.bs * d_.bs), d_.s);
        CB_CHECK(hipMemsetAsync(sm, 0, sizeof(double) * size_t(n), d_.s));
        mass(sm);
    }
    LinearNewmarkArgs A;
    A.P = hist; A.ld = m_.ntstpsinpt; A.am = am_; A.af = af_;
    A.a0 = 1 / (alpha * std::pow(dtime, 2));
    A.a2 = 1 / (alpha * dtime);
    A.a3 = 1 / (2 * alpha) - 1;
    A.a6 = dtime * (1 - delta);
    A.a7 = delta * dtime;
    add_mass_diag(d_, sm, A.a0 * (1 - am_), 1 - af_);
    int det = 0;
    factor_current(&det);                 // status ignored, as solve() does
    h2d(um, m_.um.data(), size_t(n), d_.s);
    h2d(vm, m_.vm.data(), size_t(n), d_.s);
    h2d(amv, m_.am.data(), size_t(n), d_.s);
    const bool ckpt = !o_.checkpoint_path.empty();
    long k = 0;
    if (m_.rflag == 1) {
        if (!ckpt) FAIL(1, "restart (RFLAG 1) without a checkpoint file");
        long tstp = 0;
        std::vector<double> u0, v0, a0v;
        if (read_checkpoint_linear(o_.checkpoint_path, n, &tstp, u0, v0, a0v))
            FAIL(1, "unable to read the checkpoint file");
        h2d(um, u0.data(), size_t(n), d_.s);
        h2d(vm, v0.data(), size_t(n), d_.s);
        h2d(amv, a0v.data(), size_t(n), d_.s);
        if (tstp + 1 == ntst) record(double(tstp) * dtime, 0, um);
        k = tstp + 1;
    }
    A.sm = sm; A.u = um; A.v = vm; A.a = amv; A.r = r;
    A.unew = uc; A.vnew = vc; A.anew = ac;
    for (; k < ntst; ++k) {
        A.k = k;
        A.Ku = nullptr;
        if (af_ != 0) {
            // K u with the tangent kept before the mass term went onto the diagonal
            to_slots(d_, eb, um);
            std::swap(d_.K, kcopy);
            bsr_spmv(d_, eb, ex);
            std::swap(d_.K, kcopy);
            from_slots(d_, dyn, ex);
            A.Ku = dyn;
        }
        linear_newmark_vec(d_, LNM_REFF, A);
        linear_solve(r, uc);
        linear_newmark_vec(d_, LNM_UPDATE, A);
        st_->newton++;
        record(double(k) * dtime, 0, uc);
        vcopy(d_, um, uc, n);
        vcopy(d_, vm, vc, n);
        vcopy(d_, amv, ac, n);
        if (ckpt && k % m_.chkpt == 0 && k != 0) {
            const size_t nu = static_cast<size_t>(n);
            std::vector<double> hu(nu), hv(nu), ha(nu);
            d2h(hu.data(), um, size_t(n), d_.s);
            d2h(hv.data(), vm, size_t(n), d_.s);
            d2h(ha.data(), amv, size_t(n), d_.s);
            CB_CHECK(hipStreamSynchronize(d_.s));
            if (!write_checkpoint_linear(o_.checkpoint_path, k, hu, hv, ha, o_.checkpoint_exact != 0))
                FAIL(1, "unable to write the checkpoint file");
        }
        if (long(lpf.size()) >= o_.max_steps) return 0;
    }
    record(m_.lpfmax, 0, dv);             // dv is zero: the reference's d on this path
    return 0;
}

}  // namespace cb
