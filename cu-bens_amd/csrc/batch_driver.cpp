// Batched nonlinear Newmark over many frame models (the host side of batch.h).
//
// Each model keeps the control flow of the single-model driver
// (Analysis::run_newmark, which restates main.c:3305-3960): time steps, the
// load-factor loop, Newton iterations, time-step halving.  Here that flow is a
// resumable state machine per model.  A round is one Newton iteration for
// every running model; between rounds each model advances its own machine
// from the convergence sums of its last iteration to the start of its next
// one.  On the way it may commit a converged step, record it, and restart
// from the converged state, always in that order; those device copies are
// applied to the models that asked for them by masked passes over the batch.
#include "batch.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <thread>

namespace cb {
namespace {

constexpr long TILE = 64;

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

template <class V> void append(V& dst, const V& src) { dst.insert(dst.end(), src.begin(), src.end()); }

// The Newmark control state of one model.  Labels follow the loops of
// Analysis::run_newmark: L1 time steps, L2 time-step halving, L3 load factor,
// L4 Newton iterations.
struct Machine {
    const Model* m = nullptr;
    // constants of the integration scheme (main.c:3447-3462)
    double am = 0, af = 0, alpha = 0, delta = 0, dtime = 0;
    long ntst = 0;
    long max_rows = 0;
    // state
    enum Pc { L1_HEAD, L2_HEAD, L3_HEAD, AFTER_ITER, DONE } pc = L1_HEAD;
    long k = 0;
    double ddt = 1, sub_dt = 1, dt_temp = 0;
    int tsflag = 0, convchk = 1, itecnt = 0, solcnt = 0, subcnt = 0;
    double lam = 0, dlam = 0, dlamp = 0, intener1 = 0;
    double a0 = 0, a2 = 0, a3 = 0, a6 = 0, a7 = 0;
    int status = 0;
    std::string msg;
    long rows = 0, newton = 0;

    // Runs from the current label to the start of the next Newton iteration
    // (returns true, with p filled for that iteration) or to the end of the
    // analysis (false).  s: the sums of the iteration just taken (AFTER_ITER).
    bool advance(const double* s, NmParam& p, std::vector<double>& times, std::vector<int>& ites) {
        p.commit = 0;
        p.start = 0;
        p.rec = -1;
        for (;;) {
            switch (pc) {
            case L1_HEAD:
                dt_temp = ddt * dtime;
                sub_dt = ddt;
                tsflag = 0;
                pc = L2_HEAD;
                break;
            case L2_HEAD:
                a0 = 1 / (alpha * std::pow(dt_temp, 2));
                a2 = 1 / (alpha * dt_temp);
                a3 = 1 / (2 * alpha) - 1;
                a6 = dt_temp * (1 - delta);
                a7 = delta * dt_temp;
                solcnt = 0;
                subcnt = 0;
                lam = m->lpf0;
                dlam = m->dlpf0;
                pc = L3_HEAD;
                break;
            case L3_HEAD:
                if (lam > m->lpfmax) lam = m->lpfmax;
                dlamp = dlam;
                p.start = 1;   // fp = f, dt = dv, ft = f, state <- converged, um/vm/amv/vci/aci
                itecnt = 0;
                fill(p, k == 0 ? NM_PREDICT0 : NM_PREDICT);
                pc = AFTER_ITER;
                return true;
            case AFTER_ITER: {
                if (itecnt == 0) intener1 = s[5];
                convchk = 0;
                // misc.c:187-250 on the sums
                bool zero = false;
                if (m->toldisp < 1) {
                    if (s[1] == 0) zero = true;
                    else if (std::sqrt(s[0]) / std::sqrt(s[1]) > m->toldisp) convchk += 10;
                }
                if (!zero && m->tolforc < 1) {
                    if (s[3] == 0) zero = true;
                    else if (std::sqrt(s[2]) / std::sqrt(s[3]) > m->tolforc) convchk += 100;
                }
                if (!zero && m->tolener < 1) {
                    if (intener1 == 0) zero = true;
                    else if (std::fabs(s[4] / intener1) > m->tolener) convchk += 1000;
                }
                if (zero) return finish(1, "zero norm in convergence test");
                itecnt++;
                newton++;
                if (convchk != 0 && itecnt <= m->itemax) {
                    fill(p, NM_CORRECT);
                    return true;
                }
                // the load-factor loop (main.c:3783-3821)
                bool exit_l3 = false;
                if (convchk != 0 && subcnt <= m->submax) {
                    if (lam == m->lpfmax || dlamp == m->dlpfmin) {
                        exit_l3 = true;
                    } else {
                        dlam = dlamp / 2;
                        if (dlam < m->dlpfmin) dlam = m->dlpfmin;
                        lam = lam - dlamp + dlam;
                        subcnt++;
                        solcnt = 0;
                    }
                } else if (subcnt > m->submax) {
                    exit_l3 = true;
                } else {
                    solcnt++;
                    subcnt = 0;
                    if (solcnt >= m->solmin) { dlam *= 2; solcnt = 0; }
                    lam += dlam;
                }
                if (!exit_l3 && lam <= m->lpfmax) {
                    pc = L3_HEAD;
                    break;
                }
                // time-step halving (main.c:3826-3846)
                if (convchk != 0) {
                    ddt = ddt / 2;
                    dt_temp = ddt * dtime;
                    if (tsflag == 0) sub_dt = ddt;
                    else if (tsflag == 1) sub_dt = sub_dt - ddt;
                } else {
                    p.commit = 1;   // dv = dt, f = ft, uc = dt, vc = vci, ac = aci, converged <- current
                }
                if (ddt < 1) {
                    if (sub_dt <= 1) { sub_dt = sub_dt + ddt; tsflag = 1; }
                    if (sub_dt > 1) { ddt = ddt * 2; tsflag = 2; }
                } else if (ddt == 1) {
                    tsflag = 2;
                }
                if (ddt >= 0.0001 && tsflag != 2) {
                    pc = L2_HEAD;
                    break;
                }
                if (convchk != 0) return finish(3, "minimum time increment interval reached without convergence");
                p.rec = int(rows);
                times.push_back(double(k) * dtime);
                ites.push_back(itecnt);
                rows++;
                if (rows >= max_rows) return finish(0, "");
                ++k;
                if (k >= ntst) return finish(0, "");
                pc = L1_HEAD;
                break;
            }
            case DONE:
                return false;
            }
        }
    }

    void fill(NmParam& p, int op) const {
        p.sub_dt = sub_dt;
        p.lpf = lam;
        p.a0 = a0; p.a2 = a2; p.a3 = a3; p.a6 = a6; p.a7 = a7;
        p.am = am; p.af = af;
        p.cnum = a0 * (1 - am);
        p.cden = 1 - af;
        p.k = k;
        p.op = op;
        p.flip = (op == NM_CORRECT) ? 1 : 0;
    }

    bool finish(int st, const char* text) {
        status = st;
        msg = text;
        pc = DONE;
        return false;
    }
};

}  // namespace

Batch::Batch(const std::vector<std::string>& paths) {
    if (paths.empty()) throw std::runtime_error("a batch needs at least one model");
    models_.resize(paths.size());
    // the model files are read by several host threads
    std::vector<std::string> errs(paths.size());
    {
        const size_t nth = std::min<size_t>(paths.size(), std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
        std::vector<std::thread> pool;
        for (size_t t = 0; t < nth; ++t)
            pool.emplace_back([&, t] {
                for (size_t b = t; b < paths.size(); b += nth)
                    if (!load_model(paths[b].c_str(), models_[b], errs[b]) && errs[b].empty()) errs[b] = "unreadable";
            });
        for (std::thread& th : pool) th.join();
    }
    for (size_t b = 0; b < paths.size(); ++b) {
        if (!errs[b].empty()) throw std::runtime_error(paths[b] + ": " + errs[b]);
        const Model& m = models_[b];
        if (m.algflag != 5)
            throw std::runtime_error(paths[b] + ": a batch runs the nonlinear Newmark driver (ALGFLAG 5) only");
        if (m.anaflag != 2)
            throw std::runtime_error(paths[b] + ": a batch runs elastic members (ANAFLAG 2) only");
        if (m.ntr || m.nsh || !m.nfr) throw std::runtime_error(paths[b] + ": a batch takes frame-only models");
        for (int v : m.osflag)
            if (v) throw std::runtime_error(paths[b] + ": member-end offsets are not carried by the batched frame kernels");
        for (int32_t v : m.relmask)
            if (v) throw std::runtime_error(paths[b] + ": member end releases are not carried by the batched frame kernels");
        for (long j = 0; j < m.nj; ++j)
            if (m.jcode[j * 7 + 6] > 0)
                throw std::runtime_error(paths[b] + ": free warping DOFs are not carried by the device layout");
    }
    // concatenation: joints, members, equations (each model's first equation
    // on a tile boundary)
    Model& A = all_;
    const Model& m0 = models_[0];
    A.anaflag = m0.anaflag; A.algflag = m0.algflag; A.slvflag = m0.slvflag; A.optflag = m0.optflag;
    const size_t nb = models_.size();
    jb_.resize(nb);
    eb_.resize(nb);
    segs_.resize(nb);
    std::vector<BandSeg> bseg(nb);
    long jb = 0, eb = 0, qb = 0;
    int wmax = 0;
    for (size_t b = 0; b < nb; ++b) {
        const Model& m = models_[b];
        jb_[b] = jb;
        eb_[b] = eb;
        const long ntb = std::max<long>(1, (m.neq + TILE - 1) / TILE);
        segs_[b].qb = qb;
        segs_[b].neq = m.neq;
        bseg[b].t0 = qb / TILE;
        bseg[b].nt = ntb;
        bseg[b].neq = m.neq;
        bseg[b].bw = std::max(m.bandwidth, 0L);
        // the single-model band's tiles below the diagonal (BandLDLT::setup)
        int wb = int((std::max(m.bandwidth, 1L) + TILE - 1) / TILE);
        if (wb > ntb - 1) wb = int(std::max(ntb - 1, 0L));
        bseg[b].w = wb;
        wmax = std::max(wmax, wb);
        for (long v : m.minc) A.minc.push_back(v + jb);
        for (long v : m.jcode) A.jcode.push_back(v > 0 ? v + qb : v);
        for (long v : m.mcode) A.mcode.push_back(v > 0 ? v + qb : v);
        append(A.wrpres, m.wrpres);
        append(A.x, m.x);
        append(A.emod, m.emod); append(A.yield, m.yield); append(A.carea, m.carea);
        append(A.llength, m.llength); append(A.gmod, m.gmod); append(A.dens_fr, m.dens_fr);
        append(A.istrong, m.istrong); append(A.iweak, m.iweak); append(A.ipolar, m.ipolar);
        append(A.iwarp, m.iwarp); append(A.zstrong, m.zstrong); append(A.zweak, m.zweak);
        append(A.offset, m.offset); append(A.auxpt, m.auxpt); append(A.xfr, m.xfr);
        append(A.osflag, m.osflag);
        A.dens.insert(A.dens.end(), m.dens.begin(), m.dens.begin() + m.nfr);
        append(A.c1, m.c1); append(A.c2, m.c2); append(A.c3, m.c3);
        jb += m.nj;
        eb += m.nfr;
        qb += ntb * TILE;
        A.bandwidth = std::max(A.bandwidth, m.bandwidth);
    }
    A.nj = jb; A.nfr = eb; A.ntr = 0; A.nsh = 0; A.neq = qb;
    n_ = qb;
    A.q.assign(size_t(n_), 0.0);
    ld_ = 1;
    for (const Model& m : models_) ld_ = std::max(ld_, m.ntstpsinpt);
    A.ntstpsinpt = ld_;
    A.pinpt.assign(size_t(n_ * ld_), 0.0);
    for (size_t b = 0; b < nb; ++b) {
        const Model& m = models_[b];
        for (long i = 0; i < m.neq; ++i) {
            A.q[size_t(segs_[b].qb + i)] = m.q[size_t(i)];
            for (long t = 0; t < m.ntstpsinpt; ++t)
                A.pinpt[size_t((segs_[b].qb + i) * ld_ + t)] = m.pinpt[size_t(i * m.ntstpsinpt + t)];
        }
    }
    build_pattern(A);
    dev_create(d_, A);
    // owners and the equations of each model's last tile that no slot maps to
    std::vector<int32_t> oq(static_cast<size_t>(n_)), oj(static_cast<size_t>(A.nj)),
        oe(static_cast<size_t>(std::max(A.nfr, 1L)), 0), gaps;
    for (size_t b = 0; b < nb; ++b) {
        const Model& m = models_[b];
        const long end = segs_[b].qb + bseg[b].nt * TILE;
        for (long i = segs_[b].qb; i < end; ++i) oq[size_t(i)] = int32_t(b);
        for (long i = segs_[b].qb + m.neq; i < end; ++i) gaps.push_back(int32_t(i));
        for (long j = 0; j < m.nj; ++j) oj[size_t(jb_[b] + j)] = int32_t(b);
        for (long e = 0; e < m.nfr; ++e) oe[size_t(eb_[b] + e)] = int32_t(b);
    }
    own_eq_ = dalloc<int32_t>(oq.size());
    own_joint_ = dalloc<int32_t>(oj.size());
    own_el_ = dalloc<int32_t>(oe.size());
    h2d(own_eq_, oq.data(), oq.size(), d_.s);
    h2d(own_joint_, oj.data(), oj.size(), d_.s);
    h2d(own_el_, oe.data(), oe.size(), d_.s);
    band_.setup(bseg, n_, wmax, gaps, d_.s);
    segs_dev_ = dalloc<BatchSeg>(nb);
    prm_dev_ = dalloc<NmParam>(nb);
    CB_CHECK(hipHostMalloc(&prm_host_, sizeof(NmParam) * nb, hipHostMallocDefault));
    sums_dev_ = dalloc<double>(nb * 6);
    CB_CHECK(hipHostMalloc(&sums_host_, sizeof(double) * nb * 6, hipHostMallocDefault));
    double** vs[] = {&dv, &dt, &dd, &f, &fp, &ft, &fip, &r, &tmp, &sm, &qtot, &uc,
                     &vc, &ac, &vci, &aci, &um, &vm, &amv, &dyn};
    for (double** p : vs) *p = dalloc<double>(size_t(n_));
    hist = dalloc<double>(size_t(n_ * ld_));
    h2d(hist, A.pinpt.data(), A.pinpt.size(), d_.s);
    build_jobs();
    CB_CHECK(hipStreamSynchronize(d_.s));
}

void Batch::build_jobs() {
    const long nfr = all_.nfr, nj = all_.nj;
    // commit: dv = dt, f = ft, uc = dt, vc = vci, ac = aci; converged <- current
    // (state and coordinates)
    std::vector<CopyJob> cj = {
        {dv, dt, n_, 1, 0}, {f, ft, n_, 1, 0}, {uc, dt, n_, 1, 0}, {vc, vci, n_, 1, 0}, {ac, aci, n_, 1, 0},
    };
    // restart: fp = f, dt = dv, ft = f; current and previous <- converged;
    // um = uc, vm = vc, amv = ac, vci = vc, aci = ac
    std::vector<CopyJob> sj = {
        {fp, f, n_, 1, 0},  {dt, dv, n_, 1, 0}, {ft, f, n_, 1, 0},   {um, uc, n_, 1, 0},
        {vm, vc, n_, 1, 0}, {amv, ac, n_, 1, 0}, {vci, vc, n_, 1, 0}, {aci, ac, n_, 1, 0},
    };
    auto state = [&](std::vector<CopyJob>& v, int dst, int src, bool with_x) {
        v.push_back({d_.c1[dst], d_.c1[src], nfr * 3, 3, 2});
        v.push_back({d_.c2[dst], d_.c2[src], nfr * 3, 3, 2});
        v.push_back({d_.c3[dst], d_.c3[src], nfr * 3, 3, 2});
        v.push_back({d_.ef[dst], d_.ef[src], nfr * 14, 14, 2});
        v.push_back({d_.defllen[dst], d_.defllen[src], nfr, 1, 2});
        if (with_x) v.push_back({d_.x[dst], d_.x[src], nj * 3, 3, 1});
    };
    state(cj, G_BASE, G_I, true);
    state(sj, G_I, G_BASE, true);
    state(sj, G_IP, G_BASE, false);
    ncommit_ = int(cj.size());
    nstart_ = int(sj.size());
    commit_jobs_ = dalloc<CopyJob>(cj.size());
    start_jobs_ = dalloc<CopyJob>(sj.size());
    h2d(commit_jobs_, cj.data(), cj.size(), d_.s);
    h2d(start_jobs_, sj.data(), sj.size(), d_.s);
}

Batch::~Batch() {
    try {
        if (gexec_) (void)hipGraphExecDestroy(gexec_);
        band_.release();
        void* ps[] = {own_eq_, own_joint_, own_el_, segs_dev_, prm_dev_, sums_dev_, commit_jobs_, start_jobs_,
                      dv, dt, dd, f, fp, ft, fip, r, tmp, sm, qtot, uc, vc, ac, vci, aci, um, vm, amv, dyn, hist,
                      pathv};
        for (void* p : ps)
            if (p) (void)hipFree(p);
        if (prm_host_) (void)hipHostFree(prm_host_);
        if (sums_host_) (void)hipHostFree(sums_host_);
        dev_destroy(d_);
    } catch (...) {
    }
}

// One Newton iteration of every model (Analysis::run_newmark's L4 body): a
// fixed sequence of launches whose per-round values (each model's time step,
// predictor or corrector, Keff coefficient) live in prm_dev_, so that it can
// be captured once and replayed.
void Batch::iteration(NewmarkArgs& A) {
    batch_nm_vec(d_.s, NM_MODEL_OP, A, prm_dev_, own_eq_, n_);   // predictor or corrector
    // the tangent goes straight into the band (the copy the single-model
    // solver makes from the joint blocks), then Keff's mass term
    band_.zero();
    launch_frame_tangent_band(d_, G_IP, band_.tiles, band_.w, band_.flag, fp32_el_);
    vfill(d_, sm, 0.0, n_);
    launch_mass(d_, sm, 1);
    batch_add_mass_band(d_, band_.tiles, band_.w, sm, prm_dev_, own_joint_);
    band_.factor_filled(false);
    A.out = tmp;
    batch_nm_vec(d_.s, NM_REFF, A, prm_dev_, own_eq_, n_);
    band_.solve(tmp, dd);
    batch_nm_vec(d_.s, NM_UPDATE, A, prm_dev_, own_eq_, n_);
    launch_updatc(d_, G_I, dd);
    launch_forces(d_, G_I, ft, dd);
    A.out = dyn;
    batch_nm_vec(d_.s, NM_DYN, A, prm_dev_, own_eq_, n_);
    batch_test_sums(d_.s, segs_dev_, int(models_.size()), dd, dt, dyn, ft, fp, fip, sums_dev_);
    copy_gen(d_, G_IP, G_I, true, false);
}

namespace {
// Tuned helper-pair counts by batch shape: the device's CU count and every
// model's tile rows and band width.
std::mutex tune_mu;
std::map<std::string, int> tune_cache;
bool tune_file_read = false;

std::string band_shape_key(const BandBatch& b) {
    int dev = 0, ncu = 0;
    CB_CHECK(hipGetDevice(&dev));
    CB_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    uint64_t h = 1469598103934665603ull;   // FNV-1a over (nt, w) of every model
    for (const BandSeg& g : b.segs_host)
        for (long v : {g.nt, long(g.w)}) {
            h ^= uint64_t(v);
            h *= 1099511628211ull;
        }
    std::ostringstream k;
    k << "batch-pairs cu" << ncu << " models" << b.nseg << " nt" << b.nt << " w" << b.w << " h" << std::hex << h;
    return k.str();
}
}  // namespace

// The band of the initial tangent (the first round's, without the mass term:
// the factor's time does not depend on the values) is factored for 0, 1/4,
// 1/2, 3/4 and all of the allowed helper pairs; the fastest count stays.
void Batch::tune_band() {
    const std::string key = band_shape_key(band_);
    const char* file = std::getenv("CUBENS_TUNE_FILE");
    std::lock_guard<std::mutex> lk(tune_mu);
    if (file && !tune_file_read) {
        std::ifstream in(file);
        std::string line;
        while (std::getline(in, line)) {
            const size_t sp = line.rfind(' ');
            if (sp != std::string::npos) tune_cache[line.substr(0, sp)] = std::atoi(line.c_str() + sp + 1);
        }
        tune_file_read = true;
    }
    auto it = tune_cache.find(key);
    if (it != tune_cache.end()) {
        band_.set_pairs(it->second);
        return;
    }
    band_.zero();
    launch_frame_tangent_band(d_, G_IP, band_.tiles, band_.w, band_.flag);
    const int mp = band_.max_pairs;
    std::vector<int> cand = {0, mp / 4, mp / 2, 3 * mp / 4, mp};
    cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
    tuning = band_.tune_pairs(cand);
    band_.zero();
    tune_cache[key] = band_.npair;
    if (file) {
        std::ofstream out(file, std::ios::app);
        out << key << ' ' << band_.npair << '\n';
    }
}

void Batch::run(int max_steps, bool keep_path, BatchStats& st, bool timing, bool graph, bool autotune,
                bool fp32_elements) {
    const size_t nb = models_.size();
    const double t0 = now_s();
    band_.timing = timing;
    double ms_factor = 0, ms_solve = 0;
    std::vector<Machine> mc(nb);
    status.assign(nb, 0);
    message.assign(nb, "");
    times.assign(nb, {});
    itecnt.assign(nb, {});
    // path rows on the device: each model keeps room for its steps
    long pbase = 0;
    for (size_t b = 0; b < nb; ++b) {
        const Model& m = models_[b];
        Machine& s = mc[b];
        s.m = &m;
        double am = 0, af = 0;
        const double rho = m.spectrds;
        switch (m.numopt) {
        case 0:
            if (rho != 1) s.finish(1, "invalid spectral radius for Newmark without dissipation");
            break;
        case 1: am = (2 * rho - 1) / (rho + 1); af = rho / (rho + 1); break;
        case 2: af = (1 - rho) / (1 + rho); break;
        case 3: am = (rho - 1) / (rho + 1); break;
        default: s.finish(1, "unknown Newmark dissipation option");
        }
        s.am = am;
        s.af = af;
        s.alpha = std::pow(1 - am + af, 2) / 4;
        s.delta = 0.5 - am + af;
        s.dtime = m.dt;
        s.ntst = long(m.ttot / m.dt + 1);
        if (s.pc != Machine::DONE && s.ntst > m.ntstpsinpt) s.finish(1, "time steps exceed the load history");
        s.max_rows = std::max<long>(0, std::min<long>(max_steps, s.ntst));
        if (s.max_rows == 0 && s.pc != Machine::DONE) s.finish(0, "");
        segs_[b].pbase = pbase;
        if (keep_path) pbase += s.max_rows * m.neq;
    }
    if (keep_path && size_t(pbase) > path_len_) {
        if (pathv) (void)hipFree(pathv);
        pathv = dalloc<double>(size_t(pbase));
        path_len_ = size_t(pbase);
    }
    h2d(segs_dev_, segs_.data(), nb, d_.s);
    // initial state of every model (Analysis::run, main.c:3443-3445)
    dev_reset_state(d_, all_);
    double* zero[] = {dv, dt, dd, f, fp, ft, fip, r, tmp, sm, qtot, uc, vc, ac, vci, aci, um, vm, amv, dyn};
    for (double* v : zero) vfill(d_, v, 0.0, n_);
    copy_gen(d_, G_I, G_BASE, true, true);
    copy_gen(d_, G_IP, G_BASE, true, true);
    NewmarkArgs A;
    A.P = hist; A.ld = ld_;
    A.qtot = qtot; A.r = r; A.ft = ft; A.fip = fip; A.dt = dt; A.dd = dd; A.sm = sm;
    A.ac = aci; A.vc = vci; A.um = um; A.vm = vm; A.amv = amv;
    std::vector<NmParam> prm(nb);
    long rounds = 0;
    long ops = 0;
    if (autotune && !tuned_) {
        const double tt = now_s();
        tune_band();
        tuned_ = true;
        st.tune_s = now_s() - tt;
        // the factor's grid changed: capture the iteration again
        if (gexec_) {
            CB_CHECK(hipGraphExecDestroy(gexec_));
            gexec_ = nullptr;
        }
        body_ops_ = 0;
    }
    if (fp32_elements != fp32_el_) {
        // another element tangent: capture the iteration again
        fp32_el_ = fp32_elements;
        if (gexec_) {
            CB_CHECK(hipGraphExecDestroy(gexec_));
            gexec_ = nullptr;
        }
        body_ops_ = 0;
    }
    st.pairs = band_.npair;
    st.pairs_max = band_.max_pairs;
    if (!body_ops_ || (graph && !gexec_)) {
        // capture the iteration (without the timing events) once: its node
        // count, and the executable graph when it is to be replayed
        const bool keep_timing = band_.timing;
        band_.timing = false;
        hipGraph_t g = nullptr;
        CB_CHECK(hipStreamBeginCapture(d_.s, hipStreamCaptureModeThreadLocal));
        iteration(A);
        CB_CHECK(hipStreamEndCapture(d_.s, &g));
        size_t nn = 0;
        CB_CHECK(hipGraphGetNodes(g, nullptr, &nn));
        body_ops_ = long(nn);
        if (graph && !gexec_) CB_CHECK(hipGraphInstantiate(&gexec_, g, nullptr, nullptr, 0));
        CB_CHECK(hipGraphDestroy(g));
        band_.timing = keep_timing;
    }
    const bool replay = graph && !timing && gexec_;
    for (;;) {
        // every model runs its machine to its next Newton iteration
        bool any = false, any_commit = false, any_rec = false, any_start = false;
        for (size_t b = 0; b < nb; ++b) {
            Machine& s = mc[b];
            NmParam& p = prm[b];
            if (s.pc == Machine::DONE) {
                p.commit = p.start = 0;
                p.rec = -1;
                continue;
            }
            const bool more = s.advance(sums_host_ + 6 * b, p, times[b], itecnt[b]);
            if (!keep_path) p.rec = -1;
            any_commit |= p.commit != 0;
            any_rec |= p.rec >= 0;
            any_start |= p.start != 0;
            if (more) any = true;
        }
        if (!any && !any_commit && !any_rec) break;
        std::memcpy(prm_host_, prm.data(), sizeof(NmParam) * nb);
        h2d(prm_dev_, prm_host_, nb, d_.s);
        ++ops;
        const long nmax = std::max<long>(n_, all_.nfr * 14);
        if (any_commit) {
            batch_masked_copy(d_.s, commit_jobs_, ncommit_, nmax, own_eq_, own_joint_, own_el_, prm_dev_, 0);
            ++ops;
        }
        if (any_rec) {
            batch_record(d_.s, dv, pathv, segs_dev_, prm_dev_, own_eq_, n_);
            ++ops;
        }
        if (!any) break;
        if (any_start) {
            batch_masked_copy(d_.s, start_jobs_, nstart_, nmax, own_eq_, own_joint_, own_el_, prm_dev_, 1);
            ++ops;
        }
        if (replay) {
            CB_CHECK(hipGraphLaunch(gexec_, d_.s));
            ++ops;
        } else {
            iteration(A);
            ops += body_ops_;
        }
        band_.fetch_flag();
        d2h(sums_host_, sums_dev_, nb * 6, d_.s);
        ops += 2;
        CB_CHECK(hipStreamSynchronize(d_.s));
        if (*band_.flag_host & 1) throw DeviceError("batched band factorization: a wait exceeded its bound");
        if (*band_.flag_host & 2) throw DeviceError("batched band factorization: an entry outside a model's band");
        if (timing) {
            float a = 0, c = 0;
            CB_CHECK(hipEventElapsedTime(&a, band_.ev[0], band_.ev[1]));
            CB_CHECK(hipEventElapsedTime(&c, band_.ev[2], band_.ev[3]));
            ms_factor += a;
            ms_solve += c;
        }
        ++rounds;
    }
    CB_CHECK(hipStreamSynchronize(d_.s));
    st.t_total = now_s() - t0;
    st.rounds = rounds;
    st.ops_per_round = rounds ? double(ops) / double(rounds) : 0.0;
    st.body_ops = body_ops_;
    st.ms_factor = ms_factor;
    st.ms_solve = ms_solve;
    // algorithmic work of one factor / solve launch over all models
    double fl = 0, by = 0;
    const double t3 = double(TILE * TILE * TILE);
    for (const BandSeg& g : band_.segs_host) {
        for (long K = 0; K < g.nt; ++K) {
            const double m = double(std::min<long>(g.w, g.nt - 1 - K));
            fl += t3 + 2 * t3 * m + t3 * m * (m + 1);
        }
        // the two sweeps read every off-diagonal factor tile and every Linv
        // tile once, except the 16-entry chunks that are structurally zero
        // (beyond the half-bandwidth, above Linv's diagonal)
        double chunks = 0;
        for (long d = 1; d <= g.w; ++d) {
            long fwd = 0, bwd = 0;
            for (int r = 0; r < TILE; ++r)
                for (int q = 0; q < 4; ++q) {
                    fwd += (TILE * d + r - 16 * q - 15 <= g.bw);
                    bwd += (TILE * d + 16 * q - r <= g.bw);
                }
            chunks += double(fwd + bwd) * double(g.nt - d);
        }
        long lfwd = 0, lbwd = 0;
        for (int r = 0; r < TILE; ++r)
            for (int q = 0; q < 4; ++q) {
                lfwd += (16 * q <= r);
                lbwd += (16 * q + 15 >= r);
            }
        chunks += double(lfwd + lbwd) * double(g.nt);
        by += chunks * 16 * 8;
    }
    st.factor_flops = fl;
    st.solve_bytes = by;
    st.newton = 0;
    st.steps = 0;
    st.failed = 0;
    for (size_t b = 0; b < nb; ++b) {
        status[b] = mc[b].status;
        message[b] = mc[b].msg;
        st.newton += mc[b].newton;
        st.steps += long(times[b].size());
        if (mc[b].status) st.failed++;
    }
    kept_path_ = keep_path;
}

void Batch::path(int b, double* out) const {
    if (!kept_path_ || !pathv) throw std::runtime_error("the last batch run kept no displacement path");
    const size_t len = times[size_t(b)].size() * size_t(segs_[size_t(b)].neq);
    d2h(out, pathv + segs_[size_t(b)].pbase, len, d_.s);
    CB_CHECK(hipStreamSynchronize(d_.s));
}

}  // namespace cb
