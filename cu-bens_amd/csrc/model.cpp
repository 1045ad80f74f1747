// Host model: input, DOF numbering, skyline profile, initial geometry, pattern.
//
// Input order and semantics follow the reference: header main.c:340-427;
// incidences/constraints model.c:95-264; joint DOF rules model.c:267-285;
// numbering codes() model.c:937-1050; skyline profile skylin() model.c:1145-1283;
// coordinates main.c:1394-1396; properties truss.c:41-80, frame.c:43-224,
// shell.c:43-108 (+ mem_coord shell.c:2402-2446); loads model.c:1285-1606;
// solver parameters main.c:1809-1812, 2240-2245 (arc.c:41-68), 3320-3326.
#include "model.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

namespace cb {
namespace {

class Tokens {
public:
    explicit Tokens(std::string s) : s_(std::move(s)) {}
    bool bad() const { return bad_; }
    double d() {
        skip();
        if (p_ >= s_.size()) { bad_ = true; return 0; }
        char* e;
        double v = std::strtod(s_.c_str() + p_, &e);
        if (e == s_.c_str() + p_) { bad_ = true; return 0; }
        p_ = size_t(e - s_.c_str());
        return v;
    }
    long l() {
        skip();
        if (p_ >= s_.size()) { bad_ = true; return 0; }
        char* e;
        long v = std::strtol(s_.c_str() + p_, &e, 10);
        if (e == s_.c_str() + p_) { bad_ = true; return 0; }
        p_ = size_t(e - s_.c_str());
        return v;
    }
    // whether a comma follows the last value, past white space: the literal
    // "," of a scanf format that continues the line (model.c:252)
    bool comma_next() const {
        size_t p = p_;
        while (p < s_.size() && std::isspace((unsigned char)s_[p])) ++p;
        return p < s_.size() && s_[p] == ',';
    }

private:
    void skip() {
        while (p_ < s_.size() && (std::isspace((unsigned char)s_[p_]) || s_[p_] == ',')) ++p_;
    }
    std::string s_;
    size_t p_ = 0;
    bool bad_ = false;
};

double norm3(const double* a) { return std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }

void cross3(const double* a, const double* b, double* c, bool unit) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
    if (unit) {
        double l = norm3(c);
        c[0] /= l; c[1] /= l; c[2] /= l;
    }
}

// struc's last steps (model.c:267-285): per joint the members framing in
// (warping restraint counts), and joints missing an element type lose the
// slots only that type uses.
void autofix_slots(Model& m) {
    std::vector<long> cnt(size_t(m.nj) * 3, 0);
    for (long e = 0; e < m.ntr; ++e)
        for (int k = 0; k < 2; ++k) cnt[(m.minc[e * 2 + k] - 1) * 3 + 0]++;
    long off = m.ntr * 2;
    for (long e = 0; e < m.nfr; ++e)
        for (int k = 0; k < 2; ++k) {
            long j = m.minc[off + e * 2 + k] - 1;
            cnt[j * 3 + 1]++;
            m.wrpres[j * 3 + 1]++;
        }
    off += m.nfr * 2;
    for (long e = 0; e < m.nsh; ++e)
        for (int k = 0; k < 3; ++k) cnt[(m.minc[off + e * 3 + k] - 1) * 3 + 2]++;
    for (long j = 0; j < m.nj; ++j) {
        long* jc = &m.jcode[j * 7];
        if (cnt[j * 3 + 1]) continue;
        if (cnt[j * 3 + 2]) { jc[6] = 0; continue; }
        if (cnt[j * 3 + 0]) { std::fill(jc + 3, jc + 7, 0L); continue; }
        std::fill(jc, jc + 7, 0L);
    }
}

// Joint renumbering for a smaller bandwidth, OPTFLAG 2 (graph, optnum and
// updatenum, model.c:649-935).  Every joint in turn starts a numbering that
// gives the unnumbered neighbours of each numbered joint the next numbers,
// in the order of its adjacency list: per joint its trusses, frames and
// shells in element order (a shell adds its other two vertices).  An attempt
// stops extending a joint's neighbours once one lands farther than the best
// bandwidth so far; a completed attempt with a smaller bandwidth, or an equal
// one and a smaller profile (the running sum of the bandwidth), replaces the
// numbering.  jcode rows and incidences are renumbered; wrpres keeps the
// input numbering, as updatenum leaves it (model.c:893-898).
void renumber_joints(Model& m) {
    const long nj = m.nj, ntr = m.ntr, nfr = m.nfr, nsh = m.nsh;
    const size_t nju = static_cast<size_t>(nj);
    std::vector<std::vector<long>> tr(nju), fr(nju), sh(nju);
    for (long e = 0; e < ntr; ++e)
        for (int a = 0; a < 2; ++a) tr[size_t(m.minc[e * 2 + a] - 1)].push_back(e);
    for (long e = 0; e < nfr; ++e)
        for (int a = 0; a < 2; ++a) fr[size_t(m.minc[ntr * 2 + e * 2 + a] - 1)].push_back(e);
    const long so = ntr * 2 + nfr * 2;
    for (long e = 0; e < nsh; ++e)
        for (int a = 0; a < 3; ++a) sh[size_t(m.minc[so + e * 3 + a] - 1)].push_back(e);
    std::vector<std::vector<long>> adj(nju);
    for (long i = 0; i < nj; ++i) {
        auto& a = adj[size_t(i)];
        for (long e : tr[size_t(i)]) {
            const long l = m.minc[e * 2] - 1, r = m.minc[e * 2 + 1] - 1;
            a.push_back(l == i ? r : l);
        }
        for (long e : fr[size_t(i)]) {
            const long l = m.minc[ntr * 2 + e * 2] - 1, r = m.minc[ntr * 2 + e * 2 + 1] - 1;
            a.push_back(l == i ? r : l);
        }
        for (long e : sh[size_t(i)]) {
            const long v0 = m.minc[so + e * 3] - 1, v1 = m.minc[so + e * 3 + 1] - 1, v2 = m.minc[so + e * 3 + 2] - 1;
            if (v0 == i) { a.push_back(v1); a.push_back(v2); }
            else if (v1 == i) { a.push_back(v0); a.push_back(v2); }
            else { a.push_back(v0); a.push_back(v1); }
        }
    }
    long band = nj * 7, prof = band * band;
    bool changed = false;
    std::vector<long> newjt(nju), joint(nju);
    for (long ik = 0; ik < nj; ++ik) {
        std::fill(newjt.begin(), newjt.end(), -1L);
        std::fill(joint.begin(), joint.end(), -1L);
        long mx = 0, nprof = 0, i = 0, k = 0, kmax = 0;
        newjt[0] = ik;
        joint[size_t(ik)] = 0;
        do {
            if (newjt[size_t(i)] != -1) {
                for (long k5 : adj[size_t(newjt[size_t(i)])]) {
                    if (joint[size_t(k5)] != -1) continue;
                    ++k;
                    ++kmax;
                    newjt[size_t(k)] = k5;
                    joint[size_t(k5)] = k;
                    const long nb = std::labs(i - k);
                    if (nb > band) break;
                    if (nb > mx) mx = nb;
                }
            } else if (k != 0) {
                // a joint no numbered joint reaches: a placeholder entry
                // that is never visited, as in optnum
                ++k;
                newjt[size_t(k)] = newjt[size_t(k - 1)] - 1;
            } else {
                break;
            }
            nprof += mx;
            ++i;
        } while (k < nj - 1);
        if (k != nj - 1) continue;
        if (mx < band || (mx == band && nprof < prof)) {
            changed = true;
            band = mx;
            prof = nprof;
            for (long j = 0; j < nj; ++j) m.jnt[size_t(j)] = joint[size_t(j)] != -1 ? joint[size_t(j)] : ++kmax;
        }
    }
    if (!changed) return;
    std::vector<long> jc(m.jcode);
    for (long i = 0; i < nj; ++i)
        for (int k = 0; k < 7; ++k) m.jcode[size_t(m.jnt[size_t(i)] * 7 + k)] = jc[size_t(i * 7 + k)];
    for (auto& v : m.minc) v = m.jnt[size_t(v - 1)] + 1;
}

// codes (model.c:937-1050): sequential numbering joint by joint, slots 1-6
// then warping, and the element DOF maps.
void number_dofs(Model& m) {
    long off = 0;
    long neq = 0;
    for (long j = 0; j < m.nj; ++j) {
        long* jc = &m.jcode[j * 7];
        int* wr = &m.wrpres[j * 3];
        wr[2] = wr[1];
        for (int k = 0; k < 6; ++k)
            if (jc[k]) jc[k] = ++neq;
        if (jc[6]) {
            jc[6] = ++neq;
        } else if (wr[0] == 1) {
            jc[6] = ++neq;
            neq += wr[1] - 1;
        }
    }
    m.neq = neq;
    for (long e = 0; e < m.ntr; ++e)
        for (int s = 0; s < 2; ++s) {
            long j = m.minc[e * 2 + s] - 1;
            for (int k = 0; k < 3; ++k) m.mcode[e * 6 + s * 3 + k] = m.jcode[j * 7 + k];
        }
    long mo = m.ntr * 6;
    off = m.ntr * 2;
    for (long e = 0; e < m.nfr; ++e)
        for (int s = 0; s < 2; ++s) {
            long j = m.minc[off + e * 2 + s] - 1;
            long* mc = &m.mcode[mo + e * 14 + s * 7];
            for (int k = 0; k < 6; ++k) mc[k] = m.jcode[j * 7 + k];
            int* wr = &m.wrpres[j * 3];
            if (wr[0] == 1) {
                mc[6] = m.jcode[j * 7 + 6] + wr[1] - wr[2];
                wr[2]--;
            } else {
                mc[6] = m.jcode[j * 7 + 6];
            }
        }
    mo += m.nfr * 14;
    off += m.nfr * 2;
    for (long e = 0; e < m.nsh; ++e)
        for (int v = 0; v < 3; ++v) {
            long j = m.minc[off + e * 3 + v] - 1;
            for (int k = 0; k < 6; ++k) m.mcode[mo + e * 18 + v * 6 + k] = m.jcode[j * 7 + k];
        }
}

// Column heights kht and 1-based diagonal addresses maxa.
void profile(Model& m) {
    m.kht.assign(size_t(std::max(m.neq, 1L)), 0);
    m.maxa.assign(size_t(m.neq) + 1, 0);
    struct G { long n; int w; long off; } g[3] = {
        {m.ntr, 6, 0}, {m.nfr, 14, m.ntr * 6}, {m.nsh, 18, m.ntr * 6 + m.nfr * 14}};
    for (auto& gg : g)
        for (long e = 0; e < gg.n; ++e) {
            const long* mc = &m.mcode[gg.off + e * gg.w];
            long lo = m.neq;
            for (int k = 0; k < gg.w; ++k)
                if (mc[k] > 0 && mc[k] < lo) lo = mc[k];
            for (int k = 0; k < gg.w; ++k) {
                long c = mc[k];
                if (c && c - lo > m.kht[c - 1]) m.kht[c - 1] = c - lo;
            }
        }
    m.maxa[0] = 1;
    m.bandwidth = 0;
    for (long i = 0; i < m.neq; ++i) {
        m.maxa[i + 1] = m.maxa[i] + m.kht[i] + 1;
        m.bandwidth = std::max(m.bandwidth, m.kht[i]);
    }
    m.lss = m.maxa[m.neq] - 1;
}

}  // namespace

bool load_model(const char* path, Model& m, std::string& err) {
    std::ifstream in(path, std::ios::binary);
    if (!in) { err = std::string("cannot open ") + path; return false; }
    std::stringstream ss;
    ss << in.rdbuf();
    Tokens t(ss.str());
#define FAIL(msg) do { err = (msg); return false; } while (0)
    m.anaflag = int(t.l());
    if (m.anaflag == 4) FAIL("ANAFLAG 4 (acoustic FSI) is not on this path");
    m.algflag = int(t.l());
    if (m.algflag > 3) { m.chkpt = int(t.l()); m.rflag = int(t.l()); }
    m.slvflag = int(t.l());
    m.optflag = int(t.l());
    m.nj = t.l();
    m.ntr = t.l(); m.nfr = t.l(); m.nsh = t.l();
    long nsbr = t.l(), nfbr = t.l();
    if (t.bad()) FAIL("truncated header");
    if (nsbr || nfbr) FAIL("brick elements are not on this path");
    if (m.optflag != 1 && m.optflag != 2) FAIL("OPTFLAG must be 1 or 2");
    if (m.anaflag < 1 || m.anaflag > 3)
        FAIL("ANAFLAG must be 1 (linear elastic), 2 (nonlinear elastic) or 3 (inelastic)");
    if (m.anaflag == 3 && m.nsh > 0)
        FAIL("ANAFLAG 3 shells (Ivanov's yield criterion, shell.c:842-1503) are not on this path");
    if (m.algflag == 4 && m.anaflag != 1)
        FAIL("invalid entry for analysis type: ALGFLAG 4 (linear Newmark) takes ANAFLAG 1");
    if (m.algflag < 1 || m.algflag > 5) FAIL("ALGFLAG must be 1 to 5");
    if (m.anaflag == 1 && (m.algflag == 3 || m.algflag == 5))
        FAIL("linear analysis (ANAFLAG 1) runs under ALGFLAG 1, 2 (static) or 4 (linear Newmark)");
    if (m.nj <= 0 || m.ntr < 0 || m.nfr < 0 || m.nsh < 0) FAIL("bad model sizes");
    const long nj = m.nj, ntr = m.ntr, nfr = m.nfr, nsh = m.nsh;
    m.minc.assign(size_t(ntr * 2 + nfr * 2 + nsh * 3), 0);
    m.jcode.assign(size_t(nj * 7), -1);
    m.mcode.assign(size_t(ntr * 6 + nfr * 14 + nsh * 18), 0);
    m.wrpres.assign(size_t(nj * 3), 0);
    for (auto& v : m.minc) {
        v = t.l();
        if (v < 1 || v > nj) FAIL("member incidence out of range");
    }
    for (long i = 0; i < nj * 7; ++i) {
        long j = t.l(), k = t.l();
        if (t.bad()) FAIL("truncated constraints");
        if (j == 0) break;
        if (k < 1 || k > 7 || j < 1 || j > nj) FAIL("joint constraint input not recognized");
        // "j,7,restraint"; a bare "j,7" leaves the restraint 0, as the
        // reference's fscanf(",%d") matches nothing then (model.c:251-254)
        if (k == 7 && t.comma_next()) m.wrpres[(j - 1) * 3] = int(t.l());
        m.jcode[(j - 1) * 7 + (k - 1)] = 0;
    }
    m.jnt.resize(size_t(nj));
    for (long j = 0; j < nj; ++j) m.jnt[size_t(j)] = j;
    autofix_slots(m);
    if (m.optflag == 2) renumber_joints(m);
    number_dofs(m);
    // a free warping DOF shared by the members at its joint is slot 6 of a
    // 7-slot joint; per-member warping DOFs (restraint 1, model.c:951-957,
    // 1017-1031) have no joint slot
    for (long j = 0; j < nj; ++j) {
        if (m.wrpres[j * 3] == 1 && m.jcode[j * 7 + 6] != 0)
            FAIL("per-member warping DOFs (warping restraint 1) are not carried by the device layout");
        if (m.jcode[j * 7 + 6] != 0 && nsh > 0)
            FAIL("free warping DOFs in a model with shell elements are not carried by the device layout");
    }
    if (m.algflag > 3) {
        long j = t.l(), k = t.l();
        if (j != 0 || k != 0) FAIL("prescribed displacement DOFs (NBC > 0) are not on this path");
    }
    profile(m);
    m.x.assign(size_t(nj * 3), 0);
    for (long j = 0; j < nj; ++j)   // coordinates of input joint j land at its new number (main.c:1394-1396)
        for (int k = 0; k < 3; ++k) m.x[size_t(m.jnt[size_t(j)] * 3 + k)] = t.d();

    const long np = ntr + nfr + nsh, ntri = ntr + nfr * 3 + nsh * 3;
    m.emod.assign(size_t(np), 0); m.yield.assign(size_t(np), 0);
    m.carea.assign(size_t(ntr + nfr), 0); m.llength.assign(size_t(ntr + nfr), 0);
    m.dens_tr.assign(size_t(ntr), 0);
    m.c1.assign(size_t(ntri), 0); m.c2.assign(size_t(ntri), 0); m.c3.assign(size_t(ntri), 0);
    for (long e = 0; e < ntr; ++e) {
        const double* a = &m.x[(m.minc[e * 2] - 1) * 3];
        const double* b = &m.x[(m.minc[e * 2 + 1] - 1) * 3];
        double el[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        double L = norm3(el);
        m.llength[e] = L;
        m.c1[e] = el[0] / L; m.c2[e] = el[1] / L; m.c3[e] = el[2] / L;
        m.emod[e] = t.d(); m.carea[e] = t.d(); m.dens_tr[e] = t.d(); m.yield[e] = t.d();
    }
    m.gmod.assign(size_t(nfr), 0); m.dens_fr.assign(size_t(nfr), 0);
    m.istrong.assign(size_t(nfr), 0); m.iweak.assign(size_t(nfr), 0);
    m.ipolar.assign(size_t(nfr), 0); m.iwarp.assign(size_t(nfr), 0);
    m.zstrong.assign(size_t(nfr), 0); m.zweak.assign(size_t(nfr), 0);
    m.offset.assign(size_t(nfr * 6), 0); m.auxpt.assign(size_t(nfr * 3), 0);
    m.xfr.assign(size_t(nfr * 6), 0); m.osflag.assign(size_t(nfr), 0);
    if (nfr > 0) {
        for (long e = 0; e < nfr; ++e) {
            m.emod[ntr + e] = t.d(); m.gmod[e] = t.d(); m.dens_fr[e] = t.d();
            m.carea[ntr + e] = t.d(); m.istrong[e] = t.d(); m.iweak[e] = t.d();
            m.ipolar[e] = t.d(); m.iwarp[e] = t.d();
        }
        for (;;) {
            long j = t.l();
            double o[6];
            for (double& v : o) v = t.d();
            if (t.bad()) FAIL("truncated frame offsets");
            if (j == 0) break;
            if (j < 1 || j > nfr) FAIL("frame offset member out of range");
            m.osflag[j - 1] = 1;
            std::copy(o, o + 6, &m.offset[(j - 1) * 6]);
        }
        for (long e = 0; e < nfr; ++e) {
            const double* a = &m.x[(m.minc[ntr * 2 + e * 2] - 1) * 3];
            const double* b = &m.x[(m.minc[ntr * 2 + e * 2 + 1] - 1) * 3];
            double* xe = &m.xfr[e * 6];
            for (int k = 0; k < 3; ++k) {
                xe[k] = a[k] + (m.osflag[e] ? m.offset[e * 6 + k] : 0.0);
                xe[3 + k] = b[k] + (m.osflag[e] ? m.offset[e * 6 + 3 + k] : 0.0);
            }
            double el[3] = {xe[3] - xe[0], xe[4] - xe[1], xe[5] - xe[2]};
            double L = norm3(el);
            m.llength[ntr + e] = L;
            for (int k = 0; k < 3; ++k) m.auxpt[e * 3 + k] = t.d();
            double lx[3] = {el[0] / L, el[1] / L, el[2] / L}, tmp[3], ly[3], lz[3];
            for (int k = 0; k < 3; ++k) tmp[k] = m.auxpt[e * 3 + k] - xe[k];
            cross3(lx, tmp, lz, true);
            cross3(lz, lx, ly, true);
            for (int k = 0; k < 3; ++k) {
                m.c1[ntr + e * 3 + k] = lx[k];
                m.c2[ntr + e * 3 + k] = ly[k];
                m.c3[ntr + e * 3 + k] = lz[k];
            }
        }
        for (long e = 0; e < nfr; ++e) {
            m.yield[ntr + e] = t.d(); m.zstrong[e] = t.d(); m.zweak[e] = t.d();
        }
        // member end bending releases (frame.c:193-220): strong axis end 1,
        // weak axis end 1, strong axis end 2, weak axis end 2
        m.relmask.assign(size_t(nfr), 0);
        for (;;) {
            long j = t.l();
            long r[4];
            for (long& v : r) v = t.l();
            if (t.bad()) FAIL("truncated frame releases");
            if (j == 0) break;
            if (j < 1 || j > nfr) FAIL("frame release member out of range");
            int32_t mask = 0;
            for (int k = 0; k < 4; ++k) {
                if (r[k] != 0 && r[k] != 1) FAIL("frame release flags are 0 or 1");
                mask |= int32_t(r[k]) << k;
            }
            m.relmask[size_t(j - 1)] = mask;
        }
    }
    m.nu.assign(size_t(nsh), 0); m.thick.assign(size_t(nsh), 0); m.dens_sh.assign(size_t(nsh), 0);
    m.xlocal.assign(size_t(nsh * 3), 0); m.farea.assign(size_t(nsh), 0);
    m.slength.assign(size_t(nsh * 3), 0);
    const long so = ntr * 2 + nfr * 2, to = ntr + nfr * 3, po = ntr + nfr;
    for (long e = 0; e < nsh; ++e) {
        m.emod[po + e] = t.d(); m.nu[e] = t.d(); m.thick[e] = t.d();
        m.dens_sh[e] = t.d(); m.yield[po + e] = t.d();
        const double* p1 = &m.x[(m.minc[so + e * 3] - 1) * 3];
        const double* p2 = &m.x[(m.minc[so + e * 3 + 1] - 1) * 3];
        const double* p3 = &m.x[(m.minc[so + e * 3 + 2] - 1) * 3];
        double e12[3], e23[3], e31[3], nrm[3];
        for (int k = 0; k < 3; ++k) {
            e12[k] = p2[k] - p1[k]; e23[k] = p3[k] - p2[k]; e31[k] = p3[k] - p1[k];
        }
        m.slength[e * 3] = norm3(e12);
        m.slength[e * 3 + 1] = norm3(e23);
        m.slength[e * 3 + 2] = norm3(e31);
        cross3(e12, e31, nrm, false);
        m.farea[e] = 0.5 * norm3(nrm);
        double lx[3], ly[3], lz[3];
        for (int k = 0; k < 3; ++k) {
            lx[k] = e12[k] / m.slength[e * 3];
            lz[k] = nrm[k] / (2 * m.farea[e]);
        }
        cross3(lz, lx, ly, true);
        for (int k = 0; k < 3; ++k) {
            m.c1[to + e * 3 + k] = lx[k];
            m.c2[to + e * 3 + k] = ly[k];
            m.c3[to + e * 3 + k] = lz[k];
        }
        m.xlocal[e * 3] = lx[0] * e12[0] + lx[1] * e12[1] + lx[2] * e12[2];
        m.xlocal[e * 3 + 1] = lx[0] * e31[0] + lx[1] * e31[1] + lx[2] * e31[2];
        m.xlocal[e * 3 + 2] = ly[0] * e31[0] + ly[1] * e31[1] + ly[2] * e31[2];
    }
    // one density array, filled by the truss, frame and shell readers in turn,
    // each from index 0 (truss.c:66, frame.c:61, shell.c:61)
    m.dens.assign(size_t(np > 0 ? np : 1), 0.0);
    for (long e = 0; e < ntr; ++e) m.dens[e] = m.dens_tr[e];
    for (long e = 0; e < nfr; ++e) m.dens[e] = m.dens_fr[e];
    for (long e = 0; e < nsh; ++e) m.dens[e] = m.dens_sh[e];
    if (t.bad()) FAIL("truncated element properties");

    const long neq = m.neq;
    m.q.assign(size_t(std::max(neq, 1L)), 0);
    m.um.assign(size_t(std::max(neq, 1L)), 0);
    m.vm = m.um;
    m.am = m.um;
    if (m.algflag > 3) {
        long ns = t.l();
        m.ttot = t.d();
        if (ns <= 0) FAIL("bad number of time steps");
        m.dt = m.ttot / double(ns);
        m.ntstpsinpt = ns + 1;
        if (m.chkpt == 0) m.chkpt = int(m.ntstpsinpt) + 1;
        m.pinpt.assign(size_t(neq * m.ntstpsinpt), 0);
        long prev = 0, i = 0;
        for (;;) {
            long jt = t.l();
            long dir = t.l();
            double mag = t.d();
            if (t.bad()) FAIL("truncated load histories");
            if (jt == 0) break;
            if (jt < 1 || jt > nj || dir < 1 || dir > 7) FAIL("load history joint out of range");
            // load histories keep the input joint number even under OPTFLAG 2:
            // model.c:1524 indexes jcode with it directly
            if (m.algflag == 5 && mag == 0) mag = 1e-30;
            long ks = m.jcode[(jt - 1) * 7 + dir - 1];
            if (ks != prev) i = 0;
            if (ks != 0 && i < m.ntstpsinpt) m.pinpt[(ks - 1) * m.ntstpsinpt + i] = mag;
            prev = ks;
            ++i;
        }
        long jt = t.l(); t.l(); t.d();
        if (jt != 0) FAIL("prescribed support displacements are not on this path");
        for (;;) {
            long j = t.l();
            long dir = t.l();
            double u0 = t.d(), v0 = t.d(), a0 = t.d();
            if (t.bad()) FAIL("truncated initial conditions");
            if (j == 0) break;
            if (j < 1 || j > nj || dir < 1 || dir > 7) FAIL("initial condition joint out of range");
            long k = m.jcode[(j - 1) * 7 + dir - 1];
            if (k) {
                if (u0 != 0) m.um[k - 1] = u0;
                if (v0 != 0) m.vm[k - 1] = v0;
                if (a0 != 0) m.am[k - 1] = a0;
            }
        }
    } else {
        bool any = false;
        for (;;) {
            long jt = t.l();
            long dir = t.l();
            double mag = t.d();
            if (t.bad()) FAIL("truncated joint loads");
            if (jt == 0) break;
            if (jt < 1 || jt > nj || dir < 1 || dir > 7) FAIL("joint load out of range");
            any = true;
            jt = m.jnt[size_t(jt - 1)] + 1;   // model.c:1318
            long k = m.jcode[(jt - 1) * 7 + dir - 1];
            if (k) m.q[k - 1] = mag;
        }
        if (nfr > 0) {
            long fr = t.l(); t.l(); t.d();
            if (fr != 0) FAIL("frame distributed loads are not on this path");
        }
        if (!any) FAIL("no forces present in input file");
    }
    if ((m.algflag == 1 || m.algflag == 2) && m.anaflag == 1) {
        m.lpfmax = t.d();                                   // main.c:1717
    } else if (m.algflag == 4) {
        m.numopt = int(t.d()); m.spectrds = t.d();          // main.c:3184-3197
        m.lpfmax = t.d();
    } else if (m.algflag == 1 || m.algflag == 2 || m.algflag == 5) {
        if (m.algflag == 5) { m.numopt = int(t.d()); m.spectrds = t.d(); }
        m.lpfmax = t.d(); m.lpf0 = t.d(); m.dlpf0 = t.d(); m.dlpfmax = t.d(); m.dlpfmin = t.d();
        m.itemax = int(t.l()); m.submax = int(t.l()); m.solmin = int(t.l());
        m.toldisp = t.d(); m.tolforc = t.d(); m.tolener = t.d();
    } else {
        long jt = t.l();
        long dir = t.l();
        m.dk = t.d();
        if (jt < 1 || jt > nj || dir < 1 || dir > 7) FAIL("bad arc-length control DOF");
        jt = m.jnt[size_t(jt - 1)] + 1;       // arc.c:49
        long k = m.jcode[(jt - 1) * 7 + dir - 1];
        if (!k) FAIL("initially prescribed displacement at fixed degree of freedom");
        m.dkdof = k - 1;
        m.alpha = t.d(); m.psi_thresh = t.d(); m.iteopt = int(t.l());
        m.lpfmax = t.d(); m.dkimax = t.d();
        m.itemax = int(t.l()); m.submax = int(t.l()); m.imagmax = int(t.l()); m.negmax = int(t.l());
        m.toldisp = t.d(); m.tolforc = t.d(); m.tolener = t.d();
    }
    if (t.bad()) FAIL("truncated solver parameters");
    if (neq == 0) FAIL("model has no free degrees of freedom");
#undef FAIL
    return true;
}

void build_pattern(Model& m) {
    // Joints carry 3 translation slots in truss-only models, 7 in frame models
    // with a free (shared) warping DOF, else 6.
    bool warping = false;
    for (long j = 0; j < m.nj; ++j) warping = warping || m.jcode[j * 7 + 6] != 0;
    m.bs = (m.nfr == 0 && m.nsh == 0) ? 3 : (warping ? 7 : 6);
    const int bs = m.bs;
    m.eq_of_slot.assign(size_t(m.nj * bs), -1);
    for (long j = 0; j < m.nj; ++j)
        for (int k = 0; k < bs; ++k) {
            long e = m.jcode[j * 7 + k];
            m.eq_of_slot[j * bs + k] = e ? int32_t(e - 1) : -1;
        }
    // joint adjacency through shared elements (including the joint itself)
    std::vector<std::vector<int32_t>> adj(size_t(m.nj));
    for (long j = 0; j < m.nj; ++j) adj[j].push_back(int32_t(j));
    auto link = [&](const long* v, int n) {
        for (int a = 0; a < n; ++a)
            for (int b = 0; b < n; ++b)
                if (a != b) adj[v[a] - 1].push_back(int32_t(v[b] - 1));
    };
    for (long e = 0; e < m.ntr; ++e) link(&m.minc[e * 2], 2);
    for (long e = 0; e < m.nfr; ++e) link(&m.minc[m.ntr * 2 + e * 2], 2);
    for (long e = 0; e < m.nsh; ++e) link(&m.minc[m.ntr * 2 + m.nfr * 2 + e * 3], 3);
    m.rowptr.assign(size_t(m.nj) + 1, 0);
    m.colidx.clear();
    m.diagblk.assign(size_t(m.nj), 0);
    for (long j = 0; j < m.nj; ++j) {
        auto& a = adj[j];
        std::sort(a.begin(), a.end());
        a.erase(std::unique(a.begin(), a.end()), a.end());
        m.rowptr[j + 1] = m.rowptr[j] + int32_t(a.size());
        for (int32_t c : a) {
            if (c == j) m.diagblk[j] = int32_t(m.colidx.size());
            m.colidx.push_back(c);
        }
    }
    auto blk = [&](long r, long c) -> int32_t {
        const int32_t* b = m.colidx.data() + m.rowptr[r];
        const int32_t* e = m.colidx.data() + m.rowptr[r + 1];
        return int32_t(std::lower_bound(b, e, int32_t(c)) - m.colidx.data());
    };
    auto fill_blocks = [&](long ne, int nn, long ioff, std::vector<int32_t>& out) {
        out.assign(size_t(ne * nn * nn), 0);
        for (long e = 0; e < ne; ++e)
            for (int a = 0; a < nn; ++a)
                for (int b = 0; b < nn; ++b)
                    out[e * nn * nn + a * nn + b] =
                        blk(m.minc[ioff + e * nn + a] - 1, m.minc[ioff + e * nn + b] - 1);
    };
    fill_blocks(m.ntr, 2, 0, m.blk_tr);
    fill_blocks(m.nfr, 2, m.ntr * 2, m.blk_fr);
    fill_blocks(m.nsh, 3, m.ntr * 2 + m.nfr * 2, m.blk_sh);

    // Greedy colouring in element order: an element takes the smallest colour not
    // used by any element already coloured at one of its joints.  The colour set
    // of a joint is a bitset of as many 64-bit words as the colours need, so a
    // joint of any valence (a hub of many bars) gets distinct colours for all
    // its elements.
    auto colour = [&](long ne, int nn, long ioff, std::vector<int32_t>& ptr, std::vector<int32_t>& els) {
        std::vector<std::vector<uint64_t>> used(size_t(m.nj));   // colour bitset per joint
        std::vector<std::vector<int32_t>> cls;
        std::vector<uint64_t> busy;
        for (long e = 0; e < ne; ++e) {
            busy.assign(cls.size() / 64 + 1, 0);
            for (int a = 0; a < nn; ++a) {
                const auto& u = used[size_t(m.minc[ioff + e * nn + a] - 1)];
                for (size_t w = 0; w < u.size(); ++w) busy[w] |= u[w];
            }
            size_t c = 0;
            while (busy[c / 64] >> (c % 64) & 1) ++c;   // busy has a free bit: cls.size() is free
            if (c >= cls.size()) cls.resize(c + 1);
            cls[c].push_back(int32_t(e));
            for (int a = 0; a < nn; ++a) {
                auto& u = used[size_t(m.minc[ioff + e * nn + a] - 1)];
                if (u.size() <= c / 64) u.resize(c / 64 + 1, 0);
                u[c / 64] |= uint64_t(1) << (c % 64);
            }
        }
        ptr.assign(1, 0);
        els.clear();
        for (auto& v : cls) {
            els.insert(els.end(), v.begin(), v.end());
            ptr.push_back(int32_t(els.size()));
        }
    };
    colour(m.ntr, 2, 0, m.col_ptr_tr, m.col_el_tr);
    colour(m.nfr, 2, m.ntr * 2, m.col_ptr_fr, m.col_el_fr);
    colour(m.nsh, 3, m.ntr * 2 + m.nfr * 2, m.col_ptr_sh, m.col_el_sh);

    // block rows, and the frame contributions of each block in colour order
    // (the order in which the colour launches add them)
    const size_t nnzb = m.colidx.size();
    m.blkrow.assign(nnzb, 0);
    for (long j = 0; j < m.nj; ++j)
        for (int32_t k = m.rowptr[j]; k < m.rowptr[j + 1]; ++k) m.blkrow[size_t(k)] = int32_t(j);
    m.gat_ptr_fr.assign(nnzb + 1, 0);
    for (long e = 0; e < m.nfr; ++e)
        for (int q = 0; q < 4; ++q) m.gat_ptr_fr[size_t(m.blk_fr[size_t(e * 4 + q)]) + 1]++;
    for (size_t k = 0; k < nnzb; ++k) m.gat_ptr_fr[k + 1] += m.gat_ptr_fr[k];
    m.gat_item_fr.assign(size_t(m.nfr * 4), 0);
    std::vector<int32_t> fill(m.gat_ptr_fr.begin(), m.gat_ptr_fr.end() - 1);
    for (int32_t e : m.col_el_fr)
        for (int q = 0; q < 4; ++q)
            m.gat_item_fr[size_t(fill[size_t(m.blk_fr[size_t(e) * 4 + q])]++)] = e * 4 + q;
    const long fo = m.ntr * 2;
    m.jnt_ptr_fr.assign(size_t(m.nj) + 1, 0);
    for (long e = 0; e < m.nfr; ++e)
        for (int a = 0; a < 2; ++a) m.jnt_ptr_fr[size_t(m.minc[size_t(fo + e * 2 + a)])]++;   // 1-based joint
    for (long j = 0; j < m.nj; ++j) m.jnt_ptr_fr[size_t(j) + 1] += m.jnt_ptr_fr[size_t(j)];
    m.jnt_item_fr.assign(size_t(m.nfr * 2), 0);
    std::vector<int32_t> jf(m.jnt_ptr_fr.begin(), m.jnt_ptr_fr.end() - 1);
    for (int32_t e : m.col_el_fr)
        for (int a = 0; a < 2; ++a)
            m.jnt_item_fr[size_t(jf[size_t(m.minc[size_t(fo + long(e) * 2 + a)] - 1)]++)] = e * 2 + a;
}

}  // namespace cb
