// Checkpoint file of the nonlinear Newmark driver (checkPoint / restartStep,
// misc.c:494-716).  Lines, in order (misc.c:508-598):
//   tstep
//   NEQ x "uc,vc,ac", "0,0,0"
//   lss x "ss,sm", "0,0"
//   NEQ x "d,f"
//   ef and x one value a line; NE_TR+NE_FR*3+NE_SH*3 x "c1,c2,c3"
//   NE_TR x defllen; NE_FR x "defllen,llength"; efFE; xfr; yldflag (integers)
//   deffarea; defslen (the ANAFLAG 2 shell section); "0,0,0"
// then, in the full-precision file only, "cubens-exact", ddt, the working
// lengths, face areas and side lengths.
#include "checkpoint.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>

namespace cb {

void Checkpoint::size_for(const Model& m) {
    const size_t neq = size_t(m.neq), lss = size_t(m.lss), ntr = size_t(m.ntr), nfr = size_t(m.nfr);
    const size_t nsh = size_t(m.nsh);
    for (auto* v : {&uc, &vc, &ac, &d, &f}) v->assign(neq, 0.0);
    ss.assign(lss, 0.0);
    sm.assign(lss, 0.0);
    ef.assign(ntr * 2 + nfr * 14 + nsh * 18, 0.0);
    x.assign(size_t(m.nj) * 3, 0.0);
    for (auto* v : {&c1, &c2, &c3}) v->assign(ntr + nfr * 3 + nsh * 3, 0.0);
    defllen.assign(ntr + nfr, 0.0);
    llength.assign(ntr + nfr, 0.0);
    efFE.assign(nfr * 14, 0.0);
    xfr.assign(nfr * 6, 0.0);
    yldflag.assign(nfr * 2, 0);
    deffarea.assign(nsh, 0.0);
    defslen.assign(nsh * 3, 0.0);
    llen_w.assign(ntr + nfr, 0.0);
    farea_w.assign(nsh, 0.0);
    slen_w.assign(nsh * 3, 0.0);
}

namespace {

struct FileCloser {
    void operator()(FILE* f) const { if (f) std::fclose(f); }
};

void line(FILE* f, bool exact, std::initializer_list<double> v) {
    bool first = true;
    for (double x : v) {
        if (!first) std::fputc(',', f);
        std::fprintf(f, exact ? "%.17g" : "%e", x);
        first = false;
    }
    std::fputc('\n', f);
}

struct Reader {
    const char* p;
    bool bad = false;
    void skip() {
        while (*p == ',' || *p == '\n' || *p == '\r' || *p == ' ' || *p == '\t') ++p;
    }
    double num() {
        skip();
        char* e = nullptr;
        double v = std::strtod(p, &e);
        if (e == p) { bad = true; return 0; }
        p = e;
        return v;
    }
    void zeros(int n) {
        for (int i = 0; i < n; ++i)
            if (num() != 0) bad = true;
    }
};

}  // namespace

bool write_checkpoint(const std::string& path, const Model& m, const Checkpoint& c, bool exact) {
    std::unique_ptr<FILE, FileCloser> fp(std::fopen(path.c_str(), "w"));
    if (!fp) return false;
    FILE* f = fp.get();
    const long neq = m.neq, ntr = m.ntr, nfr = m.nfr, nsh = m.nsh;
    std::fprintf(f, "%ld\n", c.tstep);
    for (long i = 0; i < neq; ++i) line(f, exact, {c.uc[i], c.vc[i], c.ac[i]});
    std::fprintf(f, "0,0,0\n");
    for (long i = 0; i < m.lss; ++i) line(f, exact, {c.ss[i], c.sm[i]});
    std::fprintf(f, "0,0\n");
    for (long i = 0; i < neq; ++i) line(f, exact, {c.d[i], c.f[i]});
    for (double v : c.ef) line(f, exact, {v});
    for (double v : c.x) line(f, exact, {v});
    for (size_t i = 0; i < c.c1.size(); ++i) line(f, exact, {c.c1[i], c.c2[i], c.c3[i]});
    for (long i = 0; i < ntr; ++i) line(f, exact, {c.defllen[i]});
    for (long i = 0; i < nfr; ++i) line(f, exact, {c.defllen[ntr + i], c.llength[ntr + i]});
    for (double v : c.efFE) line(f, exact, {v});
    for (double v : c.xfr) line(f, exact, {v});
    for (int v : c.yldflag) std::fprintf(f, "%d\n", v);
    for (long i = 0; i < nsh; ++i) line(f, exact, {c.deffarea[i]});
    for (long i = 0; i < nsh * 3; ++i) line(f, exact, {c.defslen[i]});
    std::fprintf(f, "0,0,0\n");
    if (exact) {
        std::fprintf(f, "cubens-exact\n");
        line(f, true, {c.ddt});
        for (double v : c.llen_w) line(f, true, {v});
        for (double v : c.farea_w) line(f, true, {v});
        for (double v : c.slen_w) line(f, true, {v});
    }
    return std::fflush(f) == 0 && !std::ferror(f);
}

int read_checkpoint(const std::string& path, const Model& m, Checkpoint& c) {
    std::unique_ptr<FILE, FileCloser> fp(std::fopen(path.c_str(), "rb"));
    if (!fp) return 1;
    std::string text;
    char buf[1 << 16];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof buf, fp.get())) > 0) text.append(buf, got);
    c.size_for(m);
    Reader r{text.c_str()};
    const long neq = m.neq, ntr = m.ntr, nfr = m.nfr, nsh = m.nsh;
    c.tstep = long(r.num());
    for (long i = 0; i < neq; ++i) { c.uc[i] = r.num(); c.vc[i] = r.num(); c.ac[i] = r.num(); }
    r.zeros(3);
    for (long i = 0; i < m.lss; ++i) { c.ss[i] = r.num(); c.sm[i] = r.num(); }
    r.zeros(2);
    for (long i = 0; i < neq; ++i) { c.d[i] = r.num(); c.f[i] = r.num(); }
    for (double& v : c.ef) v = r.num();
    for (double& v : c.x) v = r.num();
    for (size_t i = 0; i < c.c1.size(); ++i) { c.c1[i] = r.num(); c.c2[i] = r.num(); c.c3[i] = r.num(); }
    for (long i = 0; i < ntr; ++i) c.defllen[i] = r.num();
    for (long i = 0; i < nfr; ++i) { c.defllen[ntr + i] = r.num(); c.llength[ntr + i] = r.num(); }
    for (double& v : c.efFE) v = r.num();
    for (double& v : c.xfr) v = r.num();
    for (int& v : c.yldflag) v = int(r.num());
    for (long i = 0; i < nsh; ++i) c.deffarea[i] = r.num();
    for (long i = 0; i < nsh * 3; ++i) c.defslen[i] = r.num();
    r.zeros(3);
    if (r.bad) return 2;
    r.skip();
    c.has_ext = std::strncmp(r.p, "cubens-exact", 12) == 0;
    if (c.has_ext) {
        r.p += 12;
        c.ddt = r.num();
        for (double& v : c.llen_w) v = r.num();
        for (double& v : c.farea_w) v = r.num();
        for (double& v : c.slen_w) v = r.num();
    }
    return r.bad ? 2 : 0;
}

bool write_checkpoint_linear(const std::string& path, long k, const std::vector<double>& u,
                             const std::vector<double>& v, const std::vector<double>& a, bool exact) {
    std::unique_ptr<FILE, FileCloser> fp(std::fopen(path.c_str(), "w"));
    if (!fp) return false;
    FILE* f = fp.get();
    std::fprintf(f, "%ld\n", k);
    for (size_t i = 0; i < u.size(); ++i) line(f, exact, {u[i], v[i], a[i]});
    std::fprintf(f, "0,0,0\n");
    return std::fflush(f) == 0 && !std::ferror(f);
}

int read_checkpoint_linear(const std::string& path, long neq, long* k, std::vector<double>& u,
                           std::vector<double>& v, std::vector<double>& a) {
    std::unique_ptr<FILE, FileCloser> fp(std::fopen(path.c_str(), "rb"));
    if (!fp) return 1;
    std::string text;
    char buf[1 << 16];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof buf, fp.get())) > 0) text.append(buf, got);
    Reader r{text.c_str()};
    u.assign(size_t(neq), 0.0);
    v.assign(size_t(neq), 0.0);
    a.assign(size_t(neq), 0.0);
    *k = long(r.num());
    for (long i = 0; i < neq; ++i) { u[size_t(i)] = r.num(); v[size_t(i)] = r.num(); a[size_t(i)] = r.num(); }
    r.zeros(3);
    return r.bad ? 2 : 0;
}

}  // namespace cb
