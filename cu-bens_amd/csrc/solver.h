// Linear solvers standing in for the reference's solve() (solve.c:51-89):
// device band LDL^T, and flexible GMRES with block-Jacobi, band-FP32 or learned
// preconditioners.
#pragma once

#include <functional>
#include <vector>

#include "band.h"
#include "dev.h"

namespace cb {

struct KrylovStats {
    long solves = 0, iters = 0, fallbacks = 0;
    double last_relres = 0, worst_relres = 0;
};

// Flexible GMRES(m) on the slot-layout system K x = b (right preconditioned,
// classical Gram-Schmidt with one re-orthogonalization).
struct Fgmres {
    long n = 0;
    int m = 30;
    double* V = nullptr;   // (m+1) x n
    double* Z = nullptr;   // m x n
    double* w = nullptr;
    double* hd = nullptr;  // device Hessenberg column scratch
    double* part = nullptr;
    double* hh = nullptr;  // pinned
    void setup(long n_, int m_);
    void release();
    // x <- approx K^{-1} b; returns iterations, relres via *relres.
    int solve(Dev& d, const std::function<void(const double*, double*)>& prec, const double* b,
              double* x, double tol, int maxit, double* relres);
};

// 6x6 (bs x bs) joint-block Jacobi.
struct BlockJacobi {
    double* inv = nullptr;   // nj * bs * bs
    void setup(Dev& d);
    void apply(Dev& d, const double* v, double* z);
    void release();
};

// Fused reductions used by the drivers (deterministic two-level sums).
// test(): |dd|^2, |d|^2, |lam q - f|^2, |lam q - fp|^2, dd.(lam q - f_ip)
void test_sums(Dev& d, const double* dd, const double* dt, const double* q, double lam,
               const double* f, const double* fp, const double* fip, long n, double out[5]);
// quad(): |dt-dp|^2, (dt-dp).ddr, (dt-dp).ddq
void quad_sums(Dev& d, const double* dt, const double* dp, const double* ddr, const double* ddq,
               long n, double out[3]);
// arc-length root update: dd = ddr + root ddq; dt = (dt - dp + dd) + dp
void quad_update(Dev& d, double* dt, const double* dp, const double* ddr, const double* ddq,
                 double* dd, double root, long n);

}  // namespace cb
