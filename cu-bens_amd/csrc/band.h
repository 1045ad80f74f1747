// Device band LDL^T of the tangent in the reference's equation order.
#pragma once

#include <vector>

#include "dev.h"

namespace cb {

template <class R>
struct BandLDLT {
    long n = 0, nt = 0;
    int w = 0;
    hipStream_t s = nullptr;
    R* tiles = nullptr;
    R* linv = nullptr;
    double* piv = nullptr;   // pivots D (padded to nt*T)
    double* xw = nullptr;    // work vectors (padded)
    double* yw = nullptr;
    int* flag = nullptr;       // error flag (bad band entry / wait bound hit)
    int* flags = nullptr;      // per tile row completion epochs: forward [nt], backward [nt]
    int epoch = 0;
    std::vector<double> piv_host;

    void setup(long n_, long bandwidth, hipStream_t s_);
    void release();
    // Copies the lower band of d.K (equation order) and factorizes.
    // 0 ok, 1 zero or non-finite pivot, 2 entry outside the band.
    int factor(const Dev& d);
    // out = (L D L^T)^{-1} rhs, both device NEQ vectors (may alias).
    void solve(const double* rhs, double* out);
    size_t bytes() const { return sizeof(R) * size_t(nt) * size_t(w + 1) * 64 * 64; }
};

}  // namespace cb
