// Device band LDL^T of the tangent in the reference's equation order.
#pragma once

#include <utility>
#include <vector>

#include "dev.h"

namespace cb {

// Kernel families of the band solver whose launches can be timed with HIP
// events: KT_COLUMN is the fused per-column factorization launch; KT_DIAG,
// KT_PANEL and KT_UPDATE are the launches of the split schedule.
// KT_PREP times the launch that forms the sweeps' chain matrices after a
// factorization.
enum { KT_COLUMN = 0, KT_DIAG = 1, KT_PANEL = 2, KT_UPDATE = 3, KT_FWD = 4, KT_BWD = 5, KT_PREP = 6, KT_N = 7 };

// Launches of a selected family are bracketed by two events.  The fused
// factorization's nt launches are bracketed as one span (its launch period:
// kernel time plus the gaps between the launches, without event overhead per
// launch); of the split schedule's launches only every `every`-th is timed;
// the sweeps are always timed.  launches counts every launch issued and ms is
// the timed total scaled to all of them, so ms / launches is the average
// launch duration.
struct KernelClock {
    int mask = 0;                        // bit f: time family f
    int every = 1;
    double ms[KT_N] = {};
    long launches[KT_N] = {};
    double ms_timed[KT_N] = {};
    long n_timed[KT_N] = {};
    std::vector<hipEvent_t> ev;
    std::vector<int> fam;                // family of each recorded pair
    std::vector<long> count;             // launches covered by each recorded pair
    size_t used = 0;
    bool on(int f) const { return (mask >> f) & 1; }
    // whether to time this launch of family f
    bool take(int f) {
        if (!on(f)) return false;
        const long k = launches[f]++;
        const long e = (f == KT_FWD || f == KT_BWD || f == KT_PREP || every < 1) ? 1 : every;
        return k % e == 0;
    }
    void begin(int f, hipStream_t s, long n = 1);
    void end(hipStream_t s);
    void collect();                      // after a stream sync
    void release();
};

template <class R>
struct BandLDLT {
    KernelClock clock;
    long n = 0, nt = 0;
    int w = 0;
    hipStream_t s = nullptr;
    R* tiles = nullptr;
    R* tiles_t = nullptr;    // the factor's off-diagonal tiles transposed (row-major), for the backward sweep
    R* linv = nullptr;
    double* piv = nullptr;   // pivots D (padded to nt*T)
    // chain matrices of the sweeps, formed after each factorization: M_I =
    // Linv_I L_{I,I-1} and G_I = L_{I+1,I} Linv_I, [nt][T*T] each
    R* mch = nullptr;
    R* gch = nullptr;
    bool chained = true;     // sweeps with the chain matrices (default) or the two-product row step
    // the forward sweep two rows ahead (with chained): L' and W of the
    // look-ahead system [nt][w] and [nt] tiles, b' [nt * 64]; allocated at the
    // first factorization that needs them
    int look = 0;            // look-ahead depth of the sweeps (with chained): 0 none, 2 or 3
    R* l2 = nullptr;
    R* w2 = nullptr;
    double* bp = nullptr;
    // the backward sweep's look-ahead system: B [nt][w] and V [nt] tiles, z' [nt * 64]
    R* b2 = nullptr;
    R* v2 = nullptr;
    double* zp = nullptr;
    // depth 3: both systems deepened by one row ([nt][w+1] tiles each) and
    // their right-hand sides
    R* l3 = nullptr;
    R* b3 = nullptr;
    double* bq = nullptr;
    double* zq = nullptr;
    double* xw = nullptr;    // work vectors (padded)
    double* yw = nullptr;
    int* flag = nullptr;       // error bits: 1 a wait bound hit, 2 an entry outside the band
    int* pflags = nullptr;     // panel completion epochs of the fused factorization launch
    int cepoch = 0;            // epoch of the last fused launch
    // factorization schedule: one fused launch per tile column (default), or
    // separate diag/panel/update launches
    bool split = false;
    // the tangent comes from d.K32 (the FP32 joint blocks) instead of d.K
    bool src32 = false;
    std::vector<double> piv_host;

    void setup(long n_, long bandwidth, hipStream_t s_);
    void release();
    // Copies the lower band of d.K (d.K32 with src32; equation order) and factorizes.
    // 0 ok, 1 zero or non-finite pivot, 2 entry outside the band.
    int factor(const Dev& d);
    // out = (L D L^T)^{-1} rhs, both device NEQ vectors (may alias).
    void solve(const double* rhs, double* out);
    size_t bytes() const { return sizeof(R) * size_t(nt) * size_t(2 * (w + 1) + 2) * 64 * 64; }
};

// A batch of independent models sharing one FP64 band: model b owns the tile
// rows [t0, t0 + nt) (its equations start on a tile boundary, the unused
// equations of its last tile are identity rows) and has w tiles below the
// diagonal, w <= the band's.  One workgroup factors and solves one model, so
// the models of a batch run side by side in one launch.
struct BandSeg {
    long t0 = 0, nt = 0;
    long neq = 0;   // equations of the model (t0 * 64 is its first)
    long bw = 0;    // half-bandwidth: entries (i, j) with i - j > bw are zero
    int w = 0;
    int pad_ = 0;
};

struct BandBatch {
    long n = 0, nt = 0;
    int w = 0;
    int nseg = 0;
    hipStream_t s = nullptr;
    double* tiles = nullptr;
    double* linv = nullptr;
    double* piv = nullptr;
    double* yw = nullptr;           // forward-sweep values
    int* flag = nullptr;            // 1: a factor wait exceeded its bound; 2: an entry outside the band
    int* flag_host = nullptr;       // pinned copy, valid after the stream syncs
    BandSeg* segs = nullptr;        // device copy
    int* order = nullptr;           // factor workgroup -> model: pairs (main, helper) first, most work first
    int* role = nullptr;            // factor workgroup -> single, main or helper
    int* sorder = nullptr;          // solve workgroup -> model, most work first
    int* prog = nullptr;            // progress counters of the paired models
    int nwg = 0;                    // factor workgroups
    // models in decreasing order of their tile work; the helper pairs the
    // factor may use (models with two tiles below the diagonal, as many as
    // keep main and helper resident together) and the pairs it uses
    std::vector<int> work_order;
    int max_pairs = 0, npair = 0;
    int32_t* gaps = nullptr;        // equations that no DOF slot maps to
    long ngap = 0;
    std::vector<BandSeg> segs_host;
    // HIP events around the factor and the solve kernel of the last round
    // (recorded when timing is on): [0, 1] factor, [2, 3] solve
    bool timing = false;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};

    void setup(const std::vector<BandSeg>& sg, long n_total, int w_band, const std::vector<int32_t>& gap_eqs,
               hipStream_t s_);
    void release();
    // Copies the lower band of d.K and factors every model; enqueued only.
    void factor(const Dev& d);
    // The same in two parts, for a caller that writes the band itself:
    // zero() clears the tiles and the error bits; factor_filled() adds the
    // identity rows of the unused equations and factors, and copies the error
    // bits to flag_host unless fetch_flag is false (a caller that captures
    // the launches in a graph then calls fetch_flag() after it).  All device
    // work here is kernels: no memset node enters a captured graph.
    void zero();
    void factor_filled(bool fetch_flag = true);
    void fetch_flag();
    // The factor workgroups with np helper pairs (clamped to max_pairs): the
    // pairs first, then the other models, each part in decreasing tile work.
    // Every model's factor is the same whatever np is.
    void set_pairs(int np);
    // Times the factor of the band as it stands (filled, not yet factored) for
    // each candidate pair count, restoring the band from a copy before each
    // trial and taking the better of two runs, and keeps the fastest; returns
    // the (pairs, ms) of every trial (ms < 0: a wait bound was hit).
    std::vector<std::pair<int, double>> tune_pairs(const std::vector<int>& cand);
    // out = (L D L^T)^{-1} rhs for every model (device vectors of n, no alias).
    void solve(const double* rhs, double* out);
    size_t bytes() const { return sizeof(double) * size_t(nt) * size_t(w + 1) * 64 * 64; }
};

}  // namespace cb
