// Band LDL^T factorization and solves on the device: the reference's skyline
// solver (skyfact/skysolve, solve.c:539-698) recast for the GPU.
//
// The matrix is taken in the reference's equation order, so the pivots D are
// those of the reference's active-column factorization (same matrix, same
// order, no pivoting) up to rounding; they drive the arc-length predictor sign
// and the psi switch exactly as ssd does (solve.c:612-621, main.c:2636-2642).
//
// Storage: the lower band (half-bandwidth b = max column height) in dense T x T
// tiles; tile row I holds tiles J = I-w .. I (w = ceil(b/T)), column-major.
// Factorization is right-looking by tile column K:
//   diag : D_K and Linv_K = L_KK^{-1} of A_KK = L_KK D_K L_KK^T, in registers
//   panel: L_IK = A_IK Linv_K^T D_K^{-1},            K < I <= K+w
//   update: A_IJ -= (L_IK D_K) L_JK^T,               K < J <= I <= K+w
// Solves: pipelined forward and backward sweeps, one workgroup per tile row,
// with Linv_I for the diagonal blocks and pivot scaling folded into the
// backward sweep.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <type_traits>

#include "band.h"

namespace cb {
namespace {

constexpr int T = 64;
constexpr int TT = T * T;

template <class R, class KS = double>
__global__ void k_band_from_bsr(R* __restrict__ tiles, const KS* __restrict__ K,
                                const int32_t* __restrict__ rowptr, const int32_t* __restrict__ colidx,
                                const int32_t* __restrict__ eq, long nj, int bs, int w, int* __restrict__ bad) {
    long r = blockIdx.x * (long)blockDim.x + threadIdx.x;   // slot row
    if (r >= nj * bs) return;
    long A = r / bs;
    int p = int(r % bs);
    int32_t i = eq[r];
    if (i < 0) return;
    for (int32_t k = rowptr[A]; k < rowptr[A + 1]; ++k) {
        long B = colidx[k];
        for (int q = 0; q < bs; ++q) {
            int32_t j = eq[B * bs + q];
            if (j < 0 || j > i) continue;
            long I = i / T, J = j / T;
            if (I - J > w) { *bad = 2; continue; }
            tiles[(I * (w + 1) + (J - I + w)) * (long)TT + (j % T) * T + (i % T)] =
                R(K[(long)k * bs * bs + p * bs + q]);
        }
    }
}

template <class R>
__global__ void k_band_pad(R* __restrict__ tiles, long n, long nt, int w) {
    long i = n + blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= nt * T) return;
    long I = i / T;
    tiles[(I * (w + 1) + w) * (long)TT + (i % T) * T + (i % T)] = R(1);
}

// ---- factorization --------------------------------------------------------
// Tile tasks of column K:
//   D_K          D_K and Linv_K of tile (K,K)
//   P_{I,K}      L_IK = A_IK Linv_K^T D_K^{-1},            K < I <= K+m
//   U_{I,J,K}    A_IJ -= (L_IK D_K) L_JK^T,               K < J <= I <= K+m
// (m = min(w, nt-1-K)).  Two schedules run the same tile arithmetic:
//  - fused (default): one launch per column, k_band_column(K) for K = -1 ..
//    nt-2.  Workgroup 0 runs the critical chain P_{K+1,K} -> U_{K+1,K+1,K} ->
//    D_{K+1} entirely in its LDS; workgroups 1..m-1 run the other panels and
//    the rest run the updates, each waiting on the flags of the two panels it
//    reads.  A workgroup only waits on workgroups with a lower index, which
//    were dispatched before it, so every wait ends; waits are bounded anyway.
//  - split: three launches per column (diag, panels, updates), for timing
//    the tile tasks one by one.
// Tiles handed between workgroups of one launch (the panels L_IK) are stored
// and loaded with system-scope coherent accesses (written through to memory,
// read past the XCD caches): a producer waits for its stores to complete
// (vmcnt 0) before the flag store, a consumer reads the data after it has
// seen the flag.  Everything else crosses a launch boundary.
constexpr int NTH = 512;   // threads of every factorization workgroup (8 wavefronts)
constexpr int NWAVE = NTH / 64;

__device__ inline long tile_index(long I, long J, int w) { return I * (w + 1) + (J - I + w); }

template <bool CO, class R> __device__ __forceinline__ R cld(const R* p) {
    if constexpr (CO) return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    else return *p;
}
template <bool CO, class R> __device__ __forceinline__ void cst(R* p, R v) {
    if constexpr (CO) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    else *p = v;
}

// Panel flags sit one per 256-byte line so that pollers of different panels
// do not share a memory channel.
constexpr long FLAG_STRIDE = 64;

// thread 0 waits until flag i holds epoch (or an error is raised anywhere)
__device__ inline bool wait_flag_bounded(const int* flags, long i, int epoch, long spin_limit, int* err,
                                         int* ok_s) {
    if (threadIdx.x == 0) {
        long n = 0;
        bool good = true;
        const int* f = flags + i * FLAG_STRIDE;
        while (__hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) != epoch) {
            if (++n >= spin_limit) { good = false; break; }
            if ((n & 63) == 0 && (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) & 1)) {
                good = false;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        if (!good) atomicOr(err, 1);
        *ok_s = good;
    }
    __syncthreads();
    return *ok_s != 0;
}

__device__ inline void raise_flag(int* flags, long i, int epoch) {
    __builtin_amdgcn_s_waitcnt(0);   // this wave's coherent stores have completed
    __syncthreads();
    if (threadIdx.x == 0)
        __hip_atomic_store(flags + i * FLAG_STRIDE, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// ---- tile product C = X Y^T -------------------------------------------------
// Operands staged in LDS column-major: xs[k][row] = X[row][k], ys[k][col] =
// Y[col][k]; the result stays in registers and is handed out entry by entry.
template <class R> struct TileAcc;

typedef double d4 __attribute__((ext_vector_type(4)));

// FP64: FP64 MFMA 16x16x4, two 16x16 output blocks per wavefront.  The MFMA
// is fed Y as its A operand and X as its B operand (lane l: Y[16bj + l%16][k]
// and X[16bi + l%16][k], k = k0 + l/16), so element e of lane l is
// C[16bi + l%16][16bj + l/16 + 4e]: sixteen consecutive rows per column, which
// makes the column-major tile stores coalesced.
template <> struct TileAcc<double> {
    d4 acc[2];
    __device__ __forceinline__ void product(const double (*xs)[T], const double (*ys)[T]) {
        const int l = threadIdx.x & 63, wv = threadIdx.x >> 6;
        acc[0] = d4{0.0, 0.0, 0.0, 0.0};
        acc[1] = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
        for (int k0 = 0; k0 < T; k0 += 4) {
            const int k = k0 + (l >> 4);
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const int blk = 2 * wv + b, bi = blk & 3, bj = blk >> 2;
                acc[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ys[k][16 * bj + (l & 15)], xs[k][16 * bi + (l & 15)],
                                                             acc[b], 0, 0, 0);
            }
        }
    }
    // The same product for a lower-triangular Y (Y[col][k] = 0 for k > col):
    // the k-steps past the last column of the output block column add exact
    // zeros and are skipped.  Both blocks of wavefront wv lie in block column
    // wv / 2.
    __device__ __forceinline__ void product_ylower(const double (*xs)[T], const double (*ys)[T]) {
        const int l = threadIdx.x & 63, wv = threadIdx.x >> 6;
        const int nkb = (__builtin_amdgcn_readfirstlane(wv) >> 1) + 1;   // 16-wide k blocks
        acc[0] = d4{0.0, 0.0, 0.0, 0.0};
        acc[1] = d4{0.0, 0.0, 0.0, 0.0};
        for (int kb = 0; kb < nkb; ++kb) {
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int k = 16 * kb + 4 * t + (l >> 4);
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    const int blk = 2 * wv + b, bi = blk & 3, bj = blk >> 2;
                    acc[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ys[k][16 * bj + (l & 15)],
                                                                 xs[k][16 * bi + (l & 15)], acc[b], 0, 0, 0);
                }
            }
        }
    }
    static constexpr int N = 8;   // entries per thread
    template <class F> __device__ __forceinline__ void each_i(F&& f) const {
        const int l = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int blk = 2 * wv + b, bi = blk & 3, bj = blk >> 2;
#pragma unroll
            for (int e = 0; e < 4; ++e) f(16 * bi + (l & 15), 16 * bj + (l >> 4) + 4 * e, acc[b][e], 4 * b + e);
        }
    }
    template <class F> __device__ __forceinline__ void each(F&& f) const {
        each_i([&](int row, int col, double v, int) { f(row, col, v); });
    }
};

// FP32: vector FMA, thread (tr, tc) owns rows tr + 16u and columns tc + 32v.
template <> struct TileAcc<float> {
    float acc[4][2];
    __device__ __forceinline__ void product(const float (*xs)[T], const float (*ys)[T]) {
        const int tr = threadIdx.x & 15, tc = threadIdx.x >> 4;
#pragma unroll
        for (int u = 0; u < 4; ++u) acc[u][0] = acc[u][1] = 0.0f;
        for (int k = 0; k < T; ++k) {
            float av[4], bv[2];
#pragma unroll
            for (int u = 0; u < 4; ++u) av[u] = xs[k][tr + 16 * u];
#pragma unroll
            for (int v = 0; v < 2; ++v) bv[v] = ys[k][tc + 32 * v];
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 2; ++v) acc[u][v] += av[u] * bv[v];
        }
    }
    __device__ __forceinline__ void product_ylower(const float (*xs)[T], const float (*ys)[T]) { product(xs, ys); }
    static constexpr int N = 8;   // entries per thread
    template <class F> __device__ __forceinline__ void each_i(F&& f) const {
        const int tr = threadIdx.x & 15, tc = threadIdx.x >> 4;
#pragma unroll
        for (int v = 0; v < 2; ++v)
#pragma unroll
            for (int u = 0; u < 4; ++u) f(tr + 16 * u, tc + 32 * v, acc[u][v], 4 * v + u);
    }
    template <class F> __device__ __forceinline__ void each(F&& f) const {
        each_i([&](int row, int col, float v, int) { f(row, col, v); });
    }
};

// The tile entries at the positions TileAcc<R>::each_i hands out, in its order.
template <class R>
__device__ __forceinline__ void gather_tile(const R* __restrict__ tile, R* buf) {
    TileAcc<R> pos;
    pos.each_i([&](int row, int col, R, int i) { buf[i] = tile[col * T + row]; });
}

// C = X Y^T (operands as for TileAcc) for a symmetric C of which only the ten
// 16x16 blocks on and below the diagonal are wanted: one or two blocks per
// wavefront, three blocks' work on two of the four SIMDs and two on the
// others (wavefronts wv and wv + 4 share a SIMD), so the product takes 48
// MFMA steps per SIMD instead of 64.  Element e of slot s of lane l is
// C[16 bi_s + l % 16][16 bj_s + l / 16 + 4e], as in TileAcc<double>.
struct LowerAcc {
    d4 acc[2];
    int bi[2], bj[2], n;
    __device__ __forceinline__ LowerAcc() {
        const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        // per wavefront: (bi, bj) of slot 0, then of slot 1 (-1: none)
        constexpr signed char tab[8][4] = {{0, 0, 3, 3}, {1, 1, 2, 2}, {2, 1, -1, -1}, {3, 1, -1, -1},
                                           {1, 0, -1, -1}, {2, 0, -1, -1}, {3, 0, -1, -1}, {3, 2, -1, -1}};
        bi[0] = tab[wv][0];
        bj[0] = tab[wv][1];
        bi[1] = tab[wv][2];
        bj[1] = tab[wv][3];
        n = bi[1] >= 0 ? 2 : 1;
    }
    __device__ __forceinline__ void product(const double (*xs)[T], const double (*ys)[T]) {
        const int l = threadIdx.x & 63;
        acc[0] = d4{0.0, 0.0, 0.0, 0.0};
        acc[1] = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
        for (int k0 = 0; k0 < T; k0 += 4) {
            const int k = k0 + (l >> 4);
            acc[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(ys[k][16 * bj[0] + (l & 15)], xs[k][16 * bi[0] + (l & 15)],
                                                         acc[0], 0, 0, 0);
            if (n == 2)
                acc[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(ys[k][16 * bj[1] + (l & 15)],
                                                             xs[k][16 * bi[1] + (l & 15)], acc[1], 0, 0, 0);
        }
    }
    // f(row, col, value, i), i = 4 s + e
    template <class F> __device__ __forceinline__ void each_i(F&& f) const {
        const int l = threadIdx.x & 63;
#pragma unroll
        for (int s = 0; s < 2; ++s)
            if (s < n) {
#pragma unroll
                for (int e = 0; e < 4; ++e) f(16 * bi[s] + (l & 15), 16 * bj[s] + (l >> 4) + 4 * e, acc[s][e], 4 * s + e);
            }
    }
};

// ---- D_K --------------------------------------------------------------------
// The task takes the 64x64 tile S (lower part; the part above the diagonal is
// never read) in LDS as xs[col][row], writes D to piv and Linv = L^{-1} (zeros
// above the diagonal) to lout, and overwrites xs.  L_KK itself is read by
// neither the panels and updates nor the solves and is not stored.  FP32
// tiles go through the FP64 task (below).
template <class R> struct DiagScratch;
template <> struct DiagScratch<double> { double tsc[5][256]; double ub[16]; };
// FP64: blocked by 16.  For each column block k:
//  - wavefront 0 factorizes the 64x16 panel (rows 16k..63, lane = row):
//    16 pivot steps, each broadcasting column c of the block through LDS; the
//    pivot and the next step's critical entry travel by readlane, and the
//    reciprocal of the next pivot is formed while the rest of the step's
//    updates are issued.  The block lanes also form column p of Linv_kk by
//    the same elimination applied to the identity.  Multipliers are
//    l = a * (1/d) with the reciprocal refined by two Newton steps.
//  - the other wavefronts meanwhile form the off-diagonal blocks of Linv
//    that are ready, Linv_ij = -Linv_ii sum_{k=j}^{i-1} L_ik Linv_kj, as
//    16x16x16 FP64 MFMA products, avoiding wavefront 4 (which shares a SIMD
//    with wavefront 0);
//  - after a barrier, up to six wavefronts update the trailing blocks
//    A_ij -= L_ik D_k L_jk^T (16x16x16 FP64 MFMA).
// Linv is kept in ys column-major with the row index swizzled by the column
// (ysw), which makes both row and column accesses free of bank conflicts.

__device__ __forceinline__ int ysw(int col, int row) { return col * T + (row ^ (col & 15)); }

__device__ __forceinline__ double lane_value(double x, int lane) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, x);
    const unsigned lo = __builtin_amdgcn_readlane(unsigned(u), lane);
    const unsigned hi = __builtin_amdgcn_readlane(unsigned(u >> 32), lane);
    return __builtin_bit_cast(double, (unsigned long long)hi << 32 | lo);
}

__device__ __forceinline__ double recip(double d) {
    double r = __builtin_amdgcn_rcp(d);
    double e = __builtin_fma(-d, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-d, r, 1.0);
    return __builtin_fma(r, e, r);
}

// D[m][n] = sum_t A(m, t) B(t, n) over 16x16 blocks on one wavefront (FP64
// MFMA 16x16x4: lane l supplies A(l%16, t) and B(t, l%16), t = t0 + l/16);
// element e of lane l is D[l/16 + 4e][l%16].
template <class FA, class FB>
__device__ __forceinline__ d4 mm16(FA A, FB B, d4 acc) {
    const int l = threadIdx.x & 63, rr = l & 15, lg = l >> 4;
#pragma unroll
    for (int t0 = 0; t0 < 16; t0 += 4) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(A(rr, t0 + lg), B(t0 + lg, rr), acc, 0, 0, 0);
    return acc;
}

struct NoHook {
    __device__ void operator()() const {}
};

// tail_hook runs once the pivot steps are done (every thread), where few
// registers are live: a caller issues loads there that it consumes after the
// task.
template <class Hook = NoHook, class O = double>
__device__ __forceinline__ void diag_task(double (*xs)[T], double* ys, double* dd, DiagScratch<double>& sm,
                                          O* __restrict__ lout, double* __restrict__ piv, long Kt,
                                          double* lds_out = nullptr, Hook tail_hook = Hook()) {
    const int wv = threadIdx.x >> 6, l = threadIdx.x & 63, rr = l & 15, lg = l >> 4;
    double(*tsc)[256] = sm.tsc;
    // operand views: L_ik (rows r, cols t), Linv_ij as (t, c) and as (r, t), T slot s as (t, c)
    auto Lb = [&](int i, int k) { return [&, i, k](int r, int t) { return xs[16 * k + t][16 * i + r]; }; };
    auto Vb = [&](int i, int j) { return [&, i, j](int t, int c) { return ys[ysw(16 * j + c, 16 * i + t)]; }; };
    auto Vr = [&](int i, int j) { return [&, i, j](int r, int t) { return ys[ysw(16 * j + t, 16 * i + r)]; }; };
    auto Tb = [&](int s) { return [&, s](int t, int c) { return tsc[s][t * 16 + c]; }; };
    auto getT = [&](int s) {
        d4 acc;
#pragma unroll
        for (int e = 0; e < 4; ++e) acc[e] = tsc[s][(lg + 4 * e) * 16 + rr];
        return acc;
    };
    auto putT = [&](int s, d4 acc) {
#pragma unroll
        for (int e = 0; e < 4; ++e) tsc[s][(lg + 4 * e) * 16 + rr] = acc[e];
        __builtin_amdgcn_wave_barrier();
    };
    auto putV = [&](int i, int j, d4 acc) {   // Linv_ij = -acc
#pragma unroll
        for (int e = 0; e < 4; ++e) ys[ysw(16 * j + rr, 16 * i + lg + 4 * e)] = -acc[e];
        __builtin_amdgcn_wave_barrier();
    };
    const d4 z = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 1
    for (int k = 0; k < 4; ++k) {
        if (wv == 0) {
            const int r = l, base = 16 * k, p = r - base;
            double v[16], bcol[16];
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                v[q] = xs[base + q][r];
                bcol[q] = (q == p) ? 1.0 : 0.0;
            }
            if (p >= 0 && p < 16) sm.ub[p] = v[0];
            __builtin_amdgcn_wave_barrier();
            double d = lane_value(v[0], base);
            double rd = recip(d);
            double dr = 0.0;
            double U[16];
#pragma unroll
            for (int q = 0; q < 16; ++q) U[q] = sm.ub[q];
#pragma unroll
            for (int c = 0; c < 16; ++c) {
                int po = p;
                __asm__ volatile("" : "+v"(po));   // keeps the per-step lane tests in the step
                const bool below = po > c;
                dr = (po == c) ? d : dr;
                const double le = below ? v[c] * rd : 0.0;
                const double fb = below ? 0.0 : bcol[c] * rd;
                double dn = 0.0, rdn = 0.0, un = 0.0;
                if (c + 1 < 16) {
                    v[c + 1] = __builtin_fma(-le, U[c + 1], v[c + 1]);
                    __builtin_amdgcn_wave_barrier();
                    if (po >= 0 && po < 16) sm.ub[po] = v[c + 1];
                    __builtin_amdgcn_wave_barrier();
                    dn = lane_value(v[c + 1], base + c + 1);
                    if (c + 2 < 16) un = lane_value(v[c + 1], base + c + 2);
                    rdn = recip(dn);
                }
#pragma unroll
                for (int q = c + 2; q < 16; ++q) v[q] = __builtin_fma(-le, U[q], v[q]);
#pragma unroll
                for (int q = c + 1; q < 16; ++q) bcol[q] = __builtin_fma(-fb, U[q], bcol[q]);
                v[c] = below ? le : v[c];
                if (c + 1 < 16) {
#pragma unroll
                    for (int q = 0; q < 16; ++q) U[q] = (q == c + 2) ? un : sm.ub[q];
                    d = dn;
                    rd = rdn;
                }
                // keep every update in its own step (no deferred chains)
#pragma unroll
                for (int q = 0; q < 16; ++q) {
                    __asm__ volatile("" : "+v"(v[q]));
                    __asm__ volatile("" : "+v"(bcol[q]));
                }
            }
            if (p >= 0) {
#pragma unroll
                for (int q = 0; q < 16; ++q) xs[base + q][r] = v[q];
            }
            if (p >= 0 && p < 16) {
                dd[r] = dr;
                piv[Kt * T + r] = dr;
#pragma unroll
                for (int q = 0; q < 16; ++q) ys[ysw(base + p, base + q)] = (q >= p) ? bcol[q] : 0.0;
            }
        } else if (k == 1) {
            if (wv == 1) putT(0, mm16(Lb(1, 0), Vb(0, 0), z));                      // T_10
        } else if (k == 2) {
            if (wv == 1) {
                putV(1, 0, mm16(Vr(1, 1), Tb(0), z));                                // Linv_10
                putT(1, mm16(Lb(2, 1), Vb(1, 0), mm16(Lb(2, 0), Vb(0, 0), z)));      // T_20
            } else if (wv == 2) {
                putT(2, mm16(Lb(2, 1), Vb(1, 1), z));                                // T_21
            } else if (wv == 3) {
                putT(3, mm16(Lb(3, 0), Vb(0, 0), z));                                // T_30, first term
            } else if (wv == 5) {
                putT(4, mm16(Lb(3, 1), Vb(1, 1), z));                                // T_31, first term
            }
        } else if (k == 3) {
            if (wv == 1) putV(2, 0, mm16(Vr(2, 2), Tb(1), z));                       // Linv_20
            else if (wv == 2) putV(2, 1, mm16(Vr(2, 2), Tb(2), z));                  // Linv_21
            else if (wv == 3) putT(3, mm16(Lb(3, 1), Vb(1, 0), getT(3)));            // T_30, second term
            else if (wv == 5) putT(0, mm16(Lb(3, 2), Vb(2, 2), z));                  // T_32
        }
        __syncthreads();
        if (k == 3) break;
        // trailing update of the blocks k < j <= i <= 3, one per wavefront
        {
            const int nb = 3 - k;
            int pp = wv, i = -1, j = -1;
            for (int a = 0; a < nb && i < 0; ++a)
                for (int b = 0; b <= a; ++b) {
                    if (pp == 0) { i = k + 1 + a; j = k + 1 + b; break; }
                    --pp;
                }
            if (i >= 0) {
                // D[m = c][n = r] = sum_t L_jk[c][t] (L_ik[r][t] d_t)
                const d4 acc = mm16([&](int c, int t) { return xs[16 * k + t][16 * j + c]; },
                                    [&](int t, int r) { return xs[16 * k + t][16 * i + r] * dd[16 * k + t]; }, z);
#pragma unroll
                for (int e = 0; e < 4; ++e) xs[16 * j + lg + 4 * e][16 * i + rr] -= acc[e];
            }
        }
        __syncthreads();
    }
    tail_hook();
    // the last row of blocks: T_3j completed, Linv_3j = -Linv_33 T_3j
    if (wv == 0) {
        putT(1, mm16(Lb(3, 2), Vb(2, 0), getT(3)));
        putV(3, 0, mm16(Vr(3, 3), Tb(1), z));
    } else if (wv == 1) {
        putT(2, mm16(Lb(3, 2), Vb(2, 1), getT(4)));
        putV(3, 1, mm16(Vr(3, 3), Tb(2), z));
    } else if (wv == 2) {
        putV(3, 2, mm16(Vr(3, 3), Tb(0), z));
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < TT; idx += NTH) {
        const int col = idx / T, row = idx % T;
        const double v = (row >= col) ? ys[ysw(col, row)] : 0.0;
        lout[idx] = O(v);
        if (lds_out) lds_out[idx] = v;
    }
}

// FP32 tiles: the FP64 task on the tile widened in LDS, with Linv rounded once
// to FP32 and the pivots kept in FP64.  The FP64 task is blocked by 16 and
// carries each panel on one wavefront; a 64-step FP32 chain with a workgroup
// barrier per pivot made the FP32 column launch slower than the FP64 one.
template <> struct DiagScratch<float> { double xd[T][T]; double yd[TT]; DiagScratch<double> in; };

__device__ __forceinline__ void diag_task(float (*xs)[T], float* /*ys*/, double* dd, DiagScratch<float>& sm,
                                          float* __restrict__ lout, double* __restrict__ piv, long Kt) {
    for (int idx = threadIdx.x; idx < TT; idx += NTH) (&sm.xd[0][0])[idx] = double((&xs[0][0])[idx]);
    __syncthreads();
    diag_task(sm.xd, sm.yd, dd, sm.in, lout, piv, Kt);
}

// ---- staging ----------------------------------------------------------------
template <bool CO, class R>
__device__ __forceinline__ void stage(R (*dst)[T], const R* __restrict__ src) {
    for (int idx = threadIdx.x; idx < TT; idx += NTH) dst[idx / T][idx % T] = cld<CO>(&src[idx]);
}
// dst[k][row] = src[k][row] * d[k]
template <bool CO, class R>
__device__ __forceinline__ void stage_scaled(R (*dst)[T], const R* __restrict__ src, const double* __restrict__ d) {
    for (int idx = threadIdx.x; idx < TT; idx += NTH) dst[idx / T][idx % T] = cld<CO>(&src[idx]) * R(d[idx / T]);
}

// P_{I,K}: out = A_IK Linv_K^T D_K^{-1} (out may alias a); tout receives the
// transposed tile (row-major), which the backward sweep reads
template <bool CO, class R>
__device__ __forceinline__ void panel_task(R* __restrict__ out, const R* __restrict__ a, const R* __restrict__ lv,
                                           const double* __restrict__ dk, R (*xs)[T], R (*ys)[T],
                                           R* __restrict__ tout) {
    stage<false>(xs, a);
    stage<false>(ys, lv);
    __syncthreads();
    TileAcc<R> acc;
    acc.product(xs, ys);
    acc.each([&](int row, int col, R v) {
        const R p = v * (R(1) / R(dk[col]));
        cst<CO>(&out[col * T + row], p);
        tout[row * T + col] = p;
    });
}

// U_{I,J,K}: out -= (L_IK D_K) L_JK^T
template <bool CO, class R>
__device__ __forceinline__ void update_task(R* __restrict__ out, const R* __restrict__ lik, const R* __restrict__ ljk,
                                            const double* __restrict__ dk, R (*xs)[T], R (*ys)[T]) {
    stage_scaled<CO>(xs, lik, dk);
    stage<CO>(ys, ljk);
    __syncthreads();
    TileAcc<R> acc;
    acc.product(xs, ys);
    acc.each([&](int row, int col, R v) { out[col * T + row] -= v; });
}

// ---- split schedule ----------------------------------------------------------
template <class R>
__global__ void __launch_bounds__(NTH) k_tile_diag(const R* __restrict__ tiles, R* __restrict__ linv,
                                                   double* __restrict__ piv, long Kt, int w) {
    __shared__ R xs[T][T];
    __shared__ R ys[TT];
    __shared__ double dd[T];
    __shared__ DiagScratch<R> sm;
    stage<false>(xs, tiles + tile_index(Kt, Kt, w) * TT);
    __syncthreads();
    diag_task(xs, ys, dd, sm, linv + Kt * (long)TT, piv, Kt);
}

// P_{I,K} for I = Kt+1+blockIdx.x
template <class R>
__global__ void __launch_bounds__(NTH) k_tile_panel(R* __restrict__ tiles, R* __restrict__ tiles_t,
                                                    const R* __restrict__ linv, const double* __restrict__ piv,
                                                    long Kt, int w) {
    __shared__ R xs[T][T];
    __shared__ R ys[T][T];
    const long ti = tile_index(Kt + 1 + blockIdx.x, Kt, w) * TT;
    panel_task<false, R>(tiles + ti, tiles + ti, linv + Kt * (long)TT, piv + Kt * T, xs, ys, tiles_t + ti);
}

// (r, s), s <= r, of entry t of the triangle enumerated row by row
__device__ inline void tri_pair(long t, long& r, long& s) {
    r = long((sqrt(8.0 * double(t) + 1.0) - 1.0) / 2.0);
    while ((r + 1) * (r + 2) / 2 <= t) ++r;
    while (r * (r + 1) / 2 > t) --r;
    s = t - r * (r + 1) / 2;
}

// U_{I,J,K} for the pairs K < J <= I <= K+m, blockIdx.x -> (I, J) row by row
template <class R>
__global__ void __launch_bounds__(NTH) k_tile_update(R* __restrict__ tiles, const double* __restrict__ piv,
                                                     long Kt, int w) {
    __shared__ R xs[T][T];
    __shared__ R ys[T][T];
    long r, s;
    tri_pair(blockIdx.x, r, s);
    const long I = Kt + 1 + r, J = Kt + 1 + s;
    update_task<false, R>(tiles + tile_index(I, J, w) * TT, tiles + tile_index(I, Kt, w) * TT,
                          tiles + tile_index(J, Kt, w) * TT, piv + Kt * T, xs, ys);
}

// ---- fused schedule -----------------------------------------------------------
// Launch K (K = -1 .. nt-2), m = min(w, nt-1-K) (0 for K = -1):
//   workgroup 0          P_{K+1,K} and U_{K+1,K+1,K} in LDS only, then D_{K+1};
//                        raises flag w once it has read A_{K+1,K}
//   workgroups 1..m-1    P_{K+1+b,K}; raise flag b
//   workgroup m          P_{K+1,K} again (the same arithmetic), stored once
//                        flag w is up; raises flag 0
//   the rest             U_{I,J,K} for the pairs other than (K+1,K+1), two tiles
//                        of one tile row I per workgroup (they share L_IK)
// Grid: m + 1 + sum_{r=1}^{m-1} ceil((r+1)/2) workgroups for m > 0 (194 at
// m = 25, so that all of them are resident at once), 1 for m = 0.  Workgroup 0 neither
// stores nor waits before its diagonal step.
template <class R>
__global__ void __launch_bounds__(NTH) k_band_column(R* __restrict__ tiles, R* __restrict__ tiles_t,
                                                     R* __restrict__ linv,
                                                     double* __restrict__ piv, long K, int w, long nt,
                                                     int* flags, int epoch, long spin_limit, int* err) {
    __shared__ R xs[T][T];
    __shared__ R ys[T][T];
    __shared__ R zs[T][T];   // an update workgroup's second L_JK
    __shared__ double dk[T];
    // wait results: one slot per flag waited on
    __shared__ int ok_s[3];
    const long m = (K < 0) ? 0 : ((nt - 1 - K < w) ? nt - 1 - K : w);
    const long b = blockIdx.x;
    if (m > 0 && threadIdx.x < T) dk[threadIdx.x] = piv[K * T + threadIdx.x];
    if (b == 0) {
        __shared__ DiagScratch<R> sm;
        const long D = K + 1;
        const R* dtile = tiles + tile_index(D, D, w) * TT;
        if (m == 0) {
            stage<false>(xs, dtile);
            __syncthreads();
            diag_task(xs, &ys[0][0], dk, sm, linv + D * (long)TT, piv, D);
            return;
        }
        R pre[TileAcc<R>::N];
        if constexpr (std::is_same<R, double>::value) {
            LowerAcc pos;
            pos.each_i([&](int row, int col, double, int i) { pre[i] = dtile[col * T + row]; });
        } else {
            gather_tile(dtile, pre);
        }
        stage<false>(xs, tiles + tile_index(D, K, w) * TT);
        stage<false>(ys, linv + K * (long)TT);
        raise_flag(flags, w, epoch);   // A_{K+1,K} has been read (includes the barrier)
        TileAcc<R> acc;
        acc.product_ylower(xs, ys);   // Linv_K is lower triangular
        __syncthreads();
        // (L D, L) of L = P_{K+1,K} to LDS for U_{K+1,K+1,K}
        acc.each([&](int row, int col, R v) {
            const R lv = v * (R(1) / R(dk[col]));
            xs[col][row] = lv * R(dk[col]);
            ys[col][row] = lv;
        });
        __syncthreads();
        // U_{K+1,K+1,K} is symmetric and the diagonal task reads only its
        // lower part
        if constexpr (std::is_same<R, double>::value) {
            LowerAcc u;
            u.product(xs, ys);
            __syncthreads();
            u.each_i([&](int row, int col, double v, int i) { xs[col][row] = pre[i] - v; });
        } else {
            acc.product(xs, ys);
            __syncthreads();
            acc.each_i([&](int row, int col, R v, int i) { xs[col][row] = pre[i] - v; });
        }
        __syncthreads();
        diag_task(xs, &ys[0][0], dk, sm, linv + D * (long)TT, piv, D);
        return;
    }
    __syncthreads();   // dk
    if (b <= m) {
        const long pb = (b == m) ? 0 : b;
        const long I = K + 1 + pb;
        R* tile = tiles + tile_index(I, K, w) * TT;
        R* ttile = tiles_t + tile_index(I, K, w) * TT;
        // stage (workgroup m then waits until workgroup 0 has read the tile
        // before overwriting it), product, coherent stores, flag; the
        // transposed copy for the backward sweep is stored after the flag
        stage<false>(xs, tile);
        stage<false>(ys, linv + K * (long)TT);
        bool ok = true;
        if (pb == 0) ok = wait_flag_bounded(flags, w, epoch, spin_limit, err, &ok_s[0]);
        else __syncthreads();
        TileAcc<R> acc;
        acc.product_ylower(xs, ys);   // Linv_K is lower triangular
        if (ok) acc.each([&](int row, int col, R v) { cst<true>(&tile[col * T + row], v * (R(1) / R(dk[col]))); });
        raise_flag(flags, pb, epoch);
        if (ok) acc.each([&](int row, int col, R v) { ttile[row * T + col] = v * (R(1) / R(dk[col])); });
        return;
    }
    // row r >= 1 of the update triangle holds ceil((r+1)/2) workgroups
    long t = b - (m + 1), r = 1;
    for (long c = (r + 2) / 2; t >= c; c = (r + 2) / 2) {
        t -= c;
        ++r;
    }
    const long s0 = 2 * t, ns = (s0 + 1 <= r) ? 2 : 1;
    // the (up to three) panel flags of the pair, polled side by side: lane i
    // of wavefront 0 polls flag i, so the waits cost one round trip together
    if (threadIdx.x < 3) {
        const int i = threadIdx.x;
        const long f = (i == 0) ? r : s0 + (i - 1);
        const bool need = (i == 0) || (i - 1 < ns && f != r);
        bool good = true;
        if (need) {
            const int* fp = flags + f * FLAG_STRIDE;
            long n = 0;
            while (__hip_atomic_load(fp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) != epoch) {
                if (++n >= spin_limit) { good = false; break; }
                if ((n & 63) == 0 && (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) & 1)) {
                    good = false;
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
            if (!good) atomicOr(err, 1);
        }
        ok_s[i] = good;
    }
    __syncthreads();
    if (!(ok_s[0] && ok_s[1] && ok_s[2])) return;
    const long I = K + 1 + r;
    // L_IK D_K and both L_JK staged together
    stage_scaled<true>(xs, tiles + tile_index(I, K, w) * TT, dk);
    stage<true>(ys, tiles + tile_index(K + 1 + s0, K, w) * TT);
    if (ns == 2) stage<true>(zs, tiles + tile_index(K + 2 + s0, K, w) * TT);
    __syncthreads();
    for (long j = 0; j < ns; ++j) {
        TileAcc<R> acc;
        acc.product(xs, j == 0 ? ys : zs);
        R* out = tiles + tile_index(I, K + 1 + s0 + j, w) * TT;
        acc.each([&](int row, int col, R v) { out[col * T + row] -= v; });
    }
}

// ---- pipelined sweeps ---------------------------------------------------------
// One workgroup per tile row, launched in tile order (forward) or reverse tile
// order (backward).  A workgroup waits only on workgroups with a lower index,
// which the hardware dispatched before it, so the oldest unfinished workgroup
// is always resident and the sweep makes progress.  A tile row's values are
// their own completion signal: before a sweep every slot holds SENT, a
// signalling-NaN bit pattern that no arithmetic produces; the producer writes
// the 64 values with system-scope stores (written through to memory) and a
// consumer polls them with system-scope loads until none is SENT.  There is
// no flag and no cache maintenance on the critical path.  Every lane polls
// its own value, one load at a time (several in flight, which sees a value
// sooner, made the sweeps slower: about 26 consumers poll each row, so the
// polls crowd the row's memory channel); polling slows down once a wait has
// lasted long, so the workgroups far ahead of the front load memory little.
// Each wait is bounded: past spin_limit polls the value reads as
// a quiet NaN and *err is raised, so every workgroup still finishes and the
// sweep drains.  Contributions of earlier rows are applied while the newest
// one is still pending, four rows at a time: each wavefront polls one of
// them, so a workgroup that is catching up pays one memory round trip per
// four rows rather than per row (one per row made the workgroups finish
// their older rows only just before the newest one arrived).  The newest
// off-diagonal tile and Linv stay in registers, so the critical path per row
// is one store-to-load hand-off and two 64x64 products.
constexpr unsigned long long SENT = 0x7FF4C0DE5E471E11ull;

__device__ __forceinline__ double poll_value(const double* p, long spin_limit, int* err) {
    const unsigned long long* q = reinterpret_cast<const unsigned long long*>(p);
    auto ld = [q] { return __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM); };
    for (long n = 0;; ++n) {
        const unsigned long long a = ld();
        if (a != SENT) return __builtin_bit_cast(double, a);
        if (n >= spin_limit) {
            atomicOr(err, 1);
            return __builtin_nan("");
        }
        if (n >= 256) __builtin_amdgcn_s_sleep(4);   // far from the front: poll less often
    }
}

// vin[0..63] = the values of tile row J of v (wave 0 polls, a value per lane)
__device__ __forceinline__ void fetch_row(const double* v, long J, double* vin, long spin_limit, int* err) {
    if (threadIdx.x < T) vin[threadIdx.x] = poll_value(v + J * T + threadIdx.x, spin_limit, err);
    __syncthreads();
}

// Rows that the sweeps' older contributions need, polled side by side:
// wavefront c polls row J + dir * c (c < nc) into vch[c], so a chunk costs
// one memory round trip instead of one per row.
constexpr int NCH = 4;
__device__ __forceinline__ void fetch_rows(const double* v, long J, int dir, int nc, double (*vch)[T],
                                           long spin_limit, int* err) {
    const int c = threadIdx.x >> 6;
    if (c < nc) vch[c][threadIdx.x & 63] = poll_value(v + (J + dir * c) * T + (threadIdx.x & 63), spin_limit, err);
    __syncthreads();
}

__device__ __forceinline__ void put_value(double* p, double x) {
    __hip_atomic_store(p, x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// y_I = Linv_I (r_I - sum_{J<I} L_IJ y_J).  r is read from xr, whose slot I
// is set to SENT for the backward sweep; y must hold SENT.
template <class R>
__global__ void __launch_bounds__(256) k_fwd_sweep(const R* __restrict__ tiles, const R* __restrict__ linv,
                                                   double* __restrict__ xr, double* y, long nt, int w,
                                                   long spin_limit, int* err) {
    __shared__ double vin[T];
    __shared__ double vch[NCH][T];
    __shared__ double part[4][T];
    const long I = blockIdx.x;
    const int tid = threadIdx.x, row = tid % T, q = tid / T, k0 = 16 * q;
    R lv[16], ln[16];
    const R* lvp = linv + I * (long)TT;
#pragma unroll
    for (int k = 0; k < 16; ++k) lv[k] = lvp[(k0 + k) * T + row];
    const long J0 = I - w > 0 ? I - w : 0;
    const bool has_prev = J0 <= I - 1;
    if (has_prev) {
        const R* lnp = tiles + (I * (w + 1) + (w - 1)) * (long)TT;   // tile (I, I-1)
#pragma unroll
        for (int k = 0; k < 16; ++k) ln[k] = lnp[(k0 + k) * T + row];
    }
    double acc = 0.0;
    if (q == 0) {
        acc = xr[I * T + row];
        xr[I * T + row] = __builtin_bit_cast(double, SENT);
    }
    // the older rows J < I - 1 in chunks of NCH, wavefront c polling row J + c
    for (long J = J0; J < I - 1; J += NCH) {
        const int nc = (I - 1 - J < NCH) ? int(I - 1 - J) : NCH;
        R lx[NCH][16];
#pragma unroll
        for (int c = 0; c < NCH; ++c)
            if (c < nc) {
                const R* l = tiles + (I * (w + 1) + (J + c - I + w)) * (long)TT;
#pragma unroll
                for (int k = 0; k < 16; ++k) lx[c][k] = l[(k0 + k) * T + row];
            }
        fetch_rows(y, J, 1, nc, vch, spin_limit, err);
#pragma unroll
        for (int c = 0; c < NCH; ++c)
            if (c < nc) {
#pragma unroll
                for (int k = 0; k < 16; ++k) acc -= double(lx[c][k]) * vch[c][k0 + k];
            }
        __syncthreads();
    }
    if (has_prev) {
        fetch_row(y, I - 1, vin, spin_limit, err);
#pragma unroll
        for (int k = 0; k < 16; ++k) acc -= double(ln[k]) * vin[k0 + k];
    }
    // every thread has read vin before it passes this barrier, so vin may be
    // overwritten after it
    part[q][row] = acc;
    __syncthreads();
    if (tid < T) vin[tid] = part[0][tid] + part[1][tid] + part[2][tid] + part[3][tid];
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < 16; ++k) s += double(lv[k]) * vin[k0 + k];
    part[q][row] = s;
    __syncthreads();
    if (tid < T) put_value(&y[I * T + tid], part[0][tid] + part[1][tid] + part[2][tid] + part[3][tid]);
}

// x_I = Linv_I^T (y_I / D_I - sum_{J>I} L_JI^T x_J); workgroup g handles
// I = nt-1-g.  x must hold SENT.  The factor tiles are read from their
// transposed copies, so that consecutive lanes read consecutive entries.
template <class R>
__global__ void __launch_bounds__(256) k_bwd_sweep(const R* __restrict__ tiles_t, const R* __restrict__ linv,
                                                   const double* __restrict__ y, const double* __restrict__ piv,
                                                   double* x, long nt, int w, long spin_limit, int* err) {
    __shared__ double vin[T];
    __shared__ double vch[NCH][T];
    __shared__ double part[4][T];
    const long I = nt - 1 - blockIdx.x;
    const int tid = threadIdx.x, col = tid % T, q = tid / T, k0 = 16 * q;
    R lv[16], ln[16];
    const R* lvp = linv + I * (long)TT;
#pragma unroll
    for (int k = 0; k < 16; ++k) lv[k] = lvp[col * T + k0 + k];   // Linv[k][col]
    const long J1 = (I + w < nt - 1) ? I + w : nt - 1;
    const bool has_next = I + 1 <= J1;
    if (has_next) {
        const R* lnp = tiles_t + ((I + 1) * (w + 1) + (w - 1)) * (long)TT;   // tile (I+1, I), transposed
#pragma unroll
        for (int k = 0; k < 16; ++k) ln[k] = lnp[(k0 + k) * T + col];
    }
    double acc = (q == 0) ? y[I * T + col] / piv[I * T + col] : 0.0;
    // the later rows J > I + 1 in chunks of NCH, wavefront c polling row J - c
    for (long J = J1; J > I + 1; J -= NCH) {
        const int nc = (J - (I + 1) < NCH) ? int(J - (I + 1)) : NCH;
        R lx[NCH][16];
#pragma unroll
        for (int c = 0; c < NCH; ++c)
            if (c < nc) {
                const R* l = tiles_t + ((J - c) * (w + 1) + (I - (J - c) + w)) * (long)TT;
#pragma unroll
                for (int k = 0; k < 16; ++k) lx[c][k] = l[(k0 + k) * T + col];
            }
        fetch_rows(x, J, -1, nc, vch, spin_limit, err);
#pragma unroll
        for (int c = 0; c < NCH; ++c)
            if (c < nc) {
#pragma unroll
                for (int k = 0; k < 16; ++k) acc -= double(lx[c][k]) * vch[c][k0 + k];
            }
        __syncthreads();
    }
    if (has_next) {
        fetch_row(x, I + 1, vin, spin_limit, err);
#pragma unroll
        for (int k = 0; k < 16; ++k) acc -= double(ln[k]) * vin[k0 + k];
    }
    part[q][col] = acc;
    __syncthreads();
    if (tid < T) vin[tid] = part[0][tid] + part[1][tid] + part[2][tid] + part[3][tid];
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < 16; ++k) s += double(lv[k]) * vin[k0 + k];
    part[q][col] = s;
    __syncthreads();
    if (tid < T) put_value(&x[I * T + tid], part[0][tid] + part[1][tid] + part[2][tid] + part[3][tid]);
}

// ---- sweeps with a precomputed chain matrix ----------------------------------
// In the sweeps above the last step of a tile row waits for its predecessor
// and then applies two dependent 64x64 products (L_{I,I-1} y_{I-1}, then Linv_I).
// After each factorization k_sweep_chain forms per tile row
//   M_I = Linv_I L_{I,I-1}            (column-major, for the forward sweep)
//   G_I = L_{I+1,I} Linv_I            (row-major, for the backward sweep)
// so that y_I = U_I - M_I y_{I-1}, U_I = Linv_I (b_I - sum_{J <= I-2} L_IJ y_J),
// and x_I = U'_I - G_I^T x_{I+1}, U'_I = Linv_I^T (y_I / D_I - sum_{J >= I+2} L_JI^T x_J).
// U_I needs the rows up to I-2 only, which arrive one chain step before
// y_{I-1}; once y_{I-1} lands the row takes one 64x64 product and one
// reduction.  Each workgroup still reads only its own tile row (the older
// contributions' bandwidth per workgroup is that of the sweeps above).

// tile (I, J) of the factor, I > J, read as doubles; zero outside the band
template <class R> struct TileRd {
    const R* p;
    __device__ double operator()(int r, int c) const { return p ? double(p[c * T + r]) : 0.0; }
};
template <class R> __device__ inline TileRd<R> factor_tile(const R* tiles, long I, long J, long nt, int w) {
    if (J < 0 || I >= nt || I - J > w || I - J < 1) return TileRd<R>{nullptr};
    return TileRd<R>{tiles + tile_index(I, J, w) * TT};
}
template <class R> __device__ inline TileRd<R> linv_tile(const R* linv, long I, long nt) {
    return TileRd<R>{(I >= 0 && I < nt) ? linv + I * TT : nullptr};
}

// acc[j] += sum_k A(r, k) B(k, c0 + j), j < 16 (B's reads are uniform over a wavefront)
template <class FA, class FB> __device__ inline void mm64(const FA& A, const FB& B, int r, int c0, double* acc) {
    for (int k = 0; k < T; ++k) {
        const double a = A(r, k);
#pragma unroll
        for (int j = 0; j < 16; ++j) acc[j] += a * B(k, c0 + j);
    }
}

// M_I and G_I for every tile row: one workgroup per row, thread (r, c0)
// forming entries (r, c0..c0+15) of each product.
template <class R>
__global__ void __launch_bounds__(256) k_sweep_chain(const R* __restrict__ tiles, const R* __restrict__ linv,
                                                     R* __restrict__ mch, R* __restrict__ gch, long nt, int w) {
    const long I = blockIdx.x;
    const int r = threadIdx.x % T, c0 = 16 * (threadIdx.x / T);
    double acc[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) acc[j] = 0.0;
    mm64(linv_tile(linv, I, nt), factor_tile(tiles, I, I - 1, nt, w), r, c0, acc);
#pragma unroll
    for (int j = 0; j < 16; ++j) mch[I * TT + (c0 + j) * T + r] = R(acc[j]);
#pragma unroll
    for (int j = 0; j < 16; ++j) acc[j] = 0.0;
    mm64(factor_tile(tiles, I + 1, I, nt, w), linv_tile(linv, I, nt), r, c0, acc);
#pragma unroll
    for (int j = 0; j < 16; ++j) gch[I * TT + r * T + c0 + j] = R(acc[j]);
}

// y_I = U_I - M_I y_{I-1}.  r is read from xr, whose slot I is set to SENT for
// the backward sweep; y must hold SENT.
template <class R>
__global__ void __launch_bounds__(256) k_fwd_sweep_m(const R* __restrict__ tiles, const R* __restrict__ linv,
                                                     const R* __restrict__ mch, double* __restrict__ xr, double* y,
                                                     long nt, int w, long spin_limit, int* err) {
    __shared__ double vin[T];
    __shared__ double vch[NCH][T];
    __shared__ double part[4][T];
    const long I = blockIdx.x;
    const int tid = threadIdx.x, row = tid % T, q = tid / T, k0 = 16 * q;
    R lv[16], lm[16];
    const R* lvp = linv + I * (long)TT;
#pragma unroll
    for (int k = 0; k < 16; ++k) lv[k] = lvp[(k0 + k) * T + row];
    const bool has_prev = I >= 1;
    if (has_prev) {
        const R* lmp = mch + I * (long)TT;
#pragma unroll
        for (int k = 0; k < 16; ++k) lm[k] = lmp[(k0 + k) * T + row];
    }
    double acc = 0.0;
    if (q == 0) {
        acc = xr[I * T + row];
        xr[I * T + row] = __builtin_bit_cast(double, SENT);
    }
    // the older rows J <= I - 2 in chunks of NCH, wavefront c polling row J + c
    const long J0 = I - w > 0 ? I - w : 0;
    for (long J = J0; J < I - 1; J += NCH) {
        const int nc = (I - 1 - J < NCH) ? int(I - 1 - J) : NCH;
        R lx[NCH][16];
#pragma unroll
        for (int c = 0; c < NCH; ++c)
            if (c < nc) {
                const R* l = tiles + (I * (w + 1) + (J + c - I + w)) * (long)TT;
#pragma unroll
                for (int k = 0; k < 16; ++k) lx[c][k] = l[(k0 + k) * T + row];
            }
        fetch_rows(y, J, 1, nc, vch, spin_limit, err);
#pragma unroll
        for (int c = 0; c < NCH; ++c)
            if (c < nc) {
#pragma unroll
                for (int k = 0; k < 16; ++k) acc -= double(lx[c][k]) * vch[c][k0 + k];
            }
        __syncthreads();
    }
    // U_I = Linv_I S_I, formed before y_{I-1} is needed
    part[q][row] = acc;
    __syncthreads();
    if (tid < T) vin[tid] = part[0][tid] + part[1][tid] + part[2][tid] + part[3][tid];
    __syncthreads();
    double u = 0.0;
#pragma unroll
    for (int k = 0; k < 16; ++k) u += double(lv[k]) * vin[k0 + k];
    __syncthreads();
    part[q][row] = u;
    __syncthreads();
    double yi = (q == 0) ? part[0][row] + part[1][row] + part[2][row] + part[3][row] : 0.0;
    if (has_prev) {
        fetch_row(y, I - 1, vin, spin_limit, err);
        double m = 0.0;
#pragma unroll
        for (int k = 0; k < 16; ++k) m += double(lm[k]) * vin[k0 + k];
        part[q][row] = m;
        __syncthreads();
        if (q == 0) yi -= part[0][row] + part[1][row] + part[2][row] + part[3][row];
    }
    if (q == 0) put_value(&y[I * T + row], yi);
}

// x_I = U'_I - G_I^T x_{I+1}; workgroup g handles I = nt-1-g.  x must hold SENT.
template <class R>
__global__ void __launch_bounds__(256) k_bwd_sweep_m(const R* __restrict__ tiles_t, const R* __restrict__ linv,
                                                     const R* __restrict__ gch, const double* __restrict__ y,
                                                     const double* __restrict__ piv, double* x, long nt, int w,
                                                     long spin_limit, int* err) {
    __shared__ double vin[T];
    __shared__ double vch[NCH][T];
    __shared__ double part[4][T];
    const long I = nt - 1 - blockIdx.x;
    const int tid = threadIdx.x, col = tid % T, q = tid / T, k0 = 16 * q;
    R lv[16], lm[16];
    const R* lvp = linv + I * (long)TT;
#pragma unroll
    for (int k = 0; k < 16; ++k) lv[k] = lvp[col * T + k0 + k];   // Linv[k][col]
    const bool has_next = I + 1 < nt;
    if (has_next) {
        const R* lgp = gch + I * (long)TT;
#pragma unroll
        for (int k = 0; k < 16; ++k) lm[k] = lgp[(k0 + k) * T + col];   // G_I[k][col]
    }
    double acc = (q == 0) ? y[I * T + col] / piv[I * T + col] : 0.0;
    // the later rows J >= I + 2 in chunks of NCH, wavefront c polling row J - c
    const long J1 = (I + w < nt - 1) ? I + w : nt - 1;
    for (long J = J1; J > I + 1; J -= NCH) {
        const int nc = (J - (I + 1) < NCH) ? int(J - (I + 1)) : NCH;
        R lx[NCH][16];
#pragma unroll
        for (int c = 0; c < NCH; ++c)
            if (c < nc) {
                const R* l = tiles_t + ((J - c) * (w + 1) + (I - (J - c) + w)) * (long)TT;
#pragma unroll
                for (int k = 0; k < 16; ++k) lx[c][k] = l[(k0 + k) * T + col];
            }
        fetch_rows(x, J, -1, nc, vch, spin_limit, err);
#pragma unroll
        for (int c = 0; c < NCH; ++c)
            if (c < nc) {
#pragma unroll
                for (int k = 0; k < 16; ++k) acc -= double(lx[c][k]) * vch[c][k0 + k];
            }
        __syncthreads();
    }
    part[q][col] = acc;
    __syncthreads();
    if (tid < T) vin[tid] = part[0][tid] + part[1][tid] + part[2][tid] + part[3][tid];
    __syncthreads();
    double u = 0.0;
#pragma unroll
    for (int k = 0; k < 16; ++k) u += double(lv[k]) * vin[k0 + k];
    __syncthreads();
    part[q][col] = u;
    __syncthreads();
    double xi = (q == 0) ? part[0][col] + part[1][col] + part[2][col] + part[3][col] : 0.0;
    if (has_next) {
        fetch_row(x, I + 1, vin, spin_limit, err);
        double m = 0.0;
#pragma unroll
        for (int k = 0; k < 16; ++k) m += double(lm[k]) * vin[k0 + k];
        part[q][col] = m;
        __syncthreads();
        if (q == 0) xi -= part[0][col] + part[1][col] + part[2][col] + part[3][col];
    }
    if (q == 0) put_value(&x[I * T + col], xi);
}

// ---- look-ahead forward sweep -------------------------------------------------
// Row I of the block system L y = b, multiplied by Linv_I, less W_I = M_I
// Linv_{I-1} times row I-1 (multiplied by Linv_{I-1}), has no coupling to row
// I-1 any more (Linv_I L_{I,I-1} - M_I = 0):
//   y_I = b'_I - sum_{J=I-w-1}^{I-2} L'_IJ y_J,
//   b'_I = Linv_I b_I - W_I b_{I-1},  L'_IJ = Linv_I L_IJ - W_I L_{I-1,J}.
// The newest row a row waits for is then I-2, so the even and the odd rows
// form two chains that run side by side.  L' is kept per row in w slots, slot
// s for J = I-2-s (column-major tiles), formed after each factorization.

// W_I = M_I Linv_{I-1} (I >= 1), column-major.  Y^T = Linv_{I-1}: ys[k][col] =
// Linv_{I-1}[k][col], a transposed staging of the column-major tile.
template <class R>
__global__ void __launch_bounds__(NTH) k_look2_w(const R* __restrict__ mch, const R* __restrict__ linv,
                                                 R* __restrict__ w2, long nt) {
    __shared__ R xs[T][T];
    __shared__ R ys[T][T];
    const long I = blockIdx.x + 1;
    stage<false>(xs, mch + I * (long)TT);
    const R* lv = linv + (I - 1) * (long)TT;
    for (int idx = threadIdx.x; idx < TT; idx += NTH) ys[idx % T][idx / T] = lv[idx];
    __syncthreads();
    TileAcc<R> acc;
    acc.product(xs, ys);
    acc.each([&](int row, int col, R v) { w2[I * (long)TT + col * T + row] = v; });
}

// L'_{I, I-2-s} for workgroup (I, s): Linv_I L_IJ (when I - J <= w) less
// W_I L_{I-1,J}; the L tiles are staged from their row-major transposed
// copies, so that ys[k][col] = L[k][col] reads straight.
template <class R>
__global__ void __launch_bounds__(NTH) k_look2_band(const R* __restrict__ tiles_t, const R* __restrict__ linv,
                                                    const R* __restrict__ w2, R* __restrict__ l2, long nt, int w) {
    __shared__ R xs[T][T];
    __shared__ R ys[T][T];
    const long I = blockIdx.x / w;
    const int s = int(blockIdx.x % w);
    const long J = I - 2 - s;
    if (J < 0) return;
    TileAcc<R> a, b;
    const bool direct = I - J <= w;
    if (direct) {
        stage<false>(xs, linv + I * (long)TT);
        stage<false>(ys, tiles_t + tile_index(I, J, w) * TT);
        __syncthreads();
        a.product(xs, ys);
        __syncthreads();
    }
    stage<false>(xs, w2 + I * (long)TT);
    stage<false>(ys, tiles_t + tile_index(I - 1, J, w) * TT);
    __syncthreads();
    b.product(xs, ys);
    R* out = l2 + (I * w + s) * (long)TT;
    if (direct) {
        R av[TileAcc<R>::N];
        a.each_i([&](int, int, R v, int e) { av[e] = v; });
        b.each_i([&](int row, int col, R v, int e) { out[col * T + row] = av[e] - v; });
    } else {
        b.each([&](int row, int col, R v) { out[col * T + row] = -v; });
    }
}

// b'_I = Linv_I b_I - W_I b_{I-1}, one workgroup per tile row, thread (row, q)
// summing columns 16q..16q+15
template <class R>
__global__ void __launch_bounds__(256) k_look2_rhs(const R* __restrict__ linv, const R* __restrict__ w2,
                                                   const double* __restrict__ b, double* __restrict__ bp, long nt) {
    __shared__ double part[4][T];
    const long I = blockIdx.x;
    const int tid = threadIdx.x, row = tid % T, q = tid / T, k0 = 16 * q;
    const R* lv = linv + I * (long)TT;
    double u = 0.0;
#pragma unroll
    for (int k = 0; k < 16; ++k) u += double(lv[(k0 + k) * T + row]) * b[I * T + k0 + k];
    if (I >= 1) {
        const R* wv = w2 + I * (long)TT;
#pragma unroll
        for (int k = 0; k < 16; ++k) u -= double(wv[(k0 + k) * T + row]) * b[(I - 1) * T + k0 + k];
    }
    part[q][row] = u;
    __syncthreads();
    if (tid < T) bp[I * T + tid] = part[0][tid] + part[1][tid] + part[2][tid] + part[3][tid];
}

// The look-ahead sweeps take their older rows NCW at a time: a workgroup
// catching up behind the front consumes one chunk per memory round trip, so
// the front advances at most NCW rows per round trip whatever the depth of
// the system.  Eight waves, thread (row, q) holding columns 8q..8q+7 of each
// tile; a chunk's tile loads are issued before its rows are polled and share
// that round trip.
constexpr int NCW = 8;
constexpr int SWT = 512;

// y_I = b'_I - sum_{J=I-w-D+1}^{I-D} L'_IJ y_J on a look-ahead system of
// depth D (slot s of row I for J = I-D-s, w slots), the older rows in chunks
// of NCW aligned so that the last chunk ends at row I-D: a row waits once for
// the newest rows it needs and the D chains of the system (the rows of each
// residue mod D) run side by side.  y must hold SENT.
template <class R>
__global__ void __launch_bounds__(SWT) k_fwd_sweep_la(const R* __restrict__ l2, const double* __restrict__ bp,
                                                      double* y, long nt, int w, int D, long spin_limit, int* err) {
    __shared__ double vch[NCW][T];
    __shared__ double part[SWT / T][T];
    const long I = blockIdx.x;
    const int tid = threadIdx.x, row = tid % T, q = tid / T, k0 = 8 * q;
    double acc = (q == 0) ? bp[I * T + row] : 0.0;
    const long J0 = I - w - D + 1 > 0 ? I - w - D + 1 : 0;
    const long nrow = I - D + 1 - J0;   // rows J0 .. I-D
    int nc = int(nrow % NCW);
    if (nc == 0) nc = NCW;
    for (long J = J0; J <= I - D; J += nc, nc = NCW) {
        R lx[NCW][8];
#pragma unroll
        for (int c = 0; c < NCW; ++c)
            if (c < nc) {
                const R* l = l2 + (I * w + (I - D - (J + c))) * (long)TT;
#pragma unroll
                for (int k = 0; k < 8; ++k) lx[c][k] = l[(k0 + k) * T + row];
            }
        fetch_rows(y, J, 1, nc, vch, spin_limit, err);
#pragma unroll
        for (int c = 0; c < NCW; ++c)
            if (c < nc) {
#pragma unroll
                for (int k = 0; k < 8; ++k) acc -= double(lx[c][k]) * vch[c][k0 + k];
            }
        __syncthreads();
    }
    part[q][row] = acc;
    __syncthreads();
    if (q == 0) {
        double v = part[0][row];
#pragma unroll
        for (int g = 1; g < SWT / T; ++g) v += part[g][row];
        put_value(&y[I * T + row], v);
    }
}

// The backward sweep x_I = Linv_I^T (z_I - sum_{J>I} L_JI^T x_J), z = y / d,
// on the same look-ahead system: with V_I = G_I^T Linv_{I+1}^T the coupling
// to x_{I+1} cancels and
//   x_I = z'_I - sum_{J=I+2}^{I+w+1} B_IJ x_J,
//   z'_I = Linv_I^T z_I - V_I z_{I+1},  B_IJ = Linv_I^T L_JI^T - V_I L_{J,I+1}^T,
// B kept per row in w slots, slot s for J = I+2+s (column-major tiles).

// V_I (I <= nt-2): X = G_I^T (G_I row-major: xs[k][row] = G_I[k][row]),
// Y = Linv_{I+1} (ys[k][col] = Linv_{I+1}[col][k], column-major storage)
template <class R>
__global__ void __launch_bounds__(NTH) k_look2_v(const R* __restrict__ gch, const R* __restrict__ linv,
                                                 R* __restrict__ v2, long nt) {
    __shared__ R xs[T][T];
    __shared__ R ys[T][T];
    const long I = blockIdx.x;
    stage<false>(xs, gch + I * (long)TT);
    stage<false>(ys, linv + (I + 1) * (long)TT);
    __syncthreads();
    TileAcc<R> acc;
    acc.product(xs, ys);
    acc.each([&](int row, int col, R v) { v2[I * (long)TT + col * T + row] = v; });
}

// B_{I, I+2+s} for workgroup (I, s): X = Linv_I^T (a transposed staging of the
// column-major tile) with Y = L_JI, less X = V_I with Y = L_{J,I+1}; the L
// tiles are column-major, so ys[k][col] = L[col][k] stages straight.
template <class R>
__global__ void __launch_bounds__(NTH) k_look2_bband(const R* __restrict__ tiles, const R* __restrict__ linv,
                                                     const R* __restrict__ v2, R* __restrict__ b2, long nt,
                                                     int w) {
    __shared__ R xs[T][T];
    __shared__ R ys[T][T];
    const long I = blockIdx.x / w;
    const int s = int(blockIdx.x % w);
    const long J = I + 2 + s;
    if (J >= nt) return;
    TileAcc<R> a, b;
    const bool direct = J - I <= w;
    if (direct) {
        const R* lv = linv + I * (long)TT;
        for (int idx = threadIdx.x; idx < TT; idx += NTH) xs[idx % T][idx / T] = lv[idx];
        stage<false>(ys, tiles + tile_index(J, I, w) * TT);
        __syncthreads();
        a.product(xs, ys);
        __syncthreads();
    }
    stage<false>(xs, v2 + I * (long)TT);
    stage<false>(ys, tiles + tile_index(J, I + 1, w) * TT);
    __syncthreads();
    b.product(xs, ys);
    R* out = b2 + (I * w + s) * (long)TT;
    if (direct) {
        R av[TileAcc<R>::N];
        a.each_i([&](int, int, R v, int e) { av[e] = v; });
        b.each_i([&](int row, int col, R v, int e) { out[col * T + row] = av[e] - v; });
    } else {
        b.each([&](int row, int col, R v) { out[col * T + row] = -v; });
    }
}

// z'_I = Linv_I^T z_I - V_I z_{I+1}, z = y / d, one workgroup per tile row
template <class R>
__global__ void __launch_bounds__(256) k_look2_brhs(const R* __restrict__ linv, const R* __restrict__ v2,
                                                    const double* __restrict__ y, const double* __restrict__ piv,
                                                    double* __restrict__ zp, long nt) {
    __shared__ double part[4][T];
    const long I = blockIdx.x;
    const int tid = threadIdx.x, row = tid % T, q = tid / T, k0 = 16 * q;
    const R* lv = linv + I * (long)TT;
    double u = 0.0;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const long e = I * T + k0 + k;
        u += double(lv[row * T + k0 + k]) * (y[e] / piv[e]);   // Linv_I[k][row]
    }
    if (I + 1 < nt) {
        const R* vv = v2 + I * (long)TT;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const long e = (I + 1) * T + k0 + k;
            u -= double(vv[(k0 + k) * T + row]) * (y[e] / piv[e]);
        }
    }
    part[q][row] = u;
    __syncthreads();
    if (tid < T) zp[I * T + tid] = part[0][tid] + part[1][tid] + part[2][tid] + part[3][tid];
}

// x_I = z'_I - sum_{J=I+D}^{I+D+w-1} B_IJ x_J on a look-ahead system of depth
// D (slot s of row I for J = I+D+s), I = nt-1-blockIdx.x: the later rows in
// chunks of NCW, farthest first, aligned so that the last chunk ends at row
// I+D.  x must hold SENT.
template <class R>
__global__ void __launch_bounds__(SWT) k_bwd_sweep_la(const R* __restrict__ b2, const double* __restrict__ zp,
                                                      double* x, long nt, int w, int D, long spin_limit, int* err) {
    __shared__ double vch[NCW][T];
    __shared__ double part[SWT / T][T];
    const long I = nt - 1 - blockIdx.x;
    const int tid = threadIdx.x, row = tid % T, q = tid / T, k0 = 8 * q;
    double acc = (q == 0) ? zp[I * T + row] : 0.0;
    const long J1 = (I + D + w - 1 < nt - 1) ? I + D + w - 1 : nt - 1;
    const long nrow = J1 - (I + D) + 1;   // rows J1 down to I+D
    int nc = int(nrow % NCW);
    if (nc == 0) nc = NCW;
    for (long J = J1; J >= I + D; J -= nc, nc = NCW) {
        R lx[NCW][8];
#pragma unroll
        for (int c = 0; c < NCW; ++c)
            if (c < nc) {
                const R* l = b2 + (I * w + (J - c - I - D)) * (long)TT;
#pragma unroll
                for (int k = 0; k < 8; ++k) lx[c][k] = l[(k0 + k) * T + row];
            }
        fetch_rows(x, J, -1, nc, vch, spin_limit, err);
#pragma unroll
        for (int c = 0; c < NCW; ++c)
            if (c < nc) {
#pragma unroll
                for (int k = 0; k < 8; ++k) acc -= double(lx[c][k]) * vch[c][k0 + k];
            }
        __syncthreads();
    }
    part[q][row] = acc;
    __syncthreads();
    if (q == 0) {
        double v = part[0][row];
#pragma unroll
        for (int g = 1; g < SWT / T; ++g) v += part[g][row];
        put_value(&x[I * T + row], v);
    }
}

// ---- deeper look-ahead -------------------------------------------------------
// A look-ahead system of depth D (row I coupled to row I-D and rows further
// back, slot s of row I holding the tile for J = I-D-s, ws slots) is deepened
// by substituting row I-D's own equation for the coupling P_I (slot 0) to it:
//   L+_IJ = L_IJ - P_I L_{I-D,J},  b+_I = b_I - P_I b_{I-D},
// which couples row I to rows I-D-1 and further back in ws + D - 1 slots: slot
// s (J = I-D-1-s) is slot s+1 of row I less P_I times slot s+1-D of row I-D.
// The backward system deepens the same way towards the later rows (J =
// I+D+1+s, row I+D): dir = +1 forward, -1 backward.  Every slot whose row J
// exists is written (zero where neither term has a tile).
template <class R>
__global__ void __launch_bounds__(NTH) k_look_deepen(const R* __restrict__ src, int ws, R* __restrict__ dst, int wd,
                                                     long nt, int D, int dir) {
    __shared__ R xs[T][T];
    __shared__ R ys[T][T];
    const long I = blockIdx.x / wd;
    const int s = int(blockIdx.x % wd);
    const long J = I - long(dir) * (D + 1 + s);
    if (J < 0 || J >= nt) return;
    R* out = dst + (I * wd + s) * (long)TT;
    const R* keep = (s + 1 <= ws - 1) ? src + (I * ws + s + 1) * (long)TT : nullptr;
    const int sr = s + 1 - D;   // <= ws - 1 for every s < wd
    if (sr < 0) {
        for (int idx = threadIdx.x; idx < TT; idx += NTH) out[idx] = keep ? keep[idx] : R(0);
        return;
    }
    const long Ir = I - long(dir) * D;
    stage<false>(xs, src + I * ws * (long)TT);   // P_I, column-major: xs[k][row] = P_I[row][k]
    const R* y = src + (Ir * ws + sr) * (long)TT;
    for (int idx = threadIdx.x; idx < TT; idx += NTH) ys[idx % T][idx / T] = y[idx];
    __syncthreads();
    TileAcc<R> acc;
    acc.product(xs, ys);
    if (keep)
        acc.each([&](int row, int col, R v) { out[col * T + row] = keep[col * T + row] - v; });
    else
        acc.each([&](int row, int col, R v) { out[col * T + row] = -v; });
}

// b+_I = b_I - P_I b_{I-dir*D} with P_I slot 0 of row I of the depth-D
// system, one workgroup per tile row
template <class R>
__global__ void __launch_bounds__(256) k_look_deepen_rhs(const R* __restrict__ src, int ws,
                                                         const double* __restrict__ b, double* __restrict__ bo,
                                                         long nt, int D, int dir) {
    __shared__ double part[4][T];
    const long I = blockIdx.x;
    const int tid = threadIdx.x, row = tid % T, q = tid / T, k0 = 16 * q;
    const long Ir = I - long(dir) * D;
    double u = (q == 0) ? b[I * T + row] : 0.0;
    if (Ir >= 0 && Ir < nt) {
        const R* pv = src + I * ws * (long)TT;
#pragma unroll
        for (int k = 0; k < 16; ++k) u -= double(pv[(k0 + k) * T + row]) * b[Ir * T + k0 + k];
    }
    part[q][row] = u;
    __syncthreads();
    if (tid < T) bo[I * T + tid] = part[0][tid] + part[1][tid] + part[2][tid] + part[3][tid];
}

// v = SENT
__global__ void k_fill_sent(double* __restrict__ v, long n) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) v[i] = __builtin_bit_cast(double, SENT);
}

// dst = rhs padded with zeros to ntot; y = SENT
__global__ void k_sweep_prep(double* __restrict__ dst, const double* __restrict__ src, double* __restrict__ y,
                             long n, long ntot) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < ntot) {
        dst[i] = i < n ? src[i] : 0.0;
        y[i] = __builtin_bit_cast(double, SENT);
    }
}

// ---- batches of models ---------------------------------------------------------
// Identity rows for the equations no DOF slot maps to (the tail of each
// model's last tile).
// Fills by kernel rather than hipMemsetAsync, so that a captured HIP graph of a
// batch round holds no memset node.
__global__ void k_fill_f64(double* __restrict__ y, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) y[i] = 0.0;
}
__global__ void k_fill_i32(int* __restrict__ y, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) y[i] = 0;
}

__global__ void k_band_gap_pad(double* __restrict__ tiles, const int32_t* __restrict__ gaps, long ngap, int w) {
    long k = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (k >= ngap) return;
    const long i = gaps[k], I = i / T;
    tiles[tile_index(I, I, w) * TT + (i % T) * T + (i % T)] = 1.0;
}

// One workgroup factors one model, column by column, with the tile arithmetic
// of the single-model schedules (so the same factor): D_K, then P_{I,K} for
// the m tiles below it, then U_{I,J,K}.  Linv_K stays in LDS for the panels.
// While m <= 2 the panels also stay in LDS (P_{K+1,K} in ys, P_{K+2,K} in lk
// once Linv_K is no longer needed), the updates take their operands from
// there, and U_{K+1,K+1,K} is applied last and left in xs, where D_{K+1}
// starts.  Wider columns take the operands of the updates from memory.  The
// barrier after each task orders its LDS and memory writes before the next
// task reads them.
//
// A model with two tiles below the diagonal may instead be factored by a pair
// of workgroups (role ROLE_MAIN and ROLE_HELPER).  The main workgroup keeps the
// critical chain D_K -> P_{K+1,K} -> U_{K+1,K+1,K} -> D_{K+1}; the helper
// computes P_{K+2,K}, U_{K+2,K+1,K} and U_{K+2,K+2,K} meanwhile.  They hand
// over through memory with system-scope stores and loads and four progress
// counters per model (the main's D_K and P_{K+1,K}, the helper's two
// updates), each raised after the writer's stores have completed.  The main
// reads A_{K+1,K} after the helper's U_{K+1,K,K-1} and A_{K+1,K+1} after its
// U_{K+1,K+1,K-1}; the helper reads Linv_K, D_K and P_{K+1,K} after the main
// has published them.  Every task keeps its arithmetic, so the factor is the
// one a single workgroup computes.  The pairs come first in the grid, so both
// workgroups of a pair are resident together; every wait is bounded and ends
// early once any wait has failed (error bit 1).
constexpr int ROLE_SINGLE = 0, ROLE_MAIN = 1, ROLE_HELPER = 2;
enum { C_DIAG = 0, C_P1 = 1, C_U21 = 2, C_U22 = 3 };

// thread 0 waits until *c >= target (or an error is raised anywhere)
__device__ inline bool wait_count(const int* c, int target, long spin_limit, int* err, int* ok_s) {
    if (threadIdx.x == 0) {
        long n = 0;
        bool good = true;
        while (__hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) < target) {
            if (++n >= spin_limit) { good = false; break; }
            if ((n & 63) == 0 && (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) & 1)) {
                good = false;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        if (!good) atomicOr(err, 1);
        *ok_s = good;
    }
    __syncthreads();
    return *ok_s != 0;
}

// *c = value once every thread's stores have completed
__device__ inline void publish(int* c, int value) {
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_store(c, value, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

__global__ void __launch_bounds__(NTH) k_band_factor_models(double* __restrict__ tiles, double* __restrict__ linv,
                                                            double* __restrict__ piv, int w,
                                                            const BandSeg* __restrict__ segs,
                                                            const int* __restrict__ order,
                                                            const int* __restrict__ role, int* prog,
                                                            long spin_limit, int* err) {
    __shared__ double xs[T][T];
    __shared__ double ys[TT];
    __shared__ double lk[T][T];   // Linv_K as the panels' second operand (lk[k][row] = Linv(row, k))
    __shared__ double dd[T];
    __shared__ DiagScratch<double> sm;
    __shared__ int ok_s;
    const int mdl = order[blockIdx.x], rl = role[blockIdx.x];
    const BandSeg g = segs[mdl];
    int* pc = prog + long(mdl) * 4 * FLAG_STRIDE;   // the model's progress counters
    double(*yt)[T] = reinterpret_cast<double(*)[T]>(ys);
    const long Kend = g.t0 + g.nt;
    bool staged = false;   // xs already holds A_KK
    // xs = src * D_K by column (the scaled operand of the updates)
    auto scale_into_xs = [&](const double (*src)[T], const double* dk) {
        for (int idx = threadIdx.x; idx < TT; idx += NTH) xs[idx / T][idx % T] = src[idx / T][idx % T] * dk[idx / T];
    };
    if (rl == ROLE_HELPER) {
        // P_{K+2,K}, U_{K+2,K+1,K}, U_{K+2,K+2,K} for the columns with two tiles below
        for (long K = g.t0; K + 2 < Kend; ++K) {
            const int rel = int(K - g.t0);
            if (!wait_count(pc + C_DIAG * FLAG_STRIDE, rel + 1, spin_limit, err, &ok_s)) return;
            stage<true>(lk, linv + K * (long)TT);
            if (threadIdx.x < T) dd[threadIdx.x] = cld<true>(&piv[K * T + threadIdx.x]);
            double* p2 = tiles + tile_index(K + 2, K, w) * TT;
            stage<false>(xs, p2);
            __syncthreads();
            TileAcc<double> acc;
            acc.product_ylower(xs, lk);
            acc.each([&](int row, int col, double v) {
                const double p = v * (1.0 / dd[col]);
                p2[col * T + row] = p;
                yt[col][row] = p;
            });
            __syncthreads();
            if (!wait_count(pc + C_P1 * FLAG_STRIDE, rel + 1, spin_limit, err, &ok_s)) return;
            stage<true>(lk, tiles + tile_index(K + 1, K, w) * TT);   // P_{K+1,K}
            scale_into_xs(yt, dd);
            __syncthreads();
            acc.product(xs, lk);
            double* o21 = tiles + tile_index(K + 2, K + 1, w) * TT;
            acc.each([&](int row, int col, double v) { cst<true>(&o21[col * T + row], o21[col * T + row] - v); });
            publish(pc + C_U21 * FLAG_STRIDE, rel + 1);
            acc.product(xs, yt);
            double* o22 = tiles + tile_index(K + 2, K + 2, w) * TT;
            acc.each([&](int row, int col, double v) { cst<true>(&o22[col * T + row], o22[col * T + row] - v); });
            publish(pc + C_U22 * FLAG_STRIDE, rel + 1);
        }
        return;
    }
    for (long K = g.t0; K < Kend; ++K) {
        const long m = (Kend - 1 - K < g.w) ? Kend - 1 - K : g.w;
        const bool in_lds = m <= 2;
        // A workgroup that factors its model alone stages A_{K+1,K} after the
        // diagonal task; the previous column has written it, so it is loaded
        // into registers (in the staging layout) once the pivot steps are
        // done, and its latency overlaps the rest of the diagonal task.
        const bool pre = rl != ROLE_MAIN && in_lds && m >= 1;
        double pa[TT / NTH];
        if (!staged) {
            stage<false>(xs, tiles + tile_index(K, K, w) * TT);
            __syncthreads();
        }
        staged = false;
        diag_task(xs, ys, dd, sm, linv + K * (long)TT, piv, K, &lk[0][0], [&] {
            if (pre) {
                const double* t1 = tiles + tile_index(K + 1, K, w) * TT;
#pragma unroll
                for (int i = 0; i < TT / NTH; ++i) pa[i] = t1[threadIdx.x + NTH * i];
            }
        });
        __syncthreads();
        const double* dk = piv + K * T;
        if (rl == ROLE_MAIN) {
            const int rel = int(K - g.t0);
            if (m == 2) {
                // Linv_K and D_K for the helper, stored again through to memory
                for (int idx = threadIdx.x; idx < TT; idx += NTH)
                    cst<true>(&linv[K * (long)TT + idx], lk[idx / T][idx % T]);
                if (threadIdx.x < T) cst<true>(&piv[K * T + threadIdx.x], piv[K * T + threadIdx.x]);
                publish(pc + C_DIAG * FLAG_STRIDE, rel + 1);
            }
            if (m >= 1) {
                // P_{K+1,K}, after the helper's U_{K+1,K,K-1}
                if (K > g.t0 && !wait_count(pc + C_U21 * FLAG_STRIDE, rel, spin_limit, err, &ok_s)) return;
                double* p1 = tiles + tile_index(K + 1, K, w) * TT;
                stage<true>(xs, p1);
                __syncthreads();
                TileAcc<double> acc;
                acc.product_ylower(xs, lk);
                acc.each([&](int row, int col, double v) {
                    const double p = v * (1.0 / dk[col]);
                    cst<true>(&p1[col * T + row], p);
                    yt[col][row] = p;
                });
                if (m == 2) publish(pc + C_P1 * FLAG_STRIDE, rel + 1);
                else __syncthreads();
                // U_{K+1,K+1,K}, after the helper's U_{K+1,K+1,K-1}; left in xs for D_{K+1}
                if (K > g.t0 && !wait_count(pc + C_U22 * FLAG_STRIDE, rel, spin_limit, err, &ok_s)) return;
                scale_into_xs(yt, dk);
                __syncthreads();
                // symmetric, and the diagonal task reads only the lower part
                LowerAcc u;
                u.product(xs, yt);
                __syncthreads();
                const double* a11 = tiles + tile_index(K + 1, K + 1, w) * TT;
                u.each_i([&](int row, int col, double v, int) { xs[col][row] = cld<true>(&a11[col * T + row]) - v; });
                __syncthreads();
                staged = true;
            }
            continue;
        }
        for (long r = 1; r <= m; ++r) {
            // P_{K+r,K} as panel_task computes it, with Linv_K from LDS
            double* tile = tiles + tile_index(K + r, K, w) * TT;
            if (pre && r == 1) {
#pragma unroll
                for (int i = 0; i < TT / NTH; ++i) {
                    const int idx = threadIdx.x + NTH * i;
                    xs[idx / T][idx % T] = pa[i];
                }
            } else {
                stage<false>(xs, tile);
            }
            __syncthreads();
            TileAcc<double> acc;
            acc.product_ylower(xs, lk);
            if (in_lds) {
                __syncthreads();   // lk is free after the last panel's product
                double(*dst)[T] = (r == 1) ? yt : lk;
                acc.each([&](int row, int col, double v) {
                    const double p = v * (1.0 / dk[col]);
                    tile[col * T + row] = p;
                    dst[col][row] = p;
                });
            } else {
                acc.each([&](int row, int col, double v) { tile[col * T + row] = v * (1.0 / dk[col]); });
            }
            __syncthreads();
        }
        if (in_lds) {
            if (m == 2) {
                // U_{K+2,K+1,K} and U_{K+2,K+2,K}: (P_{K+2} D_K) P_{K+1}^T and (P_{K+2} D_K) P_{K+2}^T
                scale_into_xs(lk, dk);
                __syncthreads();
                TileAcc<double> acc;
                acc.product(xs, yt);
                double* o21 = tiles + tile_index(K + 2, K + 1, w) * TT;
                acc.each([&](int row, int col, double v) { o21[col * T + row] -= v; });
                acc.product(xs, lk);
                double* o22 = tiles + tile_index(K + 2, K + 2, w) * TT;
                acc.each([&](int row, int col, double v) { o22[col * T + row] -= v; });
                __syncthreads();
            }
            if (m >= 1) {
                // U_{K+1,K+1,K}, left in xs for D_{K+1}: symmetric, and the
                // diagonal task reads only the lower part
                scale_into_xs(yt, dk);
                __syncthreads();
                LowerAcc u;
                u.product(xs, yt);
                __syncthreads();
                const double* a11 = tiles + tile_index(K + 1, K + 1, w) * TT;
                u.each_i([&](int row, int col, double v, int) { xs[col][row] = a11[col * T + row] - v; });
                __syncthreads();
                staged = true;
            }
        } else {
            for (long r = 1; r <= m; ++r)
                for (long c = 1; c <= r; ++c) {
                    update_task<false, double>(tiles + tile_index(K + r, K + c, w) * TT,
                                               tiles + tile_index(K + r, K, w) * TT,
                                               tiles + tile_index(K + c, K, w) * TT, piv + K * T, xs, yt);
                    __syncthreads();
                }
        }
    }
}

// One 256-thread workgroup solves one model: the forward sweep into y, then
// the backward sweep into out, tile row by tile row.  Thread (row, q) keeps the
// partial sum over columns 16q..16q+15 and the partial sums are combined in
// the order of k_fwd_sweep / k_bwd_sweep, so a model's solution is the one
// those kernels give.  The last RING rows of the solution stay in LDS, where
// every thread reads the values it needs; the first RW off-diagonal tiles of
// a row and its Linv are loaded into registers when the row starts.  A row
// then costs three barriers.
__global__ void __launch_bounds__(256) k_band_solve_models(const double* __restrict__ tiles,
                                                           const double* __restrict__ linv,
                                                           const double* __restrict__ piv, int w,
                                                           const BandSeg* __restrict__ segs,
                                                           const int* __restrict__ order,
                                                           const double* __restrict__ rhs, double* y, double* out) {
    constexpr int RING = 8, RW = 2;
    __shared__ double vin[T];
    __shared__ double part[4][T];
    __shared__ double ring[RING][T];
    const BandSeg g = segs[order[blockIdx.x]];
    const int tid = threadIdx.x, row = tid % T, q = tid / T, k0 = 16 * q;
    const long Iend = g.t0 + g.nt, e0 = g.t0 * T;
    // acc -= sum_k a[k] * (row J of the solution)[k0 + k], from the ring when
    // row J is one of the last RING rows
    auto sub_row = [&](double& acc, const double* a, const double* v, long I, long J) {
        if (I - J < RING && J - I < RING) {
            const double* rj = &ring[J % RING][k0];
#pragma unroll
            for (int k = 0; k < 16; ++k) acc -= a[k] * rj[k];
        } else {
            const double* vj = v + J * T + k0;
#pragma unroll
            for (int k = 0; k < 16; ++k) acc -= a[k] * vj[k];
        }
    };
    // acc -> combined over q -> Linv-product (lv) -> combined; thread tid < T
    // returns entry tid of the row's result
    auto finish_row = [&](double acc, const double* lv) {
        part[q][row] = acc;
        __syncthreads();
        if (tid < T) vin[tid] = part[0][tid] + part[1][tid] + part[2][tid] + part[3][tid];
        __syncthreads();
        double t = 0.0;
#pragma unroll
        for (int k = 0; k < 16; ++k) t += lv[k] * vin[k0 + k];
        part[q][row] = t;
        __syncthreads();
        double s = 0.0;
        if (tid < T) s = part[0][tid] + part[1][tid] + part[2][tid] + part[3][tid];
        return s;
    };
    // The entries of a model's band more than bw below the diagonal are zero
    // in A and in L, and Linv is lower triangular: a thread whose 16 entries
    // of a tile are all such zeros takes zeros instead of reading them (the
    // products are the same up to the sign of zero).
    // a[k] = tile entry (r0, c0 + k) of a tile d tiles below the diagonal
    auto load16 = [&](double* a, const double* l, long d, int r0, int c0, int stride_r, int stride_c) {
        if (64 * d + r0 - (c0 + 15) > g.bw) {
#pragma unroll
            for (int k = 0; k < 16; ++k) a[k] = 0.0;
        } else {
#pragma unroll
            for (int k = 0; k < 16; ++k) a[k] = l[(c0 + k) * stride_c + r0 * stride_r];
        }
    };
    double lt[RW][16], lv[16];
    for (long I = g.t0; I < Iend; ++I) {
        const long J0 = (I - g.w > g.t0) ? I - g.w : g.t0;
        const int nJ = int(I - J0);
#pragma unroll
        for (int j = 0; j < RW; ++j)
            if (j < nJ) load16(lt[j], tiles + tile_index(I, J0 + j, w) * TT, I - J0 - j, row, k0, 1, T);
        if (k0 > row) {
#pragma unroll
            for (int k = 0; k < 16; ++k) lv[k] = 0.0;
        } else {
#pragma unroll
            for (int k = 0; k < 16; ++k) lv[k] = linv[I * (long)TT + (k0 + k) * T + row];
        }
        double acc = 0.0;
        if (q == 0) {
            const long e = I * T + row;
            acc = (e - e0 < g.neq) ? rhs[e] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < RW; ++j)
            if (j < nJ) sub_row(acc, lt[j], y, I, J0 + j);
        for (long J = J0 + RW; J < I; ++J) {
            double a[16];
            load16(a, tiles + tile_index(I, J, w) * TT, I - J, row, k0, 1, T);
            sub_row(acc, a, y, I, J);
        }
        const double s = finish_row(acc, lv);
        if (tid < T) {
            y[I * T + tid] = s;
            ring[I % RING][tid] = s;
        }
        __syncthreads();
    }
    for (long I = Iend - 1; I >= g.t0; --I) {
        const int col = row;
        const long J1 = (I + g.w < Iend - 1) ? I + g.w : Iend - 1;
        const int nJ = int(J1 - I);
        // tile (J, I) read by columns: entry (k0 + k, col), 64 d + k0 + k - col
        // below the diagonal
#pragma unroll
        for (int j = 0; j < RW; ++j)
            if (j < nJ) {
                const double* l = tiles + tile_index(J1 - j, I, w) * TT;
                const long d = J1 - j - I;
                if (64 * d + k0 - col > g.bw) {
#pragma unroll
                    for (int k = 0; k < 16; ++k) lt[j][k] = 0.0;
                } else {
#pragma unroll
                    for (int k = 0; k < 16; ++k) lt[j][k] = l[col * T + k0 + k];
                }
            }
        if (k0 + 15 < col) {
#pragma unroll
            for (int k = 0; k < 16; ++k) lv[k] = 0.0;
        } else {
#pragma unroll
            for (int k = 0; k < 16; ++k) lv[k] = linv[I * (long)TT + col * T + k0 + k];
        }
        double acc = (q == 0) ? y[I * T + col] / piv[I * T + col] : 0.0;
#pragma unroll
        for (int j = 0; j < RW; ++j)
            if (j < nJ) sub_row(acc, lt[j], out, I, J1 - j);
        for (long J = J1 - RW; J > I; --J) {
            const double* l = tiles + tile_index(J, I, w) * TT;
            double a[16];
            if (64 * (J - I) + k0 - col > g.bw) {
#pragma unroll
                for (int k = 0; k < 16; ++k) a[k] = 0.0;
            } else {
#pragma unroll
                for (int k = 0; k < 16; ++k) a[k] = l[col * T + k0 + k];
            }
            sub_row(acc, a, out, I, J);
        }
        const double s = finish_row(acc, lv);
        if (tid < T) {
            out[I * T + tid] = s;
            ring[I % RING][tid] = s;
        }
        __syncthreads();
    }
}

}  // namespace

void KernelClock::begin(int f, hipStream_t s, long n) {
    if (used + 2 > ev.size()) {
        size_t old = ev.size();
        ev.resize(old + 4096);
        for (size_t i = old; i < ev.size(); ++i) CB_CHECK(hipEventCreate(&ev[i]));
    }
    fam.push_back(f);
    count.push_back(n);
    CB_CHECK(hipEventRecord(ev[used++], s));
}

void KernelClock::end(hipStream_t s) { CB_CHECK(hipEventRecord(ev[used++], s)); }

void KernelClock::collect() {
    for (size_t i = 0; i < fam.size(); ++i) {
        float t = 0;
        CB_CHECK(hipEventElapsedTime(&t, ev[2 * i], ev[2 * i + 1]));
        ms_timed[fam[i]] += t;
        n_timed[fam[i]] += count[i];
    }
    for (int f = 0; f < KT_N; ++f)
        ms[f] = n_timed[f] > 0 ? ms_timed[f] * double(launches[f]) / double(n_timed[f]) : 0.0;
    fam.clear();
    count.clear();
    used = 0;
}

void KernelClock::release() {
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    ev.clear();
    fam.clear();
    count.clear();
    used = 0;
}

template <class R>
void BandLDLT<R>::setup(long n_, long bandwidth, hipStream_t s_) {
    release();
    n = n_;
    s = s_;
    nt = (n + T - 1) / T;
    w = int((std::max(bandwidth, 1L) + T - 1) / T);
    if (w > nt - 1) w = int(std::max(nt - 1, 0L));
    tiles = dalloc<R>(size_t(nt) * size_t(w + 1) * TT);
    tiles_t = dalloc<R>(size_t(nt) * size_t(w + 1) * TT);
    linv = dalloc<R>(size_t(nt) * TT);
    mch = dalloc<R>(size_t(nt) * TT);
    gch = dalloc<R>(size_t(nt) * TT);
    piv = dalloc<double>(size_t(nt) * T);
    xw = dalloc<double>(size_t(nt) * T);
    yw = dalloc<double>(size_t(nt) * T);
    flag = dalloc<int>(1);
    pflags = dalloc<int>(size_t(w + 1) * FLAG_STRIDE);
    CB_CHECK(hipMemsetAsync(pflags, 0, sizeof(int) * size_t(w + 1) * FLAG_STRIDE, s));
    CB_CHECK(hipMemsetAsync(flag, 0, sizeof(int), s));
    cepoch = 0;
    piv_host.resize(size_t(n));
    CB_CHECK(hipStreamSynchronize(s));
}

template <class R>
void BandLDLT<R>::release() {
    void* ps[] = {tiles, tiles_t, linv, mch, gch, piv, xw, yw, flag, pflags, l2, w2, bp, b2, v2, zp, l3, b3, bq, zq};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    tiles = nullptr; tiles_t = nullptr; linv = nullptr; piv = nullptr; xw = nullptr; yw = nullptr; flag = nullptr;
    mch = nullptr; gch = nullptr;
    l2 = nullptr; w2 = nullptr; bp = nullptr;
    b2 = nullptr; v2 = nullptr; zp = nullptr;
    l3 = nullptr; b3 = nullptr; bq = nullptr; zq = nullptr;
    pflags = nullptr;
    clock.release();
}

template <class R>
int BandLDLT<R>::factor(const Dev& d) {
    CB_CHECK(hipMemsetAsync(tiles, 0, sizeof(R) * size_t(nt) * (w + 1) * TT, s));
    CB_CHECK(hipMemsetAsync(flag, 0, sizeof(int), s));
    if (src32)
        hipLaunchKernelGGL((k_band_from_bsr<R, float>), dim3(int((d.nslot + 255) / 256)), dim3(256), 0, s, tiles,
                           d.K32, d.rowptr, d.colidx, d.eq_of_slot, d.nj, d.bs, w, flag);
    else
        hipLaunchKernelGGL((k_band_from_bsr<R>), dim3(int((d.nslot + 255) / 256)), dim3(256), 0, s, tiles,
                           d.K, d.rowptr, d.colidx, d.eq_of_slot, d.nj, d.bs, w, flag);
    if (nt * T > n)
        hipLaunchKernelGGL((k_band_pad<R>), dim3(1), dim3(T), 0, s, tiles, n, nt, w);
    CB_CHECK(hipGetLastError());
    if (!split) {
        if (cepoch + nt + 1 >= (1 << 30)) {
            CB_CHECK(hipMemsetAsync(pflags, 0, sizeof(int) * size_t(w + 1) * FLAG_STRIDE, s));
            cepoch = 0;
        }
        const long spin_limit = 1L << 22;
        const bool tk_column = clock.on(KT_COLUMN);
        if (tk_column) {
            clock.launches[KT_COLUMN] += nt;
            clock.begin(KT_COLUMN, s, nt);
        }
        for (long K = -1; K < nt - 1; ++K) {
            const long m = (K < 0) ? 0 : std::min<long>(w, nt - 1 - K);
            long grid = 1;
            if (m > 0) {
                grid = m + 1;
                for (long rr = 1; rr < m; ++rr) grid += (rr + 2) / 2;
            }
            cepoch += 1;
            hipLaunchKernelGGL((k_band_column<R>), dim3(int(grid)), dim3(NTH), 0, s, tiles, tiles_t, linv, piv, K, w,
                               nt, pflags, cepoch, spin_limit, flag + 0);
        }
        if (tk_column) clock.end(s);
    } else {
        for (long Kt = 0; Kt < nt; ++Kt) {
            const bool tk_diag = clock.take(KT_DIAG);
            if (tk_diag) clock.begin(KT_DIAG, s);
            hipLaunchKernelGGL((k_tile_diag<R>), dim3(1), dim3(NTH), 0, s, tiles, linv, piv, Kt, w);
            if (tk_diag) clock.end(s);
            const long m = std::min<long>(w, nt - 1 - Kt);
            if (m > 0) {
                const bool tk_panel = clock.take(KT_PANEL);
                if (tk_panel) clock.begin(KT_PANEL, s);
                hipLaunchKernelGGL((k_tile_panel<R>), dim3(int(m)), dim3(NTH), 0, s, tiles, tiles_t, linv, piv, Kt,
                                   w);
                if (tk_panel) clock.end(s);
                const bool tk_update = clock.take(KT_UPDATE);
                if (tk_update) clock.begin(KT_UPDATE, s);
                hipLaunchKernelGGL((k_tile_update<R>), dim3(int(m * (m + 1) / 2)), dim3(NTH), 0, s, tiles, piv,
                                   Kt, w);
                if (tk_update) clock.end(s);
            }
        }
    }
    CB_CHECK(hipGetLastError());
    if (chained) {
        const bool tk_prep = clock.take(KT_PREP);
        if (tk_prep) clock.begin(KT_PREP, s);
        hipLaunchKernelGGL((k_sweep_chain<R>), dim3(int(nt)), dim3(256), 0, s, tiles, linv, mch, gch, nt, w);
        if (look >= 2 && nt >= 2) {
            if (!l2) {
                l2 = dalloc<R>(size_t(nt) * size_t(w) * TT);
                w2 = dalloc<R>(size_t(nt) * TT);
                bp = dalloc<double>(size_t(nt) * T);
                b2 = dalloc<R>(size_t(nt) * size_t(w) * TT);
                v2 = dalloc<R>(size_t(nt) * TT);
                zp = dalloc<double>(size_t(nt) * T);
            }
            hipLaunchKernelGGL((k_look2_w<R>), dim3(int(nt - 1)), dim3(NTH), 0, s, mch, linv, w2, nt);
            hipLaunchKernelGGL((k_look2_band<R>), dim3(int(nt * w)), dim3(NTH), 0, s, tiles_t, linv, w2, l2, nt, w);
            hipLaunchKernelGGL((k_look2_v<R>), dim3(int(nt - 1)), dim3(NTH), 0, s, gch, linv, v2, nt);
            hipLaunchKernelGGL((k_look2_bband<R>), dim3(int(nt * w)), dim3(NTH), 0, s, tiles, linv, v2, b2, nt, w);
            if (look >= 3) {
                if (!l3) {
                    l3 = dalloc<R>(size_t(nt) * size_t(w + 1) * TT);
                    b3 = dalloc<R>(size_t(nt) * size_t(w + 1) * TT);
                    bq = dalloc<double>(size_t(nt) * T);
                    zq = dalloc<double>(size_t(nt) * T);
                }
                hipLaunchKernelGGL((k_look_deepen<R>), dim3(int(nt * (w + 1))), dim3(NTH), 0, s, l2, w, l3, w + 1,
                                   nt, 2, 1);
                hipLaunchKernelGGL((k_look_deepen<R>), dim3(int(nt * (w + 1))), dim3(NTH), 0, s, b2, w, b3, w + 1,
                                   nt, 2, -1);
            }
        }
        if (tk_prep) clock.end(s);
        CB_CHECK(hipGetLastError());
    }
    int bad = 0;
    CB_CHECK(hipMemcpyAsync(&bad, flag, sizeof(int), hipMemcpyDeviceToHost, s));
    CB_CHECK(hipMemcpyAsync(piv_host.data(), piv, sizeof(double) * size_t(n), hipMemcpyDeviceToHost, s));
    CB_CHECK(hipStreamSynchronize(s));
    clock.collect();
    if (bad & 1) throw DeviceError("band factorization: a panel wait exceeded its bound");
    if (bad & 2) return 2;
    for (long i = 0; i < n; ++i)
        if (piv_host[i] == 0.0 || !std::isfinite(piv_host[i])) return 1;
    return 0;
}

template <class R>
void BandLDLT<R>::solve(const double* rhs, double* out) {
    CB_CHECK(hipMemsetAsync(flag, 0, sizeof(int), s));
    hipLaunchKernelGGL(k_sweep_prep, dim3(int((nt * T + 255) / 256)), dim3(256), 0, s, xw, rhs, yw, n, nt * T);
    const long spin_limit = 1L << 22;
    const bool ahead = chained && look >= 2 && l2;
    const bool deep = ahead && look >= 3 && l3;
    if (ahead) {
        hipLaunchKernelGGL((k_look2_rhs<R>), dim3(int(nt)), dim3(256), 0, s, linv, w2, xw, bp, nt);
        if (deep)
            hipLaunchKernelGGL((k_look_deepen_rhs<R>), dim3(int(nt)), dim3(256), 0, s, l2, w, bp, bq, nt, 2, 1);
        hipLaunchKernelGGL(k_fill_sent, dim3(int((nt * T + 255) / 256)), dim3(256), 0, s, xw, nt * T);
    }
    const bool tk_fwd = clock.take(KT_FWD);
    if (tk_fwd) clock.begin(KT_FWD, s);
    if (deep)
        hipLaunchKernelGGL((k_fwd_sweep_la<R>), dim3(int(nt)), dim3(SWT), 0, s, l3, bq, yw, nt, w + 1, 3, spin_limit,
                           flag);
    else if (ahead)
        hipLaunchKernelGGL((k_fwd_sweep_la<R>), dim3(int(nt)), dim3(SWT), 0, s, l2, bp, yw, nt, w, 2, spin_limit,
                           flag);
    else if (chained)
        hipLaunchKernelGGL((k_fwd_sweep_m<R>), dim3(int(nt)), dim3(256), 0, s, tiles, linv, mch, xw, yw, nt, w,
                           spin_limit, flag);
    else
        hipLaunchKernelGGL((k_fwd_sweep<R>), dim3(int(nt)), dim3(256), 0, s, tiles, linv, xw, yw, nt, w, spin_limit,
                           flag);
    if (tk_fwd) clock.end(s);
    if (ahead)
        hipLaunchKernelGGL((k_look2_brhs<R>), dim3(int(nt)), dim3(256), 0, s, linv, v2, yw, piv, zp, nt);
    if (deep)
        hipLaunchKernelGGL((k_look_deepen_rhs<R>), dim3(int(nt)), dim3(256), 0, s, b2, w, zp, zq, nt, 2, -1);
    const bool tk_bwd = clock.take(KT_BWD);
    if (tk_bwd) clock.begin(KT_BWD, s);
    if (deep)
        hipLaunchKernelGGL((k_bwd_sweep_la<R>), dim3(int(nt)), dim3(SWT), 0, s, b3, zq, xw, nt, w + 1, 3, spin_limit,
                           flag);
    else if (ahead)
        hipLaunchKernelGGL((k_bwd_sweep_la<R>), dim3(int(nt)), dim3(SWT), 0, s, b2, zp, xw, nt, w, 2, spin_limit,
                           flag);
    else if (chained)
        hipLaunchKernelGGL((k_bwd_sweep_m<R>), dim3(int(nt)), dim3(256), 0, s, tiles_t, linv, gch, yw, piv, xw, nt, w,
                           spin_limit, flag);
    else
        hipLaunchKernelGGL((k_bwd_sweep<R>), dim3(int(nt)), dim3(256), 0, s, tiles_t, linv, yw, piv, xw, nt, w,
                           spin_limit, flag);
    if (tk_bwd) clock.end(s);
    CB_CHECK(hipGetLastError());
    CB_CHECK(hipMemcpyAsync(out, xw, sizeof(double) * size_t(n), hipMemcpyDeviceToDevice, s));
    int bad = 0;
    CB_CHECK(hipMemcpyAsync(&bad, flag, sizeof(int), hipMemcpyDeviceToHost, s));
    CB_CHECK(hipStreamSynchronize(s));
    clock.collect();
    if (bad & 1) throw DeviceError("band solve: a tile-row wait exceeded its bound");
}

template struct BandLDLT<double>;
template struct BandLDLT<float>;

void BandBatch::setup(const std::vector<BandSeg>& sg, long n_total, int w_band, const std::vector<int32_t>& gap_eqs,
                      hipStream_t s_) {
    release();
    s = s_;
    n = n_total;
    if (n % T) throw DeviceError("band batch: the equations of a batch fill whole tiles");
    nt = n / T;
    w = w_band;
    segs_host = sg;
    nseg = int(sg.size());
    for (const BandSeg& g : sg)
        if (g.t0 < 0 || g.t0 + g.nt > nt || g.w > w || g.w < 0 || g.neq > g.nt * T)
            throw DeviceError("band batch: a model's tile range does not fit the band");
    // models in decreasing order of their tile work, so the longest
    // factorizations start first; the models with two tiles below the
    // diagonal (as many as fit with their helpers into the first workgroups
    // the GPU holds at once) get a helper workgroup and come first, in pairs
    std::vector<int> mo(sg.size());
    for (size_t i = 0; i < mo.size(); ++i) mo[i] = int(i);
    auto work = [&](int i) { return double(sg[i].nt) * double((sg[i].w + 1) * (sg[i].w + 2) / 2); };
    std::stable_sort(mo.begin(), mo.end(), [&](int a, int b) { return work(a) > work(b); });
    work_order = mo;
    // the factor kernel's LDS allows one workgroup per CU: the pairs stay
    // within the first CUs - 32 workgroups of the grid
    int dev = 0, ncu = 0;
    CB_CHECK(hipGetDevice(&dev));
    CB_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    int eligible = 0;
    for (const BandSeg& g : sg) eligible += (g.w == 2 && g.nt >= 3);
    max_pairs = std::max(0, std::min({96, (ncu - 32) / 2, eligible}));
    tiles = dalloc<double>(size_t(nt) * size_t(w + 1) * TT);
    linv = dalloc<double>(size_t(nt) * TT);
    piv = dalloc<double>(size_t(nt) * T);
    yw = dalloc<double>(size_t(nt) * T);
    flag = dalloc<int>(1);
    CB_CHECK(hipHostMalloc(&flag_host, sizeof(int), hipHostMallocDefault));
    *flag_host = 0;
    segs = dalloc<BandSeg>(sg.size());
    order = dalloc<int>(sg.size() + size_t(max_pairs));
    role = dalloc<int>(sg.size() + size_t(max_pairs));
    sorder = dalloc<int>(mo.size());
    prog = dalloc<int>(sg.size() * 4 * FLAG_STRIDE);
    ngap = long(gap_eqs.size());
    gaps = dalloc<int32_t>(size_t(ngap));
    h2d(segs, sg.data(), sg.size(), s);
    set_pairs(max_pairs);
    h2d(sorder, mo.data(), mo.size(), s);
    h2d(gaps, gap_eqs.data(), gap_eqs.size(), s);
    CB_CHECK(hipMemsetAsync(piv, 0, sizeof(double) * size_t(nt) * T, s));
    for (hipEvent_t& e : ev) CB_CHECK(hipEventCreate(&e));
    CB_CHECK(hipStreamSynchronize(s));
}

void BandBatch::release() {
    for (hipEvent_t& e : ev)
        if (e) {
            (void)hipEventDestroy(e);
            e = nullptr;
        }
    void* ps[] = {tiles, linv, piv, yw, flag, segs, order, role, sorder, prog, gaps};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    if (flag_host) (void)hipHostFree(flag_host);
    tiles = linv = piv = yw = nullptr;
    flag = nullptr;
    flag_host = nullptr;
    segs = nullptr;
    order = nullptr;
    role = nullptr;
    sorder = nullptr;
    prog = nullptr;
    gaps = nullptr;
}

void BandBatch::zero() {
    const long nv = nt * long(w + 1) * TT;
    hipLaunchKernelGGL(k_fill_f64, dim3(int(std::min<long>(8192, (nv + 255) / 256))), dim3(256), 0, s, tiles, nv);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, flag, 1L);
    CB_CHECK(hipGetLastError());
}

void BandBatch::factor(const Dev& d) {
    if (!nseg) return;
    zero();
    hipLaunchKernelGGL((k_band_from_bsr<double>), dim3(int((d.nslot + 255) / 256)), dim3(256), 0, s, tiles, d.K,
                       d.rowptr, d.colidx, d.eq_of_slot, d.nj, d.bs, w, flag);
    factor_filled();
}

void BandBatch::factor_filled(bool fetch_flag) {
    if (!nseg) return;
    if (ngap)
        hipLaunchKernelGGL(k_band_gap_pad, dim3(int((ngap + 255) / 256)), dim3(256), 0, s, tiles, gaps, ngap, w);
    const long np = long(nseg) * 4 * FLAG_STRIDE;
    hipLaunchKernelGGL(k_fill_i32, dim3(int(std::min<long>(1024, (np + 255) / 256))), dim3(256), 0, s, prog, np);
    const long spin_limit = 1L << 22;
    if (timing) CB_CHECK(hipEventRecord(ev[0], s));
    hipLaunchKernelGGL(k_band_factor_models, dim3(nwg), dim3(NTH), 0, s, tiles, linv, piv, w, segs, order, role, prog,
                       spin_limit, flag);
    if (timing) CB_CHECK(hipEventRecord(ev[1], s));
    CB_CHECK(hipGetLastError());
    if (fetch_flag) this->fetch_flag();
}

void BandBatch::fetch_flag() {
    CB_CHECK(hipMemcpyAsync(flag_host, flag, sizeof(int), hipMemcpyDeviceToHost, s));
}

void BandBatch::set_pairs(int np) {
    np = std::max(0, std::min(np, max_pairs));
    std::vector<int> ord, rol;
    std::vector<char> paired(segs_host.size(), 0);
    npair = 0;
    for (int i : work_order)
        if (segs_host[size_t(i)].w == 2 && segs_host[size_t(i)].nt >= 3 && npair < np) {
            ord.push_back(i);
            rol.push_back(ROLE_MAIN);
            ord.push_back(i);
            rol.push_back(ROLE_HELPER);
            paired[size_t(i)] = 1;
            ++npair;
        }
    for (int i : work_order)
        if (!paired[size_t(i)]) {
            ord.push_back(i);
            rol.push_back(ROLE_SINGLE);
        }
    nwg = int(ord.size());
    h2d(order, ord.data(), ord.size(), s);
    h2d(role, rol.data(), rol.size(), s);
    CB_CHECK(hipStreamSynchronize(s));
}

std::vector<std::pair<int, double>> BandBatch::tune_pairs(const std::vector<int>& cand) {
    std::vector<std::pair<int, double>> out;
    if (!nseg || cand.empty()) return out;
    const size_t nbytes = sizeof(double) * size_t(nt) * size_t(w + 1) * TT;
    double* save = dalloc<double>(nbytes / sizeof(double));
    CB_CHECK(hipMemcpyAsync(save, tiles, nbytes, hipMemcpyDeviceToDevice, s));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    CB_CHECK(hipEventCreate(&e0));
    CB_CHECK(hipEventCreate(&e1));
    const bool keep_timing = timing;
    timing = false;
    int best = max_pairs;
    double best_ms = 1e300;
    for (int c : cand) {
        double ms = 1e300;
        bool bad = false;
        for (int rep = 0; rep < 2 && !bad; ++rep) {
            CB_CHECK(hipMemcpyAsync(tiles, save, nbytes, hipMemcpyDeviceToDevice, s));
            CB_CHECK(hipMemsetAsync(flag, 0, sizeof(int), s));
            set_pairs(c);
            CB_CHECK(hipEventRecord(e0, s));
            factor_filled(true);
            CB_CHECK(hipEventRecord(e1, s));
            CB_CHECK(hipStreamSynchronize(s));
            float t = 0;
            CB_CHECK(hipEventElapsedTime(&t, e0, e1));
            if (*flag_host & 1) bad = true;
            ms = std::min(ms, double(t));
        }
        out.emplace_back(npair, bad ? -1.0 : ms);
        if (!bad && ms < best_ms) {
            best_ms = ms;
            best = npair;
        }
    }
    timing = keep_timing;
    CB_CHECK(hipMemcpyAsync(tiles, save, nbytes, hipMemcpyDeviceToDevice, s));
    CB_CHECK(hipMemsetAsync(flag, 0, sizeof(int), s));
    CB_CHECK(hipStreamSynchronize(s));
    (void)hipFree(save);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    set_pairs(best);
    return out;
}

void BandBatch::solve(const double* rhs, double* out) {
    if (!nseg) return;
    if (timing) CB_CHECK(hipEventRecord(ev[2], s));
    hipLaunchKernelGGL(k_band_solve_models, dim3(nseg), dim3(256), 0, s, tiles, linv, piv, w, segs, sorder, rhs, yw,
                       out);
    if (timing) CB_CHECK(hipEventRecord(ev[3], s));
    CB_CHECK(hipGetLastError());
}

}  // namespace cb
