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
//   diag : L_KK D_K L_KK^T = A_KK in LDS, plus Linv_K = L_KK^{-1}
//   panel: L_IK = A_IK Linv_K^T D_K^{-1},            K < I <= K+w
//   update: A_IJ -= (L_IK D_K) L_JK^T,               K < J <= I <= K+w
// Solves: forward with Linv (one launch per tile row, the launch also pushes
// y_I into the next w tile rows), pivot scaling, backward likewise.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "band.h"

namespace cb {
namespace {

constexpr int T = 64;
constexpr int TT = T * T;

template <class R>
__global__ void k_band_from_bsr(R* __restrict__ tiles, const double* __restrict__ K,
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
            if (I - J > w) { *bad = 1; continue; }
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

// LDL^T of the diagonal tile in LDS; writes L (unit, strictly lower part),
// pivots D and the inverse of the unit lower factor.  at[] holds the tile
// column-major (at[col][row]) and inv[] the inverse row-major, so that both
// the column updates and the column-per-thread inversion read LDS without
// bank conflicts; 2 x 32 KB of LDS in double.
template <class R>
__global__ void __launch_bounds__(256) k_tile_diag(R* __restrict__ tiles, R* __restrict__ linv,
                                                   double* __restrict__ piv, long Kt, int w) {
    __shared__ R at[T][T];
    __shared__ R inv[T][T];
    R* tile = tiles + (Kt * (w + 1) + w) * (long)TT;
    for (int idx = threadIdx.x; idx < TT; idx += blockDim.x) at[idx / T][idx % T] = tile[idx];
    __syncthreads();
    for (int c = 0; c < T; ++c) {
        const R dc = at[c][c];
        const int m = T - 1 - c;
        // a[i][j] -= a[i][c] a[j][c] / d_c for c < j <= i; consecutive threads
        // walk down a column (at[j][i], i fastest)
        for (int idx = threadIdx.x; idx < m * m; idx += blockDim.x) {
            const int i = c + 1 + idx % m, j = c + 1 + idx / m;
            if (j <= i) at[j][i] -= at[c][i] * (at[c][j] / dc);
        }
        __syncthreads();
        for (int i = c + 1 + threadIdx.x; i < T; i += blockDim.x) at[c][i] = at[c][i] / dc;
        __syncthreads();
    }
    for (int idx = threadIdx.x; idx < TT; idx += blockDim.x) {
        const int i = idx % T, j = idx / T;
        tile[idx] = (i >= j) ? at[j][i] : R(0);
    }
    if (threadIdx.x < T) piv[Kt * T + threadIdx.x] = double(at[threadIdx.x][threadIdx.x]);
    // inverse of the unit lower factor, column j by thread j (forward
    // substitution; L entries are broadcast reads, inv[k][j] is the thread's own)
    if (threadIdx.x < T) {
        const int j = threadIdx.x;
        for (int i = 0; i < T; ++i) {
            R v = (i == j) ? R(1) : R(0);
            for (int k = j; k < i; ++k) v -= at[k][i] * inv[k][j];
            inv[i][j] = (i < j) ? R(0) : v;
        }
    }
    __syncthreads();
    R* out = linv + Kt * (long)TT;
    for (int idx = threadIdx.x; idx < TT; idx += blockDim.x) out[idx] = inv[idx % T][idx / T];
}

// LDS tiles below are column-major like the global tiles: t[col][row].  The
// 16x16 thread grid owns 4x4 outputs each; at step k a thread reads row
// entries tr+16u of column k (consecutive across threads, no bank conflicts).

// L_IK = A_IK Linv_K^T D_K^{-1} for I = Kt+1+blockIdx.x.
template <class R>
__global__ void __launch_bounds__(256) k_tile_panel(R* __restrict__ tiles, const R* __restrict__ linv,
                                                    const double* __restrict__ piv, long Kt, int w) {
    __shared__ R at[T][T];   // A_IK
    __shared__ R lt[T][T];   // Linv_K
    const long I = Kt + 1 + blockIdx.x;
    R* tile = tiles + (I * (w + 1) + (Kt - I + w)) * (long)TT;
    const R* lv = linv + Kt * (long)TT;
    for (int idx = threadIdx.x; idx < TT; idx += blockDim.x) {
        at[idx / T][idx % T] = tile[idx];
        lt[idx / T][idx % T] = lv[idx];
    }
    __syncthreads();
    // out[r][c] = sum_k A[r][k] Linv[c][k] / d[c]
    const int tr = threadIdx.x % 16, tc = threadIdx.x / 16;
    R acc[4][4] = {};
    for (int k = 0; k < T; ++k) {
        R av[4], bv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) { av[u] = at[k][tr + 16 * u]; bv[u] = lt[k][tc + 16 * u]; }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[u][v] += av[u] * bv[v];
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int c = tc + 16 * v;
        const R dinv = R(1) / R(piv[Kt * T + c]);
#pragma unroll
        for (int u = 0; u < 4; ++u) tile[c * T + tr + 16 * u] = acc[u][v] * dinv;
    }
}

// A_IJ -= (L_IK D_K) L_JK^T for the pairs K < J <= I <= K+m.
template <class R>
__global__ void __launch_bounds__(256) k_tile_update(R* __restrict__ tiles, const double* __restrict__ piv,
                                                     long Kt, int w) {
    __shared__ R at[T][T];   // L_IK D_K
    __shared__ R bt[T][T];   // L_JK
    // blockIdx.x -> (r, s) with s <= r in a triangle of side m
    const long t = blockIdx.x;
    long r = long((sqrt(8.0 * double(t) + 1.0) - 1.0) / 2.0);
    while ((r + 1) * (r + 2) / 2 <= t) ++r;
    while (r * (r + 1) / 2 > t) --r;
    const long s = t - r * (r + 1) / 2;
    const long I = Kt + 1 + r, J = Kt + 1 + s;
    const R* lik = tiles + (I * (w + 1) + (Kt - I + w)) * (long)TT;
    const R* ljk = tiles + (J * (w + 1) + (Kt - J + w)) * (long)TT;
    R* out = tiles + (I * (w + 1) + (J - I + w)) * (long)TT;
    for (int idx = threadIdx.x; idx < TT; idx += blockDim.x) {
        const int col = idx / T;
        at[col][idx % T] = lik[idx] * R(piv[Kt * T + col]);
        bt[col][idx % T] = ljk[idx];
    }
    __syncthreads();
    const int tr = threadIdx.x % 16, tc = threadIdx.x / 16;
    R acc[4][4] = {};
    for (int k = 0; k < T; ++k) {
        R av[4], bv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) { av[u] = at[k][tr + 16 * u]; bv[u] = bt[k][tc + 16 * u]; }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[u][v] += av[u] * bv[v];
    }
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int u = 0; u < 4; ++u) out[(tc + 16 * v) * T + tr + 16 * u] -= acc[u][v];
}

// Solve steps use 256 threads: the 64x64 tiles are staged through LDS with
// coalesced loads; a row of the result is formed by 4 lanes groups of 16
// columns and reduced through LDS.

// Forward step I: y_I = Linv_I r_I (every block), block 0 stores it, block j>0
// subtracts L_{I+j,I} y_I from r_{I+j}.
template <class R>
__global__ void __launch_bounds__(256) k_fwd_step(const R* __restrict__ tiles, const R* __restrict__ linv,
                                                  double* __restrict__ x, double* __restrict__ yout, long It,
                                                  int w) {
    __shared__ R ts[T][T + 1];        // ts[row][col]
    __shared__ double vin[T], y[T];
    __shared__ double part[4][T];
    const int tid = threadIdx.x, row = tid % T, q = tid / T;   // q: column quarter
    if (tid < T) vin[tid] = x[It * T + tid];
    const R* lv = linv + It * (long)TT;
    for (int idx = tid; idx < TT; idx += 256) ts[idx % T][idx / T] = lv[idx];
    __syncthreads();
    double s = 0.0;
    for (int k = 16 * q; k < 16 * q + 16; ++k) s += double(ts[row][k]) * vin[k];
    part[q][row] = s;
    __syncthreads();
    if (tid < T) y[tid] = part[0][tid] + part[1][tid] + part[2][tid] + part[3][tid];
    __syncthreads();
    if (blockIdx.x == 0) {
        if (tid < T) yout[It * T + tid] = y[tid];
        return;
    }
    const long J = It + blockIdx.x;
    const R* l = tiles + (J * (w + 1) + (It - J + w)) * (long)TT;
    for (int idx = tid; idx < TT; idx += 256) ts[idx % T][idx / T] = l[idx];
    __syncthreads();
    s = 0.0;
    for (int k = 16 * q; k < 16 * q + 16; ++k) s += double(ts[row][k]) * y[k];
    part[q][row] = s;
    __syncthreads();
    if (tid < T) x[J * T + tid] -= part[0][tid] + part[1][tid] + part[2][tid] + part[3][tid];
}

// Backward step I: x_I = Linv_I^T z_I (every block), block 0 stores it, block
// j>0 subtracts L_{I,I-j}^T x_I from z_{I-j}.  The transposes are read as
// columns of the LDS copies.
template <class R>
__global__ void __launch_bounds__(256) k_bwd_step(const R* __restrict__ tiles, const R* __restrict__ linv,
                                                  double* __restrict__ z, double* __restrict__ xout, long It,
                                                  int w) {
    __shared__ R ts[T][T + 1];        // ts[row][col]
    __shared__ double vin[T], xs[T];
    __shared__ double part[4][T];
    const int tid = threadIdx.x, col = tid % T, q = tid / T;
    if (tid < T) vin[tid] = z[It * T + tid];
    const R* lv = linv + It * (long)TT;
    for (int idx = tid; idx < TT; idx += 256) ts[idx % T][idx / T] = lv[idx];
    __syncthreads();
    // (Linv^T v)[col] = sum_k Linv[k][col] v[k]
    double s = 0.0;
    for (int k = 16 * q; k < 16 * q + 16; ++k) s += double(ts[k][col]) * vin[k];
    part[q][col] = s;
    __syncthreads();
    if (tid < T) xs[tid] = part[0][tid] + part[1][tid] + part[2][tid] + part[3][tid];
    __syncthreads();
    if (blockIdx.x == 0) {
        if (tid < T) xout[It * T + tid] = xs[tid];
        return;
    }
    const long J = It - blockIdx.x;
    const R* l = tiles + (It * (w + 1) + (J - It + w)) * (long)TT;
    for (int idx = tid; idx < TT; idx += 256) ts[idx % T][idx / T] = l[idx];
    __syncthreads();
    s = 0.0;
    for (int k = 16 * q; k < 16 * q + 16; ++k) s += double(ts[k][col]) * xs[k];
    part[q][col] = s;
    __syncthreads();
    if (tid < T) z[J * T + tid] -= part[0][tid] + part[1][tid] + part[2][tid] + part[3][tid];
}

__global__ void k_scale_piv(double* __restrict__ x, const double* __restrict__ piv, long n) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) x[i] /= piv[i];
}

__global__ void k_copy_pad(double* __restrict__ dst, const double* __restrict__ src, long n, long ntot) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < ntot) dst[i] = i < n ? src[i] : 0.0;
}

}  // namespace

template <class R>
void BandLDLT<R>::setup(long n_, long bandwidth, hipStream_t s_) {
    release();
    n = n_;
    s = s_;
    nt = (n + T - 1) / T;
    w = int((std::max(bandwidth, 1L) + T - 1) / T);
    if (w > nt - 1) w = int(std::max(nt - 1, 0L));
    tiles = dalloc<R>(size_t(nt) * size_t(w + 1) * TT);
    linv = dalloc<R>(size_t(nt) * TT);
    piv = dalloc<double>(size_t(nt) * T);
    xw = dalloc<double>(size_t(nt) * T);
    yw = dalloc<double>(size_t(nt) * T);
    flag = dalloc<int>(1);
    piv_host.resize(size_t(n));
}

template <class R>
void BandLDLT<R>::release() {
    void* ps[] = {tiles, linv, piv, xw, yw, flag};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    tiles = nullptr; linv = nullptr; piv = nullptr; xw = nullptr; yw = nullptr; flag = nullptr;
}

template <class R>
int BandLDLT<R>::factor(const Dev& d) {
    CB_CHECK(hipMemsetAsync(tiles, 0, sizeof(R) * size_t(nt) * (w + 1) * TT, s));
    CB_CHECK(hipMemsetAsync(flag, 0, sizeof(int), s));
    hipLaunchKernelGGL((k_band_from_bsr<R>), dim3(int((d.nslot + 255) / 256)), dim3(256), 0, s, tiles,
                       d.K, d.rowptr, d.colidx, d.eq_of_slot, d.nj, d.bs, w, flag);
    if (nt * T > n)
        hipLaunchKernelGGL((k_band_pad<R>), dim3(1), dim3(T), 0, s, tiles, n, nt, w);
    CB_CHECK(hipGetLastError());
    for (long Kt = 0; Kt < nt; ++Kt) {
        hipLaunchKernelGGL((k_tile_diag<R>), dim3(1), dim3(256), 0, s, tiles, linv, piv, Kt, w);
        long m = std::min<long>(w, nt - 1 - Kt);
        if (m > 0) {
            hipLaunchKernelGGL((k_tile_panel<R>), dim3(int(m)), dim3(256), 0, s, tiles, linv, piv, Kt, w);
            hipLaunchKernelGGL((k_tile_update<R>), dim3(int(m * (m + 1) / 2)), dim3(256), 0, s, tiles, piv,
                               Kt, w);
        }
    }
    CB_CHECK(hipGetLastError());
    int bad = 0;
    CB_CHECK(hipMemcpyAsync(&bad, flag, sizeof(int), hipMemcpyDeviceToHost, s));
    CB_CHECK(hipMemcpyAsync(piv_host.data(), piv, sizeof(double) * size_t(n), hipMemcpyDeviceToHost, s));
    CB_CHECK(hipStreamSynchronize(s));
    if (bad) return 2;
    for (long i = 0; i < n; ++i)
        if (piv_host[i] == 0.0 || !std::isfinite(piv_host[i])) return 1;
    return 0;
}

template <class R>
void BandLDLT<R>::solve(const double* rhs, double* out) {
    hipLaunchKernelGGL(k_copy_pad, dim3(int((nt * T + 255) / 256)), dim3(256), 0, s, xw, rhs, n, nt * T);
    for (long I = 0; I < nt; ++I) {
        int m = int(std::min<long>(w, nt - 1 - I));
        hipLaunchKernelGGL((k_fwd_step<R>), dim3(m + 1), dim3(256), 0, s, tiles, linv, xw, yw, I, w);
    }
    hipLaunchKernelGGL(k_scale_piv, dim3(int((nt * T + 255) / 256)), dim3(256), 0, s, yw, piv, nt * T);
    for (long I = nt - 1; I >= 0; --I) {
        int m = int(std::min<long>(w, I));
        hipLaunchKernelGGL((k_bwd_step<R>), dim3(m + 1), dim3(256), 0, s, tiles, linv, yw, xw, I, w);
    }
    CB_CHECK(hipGetLastError());
    CB_CHECK(hipMemcpyAsync(out, xw, sizeof(double) * size_t(n), hipMemcpyDeviceToDevice, s));
}

template struct BandLDLT<double>;
template struct BandLDLT<float>;

}  // namespace cb
