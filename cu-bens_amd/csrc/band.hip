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

// LDL^T of the diagonal tile by one wavefront: lane i keeps row i of the tile
// in registers (fully unrolled, static indices); column c is broadcast through
// LDS at step c.  The inverse of the unit lower factor is formed the same way:
// lane j builds column j by forward substitution, reading the rows of L as LDS
// broadcasts.  Writes L (unit lower), pivots D and Linv = L^{-1}.
template <class R>
__global__ void __launch_bounds__(64) k_tile_diag(R* __restrict__ tiles, R* __restrict__ linv,
                                                  double* __restrict__ piv, long Kt, int w) {
    __shared__ R col[T];
    __shared__ R Ls[T][T + 1];
    const int lane = threadIdx.x;
    R* tile = tiles + (Kt * (w + 1) + w) * (long)TT;
    R a[T];
#pragma unroll
    for (int j = 0; j < T; ++j) a[j] = tile[j * T + lane];   // (lane, j), column-major
#pragma unroll
    for (int c = 0; c < T; ++c) {
        if (lane >= c) col[lane] = a[c];
        __syncthreads();
        const R d = col[c];
        if (lane > c) {
            const R l = a[c] / d;
#pragma unroll
            for (int j = c + 1; j < T; ++j)
                if (j <= lane) a[j] -= l * col[j];
            a[c] = l;
        }
        __syncthreads();
    }
    // pivots and L (unit lower, zero upper) back to the tile and to LDS
    piv[Kt * T + lane] = double(col[lane]);   // d_lane, written by lane at step c = lane
#pragma unroll
    for (int j = 0; j < T; ++j) {
        R v = (j < lane) ? a[j] : (j == lane ? col[lane] : R(0));
        tile[j * T + lane] = v;
        Ls[lane][j] = (j < lane) ? a[j] : R(0);
    }
    __syncthreads();
    // column lane of L^{-1}: v_i = [i == lane] - sum_{k<i} L[i][k] v_k
    R v[T];
#pragma unroll
    for (int i = 0; i < T; ++i) {
        R s0 = R(0), s1 = R(0);
#pragma unroll
        for (int k = 0; k < i; ++k) {
            if (k & 1) s1 += Ls[i][k] * v[k];
            else s0 += Ls[i][k] * v[k];
        }
        v[i] = ((i == lane) ? R(1) : R(0)) - (s0 + s1);
    }
    R* out = linv + Kt * (long)TT;
#pragma unroll
    for (int i = 0; i < T; ++i) out[lane * T + i] = v[i];   // column lane, row i
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

// ---- pipelined sweeps ---------------------------------------------------------
// One workgroup per tile row, launched in tile order (forward) or reverse tile
// order (backward).  A workgroup waits only on workgroups with a lower index,
// which the hardware dispatched before it, so the oldest unfinished workgroup
// is always resident and the sweep makes progress.  Completion of a tile row
// is published by a per-row flag holding the sweep's epoch: the values are
// stored with agent-scope atomics, fenced, then the flag is released; readers
// acquire the flag and read the values with agent-scope atomics.  Each wait is
// bounded: past spin_limit polls the workgroup sets *err and leaves, so every
// wave exits whatever happens.  Contributions of earlier rows are applied
// while the newest one is still pending; the newest off-diagonal tile and
// Linv are prefetched into registers so the critical path per row is one flag
// handoff and two 64x64 products from registers.
__device__ inline bool wait_flag(int* flags, long J, int epoch, long spin_limit, int* err, int* ok_s) {
    if (threadIdx.x == 0) {
        // poll without cache maintenance; one acquire once the flag is seen
        long n = 0;
        int f;
        while ((f = __hip_atomic_load(&flags[J], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM)) != epoch &&
               ++n < spin_limit)
            __builtin_amdgcn_s_sleep(2);
        __atomic_thread_fence(__ATOMIC_ACQUIRE);
        *ok_s = (f == epoch);
        if (!*ok_s) atomicExch(err, 1);
    }
    __syncthreads();
    return *ok_s != 0;
}

__device__ inline void publish(double* v, long I, double val, int* flags, int epoch) {
    if (threadIdx.x < T) __hip_atomic_store(&v[I * T + threadIdx.x], val, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_store(&flags[I], epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// y_I = Linv_I (r_I - sum_{J<I} L_IJ y_J)
template <class R>
__global__ void __launch_bounds__(256) k_fwd_pipe(const R* __restrict__ tiles, const R* __restrict__ linv,
                                                  const double* __restrict__ r, double* y, int* flags, int epoch,
                                                  long nt, int w, long spin_limit, int* err) {
    __shared__ double vin[T];
    __shared__ double part[4][T];
    __shared__ int ok_s;
    const long I = blockIdx.x;
    const int tid = threadIdx.x, row = tid % T, q = tid / T, k0 = 16 * q;
    // prefetch: Linv_I and the newest off-diagonal tile (I, I-1), rows of this thread
    R lv[16], ln[16];
    const R* lvp = linv + I * (long)TT;
#pragma unroll
    for (int k = 0; k < 16; ++k) lv[k] = lvp[(k0 + k) * T + row];
    const bool has_prev = I > 0;
    if (has_prev) {
        const R* lnp = tiles + (I * (w + 1) + (w - 1)) * (long)TT;
#pragma unroll
        for (int k = 0; k < 16; ++k) ln[k] = lnp[(k0 + k) * T + row];
    }
    double acc = (q == 0) ? r[I * T + row] : 0.0;
    const long J0 = I - w > 0 ? I - w : 0;
    for (long J = J0; J < I; ++J) {
        if (!wait_flag(flags, J, epoch, spin_limit, err, &ok_s)) return;
        if (tid < T) vin[tid] = __hip_atomic_load(&y[J * T + tid], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        if (J == I - 1) {
#pragma unroll
            for (int k = 0; k < 16; ++k) acc -= double(ln[k]) * vin[k0 + k];
        } else {
            const R* l = tiles + (I * (w + 1) + (J - I + w)) * (long)TT;
#pragma unroll
            for (int k = 0; k < 16; ++k) acc -= double(l[(k0 + k) * T + row]) * vin[k0 + k];
        }
        __syncthreads();
    }
    part[q][row] = acc;
    __syncthreads();
    if (tid < T) vin[tid] = part[0][tid] + part[1][tid] + part[2][tid] + part[3][tid];
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < 16; ++k) s += double(lv[k]) * vin[k0 + k];
    part[q][row] = s;
    __syncthreads();
    double yi = (tid < T) ? part[0][tid] + part[1][tid] + part[2][tid] + part[3][tid] : 0.0;
    publish(y, I, yi, flags, epoch);
}

// x_I = Linv_I^T (y_I / D_I - sum_{J>I} L_JI^T x_J); workgroup g handles I = nt-1-g
template <class R>
__global__ void __launch_bounds__(256) k_bwd_pipe(const R* __restrict__ tiles, const R* __restrict__ linv,
                                                  const double* __restrict__ y, const double* __restrict__ piv,
                                                  double* x, int* flags, int epoch, long nt, int w,
                                                  long spin_limit, int* err) {
    __shared__ double vin[T];
    __shared__ double part[4][T];
    __shared__ int ok_s;
    const long I = nt - 1 - blockIdx.x;
    const int tid = threadIdx.x, col = tid % T, q = tid / T, k0 = 16 * q;
    R lv[16], ln[16];
    const R* lvp = linv + I * (long)TT;
#pragma unroll
    for (int k = 0; k < 16; ++k) lv[k] = lvp[col * T + k0 + k];   // Linv[k][col]
    const bool has_next = I + 1 < nt;
    if (has_next) {
        const R* lnp = tiles + ((I + 1) * (w + 1) + (w - 1)) * (long)TT;   // tile (I+1, I)
#pragma unroll
        for (int k = 0; k < 16; ++k) ln[k] = lnp[col * T + k0 + k];
    }
    double acc = (q == 0) ? y[I * T + col] / piv[I * T + col] : 0.0;
    const long J1 = (I + w < nt - 1) ? I + w : nt - 1;
    for (long J = J1; J > I; --J) {
        if (!wait_flag(flags, J, epoch, spin_limit, err, &ok_s)) return;
        if (tid < T) vin[tid] = __hip_atomic_load(&x[J * T + tid], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        if (J == I + 1) {
#pragma unroll
            for (int k = 0; k < 16; ++k) acc -= double(ln[k]) * vin[k0 + k];
        } else {
            const R* l = tiles + (J * (w + 1) + (I - J + w)) * (long)TT;
#pragma unroll
            for (int k = 0; k < 16; ++k) acc -= double(l[col * T + k0 + k]) * vin[k0 + k];
        }
        __syncthreads();
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
    double xi = (tid < T) ? part[0][tid] + part[1][tid] + part[2][tid] + part[3][tid] : 0.0;
    publish(x, I, xi, flags, epoch);
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

void KernelClock::begin(int f, hipStream_t s) {
    if (used + 2 > ev.size()) {
        size_t old = ev.size();
        ev.resize(old + 4096);
        for (size_t i = old; i < ev.size(); ++i) CB_CHECK(hipEventCreate(&ev[i]));
    }
    fam.push_back(f);
    CB_CHECK(hipEventRecord(ev[used++], s));
}

void KernelClock::end(hipStream_t s) { CB_CHECK(hipEventRecord(ev[used++], s)); }

void KernelClock::collect() {
    for (size_t i = 0; i < fam.size(); ++i) {
        float t = 0;
        CB_CHECK(hipEventElapsedTime(&t, ev[2 * i], ev[2 * i + 1]));
        ms[fam[i]] += t;
        launches[fam[i]] += 1;
    }
    fam.clear();
    used = 0;
}

void KernelClock::release() {
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    ev.clear();
    fam.clear();
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
    linv = dalloc<R>(size_t(nt) * TT);
    piv = dalloc<double>(size_t(nt) * T);
    xw = dalloc<double>(size_t(nt) * T);
    yw = dalloc<double>(size_t(nt) * T);
    flag = dalloc<int>(1);
    flags = dalloc<int>(size_t(2 * nt));
    CB_CHECK(hipMemsetAsync(flags, 0, sizeof(int) * size_t(2 * nt), s));
    CB_CHECK(hipMemsetAsync(flag, 0, sizeof(int), s));
    epoch = 0;
    piv_host.resize(size_t(n));
}

template <class R>
void BandLDLT<R>::release() {
    void* ps[] = {tiles, linv, piv, xw, yw, flag, flags};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    tiles = nullptr; linv = nullptr; piv = nullptr; xw = nullptr; yw = nullptr; flag = nullptr;
    flags = nullptr;
    clock.release();
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
        if (clock.on(KT_DIAG)) clock.begin(KT_DIAG, s);
        hipLaunchKernelGGL((k_tile_diag<R>), dim3(1), dim3(64), 0, s, tiles, linv, piv, Kt, w);
        if (clock.on(KT_DIAG)) clock.end(s);
        long m = std::min<long>(w, nt - 1 - Kt);
        if (m > 0) {
            if (clock.on(KT_PANEL)) clock.begin(KT_PANEL, s);
            hipLaunchKernelGGL((k_tile_panel<R>), dim3(int(m)), dim3(256), 0, s, tiles, linv, piv, Kt, w);
            if (clock.on(KT_PANEL)) clock.end(s);
            if (clock.on(KT_UPDATE)) clock.begin(KT_UPDATE, s);
            hipLaunchKernelGGL((k_tile_update<R>), dim3(int(m * (m + 1) / 2)), dim3(256), 0, s, tiles, piv,
                               Kt, w);
            if (clock.on(KT_UPDATE)) clock.end(s);
        }
    }
    CB_CHECK(hipGetLastError());
    int bad = 0;
    CB_CHECK(hipMemcpyAsync(&bad, flag, sizeof(int), hipMemcpyDeviceToHost, s));
    CB_CHECK(hipMemcpyAsync(piv_host.data(), piv, sizeof(double) * size_t(n), hipMemcpyDeviceToHost, s));
    CB_CHECK(hipStreamSynchronize(s));
    clock.collect();
    if (bad) return 2;
    for (long i = 0; i < n; ++i)
        if (piv_host[i] == 0.0 || !std::isfinite(piv_host[i])) return 1;
    return 0;
}

template <class R>
void BandLDLT<R>::solve(const double* rhs, double* out) {
    hipLaunchKernelGGL(k_copy_pad, dim3(int((nt * T + 255) / 256)), dim3(256), 0, s, xw, rhs, n, nt * T);
    const long spin_limit = 1L << 24;
    epoch += 1;
    if (clock.on(KT_FWD)) clock.begin(KT_FWD, s);
    hipLaunchKernelGGL((k_fwd_pipe<R>), dim3(int(nt)), dim3(256), 0, s, tiles, linv, xw, yw, flags, epoch, nt, w,
                       spin_limit, flag);
    if (clock.on(KT_FWD)) clock.end(s);
    if (clock.on(KT_BWD)) clock.begin(KT_BWD, s);
    hipLaunchKernelGGL((k_bwd_pipe<R>), dim3(int(nt)), dim3(256), 0, s, tiles, linv, yw, piv, xw, flags + nt,
                       epoch, nt, w, spin_limit, flag);
    if (clock.on(KT_BWD)) clock.end(s);
    CB_CHECK(hipGetLastError());
    CB_CHECK(hipMemcpyAsync(out, xw, sizeof(double) * size_t(n), hipMemcpyDeviceToDevice, s));
    int bad = 0;
    CB_CHECK(hipMemcpyAsync(&bad, flag, sizeof(int), hipMemcpyDeviceToHost, s));
    CB_CHECK(hipStreamSynchronize(s));
    clock.collect();
    if (bad) throw DeviceError("band solve: a tile-row wait exceeded its bound");
}

template struct BandLDLT<double>;
template struct BandLDLT<float>;

}  // namespace cb
