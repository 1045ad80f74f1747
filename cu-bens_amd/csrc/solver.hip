// Flexible GMRES, joint-block Jacobi and the fused driver reductions.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "solver.h"

namespace cb {
namespace {

constexpr int TPB = 256;
constexpr int RB = 256;            // partial-sum blocks of the reductions
constexpr int MAXV = 65;           // Krylov basis limit (m + 1)

inline int nblk(long n) { return int((n + TPB - 1) / TPB); }

__device__ inline double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// partial[k*gridDim.x + block] = sum over the block's chunk of V[k][i] w[i]
__global__ void __launch_bounds__(TPB) k_gemv_t(const double* __restrict__ V, const double* __restrict__ w,
                                               long n, int cnt, double* __restrict__ part) {
    __shared__ double sh[TPB / 64];
    const long chunk = (n + gridDim.x - 1) / gridDim.x;
    const long lo = blockIdx.x * chunk, hi = min(n, lo + chunk);
    for (int k = 0; k < cnt; ++k) {
        double acc = 0.0;
        const double* v = V + (long)k * n;
        for (long i = lo + threadIdx.x; i < hi; i += TPB) acc += v[i] * w[i];
        acc = wave_sum(acc);
        if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
        __syncthreads();
        if (threadIdx.x == 0) {
            double s = 0;
            for (int q = 0; q < TPB / 64; ++q) s += sh[q];
            part[k * gridDim.x + blockIdx.x] = s;
        }
        __syncthreads();
    }
}

__global__ void k_reduce_parts(const double* __restrict__ part, int nb, int cnt, double* __restrict__ out) {
    // one wave per output
    const int k = blockIdx.x;
    if (k >= cnt) return;
    double s = 0.0;
    for (int i = threadIdx.x; i < nb; i += 64) s += part[k * nb + i];
    s = wave_sum(s);
    if (threadIdx.x == 0) out[k] = s;
}

// w -= V h (cnt columns); also returns partial sums of |w|^2
__global__ void __launch_bounds__(TPB) k_gemv_n_sub(const double* __restrict__ V, const double* __restrict__ h,
                                                   double* __restrict__ w, long n, int cnt,
                                                   double* __restrict__ part) {
    __shared__ double hs[MAXV];
    __shared__ double sh[TPB / 64];
    if (threadIdx.x < cnt) hs[threadIdx.x] = h[threadIdx.x];
    __syncthreads();
    double acc = 0.0;
    for (long i = blockIdx.x * (long)TPB + threadIdx.x; i < n; i += (long)gridDim.x * TPB) {
        double s = w[i];
        for (int k = 0; k < cnt; ++k) s -= V[(long)k * n + i] * hs[k];
        w[i] = s;
        acc += s * s;
    }
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0;
        for (int q = 0; q < TPB / 64; ++q) s += sh[q];
        part[blockIdx.x] = s;
    }
}

__global__ void k_scale_copy(double* __restrict__ y, const double* __restrict__ x, double a, long n) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) y[i] = a * x[i];
}

// x += Z y
__global__ void k_update_x(double* __restrict__ x, const double* __restrict__ Z, const double* __restrict__ y,
                           int k, long n) {
    __shared__ double ys[MAXV];
    if (threadIdx.x < k) ys[threadIdx.x] = y[threadIdx.x];
    __syncthreads();
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = x[i];
    for (int j = 0; j < k; ++j) s += Z[(long)j * n + i] * ys[j];
    x[i] = s;
}

__global__ void k_sub(double* __restrict__ r, const double* __restrict__ b, const double* __restrict__ Ax, long n) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) r[i] = b[i] - Ax[i];
}

// Gauss-Jordan inverse with partial pivoting of each joint's diagonal block.
template <int BS>
__global__ void k_bj_setup(const double* __restrict__ K, const int32_t* __restrict__ diagblk,
                           double* __restrict__ inv, long nj) {
    long j = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (j >= nj) return;
    double a[BS][2 * BS];
    const double* Kb = K + (long)diagblk[j] * BS * BS;
#pragma unroll
    for (int r = 0; r < BS; ++r)
#pragma unroll
        for (int c = 0; c < BS; ++c) {
            a[r][c] = Kb[r * BS + c];
            a[r][BS + c] = (r == c) ? 1.0 : 0.0;
        }
#pragma unroll
    for (int c = 0; c < BS; ++c) {
        int p = c;
        double best = fabs(a[c][c]);
#pragma unroll
        for (int r = c + 1; r < BS; ++r)
            if (fabs(a[r][c]) > best) { best = fabs(a[r][c]); p = r; }
        if (p != c) {
#pragma unroll
            for (int q = 0; q < 2 * BS; ++q) { double t = a[c][q]; a[c][q] = a[p][q]; a[p][q] = t; }
        }
        double piv = a[c][c];
        double ip = (piv != 0.0) ? 1.0 / piv : 0.0;
#pragma unroll
        for (int q = 0; q < 2 * BS; ++q) a[c][q] *= ip;
#pragma unroll
        for (int r = 0; r < BS; ++r) {
            if (r == c) continue;
            double f = a[r][c];
#pragma unroll
            for (int q = 0; q < 2 * BS; ++q) a[r][q] -= f * a[c][q];
        }
    }
    double* o = inv + j * BS * BS;
#pragma unroll
    for (int r = 0; r < BS; ++r)
#pragma unroll
        for (int c = 0; c < BS; ++c) o[r * BS + c] = a[r][BS + c];
}

template <int BS>
__global__ void k_bj_apply(const double* __restrict__ inv, const double* __restrict__ v, double* __restrict__ z,
                           long nj) {
    long r = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (r >= nj * BS) return;
    long j = r / BS;
    int p = int(r % BS);
    const double* b = inv + j * BS * BS + p * BS;
    const double* vv = v + j * BS;
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < BS; ++q) s += b[q] * vv[q];
    z[r] = s;
}

template <class Op, int NOUT>
__global__ void __launch_bounds__(TPB) k_fused_sums(Op op, long n, double* __restrict__ part) {
    __shared__ double sh[NOUT][TPB / 64];
    double acc[NOUT];
#pragma unroll
    for (int k = 0; k < NOUT; ++k) acc[k] = 0.0;
    for (long i = blockIdx.x * (long)TPB + threadIdx.x; i < n; i += (long)gridDim.x * TPB) op(i, acc);
#pragma unroll
    for (int k = 0; k < NOUT; ++k) {
        double v = wave_sum(acc[k]);
        if ((threadIdx.x & 63) == 0) sh[k][threadIdx.x >> 6] = v;
    }
    __syncthreads();
    if (threadIdx.x < NOUT) {
        double s = 0;
        for (int q = 0; q < TPB / 64; ++q) s += sh[threadIdx.x][q];
        part[threadIdx.x * gridDim.x + blockIdx.x] = s;
    }
}

struct TestOp {
    const double *dd, *dt, *q, *f, *fp, *fip;
    double lam;
    __device__ void operator()(long i, double* a) const {
        double qt = q[i] * lam;
        double u = qt - f[i], up = qt - fp[i];
        a[0] += dd[i] * dd[i];
        a[1] += dt[i] * dt[i];
        a[2] += u * u;
        a[3] += up * up;
        a[4] += dd[i] * (qt - fip[i]);
    }
};

struct QuadOp {
    const double *dt, *dp, *ddr, *ddq;
    __device__ void operator()(long i, double* a) const {
        double e = dt[i] - dp[i];
        a[0] += e * e;
        a[1] += e * ddr[i];
        a[2] += e * ddq[i];
    }
};

__global__ void k_quad_update(double* __restrict__ dt, const double* __restrict__ dp, const double* __restrict__ ddr,
                              const double* __restrict__ ddq, double* __restrict__ dd, double root, long n) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    double di = dt[i] - dp[i] + ddr[i] + root * ddq[i];
    dt[i] = di + dp[i];
    dd[i] = ddr[i] + root * ddq[i];
}

template <class Op, int NOUT>
void fused(Dev& d, const Op& op, long n, double* out) {
    int nb = int(std::min<long>(RB, std::max<long>(1, (n + TPB - 1) / TPB)));
    hipLaunchKernelGGL((k_fused_sums<Op, NOUT>), dim3(nb), dim3(TPB), 0, d.s, op, n, d.red);
    hipLaunchKernelGGL(k_reduce_parts, dim3(NOUT), dim3(64), 0, d.s, d.red, nb, NOUT, d.red + DOTS_PARTS);
    CB_CHECK(hipGetLastError());
    CB_CHECK(hipMemcpyAsync(d.red_host, d.red + DOTS_PARTS, sizeof(double) * NOUT, hipMemcpyDeviceToHost, d.s));
    CB_CHECK(hipStreamSynchronize(d.s));
    for (int k = 0; k < NOUT; ++k) out[k] = d.red_host[k];
}

}  // namespace

void test_sums(Dev& d, const double* dd, const double* dt, const double* q, double lam,
               const double* f, const double* fp, const double* fip, long n, double out[5]) {
    TestOp op{dd, dt, q, f, fp, fip, lam};
    fused<TestOp, 5>(d, op, n, out);
}

void quad_sums(Dev& d, const double* dt, const double* dp, const double* ddr, const double* ddq,
               long n, double out[3]) {
    QuadOp op{dt, dp, ddr, ddq};
    fused<QuadOp, 3>(d, op, n, out);
}

void quad_update(Dev& d, double* dt, const double* dp, const double* ddr, const double* ddq,
                 double* dd, double root, long n) {
    hipLaunchKernelGGL(k_quad_update, dim3(nblk(n)), dim3(TPB), 0, d.s, dt, dp, ddr, ddq, dd, root, n);
}

// ------------------------------------------------------------------ FGMRES
void Fgmres::setup(long n_, int m_) {
    release();
    n = n_;
    m = std::max(2, std::min(m_, MAXV - 1));
    V = dalloc<double>(size_t(m + 1) * size_t(n));
    Z = dalloc<double>(size_t(m) * size_t(n));
    w = dalloc<double>(size_t(n));
    hd = dalloc<double>(MAXV + 8);
    part = dalloc<double>(size_t(MAXV) * RB + 64);
    CB_CHECK(hipHostMalloc(&hh, sizeof(double) * (MAXV + 8), hipHostMallocDefault));
}

void Fgmres::release() {
    void* ps[] = {V, Z, w, hd, part};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    if (hh) (void)hipHostFree(hh);
    V = Z = w = hd = part = hh = nullptr;
}

int Fgmres::solve(Dev& d, const std::function<void(const double*, double*)>& prec, const double* b,
                  double* x, double tol, int maxit, double* relres) {
    const int nb = int(std::min<long>(RB, std::max<long>(1, (n + TPB - 1) / TPB)));
    auto dots = [&](int cnt) {   // hd[0..cnt) = V[0..cnt)^T w
        hipLaunchKernelGGL(k_gemv_t, dim3(nb), dim3(TPB), 0, d.s, V, w, n, cnt, part);
        hipLaunchKernelGGL(k_reduce_parts, dim3(cnt), dim3(64), 0, d.s, part, nb, cnt, hd);
    };
    auto norm2 = [&](const double* v) {
        const double* a[1] = {v};
        double r;
        multi_dot(d, 1, a, a, n, &r);
        return r;
    };
    double bnorm = std::sqrt(norm2(b));
    CB_CHECK(hipMemsetAsync(x, 0, sizeof(double) * size_t(n), d.s));
    if (bnorm == 0.0) { *relres = 0; return 0; }
    // r0 = b (x0 = 0)
    int total = 0;
    double beta = bnorm, rel = 1.0;
    std::vector<double> H((size_t)((m + 1) * m)), cs((size_t)m), sn((size_t)m), g((size_t)(m + 1));
    bool first = true;
    while (total < maxit) {
        if (first) {
            hipLaunchKernelGGL(k_scale_copy, dim3(nblk(n)), dim3(TPB), 0, d.s, V, b, 1.0 / beta, n);
            first = false;
        } else {
            bsr_spmv(d, x, w);
            hipLaunchKernelGGL(k_sub, dim3(nblk(n)), dim3(TPB), 0, d.s, w, b, w, n);
            beta = std::sqrt(norm2(w));
            rel = beta / bnorm;
            if (rel <= tol) break;
            hipLaunchKernelGGL(k_scale_copy, dim3(nblk(n)), dim3(TPB), 0, d.s, V, w, 1.0 / beta, n);
        }
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        int j = 0;
        for (; j < m && total < maxit; ++j, ++total) {
            double* vj = V + size_t(j) * n;
            double* zj = Z + size_t(j) * n;
            prec(vj, zj);
            bsr_spmv(d, zj, w);
            // CGS2
            double hcol[MAXV + 1];
            for (int pass = 0; pass < 2; ++pass) {
                dots(j + 1);
                hipLaunchKernelGGL(k_gemv_n_sub, dim3(nb), dim3(TPB), 0, d.s, V, hd, w, n, j + 1, part);
                hipLaunchKernelGGL(k_reduce_parts, dim3(1), dim3(64), 0, d.s, part, nb, 1, hd + MAXV);
                CB_CHECK(hipMemcpyAsync(hh, hd, sizeof(double) * (MAXV + 1), hipMemcpyDeviceToHost, d.s));
                CB_CHECK(hipStreamSynchronize(d.s));
                for (int k = 0; k <= j; ++k) hcol[k] = (pass == 0 ? 0.0 : hcol[k]) + hh[k];
            }
            double hn = std::sqrt(std::max(hh[MAXV], 0.0));
            hcol[j + 1] = hn;
            if (hn > 0)
                hipLaunchKernelGGL(k_scale_copy, dim3(nblk(n)), dim3(TPB), 0, d.s, V + size_t(j + 1) * n, w,
                                   1.0 / hn, n);
            for (int k = 0; k < j; ++k) {
                double t0 = cs[k] * hcol[k] + sn[k] * hcol[k + 1];
                hcol[k + 1] = -sn[k] * hcol[k] + cs[k] * hcol[k + 1];
                hcol[k] = t0;
            }
            double den = std::hypot(hcol[j], hcol[j + 1]);
            cs[j] = den > 0 ? hcol[j] / den : 1.0;
            sn[j] = den > 0 ? hcol[j + 1] / den : 0.0;
            hcol[j] = den;
            hcol[j + 1] = 0;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            for (int k = 0; k <= j; ++k) H[size_t(k) * m + j] = hcol[k];
            rel = std::fabs(g[j + 1]) / bnorm;
            if (rel <= tol || hn == 0.0) { ++j; ++total; break; }
        }
        // back substitution, x += Z y
        std::vector<double> y((size_t)j);
        for (int k = j - 1; k >= 0; --k) {
            double s = g[k];
            for (int q = k + 1; q < j; ++q) s -= H[size_t(k) * m + q] * y[q];
            y[k] = s / H[size_t(k) * m + k];
        }
        std::copy(y.begin(), y.end(), hh);
        CB_CHECK(hipMemcpyAsync(hd, hh, sizeof(double) * size_t(j), hipMemcpyHostToDevice, d.s));
        hipLaunchKernelGGL(k_update_x, dim3(nblk(n)), dim3(TPB), 0, d.s, x, Z, hd, j, n);
        CB_CHECK(hipGetLastError());
        if (rel <= tol) break;
    }
    // true residual
    bsr_spmv(d, x, w);
    hipLaunchKernelGGL(k_sub, dim3(nblk(n)), dim3(TPB), 0, d.s, w, b, w, n);
    *relres = std::sqrt(norm2(w)) / bnorm;
    return total;
}

// ------------------------------------------------------------------ block Jacobi
void BlockJacobi::setup(Dev& d) {
    if (!inv) inv = dalloc<double>(size_t(d.nj * d.bs * d.bs));
    if (d.bs == 7)
        hipLaunchKernelGGL(k_bj_setup<7>, dim3(nblk(d.nj)), dim3(TPB), 0, d.s, d.K, d.diagblk, inv, d.nj);
    else if (d.bs == 6)
        hipLaunchKernelGGL(k_bj_setup<6>, dim3(nblk(d.nj)), dim3(TPB), 0, d.s, d.K, d.diagblk, inv, d.nj);
    else
        hipLaunchKernelGGL(k_bj_setup<3>, dim3(nblk(d.nj)), dim3(TPB), 0, d.s, d.K, d.diagblk, inv, d.nj);
    CB_CHECK(hipGetLastError());
}

void BlockJacobi::apply(Dev& d, const double* v, double* z) {
    if (d.bs == 7)
        hipLaunchKernelGGL(k_bj_apply<7>, dim3(nblk(d.nslot)), dim3(TPB), 0, d.s, inv, v, z, d.nj);
    else if (d.bs == 6)
        hipLaunchKernelGGL(k_bj_apply<6>, dim3(nblk(d.nslot)), dim3(TPB), 0, d.s, inv, v, z, d.nj);
    else
        hipLaunchKernelGGL(k_bj_apply<3>, dim3(nblk(d.nslot)), dim3(TPB), 0, d.s, inv, v, z, d.nj);
}

void BlockJacobi::release() {
    if (inv) (void)hipFree(inv);
    inv = nullptr;
}

}  // namespace cb
