// Learned block approximate-inverse preconditioner: features, the message-
// passing network (forward and backward) on MFMA in BF16 with FP32
// accumulation, the preconditioner blocks, their application inside FGMRES,
// the training loss and Adam.  See learned.h for the formulation.
#include "learned.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

namespace cb {
namespace {

constexpr int TPB = 256;
inline int nblk(long n, int tpb = TPB) { return int((n + tpb - 1) / tpb); }

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ bf16_t f2bf(float x) {
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);   // round to nearest even (finite inputs)
    return bf16_t(u >> 16);
}
__device__ __forceinline__ float bf2f(bf16_t b) { return __uint_as_float(uint32_t(b) << 16); }

// ------------------------------------------------------------ features
__global__ void k_lp_scale(const double* __restrict__ K, const int32_t* __restrict__ diagblk,
                           const int32_t* __restrict__ eq, float* __restrict__ scale,
                           float* __restrict__ mask, long nslot) {
    long s = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (s >= nslot) return;
    const long j = s / 6;
    const int p = int(s % 6);
    const double kpp = K[(long)diagblk[j] * 36 + p * 6 + p];
    const bool free_ = eq[s] >= 0;
    const double a = fabs(kpp);
    scale[s] = (free_ && a > 0) ? float(1.0 / sqrt(a)) : 1.0f;
    mask[s] = free_ ? 1.0f : 0.0f;
}

// node features: scaled diagonal block (36), 0.1 log10 |K_pp| (6), free mask (6);
// C_i = inverse Cholesky factor of the scaled block (identity if not SPD)
__global__ void k_lp_node(const double* __restrict__ K, const int32_t* __restrict__ diagblk,
                          const float* __restrict__ scale, const float* __restrict__ mask,
                          bf16_t* __restrict__ xn, float* __restrict__ cblk, long nv) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= nv) return;
    const double* Kb = K + (long)diagblk[i] * 36;
    double A[6][6];
    float s[6], m[6];
#pragma unroll
    for (int p = 0; p < 6; ++p) { s[p] = scale[i * 6 + p]; m[p] = mask[i * 6 + p]; }
#pragma unroll
    for (int p = 0; p < 6; ++p)
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            double v = double(s[p]) * Kb[p * 6 + q] * double(s[q]);
            if (m[p] == 0.f || m[q] == 0.f) v = (p == q) ? 1.0 : 0.0;
            A[p][q] = v;
            xn[i * 48 + p * 6 + q] = f2bf(float(v));
        }
#pragma unroll
    for (int p = 0; p < 6; ++p) {
        const double kpp = fabs(Kb[p * 6 + p]);
        xn[i * 48 + 36 + p] = f2bf(m[p] != 0.f ? float(0.1 * log10(kpp + 1e-300)) : 0.f);
        xn[i * 48 + 42 + p] = f2bf(m[p]);
    }
    // Cholesky A = L L^T, then C = L^{-1}
    double L[6][6] = {};
    bool ok = true;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        double d = A[c][c];
#pragma unroll
        for (int k = 0; k < c; ++k) d -= L[c][k] * L[c][k];
        if (!(d > 1e-12)) { ok = false; d = 1.0; }
        L[c][c] = sqrt(d);
#pragma unroll
        for (int r = c + 1; r < 6; ++r) {
            double v = A[r][c];
#pragma unroll
            for (int k = 0; k < c; ++k) v -= L[r][k] * L[c][k];
            L[r][c] = v / L[c][c];
        }
    }
    double Ci[6][6] = {};
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        // column c of L^{-1} by forward substitution
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            double v = (r == c) ? 1.0 : 0.0;
#pragma unroll
            for (int k = 0; k < r; ++k) v -= L[r][k] * Ci[k][c];
            Ci[r][c] = (r >= c) ? v / L[r][r] : 0.0;
        }
    }
#pragma unroll
    for (int p = 0; p < 6; ++p)
#pragma unroll
        for (int q = 0; q < 6; ++q) cblk[i * 36 + p * 6 + q] = ok ? float(Ci[p][q]) : (p == q ? 1.f : 0.f);
}

// edge features: scaled off-diagonal block (36), free masks of both joints
__global__ void k_lp_edge(const double* __restrict__ K, const int32_t* __restrict__ eblk,
                          const int32_t* __restrict__ erow, const int32_t* __restrict__ ecol,
                          const float* __restrict__ scale, const float* __restrict__ mask,
                          bf16_t* __restrict__ xe, long ne) {
    long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (e >= ne) return;
    const double* Kb = K + (long)eblk[e] * 36;
    const long i = erow[e], j = ecol[e];
#pragma unroll
    for (int p = 0; p < 6; ++p)
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            double v = double(scale[i * 6 + p]) * Kb[p * 6 + q] * double(scale[j * 6 + q]);
            if (mask[i * 6 + p] == 0.f || mask[j * 6 + q] == 0.f) v = 0.0;
            xe[e * 48 + p * 6 + q] = f2bf(float(v));
        }
#pragma unroll
    for (int p = 0; p < 6; ++p) {
        xe[e * 48 + 36 + p] = f2bf(mask[i * 6 + p]);
        xe[e * 48 + 42 + p] = f2bf(mask[j * 6 + p]);
    }
}

// ------------------------------------------------------------ GEMMs
// A row r is the concatenation of up to three BF16 segments, each gathered by
// an optional row index: seg s contributes base[(idx ? idx[r] : r) * width + k].
struct Seg {
    const bf16_t* base;
    const int32_t* idx;
    int width;
};
struct GemmArgs {
    Seg seg[3];
    int nseg;
    int K;                  // sum of segment widths, multiple of 16
    const bf16_t* Bt;       // [N][K]: the right operand, transposed
    const float* bias;      // [N] or null
    int relu;
    long M;
    float* Yf;              // act(A B + bias) [M][N] fp32, or null
    bf16_t* Yb;             // same in bf16 when no residual output, or null
    const float* R;         // residual input [M][N]: Z = R + Y
    float* Zf;
    bf16_t* Zb;
};

// C = A B (+bias, relu, residual).  256 threads = 4 waves; a block covers 64
// rows, a wave 16 rows by all N = 16*NT columns, one MFMA 16x16x16 per
// (16-wide K step, column tile).  The right operand sits in LDS.
template <int NT>
__global__ void __launch_bounds__(256) k_gemm(GemmArgs g) {
    extern __shared__ bf16_t sB[];
    constexpr int N = NT * 16;
    const int K = g.K, ldb = K + 4;
    for (int t = threadIdx.x; t < N * K; t += 256) {
        const int n = t / K, k = t - n * K;
        sB[n * ldb + k] = g.Bt[t];
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const long row0 = (long)blockIdx.x * 64 + w * 16;
    const long r = row0 + (lane & 15);
    const bool rv = r < g.M;
    f4 acc[NT];
#pragma unroll
    for (int n = 0; n < NT; ++n) acc[n] = f4{0.f, 0.f, 0.f, 0.f};
    int kbase = 0;
    for (int s = 0; s < g.nseg; ++s) {
        const Seg sg = g.seg[s];
        const bf16_t* rp = nullptr;
        if (rv) rp = sg.base + (long)(sg.idx ? sg.idx[r] : r) * sg.width;
        for (int kk = 0; kk < sg.width; kk += 16) {
            const int ko = kk + (lane >> 4) * 4;
            s4 a = s4{0, 0, 0, 0};
            if (rv) a = *reinterpret_cast<const s4*>(rp + ko);
#pragma unroll
            for (int n = 0; n < NT; ++n) {
                const s4 b = *reinterpret_cast<const s4*>(sB + (n * 16 + (lane & 15)) * ldb + kbase + ko);
                acc[n] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[n], 0, 0, 0);
            }
        }
        kbase += sg.width;
    }
#pragma unroll
    for (int n = 0; n < NT; ++n)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const long rr = row0 + (lane >> 4) * 4 + v;
            const int c = n * 16 + (lane & 15);
            if (rr >= g.M) continue;
            float y = acc[n][v] + (g.bias ? g.bias[c] : 0.f);
            if (g.relu) y = fmaxf(y, 0.f);
            if (g.Yf) g.Yf[rr * N + c] = y;
            if (g.Zf) {
                const float z = g.R[rr * N + c] + y;
                g.Zf[rr * N + c] = z;
                g.Zb[rr * N + c] = f2bf(z);
            } else if (g.Yb) {
                g.Yb[rr * N + c] = f2bf(y);
            }
        }
}

// Weight gradient: partial[blk][K][N] = sum over this block's rows of
// A[r]^T D[r]; rows in chunks of 64 staged through LDS, contraction over 16
// rows per MFMA.  Wave w owns the K tiles kt = w, w + 4, ...
template <int NT>
__global__ void __launch_bounds__(256) k_gemm_dw(GemmArgs g, const bf16_t* __restrict__ D,
                                                 float* __restrict__ partial) {
    constexpr int N = NT * 16;
    constexpr int LDA = 192 + 8, LDD = 64 + 8;
    __shared__ bf16_t sA[64 * LDA];
    __shared__ bf16_t sD[64 * LDD];
    const int K = g.K, KT = K / 16;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    f4 acc[3][NT];
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
        for (int n = 0; n < NT; ++n) acc[q][n] = f4{0.f, 0.f, 0.f, 0.f};
    const long nchunk = (g.M + 63) / 64;
    for (long c = blockIdx.x; c < nchunk; c += gridDim.x) {
        // stage 64 rows of A (4 bf16 per load) and of D
        for (int t = threadIdx.x; t < 64 * (K / 4); t += 256) {
            const int rl = t / (K / 4), k4 = (t - rl * (K / 4)) * 4;
            const long r = c * 64 + rl;
            s4 v = s4{0, 0, 0, 0};
            if (r < g.M) {
                int kb = k4, s = 0;
                while (kb >= g.seg[s].width) { kb -= g.seg[s].width; ++s; }
                const Seg sg = g.seg[s];
                v = *reinterpret_cast<const s4*>(sg.base + (long)(sg.idx ? sg.idx[r] : r) * sg.width + kb);
            }
            *reinterpret_cast<s4*>(sA + rl * LDA + k4) = v;
        }
        for (int t = threadIdx.x; t < 64 * (N / 4); t += 256) {
            const int rl = t / (N / 4), n4 = (t - rl * (N / 4)) * 4;
            const long r = c * 64 + rl;
            s4 v = s4{0, 0, 0, 0};
            if (r < g.M) v = *reinterpret_cast<const s4*>(D + r * N + n4);
            *reinterpret_cast<s4*>(sD + rl * LDD + n4) = v;
        }
        __syncthreads();
#pragma unroll
        for (int sub = 0; sub < 4; ++sub) {
            const int rb = sub * 16 + (lane >> 4) * 4;
            s4 bfr[NT];
#pragma unroll
            for (int n = 0; n < NT; ++n) {
                const int col = n * 16 + (lane & 15);
                bfr[n] = s4{short(sD[(rb + 0) * LDD + col]), short(sD[(rb + 1) * LDD + col]),
                            short(sD[(rb + 2) * LDD + col]), short(sD[(rb + 3) * LDD + col])};
            }
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const int kt = w + 4 * q;
                if (kt >= KT) continue;   // uniform across the wave
                const int col = kt * 16 + (lane & 15);
                const s4 a = s4{short(sA[(rb + 0) * LDA + col]), short(sA[(rb + 1) * LDA + col]),
                                short(sA[(rb + 2) * LDA + col]), short(sA[(rb + 3) * LDA + col])};
#pragma unroll
                for (int n = 0; n < NT; ++n)
                    acc[q][n] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bfr[n], acc[q][n], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    float* out = partial + (long)blockIdx.x * K * N;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const int kt = w + 4 * q;
        if (kt >= KT) continue;
#pragma unroll
        for (int n = 0; n < NT; ++n)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int k = kt * 16 + (lane >> 4) * 4 + v;
                out[k * N + n * 16 + (lane & 15)] = acc[q][n][v];
            }
    }
}

__global__ void k_reduce_partial(const float* __restrict__ partial, int nb, long len, float* __restrict__ out) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= len) return;
    float s = 0.f;
    for (int b = 0; b < nb; ++b) s += partial[(long)b * len + i];
    out[i] = s;
}

// column sums of Y [M][N] (bias gradients): partial per block, then reduce
__global__ void k_colsum_partial(const float* __restrict__ Y, long M, int N, float* __restrict__ partial) {
    const int c = threadIdx.x % 64, rs = threadIdx.x / 64;   // 4 row streams x 64 columns
    float s = 0.f;
    if (c < N)
        for (long r = (long)blockIdx.x * 4 + rs; r < M; r += (long)gridDim.x * 4) s += Y[r * N + c];
    __shared__ float sh[256];
    sh[threadIdx.x] = s;
    __syncthreads();
    if (rs == 0 && c < N)
        partial[(long)blockIdx.x * N + c] = sh[c] + sh[64 + c] + sh[128 + c] + sh[192 + c];
}

// ------------------------------------------------------------ graph ops
// A[i][f] = sum over the edges of row i of M[e][f] (edges grouped by row)
__global__ void k_segsum_row(const float* __restrict__ X, int ldx, int off, const int32_t* __restrict__ eptr,
                             long nv, float* __restrict__ out, int accumulate, bf16_t* __restrict__ outb) {
    long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= nv * 64) return;
    const long i = t / 64;
    const int f = int(t % 64);
    float s = 0.f;
    for (int32_t e = eptr[i]; e < eptr[i + 1]; ++e) s += X[(long)e * ldx + off + f];
    if (out) out[t] = accumulate ? out[t] + s : s;
    if (outb) outb[t] = f2bf(s);
}

// out[j][f] += sum over edges (i <- j) of X[e][f]: the edges with column j are
// the reverses of row j's edges
__global__ void k_segsum_col(const float* __restrict__ X, int ldx, int off, const int32_t* __restrict__ eptr,
                             const int32_t* __restrict__ erev, long nv, float* __restrict__ out) {
    long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= nv * 64) return;
    const long j = t / 64;
    const int f = int(t % 64);
    float s = 0.f;
    for (int32_t e = eptr[j]; e < eptr[j + 1]; ++e) s += X[(long)erev[e] * ldx + off + f];
    out[t] += s;
}

// ------------------------------------------------------------ elementwise
__global__ void k_f2b(const float* __restrict__ x, bf16_t* __restrict__ y, long n) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) y[i] = f2bf(x[i]);
}
// y = x * (a > 0): ReLU backward, fp32 and bf16 outputs
__global__ void k_relu_bwd(const float* __restrict__ x, const float* __restrict__ a, float* __restrict__ y,
                           bf16_t* __restrict__ yb, long n) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float v = a[i] > 0.f ? x[i] : 0.f;
    y[i] = v;
    yb[i] = f2bf(v);
}
// dM[e] = (dE[e] + dA[row(e)]) * (M[e] > 0)
__global__ void k_msg_bwd(const float* __restrict__ de, const float* __restrict__ dA,
                          const int32_t* __restrict__ erow, const float* __restrict__ mf,
                          float* __restrict__ y, bf16_t* __restrict__ yb, long ne) {
    long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= ne * 64) return;
    const long e = t / 64;
    const int f = int(t % 64);
    const float v = mf[t] > 0.f ? de[t] + dA[(long)erow[e] * 128 + 64 + f] : 0.f;
    y[t] = v;
    yb[t] = f2bf(v);
}
// dst[r][0:64] += src[r][off:off+64] (row widths ldd / lds)
__global__ void k_add_cols(float* __restrict__ dst, const float* __restrict__ src, int lds, int off, long rows) {
    long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= rows * 64) return;
    dst[t] += src[(t / 64) * lds + off + (t % 64)];
}
__global__ void k_set_cols(float* __restrict__ dst, const float* __restrict__ src, int lds, int off, long rows) {
    long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= rows * 64) return;
    dst[t] = src[(t / 64) * lds + off + (t % 64)];
}

// preconditioner blocks from the decoder outputs
__global__ void k_build_g(const float* __restrict__ cblk, const float* __restrict__ dd, const float* __restrict__ dof,
                          const float* __restrict__ mask, const int32_t* __restrict__ erow,
                          const int32_t* __restrict__ ecol, float* __restrict__ gd, float* __restrict__ go,
                          long nv, long ne, float alpha) {
    long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t < nv * 36) {
        const long i = t / 36;
        const int p = int(t % 36) / 6, q = int(t % 6);
        gd[t] = cblk[t] + alpha * mask[i * 6 + p] * mask[i * 6 + q] * dd[i * 48 + p * 6 + q];
    } else if (t < nv * 36 + ne * 36) {
        const long u = t - nv * 36;
        const long e = u / 36;
        const int p = int(u % 36) / 6, q = int(u % 6);
        const long i = erow[e], j = ecol[e];
        go[u] = (i > j) ? alpha * mask[i * 6 + p] * mask[j * 6 + q] * dof[e * 48 + p * 6 + q] : 0.f;
    }
}

// y = G x on slot vectors: y_i = G_ii x_i + sum_{j<i} G_ij x_j
__global__ void k_apply_g(const float* __restrict__ gd, const float* __restrict__ go,
                          const int32_t* __restrict__ eptr, const int32_t* __restrict__ ecol,
                          const double* __restrict__ x, double* __restrict__ y, long nv) {
    long s = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (s >= nv * 6) return;
    const long i = s / 6;
    const int p = int(s % 6);
    double acc = 0;
#pragma unroll
    for (int q = 0; q < 6; ++q) acc += double(gd[i * 36 + p * 6 + q]) * x[i * 6 + q];
    for (int32_t e = eptr[i]; e < eptr[i + 1]; ++e) {
        const long j = ecol[e];
        if (j >= i) continue;
#pragma unroll
        for (int q = 0; q < 6; ++q) acc += double(go[(long)e * 36 + p * 6 + q]) * x[j * 6 + q];
    }
    y[s] = acc;
}

// z = G^T y: z_j = G_jj^T y_j + sum_{i>j} G_ij^T y_i (G_ij is the reverse of j's edge to i)
__global__ void k_apply_gt(const float* __restrict__ gd, const float* __restrict__ go,
                           const int32_t* __restrict__ eptr, const int32_t* __restrict__ ecol,
                           const int32_t* __restrict__ erev, const double* __restrict__ y,
                           double* __restrict__ z, long nv) {
    long s = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (s >= nv * 6) return;
    const long j = s / 6;
    const int q = int(s % 6);
    double acc = 0;
#pragma unroll
    for (int p = 0; p < 6; ++p) acc += double(gd[j * 36 + p * 6 + q]) * y[j * 6 + p];
    for (int32_t e = eptr[j]; e < eptr[j + 1]; ++e) {
        const long i = ecol[e];
        if (i <= j) continue;
        const long er = erev[e];   // edge (i <- j)
#pragma unroll
        for (int p = 0; p < 6; ++p) acc += double(go[er * 36 + p * 6 + q]) * y[i * 6 + p];
    }
    z[s] = acc;
}

__global__ void k_scale_vec(const float* __restrict__ sc, const double* __restrict__ x, double* __restrict__ y,
                            long n, int divide) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) y[i] = divide ? x[i] / double(sc[i]) : x[i] * double(sc[i]);
}

__global__ void k_masked_diff(double* __restrict__ z, const double* __restrict__ u,
                              const float* __restrict__ mask, long n) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) z[i] = mask[i] != 0.f ? z[i] - u[i] : 0.0;
}

// y = S v (pre) / y = mask S v (post) around the BSR product
__global__ void k_scale_mask(const float* __restrict__ sc, const float* __restrict__ mask,
                             const double* __restrict__ v, double* __restrict__ y, long n, int masked) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double w = v[i] * double(sc[i]);
    y[i] = (masked && mask[i] == 0.f) ? 0.0 : w;
}

// partial sums of mask a.b and mask c.d per block
__global__ void k_dots2(const double* __restrict__ a, const double* __restrict__ b,
                        const double* __restrict__ c, const double* __restrict__ e,
                        const float* __restrict__ mask, long n, double* __restrict__ part) {
    __shared__ double s1[TPB], s2[TPB];
    double p = 0, q = 0;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        if (mask[i] == 0.f) continue;
        p += a[i] * b[i];
        q += c[i] * e[i];
    }
    s1[threadIdx.x] = p;
    s2[threadIdx.x] = q;
    __syncthreads();
    for (int w = TPB / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            s1[threadIdx.x] += s1[threadIdx.x + w];
            s2[threadIdx.x] += s2[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[blockIdx.x * 2] = s1[0];
        part[blockIdx.x * 2 + 1] = s2[0];
    }
}

__global__ void k_scale_copy(const double* __restrict__ x, double* __restrict__ y, double a, long n) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) y[i] = a * x[i];
}

__global__ void k_scale_inplace(double* __restrict__ x, double a, long n) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) x[i] *= a;
}

// dG blocks: dG_ij[p][q] += y_ip gz_jq + gy_ip x_jq (diagonal: j = i; edges i > j)
__global__ void k_dg(const double* __restrict__ y, const double* __restrict__ gz, const double* __restrict__ gy,
                     const double* __restrict__ x, const int32_t* __restrict__ erow,
                     const int32_t* __restrict__ ecol, float* __restrict__ dgd, float* __restrict__ dgo,
                     long nv, long ne) {
    long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t < nv * 36) {
        const long i = t / 36;
        const int p = int(t % 36) / 6, q = int(t % 6);
        dgd[t] += float(y[i * 6 + p] * gz[i * 6 + q] + gy[i * 6 + p] * x[i * 6 + q]);
    } else if (t < nv * 36 + ne * 36) {
        const long u = t - nv * 36;
        const long e = u / 36;
        const int p = int(u % 36) / 6, q = int(u % 6);
        const long i = erow[e], j = ecol[e];
        if (i > j) dgo[u] += float(y[i * 6 + p] * gz[j * 6 + q] + gy[i * 6 + p] * x[j * 6 + q]);
    }
}

// decoder-output gradients: dD = alpha * mask (x) mask * dG, padded to 48 columns
__global__ void k_dg_to_dout(const float* __restrict__ dgd, const float* __restrict__ dgo,
                             const float* __restrict__ mask, const int32_t* __restrict__ erow,
                             const int32_t* __restrict__ ecol, float* __restrict__ ddn,
                             float* __restrict__ dde, long nv, long ne, float alpha) {
    long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t < nv * 48) {
        const long i = t / 48;
        const int c = int(t % 48);
        float v = 0.f;
        if (c < 36) {
            const int p = c / 6, q = c % 6;
            v = alpha * mask[i * 6 + p] * mask[i * 6 + q] * dgd[i * 36 + c];
        }
        ddn[t] = v;
    } else if (t < nv * 48 + ne * 48) {
        const long u = t - nv * 48;
        const long e = u / 48;
        const int c = int(u % 48);
        float v = 0.f;
        const long i = erow[e], j = ecol[e];
        if (c < 36 && i > j) {
            const int p = c / 6, q = c % 6;
            v = alpha * mask[i * 6 + p] * mask[j * 6 + q] * dgo[e * 36 + c];
        }
        dde[u] = v;
    }
}

// bf16 copies of every weight: W^T [N][K] and W [K][N]
__global__ void k_weights_bf16(const float* __restrict__ W, int K, int N, bf16_t* __restrict__ wt,
                               bf16_t* __restrict__ wn) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= K * N) return;
    const int k = t / N, n = t % N;
    const bf16_t b = f2bf(W[t]);
    wn[t] = b;
    wt[n * K + k] = b;
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m1,
                       float* __restrict__ m2, long n, float lr, float c1, float c2) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float a = 0.9f * m1[i] + 0.1f * gi;
    const float b = 0.999f * m2[i] + 0.001f * gi * gi;
    m1[i] = a;
    m2[i] = b;
    p[i] -= lr * (a / c1) / (sqrtf(b / c2) + 1e-8f);
}

// ------------------------------------------------------------ launch helpers
template <class T> T* alloc(size_t n) { return dalloc<T>(n); }

void gemm(hipStream_t s, const GemmArgs& g, int N) {
    if (g.M <= 0) return;
    const int grid = int((g.M + 63) / 64);
    const size_t lds = size_t(N) * (g.K + 4) * sizeof(bf16_t);
    switch (N / 16) {
    case 3: hipLaunchKernelGGL(k_gemm<3>, dim3(grid), dim3(256), lds, s, g); break;
    case 4: hipLaunchKernelGGL(k_gemm<4>, dim3(grid), dim3(256), lds, s, g); break;
    case 8: hipLaunchKernelGGL(k_gemm<8>, dim3(grid), dim3(256), lds, s, g); break;
    case 12: hipLaunchKernelGGL(k_gemm<12>, dim3(grid), dim3(256), lds, s, g); break;
    default: throw DeviceError("unsupported GEMM width");
    }
    CB_CHECK(hipGetLastError());
}

GemmArgs args1(const bf16_t* a, int w) {
    GemmArgs g{};
    g.seg[0] = Seg{a, nullptr, w};
    g.nseg = 1;
    g.K = w;
    return g;
}

}  // namespace

// ============================================================== Learned
Learned::Learned(unsigned seed) {
    // weights in order: Wn, We, Wm[3], Wu[3], Wo, Wd; each followed by its bias
    const int dims[10][2] = {{FIN, H}, {FIN, H}, {3 * H, H}, {3 * H, H}, {3 * H, H},
                             {2 * H, H}, {2 * H, H}, {2 * H, H}, {3 * H, FOUT}, {H, FOUT}};
    long off = 0, boff = 0;
    for (auto& dm : dims) {
        w_.push_back(Wdesc{off, dm[0], dm[1], boff});
        off += long(dm[0]) * dm[1] + dm[1];
        boff += long(dm[0]) * dm[1];
    }
    np_ = off;
    std::vector<float> h(size_t(np_), 0.f);
    // SplitMix64 uniform in [-a, a], a = sqrt(6/(K+N)); decoder weights start at zero
    uint64_t st = 0x9E3779B97F4A7C15ull * (uint64_t(seed) + 1);
    auto next = [&]() {
        st += 0x9E3779B97F4A7C15ull;
        uint64_t z = st;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return double(z >> 11) * (1.0 / 9007199254740992.0);
    };
    for (size_t wi = 0; wi < w_.size(); ++wi) {
        const Wdesc& d = w_[wi];
        const double a = std::sqrt(6.0 / (d.K + d.N));
        for (long k = 0; k < long(d.K) * d.N; ++k)
            h[size_t(d.off + k)] = (wi >= 8) ? 0.f : float(a * (2 * next() - 1));
    }
    p_ = alloc<float>(size_t(np_));
    g_ = alloc<float>(size_t(np_));
    m1_ = alloc<float>(size_t(np_));
    m2_ = alloc<float>(size_t(np_));
    CB_CHECK(hipMemcpy(p_, h.data(), sizeof(float) * size_t(np_), hipMemcpyHostToDevice));
    CB_CHECK(hipMemset(g_, 0, sizeof(float) * size_t(np_)));
    CB_CHECK(hipMemset(m1_, 0, sizeof(float) * size_t(np_)));
    CB_CHECK(hipMemset(m2_, 0, sizeof(float) * size_t(np_)));
    wt_ = alloc<bf16_t>(size_t(boff));
    wn_ = alloc<bf16_t>(size_t(boff));
    nblk_dw_ = 256;
    partial_ = alloc<float>(size_t(nblk_dw_) * 3 * H * H);
}

Learned::~Learned() {
    release_graph();
    for (void* p : {(void*)p_, (void*)g_, (void*)m1_, (void*)m2_, (void*)wt_, (void*)wn_, (void*)partial_})
        if (p) (void)hipFree(p);
}

void Learned::release_graph() {
    void* ps[] = {erow_, ecol_, eblk_, eptr_, erev_, xn_, xe_, scale_, cblk_, gd_, go_, mask_, dof_, ddf_,
                  dh_, de_, dtmp_e_, dtmp_n_, dga_, dyb_e_, dyb_n_, dgd_, dgo_, v1_, v2_, v3_, v4_, v5_, v6_, red_};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    for (int l = 0; l <= NL; ++l) {
        void* qs[] = {hf_[l], ef_[l], mf_[l], uf_[l], hb_[l], eb_[l], ab_[l]};
        for (void* p : qs)
            if (p) (void)hipFree(p);
        hf_[l] = ef_[l] = mf_[l] = uf_[l] = nullptr;
        hb_[l] = eb_[l] = ab_[l] = nullptr;
    }
    if (red_host_) (void)hipHostFree(red_host_);
    erow_ = ecol_ = eblk_ = eptr_ = erev_ = nullptr;
    xn_ = xe_ = nullptr;
    scale_ = cblk_ = gd_ = go_ = mask_ = dof_ = ddf_ = dh_ = de_ = dtmp_e_ = dtmp_n_ = dga_ = nullptr;
    dyb_e_ = dyb_n_ = nullptr;
    dgd_ = dgo_ = nullptr;
    v1_ = v2_ = v3_ = v4_ = v5_ = v6_ = red_ = nullptr;
    red_host_ = nullptr;
    owner_ = 0;
}

void Learned::get_params(float* host) const {
    CB_CHECK(hipMemcpy(host, p_, sizeof(float) * size_t(np_), hipMemcpyDeviceToHost));
}
void Learned::set_params(const float* host) {
    CB_CHECK(hipMemcpy(p_, host, sizeof(float) * size_t(np_), hipMemcpyHostToDevice));
}
void Learned::get_grad(float* host) const {
    CB_CHECK(hipMemcpy(host, g_, sizeof(float) * size_t(np_), hipMemcpyDeviceToHost));
}

void Learned::bind(Dev& d) {
    if (owner_ == d.uid) return;
    release_graph();
    if (d.bs != 6) throw DeviceError("the learned preconditioner needs 6-slot joints");
    owner_ = d.uid;
    s_ = d.s;
    nv_ = d.nj;
    nslot_ = d.nslot;
    // edges = off-diagonal joint blocks, grouped by row
    std::vector<int32_t> rp(size_t(nv_ + 1)), ci(size_t(d.nnzb));
    d2h(rp.data(), d.rowptr, rp.size(), d.s);
    d2h(ci.data(), d.colidx, ci.size(), d.s);
    CB_CHECK(hipStreamSynchronize(d.s));
    std::vector<int32_t> er, ec, eb, ep(size_t(nv_ + 1), 0);
    for (long i = 0; i < nv_; ++i) {
        ep[size_t(i)] = int32_t(er.size());
        for (int32_t k = rp[size_t(i)]; k < rp[size_t(i + 1)]; ++k)
            if (ci[size_t(k)] != i) {
                er.push_back(int32_t(i));
                ec.push_back(ci[size_t(k)]);
                eb.push_back(k);
            }
    }
    ep[size_t(nv_)] = int32_t(er.size());
    ne_ = long(er.size());
    std::vector<int32_t> rev(size_t(ne_), -1);
    for (long e = 0; e < ne_; ++e) {
        const int32_t i = er[size_t(e)], j = ec[size_t(e)];
        for (int32_t f = ep[size_t(j)]; f < ep[size_t(j + 1)]; ++f)
            if (ec[size_t(f)] == i) { rev[size_t(e)] = f; break; }
        if (rev[size_t(e)] < 0) throw DeviceError("tangent pattern is not symmetric");
    }
    auto up = [&](const std::vector<int32_t>& v) {
        int32_t* p = alloc<int32_t>(v.size());
        h2d(p, v.data(), v.size(), d.s);
        return p;
    };
    erow_ = up(er); ecol_ = up(ec); eblk_ = up(eb); eptr_ = up(ep); erev_ = up(rev);
    const long ne = std::max(ne_, 1L), nv = std::max(nv_, 1L);
    xn_ = alloc<bf16_t>(size_t(nv * FIN));
    xe_ = alloc<bf16_t>(size_t(ne * FIN));
    scale_ = alloc<float>(size_t(nslot_));
    mask_ = alloc<float>(size_t(nslot_));
    cblk_ = alloc<float>(size_t(nv * 36));
    gd_ = alloc<float>(size_t(nv * 36));
    go_ = alloc<float>(size_t(ne * 36));
    for (int l = 0; l <= NL; ++l) {
        hf_[l] = alloc<float>(size_t(nv * H));
        hb_[l] = alloc<bf16_t>(size_t(nv * H));
        ef_[l] = alloc<float>(size_t(ne * H));
        eb_[l] = alloc<bf16_t>(size_t(ne * H));
        if (l > 0) {
            mf_[l] = alloc<float>(size_t(ne * H));
            uf_[l] = alloc<float>(size_t(nv * H));
            ab_[l] = alloc<bf16_t>(size_t(nv * H));
        }
    }
    dof_ = alloc<float>(size_t(ne * FOUT));
    ddf_ = alloc<float>(size_t(nv * FOUT));
    dh_ = alloc<float>(size_t(nv * H));
    de_ = alloc<float>(size_t(ne * H));
    dtmp_e_ = alloc<float>(size_t(ne * 3 * H));
    dtmp_n_ = alloc<float>(size_t(nv * 2 * H));
    dga_ = alloc<float>(size_t(std::max(ne * H, nv * 2 * H)));
    dyb_e_ = alloc<bf16_t>(size_t(ne * H));
    dyb_n_ = alloc<bf16_t>(size_t(nv * H));
    dgd_ = alloc<float>(size_t(nv * 36));
    dgo_ = alloc<float>(size_t(ne * 36));
    const int BMAX = 16;
    v1_ = alloc<double>(size_t(nslot_ * BMAX));
    v2_ = alloc<double>(size_t(nslot_ * BMAX));
    v3_ = alloc<double>(size_t(nslot_ * BMAX));
    v4_ = alloc<double>(size_t(nslot_ * BMAX));
    v5_ = alloc<double>(size_t(nslot_));
    v6_ = alloc<double>(size_t(nslot_));
    red_ = alloc<double>(size_t(2 * 512));
    CB_CHECK(hipHostMalloc(&red_host_, sizeof(double) * 2 * 512, hipHostMallocDefault));
    CB_CHECK(hipStreamSynchronize(d.s));
}

void Learned::refresh_bf16() {
    for (const Wdesc& w : w_)
        hipLaunchKernelGGL(k_weights_bf16, dim3(nblk(long(w.K) * w.N)), dim3(TPB), 0, s_, p_ + w.off, w.K, w.N,
                           wt_ + w.boff, wn_ + w.boff);
    CB_CHECK(hipGetLastError());
}

void Learned::setup(Dev& d) {
    bind(d);
    refresh_bf16();
    hipLaunchKernelGGL(k_lp_scale, dim3(nblk(nslot_)), dim3(TPB), 0, s_, d.K, d.diagblk, d.eq_of_slot, scale_,
                       mask_, nslot_);
    hipLaunchKernelGGL(k_lp_node, dim3(nblk(nv_)), dim3(TPB), 0, s_, d.K, d.diagblk, scale_, mask_, xn_, cblk_, nv_);
    if (ne_)
        hipLaunchKernelGGL(k_lp_edge, dim3(nblk(ne_)), dim3(TPB), 0, s_, d.K, eblk_, erow_, ecol_, scale_, mask_,
                           xe_, ne_);
    CB_CHECK(hipGetLastError());
    auto wt = [&](int i) { return wt_ + w_[size_t(i)].boff; };
    auto bias = [&](int i) { return p_ + w_[size_t(i)].off + long(w_[size_t(i)].K) * w_[size_t(i)].N; };
    // encoders
    {
        GemmArgs g = args1(xn_, FIN);
        g.Bt = wt(0); g.bias = bias(0); g.relu = 1; g.M = nv_; g.Yf = hf_[0]; g.Yb = hb_[0];
        gemm(s_, g, H);
        GemmArgs h = args1(xe_, FIN);
        h.Bt = wt(1); h.bias = bias(1); h.relu = 1; h.M = ne_; h.Yf = ef_[0]; h.Yb = eb_[0];
        gemm(s_, h, H);
    }
    for (int l = 1; l <= NL; ++l) {
        GemmArgs g{};
        g.seg[0] = Seg{hb_[l - 1], erow_, H};
        g.seg[1] = Seg{hb_[l - 1], ecol_, H};
        g.seg[2] = Seg{eb_[l - 1], nullptr, H};
        g.nseg = 3; g.K = 3 * H;
        g.Bt = wt(1 + l); g.bias = bias(1 + l); g.relu = 1; g.M = ne_;
        g.Yf = mf_[l]; g.R = ef_[l - 1]; g.Zf = ef_[l]; g.Zb = eb_[l];
        gemm(s_, g, H);
        hipLaunchKernelGGL(k_segsum_row, dim3(nblk(nv_ * H)), dim3(TPB), 0, s_, mf_[l], H, 0, eptr_, nv_,
                           (float*)nullptr, 0, ab_[l]);
        GemmArgs u{};
        u.seg[0] = Seg{hb_[l - 1], nullptr, H};
        u.seg[1] = Seg{ab_[l], nullptr, H};
        u.nseg = 2; u.K = 2 * H;
        u.Bt = wt(4 + l); u.bias = bias(4 + l); u.relu = 1; u.M = nv_;
        u.Yf = uf_[l]; u.R = hf_[l - 1]; u.Zf = hf_[l]; u.Zb = hb_[l];
        gemm(s_, u, H);
    }
    {
        GemmArgs g{};
        g.seg[0] = Seg{hb_[NL], erow_, H};
        g.seg[1] = Seg{hb_[NL], ecol_, H};
        g.seg[2] = Seg{eb_[NL], nullptr, H};
        g.nseg = 3; g.K = 3 * H;
        g.Bt = wt(8); g.bias = bias(8); g.relu = 0; g.M = ne_; g.Yf = dof_;
        gemm(s_, g, FOUT);
        GemmArgs h = args1(hb_[NL], H);
        h.Bt = wt(9); h.bias = bias(9); h.relu = 0; h.M = nv_; h.Yf = ddf_;
        gemm(s_, h, FOUT);
    }
    const long nt = nv_ * 36 + ne_ * 36;
    hipLaunchKernelGGL(k_build_g, dim3(nblk(nt)), dim3(TPB), 0, s_, cblk_, ddf_, dof_, mask_, erow_, ecol_, gd_,
                       go_, nv_, ne_, ALPHA);
    CB_CHECK(hipGetLastError());
}

void Learned::apply(Dev& d, const double* v, double* z) {
    (void)d;
    hipLaunchKernelGGL(k_scale_vec, dim3(nblk(nslot_)), dim3(TPB), 0, s_, scale_, v, v1_, nslot_, 0);
    hipLaunchKernelGGL(k_apply_g, dim3(nblk(nslot_)), dim3(TPB), 0, s_, gd_, go_, eptr_, ecol_, v1_, v2_, nv_);
    hipLaunchKernelGGL(k_apply_gt, dim3(nblk(nslot_)), dim3(TPB), 0, s_, gd_, go_, eptr_, ecol_, erev_, v2_, v1_,
                       nv_);
    hipLaunchKernelGGL(k_scale_vec, dim3(nblk(nslot_)), dim3(TPB), 0, s_, scale_, v1_, z, nslot_, 0);
    CB_CHECK(hipGetLastError());
}

double Learned::loss_grad(Dev& d, const double* const* r, const double* const* u, int B) {
    if (B < 1 || B > 16) throw DeviceError("loss_grad takes 1..16 vector pairs");
    const long n = nslot_;
    CB_CHECK(hipMemsetAsync(dgd_, 0, sizeof(float) * size_t(nv_ * 36), s_));
    CB_CHECK(hipMemsetAsync(dgo_, 0, sizeof(float) * size_t(std::max(ne_, 1L) * 36), s_));
    double loss = 0;
    const int nb = 256;
    // Loss on the error t = M~ r~ - u~ of the preconditioned solve (free
    // slots), by objective:
    //   0: |K~ t|^2 / |r~|^2      1: t.K~ t / r~.u~ (energy)      2: |t|^2 / |u~|^2
    for (int b = 0; b < B; ++b) {
        double* x = v1_ + b * n;   // r~ = S r
        double* y = v2_ + b * n;   // G x
        double* zz = v3_ + b * n;  // G^T y -> t -> dL/dt
        double* uu = v4_ + b * n;  // u~ = u / S -> scratch -> dL/dy
        hipLaunchKernelGGL(k_scale_vec, dim3(nblk(n)), dim3(TPB), 0, s_, scale_, r[b], x, n, 0);
        hipLaunchKernelGGL(k_scale_vec, dim3(nblk(n)), dim3(TPB), 0, s_, scale_, u[b], uu, n, 1);
        hipLaunchKernelGGL(k_apply_g, dim3(nblk(n)), dim3(TPB), 0, s_, gd_, go_, eptr_, ecol_, x, y, nv_);
        hipLaunchKernelGGL(k_apply_gt, dim3(nblk(n)), dim3(TPB), 0, s_, gd_, go_, eptr_, ecol_, erev_, y, zz, nv_);
        hipLaunchKernelGGL(k_masked_diff, dim3(nblk(n)), dim3(TPB), 0, s_, zz, uu, mask_, n);
        double num = 0, den = 0;
        auto fetch = [&]() {
            CB_CHECK(hipGetLastError());
            CB_CHECK(hipMemcpyAsync(red_host_, red_, sizeof(double) * 2 * nb, hipMemcpyDeviceToHost, s_));
            CB_CHECK(hipStreamSynchronize(s_));
            num = den = 0;
            for (int k = 0; k < nb; ++k) { num += red_host_[2 * k]; den += red_host_[2 * k + 1]; }
            if (!(den > 0)) den = 1;
        };
        if (objective_ == 0) {
            ktilde(d, zz, uu);                                   // e = K~ t
            hipLaunchKernelGGL(k_dots2, dim3(nb), dim3(TPB), 0, s_, uu, uu, x, x, mask_, n, red_);
            fetch();
            hipLaunchKernelGGL(k_scale_inplace, dim3(nblk(n)), dim3(TPB), 0, s_, uu, 2.0 / (B * den), n);
            ktilde(d, uu, zz);                                   // dL/dt = K~ dL/de
        } else if (objective_ == 1) {
            hipLaunchKernelGGL(k_dots2, dim3(nb), dim3(TPB), 0, s_, x, uu, x, uu, mask_, n, red_);
            fetch();
            const double rdu = num;                              // r~.u~ = u~.K~ u~
            ktilde(d, zz, uu);                                   // w = K~ t
            hipLaunchKernelGGL(k_dots2, dim3(nb), dim3(TPB), 0, s_, zz, uu, zz, uu, mask_, n, red_);
            fetch();
            den = (rdu > 0) ? rdu : 1.0;
            hipLaunchKernelGGL(k_scale_copy, dim3(nblk(n)), dim3(TPB), 0, s_, uu, zz, 2.0 / (B * den), n);
        } else {
            hipLaunchKernelGGL(k_dots2, dim3(nb), dim3(TPB), 0, s_, zz, zz, uu, uu, mask_, n, red_);
            fetch();
            hipLaunchKernelGGL(k_scale_inplace, dim3(nblk(n)), dim3(TPB), 0, s_, zz, 2.0 / (B * den), n);
        }
        loss += num / den / B;
        hipLaunchKernelGGL(k_apply_g, dim3(nblk(n)), dim3(TPB), 0, s_, gd_, go_, eptr_, ecol_, zz, uu, nv_);
        const long nt = nv_ * 36 + ne_ * 36;
        hipLaunchKernelGGL(k_dg, dim3(nblk(nt)), dim3(TPB), 0, s_, y, zz, uu, x, erow_, ecol_, dgd_, dgo_, nv_, ne_);
        CB_CHECK(hipGetLastError());
    }
    // decoder output gradients (fp32) -> dtmp buffers, bf16 copies for the GEMMs
    float* dDd = dtmp_n_;            // [nv][48]
    float* dDo = dtmp_e_;            // [ne][48]
    const long nt48 = nv_ * 48 + ne_ * 48;
    hipLaunchKernelGGL(k_dg_to_dout, dim3(nblk(nt48)), dim3(TPB), 0, s_, dgd_, dgo_, mask_, erow_, ecol_, dDd, dDo,
                       nv_, ne_, ALPHA);
    bf16_t* dDd_b = dyb_n_;
    bf16_t* dDo_b = dyb_e_;
    hipLaunchKernelGGL(k_f2b, dim3(nblk(nv_ * 48)), dim3(TPB), 0, s_, dDd, dDd_b, nv_ * 48);
    hipLaunchKernelGGL(k_f2b, dim3(nblk(ne_ * 48)), dim3(TPB), 0, s_, dDo, dDo_b, ne_ * 48);
    CB_CHECK(hipGetLastError());

    auto wn = [&](int i) { return wn_ + w_[size_t(i)].boff; };
    auto grad_w = [&](int i) { return g_ + w_[size_t(i)].off; };
    auto grad_b = [&](int i) { return g_ + w_[size_t(i)].off + long(w_[size_t(i)].K) * w_[size_t(i)].N; };
    // dW = A^T D, db = colsum(D)
    auto wgrad = [&](const GemmArgs& A, const bf16_t* Db, const float* Df, int N, int wi) {
        const int nbk = int(std::min<long>(nblk_dw_, std::max<long>(1, (A.M + 63) / 64)));
        if (A.M > 0) {
            if (N == 64) hipLaunchKernelGGL(k_gemm_dw<4>, dim3(nbk), dim3(256), 0, s_, A, Db, partial_);
            else hipLaunchKernelGGL(k_gemm_dw<3>, dim3(nbk), dim3(256), 0, s_, A, Db, partial_);
            const long len = long(A.K) * N;
            hipLaunchKernelGGL(k_reduce_partial, dim3(nblk(len)), dim3(TPB), 0, s_, partial_, nbk, len, grad_w(wi));
            hipLaunchKernelGGL(k_colsum_partial, dim3(nbk), dim3(256), 0, s_, Df, A.M, N, partial_);
            hipLaunchKernelGGL(k_reduce_partial, dim3(1), dim3(64), 0, s_, partial_, nbk, long(N), grad_b(wi));
        } else {
            CB_CHECK(hipMemsetAsync(grad_w(wi), 0, sizeof(float) * (size_t(A.K) * N + N), s_));
        }
        CB_CHECK(hipGetLastError());
    };
    // ---- decoder
    {
        GemmArgs A{};
        A.seg[0] = Seg{hb_[NL], erow_, H};
        A.seg[1] = Seg{hb_[NL], ecol_, H};
        A.seg[2] = Seg{eb_[NL], nullptr, H};
        A.nseg = 3; A.K = 3 * H; A.M = ne_;
        wgrad(A, dDo_b, dDo, FOUT, 8);
        GemmArgs B1 = args1(hb_[NL], H);
        B1.M = nv_;
        wgrad(B1, dDd_b, dDd, FOUT, 9);
    }
    // dX = dD W^T: the node part directly, the edge part [ne][192] one 64-column
    // slice at a time through dga_
    {
        // dH3 = dDd Wd^T
        GemmArgs h = args1(dDd_b, FOUT);
        h.Bt = wn(9); h.M = nv_; h.Yf = dh_;
        gemm(s_, h, H);
        // slices of dX_e = dDo Wo^T: columns [0,64) -> H[i], [64,128) -> H[j], [128,192) -> E
        for (int sl = 0; sl < 3; ++sl) {
            GemmArgs g = args1(dDo_b, FOUT);
            g.Bt = wn(8) + long(sl) * 64 * FOUT;   // rows sl*64.. of W [192][48] as Bt [64][48]
            g.M = ne_; g.Yf = dga_;
            gemm(s_, g, H);
            if (sl == 0)
                hipLaunchKernelGGL(k_segsum_row, dim3(nblk(nv_ * H)), dim3(TPB), 0, s_, dga_, H, 0, eptr_, nv_,
                                   dh_, 1, (bf16_t*)nullptr);
            else if (sl == 1)
                hipLaunchKernelGGL(k_segsum_col, dim3(nblk(nv_ * H)), dim3(TPB), 0, s_, dga_, H, 0, eptr_, erev_,
                                   nv_, dh_);
            else
                CB_CHECK(hipMemcpyAsync(de_, dga_, sizeof(float) * size_t(ne_ * H), hipMemcpyDeviceToDevice, s_));
            CB_CHECK(hipGetLastError());
        }
    }
    // ---- message-passing layers, last to first
    for (int l = NL; l >= 1; --l) {
        // node update: H_l = H_{l-1} + relu(P_l)
        float* dU = dtmp_n_;   // [nv][64] (dDd no longer needed)
        hipLaunchKernelGGL(k_relu_bwd, dim3(nblk(nv_ * H)), dim3(TPB), 0, s_, dh_, uf_[l], dU, dyb_n_, nv_ * H);
        GemmArgs U{};
        U.seg[0] = Seg{hb_[l - 1], nullptr, H};
        U.seg[1] = Seg{ab_[l], nullptr, H};
        U.nseg = 2; U.K = 2 * H; U.M = nv_;
        wgrad(U, dyb_n_, dU, H, 4 + l);
        // dX_n = dU Wu^T: [0,64) -> dH_{l-1} (added), [64,128) -> dA_l (kept in dtmp_n_ cols 64..)
        {
            GemmArgs g = args1(dyb_n_, H);
            g.Bt = wn(4 + l); g.M = nv_; g.Yf = dga_;
            // Bt [128][64] = Wu [128][64]; produce all 128 columns
            gemm(s_, g, 2 * H);
        }
        hipLaunchKernelGGL(k_add_cols, dim3(nblk(nv_ * H)), dim3(TPB), 0, s_, dh_, dga_, 2 * H, 0, nv_);
        // dA_l stays in dga_ (cols 64..127, row width 128)
        // messages: dQ = (dE + dA[row]) * (M > 0)
        float* dQ = dtmp_e_;
        hipLaunchKernelGGL(k_msg_bwd, dim3(nblk(ne_ * H)), dim3(TPB), 0, s_, de_, dga_, erow_, mf_[l], dQ, dyb_e_,
                           ne_ * H);
        GemmArgs Mg{};
        Mg.seg[0] = Seg{hb_[l - 1], erow_, H};
        Mg.seg[1] = Seg{hb_[l - 1], ecol_, H};
        Mg.seg[2] = Seg{eb_[l - 1], nullptr, H};
        Mg.nseg = 3; Mg.K = 3 * H; Mg.M = ne_;
        wgrad(Mg, dyb_e_, dQ, H, 1 + l);
        // dX_e = dQ Wm^T by 64-column slices
        for (int sl = 0; sl < 3; ++sl) {
            GemmArgs g = args1(dyb_e_, H);
            g.Bt = wn(1 + l) + long(sl) * 64 * H;
            g.M = ne_; g.Yf = dga_;
            gemm(s_, g, H);
            if (sl == 0)
                hipLaunchKernelGGL(k_segsum_row, dim3(nblk(nv_ * H)), dim3(TPB), 0, s_, dga_, H, 0, eptr_, nv_,
                                   dh_, 1, (bf16_t*)nullptr);
            else if (sl == 1)
                hipLaunchKernelGGL(k_segsum_col, dim3(nblk(nv_ * H)), dim3(TPB), 0, s_, dga_, H, 0, eptr_, erev_,
                                   nv_, dh_);
            else
                hipLaunchKernelGGL(k_add_cols, dim3(nblk(ne_ * H)), dim3(TPB), 0, s_, de_, dga_, H, 0, ne_);
            CB_CHECK(hipGetLastError());
        }
    }
    // ---- encoders
    {
        hipLaunchKernelGGL(k_relu_bwd, dim3(nblk(nv_ * H)), dim3(TPB), 0, s_, dh_, hf_[0], dtmp_n_, dyb_n_, nv_ * H);
        GemmArgs A = args1(xn_, FIN);
        A.M = nv_;
        wgrad(A, dyb_n_, dtmp_n_, H, 0);
        hipLaunchKernelGGL(k_relu_bwd, dim3(nblk(ne_ * H)), dim3(TPB), 0, s_, de_, ef_[0], dtmp_e_, dyb_e_, ne_ * H);
        GemmArgs E = args1(xe_, FIN);
        E.M = ne_;
        wgrad(E, dyb_e_, dtmp_e_, H, 1);
    }
    CB_CHECK(hipStreamSynchronize(s_));
    return loss;
}

void Learned::ktilde(Dev& d, const double* v, double* out) {
    hipLaunchKernelGGL(k_scale_mask, dim3(nblk(nslot_)), dim3(TPB), 0, s_, scale_, mask_, v, v5_, nslot_, 0);
    bsr_spmv(d, v5_, v6_);
    hipLaunchKernelGGL(k_scale_mask, dim3(nblk(nslot_)), dim3(TPB), 0, s_, scale_, mask_, v6_, out, nslot_, 1);
    CB_CHECK(hipGetLastError());
}

void Learned::adam_step(float lr) {
    ++t_;
    const float c1 = 1.f - std::pow(0.9f, float(t_)), c2 = 1.f - std::pow(0.999f, float(t_));
    hipLaunchKernelGGL(k_adam, dim3(nblk(np_)), dim3(TPB), 0, s_, p_, g_, m1_, m2_, np_, lr, c1, c2);
    CB_CHECK(hipGetLastError());
}

void Learned::get_blocks(float* gd, float* go) const {
    CB_CHECK(hipStreamSynchronize(s_));
    if (gd) CB_CHECK(hipMemcpy(gd, gd_, sizeof(float) * size_t(nv_ * 36), hipMemcpyDeviceToHost));
    if (go && ne_) CB_CHECK(hipMemcpy(go, go_, sizeof(float) * size_t(ne_ * 36), hipMemcpyDeviceToHost));
}

}  // namespace cb
