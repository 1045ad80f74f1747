// Device operations of the batched Newmark driver: state copies of the models
// that commit or restart a step this round, the Newmark vector algebra and the
// Keff diagonal with each model's own coefficients, per-model convergence sums
// and the recording of converged displacement rows.  All are HBM-bound
// elementwise passes or one-workgroup-per-model reductions.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "batch.h"
#include "nm_ops.h"

namespace cb {
namespace {

constexpr int TPB = 256;

__global__ void k_masked_copy(const CopyJob* __restrict__ jobs, const int32_t* __restrict__ own_eq,
                              const int32_t* __restrict__ own_joint, const int32_t* __restrict__ own_el,
                              const NmParam* __restrict__ prm, int which) {
    const CopyJob J = jobs[blockIdx.y];
    const int32_t* own = J.owner == 0 ? own_eq : (J.owner == 1 ? own_joint : own_el);
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < J.n; i += (long)gridDim.x * blockDim.x) {
        const NmParam& p = prm[own[i / J.per]];
        if (which == 0 ? p.commit : p.start) J.dst[i] = J.src[i];
    }
}

__global__ void k_nm_vec_models(int op, NewmarkArgs A, const NmParam* __restrict__ prm,
                                const int32_t* __restrict__ own_eq, long n) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const NmParam& p = prm[own_eq[i]];
    A.k = p.k;
    A.sub_dt = p.sub_dt;
    A.lpf = p.lpf;
    A.am = p.am;
    A.af = p.af;
    A.a0 = p.a0;
    A.a2 = p.a2;
    A.a3 = p.a3;
    A.a6 = p.a6;
    A.a7 = p.a7;
    A.flip = p.flip;
    nm_apply(NewmarkVec(op == NM_MODEL_OP ? p.op : op), A, i);
}

__global__ void k_add_mass_diag_models(double* __restrict__ K, const int32_t* __restrict__ diagblk,
                                       const int32_t* __restrict__ eq, const double* __restrict__ sm,
                                       const NmParam* __restrict__ prm, const int32_t* __restrict__ own_joint,
                                       long nslot, int bs) {
    long s = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (s >= nslot) return;
    const int32_t q = eq[s];
    if (q < 0) return;
    const long j = s / bs;
    const int k = int(s % bs);
    const NmParam& p = prm[own_joint[j]];
    K[(long)diagblk[j] * bs * bs + k * bs + k] += p.cnum * sm[q] / p.cden;
}

// the same on the band diagonal: tile (I, I) entry (q % 64, q % 64) of every
// free equation q
__global__ void k_add_mass_band_models(double* __restrict__ tiles, int w, const int32_t* __restrict__ eq,
                                       const double* __restrict__ sm, const NmParam* __restrict__ prm,
                                       const int32_t* __restrict__ own_joint, long nslot, int bs) {
    long s = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (s >= nslot) return;
    const int32_t q = eq[s];
    if (q < 0) return;
    const NmParam& p = prm[own_joint[s / bs]];
    const long I = q / 64;
    tiles[(I * (w + 1) + w) * 4096 + (q % 64) * 64 + (q % 64)] += p.cnum * sm[q] / p.cden;
}

__device__ inline double wave_sum64(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// one workgroup per model over its equations
__global__ void __launch_bounds__(TPB) k_test_sums_models(const BatchSeg* __restrict__ segs,
                                                          const double* __restrict__ dd,
                                                          const double* __restrict__ dt,
                                                          const double* __restrict__ dyn,
                                                          const double* __restrict__ f,
                                                          const double* __restrict__ fp,
                                                          const double* __restrict__ fip, double* __restrict__ out) {
    constexpr int NO = 6;
    __shared__ double sh[NO][TPB / 64];
    const BatchSeg g = segs[blockIdx.x];
    double a[NO] = {0, 0, 0, 0, 0, 0};
    for (long i = g.qb + threadIdx.x; i < g.qb + g.neq; i += TPB) {
        const double qt = dyn[i] * 1.0;
        const double u = qt - f[i], up = qt - fp[i];
        a[0] += dd[i] * dd[i];
        a[1] += dt[i] * dt[i];
        a[2] += u * u;
        a[3] += up * up;
        a[4] += dd[i] * (qt - fip[i]);
        a[5] += dd[i] * (qt - fp[i]);
    }
#pragma unroll
    for (int k = 0; k < NO; ++k) {
        const double v = wave_sum64(a[k]);
        if ((threadIdx.x & 63) == 0) sh[k][threadIdx.x >> 6] = v;
    }
    __syncthreads();
    if (threadIdx.x < NO) {
        double s = 0;
        for (int q = 0; q < TPB / 64; ++q) s += sh[threadIdx.x][q];
        out[blockIdx.x * NO + threadIdx.x] = s;
    }
}

__global__ void k_record_models(const double* __restrict__ dv, double* __restrict__ path,
                                const BatchSeg* __restrict__ segs, const NmParam* __restrict__ prm,
                                const int32_t* __restrict__ own_eq, long n) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t b = own_eq[i];
    const int rec = prm[b].rec;
    if (rec < 0) return;
    const BatchSeg g = segs[b];
    const long e = i - g.qb;
    if (e < g.neq) path[g.pbase + rec * g.neq + e] = dv[i];
}

inline int nblk(long n) { return int((n + TPB - 1) / TPB); }

}  // namespace

void batch_masked_copy(hipStream_t s, const CopyJob* jobs_dev, int njobs, long nmax, const int32_t* own_eq,
                       const int32_t* own_joint, const int32_t* own_el, const NmParam* prm, int which) {
    if (!njobs || !nmax) return;
    const int gx = int(std::min<long>(nblk(nmax), 2048));
    hipLaunchKernelGGL(k_masked_copy, dim3(gx, njobs), dim3(TPB), 0, s, jobs_dev, own_eq, own_joint, own_el, prm,
                       which);
    CB_CHECK(hipGetLastError());
}

void batch_nm_vec(hipStream_t s, int op, const NewmarkArgs& a, const NmParam* prm, const int32_t* own_eq, long n) {
    if (!n) return;
    hipLaunchKernelGGL(k_nm_vec_models, dim3(nblk(n)), dim3(TPB), 0, s, op, a, prm, own_eq, n);
    CB_CHECK(hipGetLastError());
}

void batch_add_mass_diag(Dev& d, const double* sm, const NmParam* prm, const int32_t* own_joint) {
    hipLaunchKernelGGL(k_add_mass_diag_models, dim3(nblk(d.nslot)), dim3(TPB), 0, d.s, d.K, d.diagblk, d.eq_of_slot,
                       sm, prm, own_joint, d.nslot, d.bs);
    CB_CHECK(hipGetLastError());
}

void batch_add_mass_band(Dev& d, double* tiles, int w, const double* sm, const NmParam* prm,
                         const int32_t* own_joint) {
    hipLaunchKernelGGL(k_add_mass_band_models, dim3(nblk(d.nslot)), dim3(TPB), 0, d.s, tiles, w, d.eq_of_slot, sm, prm,
                       own_joint, d.nslot, d.bs);
    CB_CHECK(hipGetLastError());
}

void batch_test_sums(hipStream_t s, const BatchSeg* segs, int nmod, const double* dd, const double* dt,
                     const double* dyn, const double* f, const double* fp, const double* fip, double* out) {
    if (!nmod) return;
    hipLaunchKernelGGL(k_test_sums_models, dim3(nmod), dim3(TPB), 0, s, segs, dd, dt, dyn, f, fp, fip, out);
    CB_CHECK(hipGetLastError());
}

void batch_record(hipStream_t s, const double* dv, double* path, const BatchSeg* segs, const NmParam* prm,
                  const int32_t* own_eq, long n) {
    if (!n) return;
    hipLaunchKernelGGL(k_record_models, dim3(nblk(n)), dim3(TPB), 0, s, dv, path, segs, prm, own_eq, n);
    CB_CHECK(hipGetLastError());
}

}  // namespace cb
