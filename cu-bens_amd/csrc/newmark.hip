// Lumped masses and the vector algebra of the nonlinear Newmark driver.
//
// Masses follow mass_tr (truss.c:381-441), mass_fr (frame.c:1314-1393) and
// mass_sh (shell.c:1505-1591): each recomputes its element geometry from the
// converged coordinates, overwrites the working lengths / areas / side lengths
// the element kernels read, and adds diagonal entries into sm[NEQ].  Elements
// of one colour share no joint, so the adds of one launch never collide.
#include <hip/hip_runtime.h>

#include <cmath>

#include "dev.h"
#include "nm_ops.h"

namespace cb {
namespace {

constexpr int TPB = 256;
inline int nblk(long n, int tpb = TPB) { return int((n + tpb - 1) / tpb); }

// consistent truss mass rho A L/6 [2 1; 1 2] per direction, summed along rows
// over free DOF pairs (truss.c:403-426)
__global__ void k_mass_truss(const int32_t* __restrict__ els, int n, const int32_t* __restrict__ minc,
                             const int32_t* __restrict__ mcode, const double* __restrict__ x,
                             const double* __restrict__ dens, const double* __restrict__ carea,
                             double* __restrict__ llen, double* __restrict__ sm) {
    int it = blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= n) return;
    const long e = els[it];
    const double* a = x + (minc[e * 2] - 1) * 3;
    const double* b = x + (minc[e * 2 + 1] - 1) * 3;
    double el0 = b[0] - a[0], el1 = b[1] - a[1], el2 = b[2] - a[2];
    const double L = sqrt(el0 * el0 + el1 * el1 + el2 * el2);
    llen[e] = L;
    const double mii = dens[e] * carea[e] * L / 3, mij = dens[e] * carea[e] * L / 6;
    const int32_t* mc = mcode + e * 6;
    for (int r = 0; r < 6; ++r) {
        if (!mc[r]) continue;
        const int partner = (r < 3) ? r + 3 : r - 3;
        // the reference adds along the row in column order
        double s = sm[mc[r] - 1];
        if (r < 3) {
            s += mii;
            if (mc[partner]) s += mij;
        } else {
            if (mc[partner]) s += mij;
            s += mii;
        }
        sm[mc[r] - 1] = s;
    }
}

// The chord length of frame member e between its end points, joint + offset
// (frame.c:1331-1345; ofs null when no member has offsets).
__device__ inline double frame_chord(const int32_t* __restrict__ minc, const double* __restrict__ x,
                                     const double* __restrict__ ofs, long e) {
    const double* xa = x + (minc[e * 2] - 1) * 3;
    const double* xb = x + (minc[e * 2 + 1] - 1) * 3;
    double a[3] = {xa[0], xa[1], xa[2]}, b[3] = {xb[0], xb[1], xb[2]};
    if (ofs)
        for (int k = 0; k < 3; ++k) {
            a[k] += ofs[e * 6 + k];
            b[k] += ofs[e * 6 + 3 + k];
        }
    const double el0 = b[0] - a[0], el1 = b[1] - a[1], el2 = b[2] - a[2];
    return sqrt(el0 * el0 + el1 * el1 + el2 * el2);
}

// frame: rho A L/2 per translation, rho A L^3/24 per rotation, 0 on warping
// (frame.c:1357-1375); single-type models only (the reference indexes the
// per-type arrays from 0 here)
__global__ void k_mass_frame(const int32_t* __restrict__ els, int n, const int32_t* __restrict__ minc,
                             const int32_t* __restrict__ mcode, const double* __restrict__ x,
                             const double* __restrict__ dens, const double* __restrict__ carea,
                             double* __restrict__ llen, double* __restrict__ sm, const double* __restrict__ ofs) {
    int it = blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= n) return;
    const long e = els[it];
    const double L = frame_chord(minc, x, ofs, e);
    llen[e] = L;
    const double mt = dens[e] * carea[e] * L / 24 * 12;
    const double mr = dens[e] * carea[e] * L / 24 * (L * L);
    const int32_t* mc = mcode + e * 14;
    for (int r = 0; r < 14; ++r) {
        if (r == 6 || r == 13 || !mc[r]) continue;
        sm[mc[r] - 1] += ((r % 7) >= 3) ? mr : mt;
    }
}

// Frame-only models: the same masses in two passes, every member in one
// launch (its length, and its translational and rotational masses into fm)...
__global__ void k_mass_frame_members(long n, const int32_t* __restrict__ minc, const double* __restrict__ x,
                                     const double* __restrict__ dens, const double* __restrict__ carea,
                                     double* __restrict__ llen, double* __restrict__ fm,
                                     const double* __restrict__ ofs) {
    const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (e >= n) return;
    const double L = frame_chord(minc, x, ofs, e);
    llen[e] = L;
    fm[e * 2] = dens[e] * carea[e] * L / 24 * 12;
    fm[e * 2 + 1] = dens[e] * carea[e] * L / 24 * (L * L);
}

// ... then one thread per free slot adds the masses of the members at its
// joint in colour order, the order in which the colour launches add them.
__global__ void k_mass_frame_gather(const int32_t* __restrict__ jptr, const int32_t* __restrict__ jitem,
                                    const int32_t* __restrict__ eq, long nslot, const double* __restrict__ fm,
                                    double* __restrict__ sm) {
    const long r = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (r >= nslot) return;
    const int32_t q = eq[r];
    if (q < 0) return;
    const long j = r / 6;
    const int s = int(r % 6);
    double acc = sm[q];
    for (int32_t it = jptr[j]; it < jptr[j + 1]; ++it) acc += fm[long(jitem[it] >> 1) * 2 + (s >= 3)];
    sm[q] = acc;
}

// shell: rho A t/3 per translation, (rho A t/3) t^2/12 per rotation
// (shell.c:1544-1573); overwrites the side lengths and face area
__global__ void k_mass_shell(const int32_t* __restrict__ els, int n, const int32_t* __restrict__ minc,
                             const int32_t* __restrict__ mcode, const double* __restrict__ x,
                             const double* __restrict__ dens, const double* __restrict__ thick,
                             double* __restrict__ farea, double* __restrict__ slength,
                             double* __restrict__ sm) {
    int it = blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= n) return;
    const long e = els[it];
    const double* p1 = x + (minc[e * 3] - 1) * 3;
    const double* p2 = x + (minc[e * 3 + 1] - 1) * 3;
    const double* p3 = x + (minc[e * 3 + 2] - 1) * 3;
    double e12[3], e23[3], e31[3];
    for (int k = 0; k < 3; ++k) {
        e23[k] = p3[k] - p2[k];
        e31[k] = p3[k] - p1[k];
        e12[k] = p2[k] - p1[k];
    }
    slength[e * 3 + 1] = sqrt(e23[0] * e23[0] + e23[1] * e23[1] + e23[2] * e23[2]);
    slength[e * 3 + 2] = sqrt(e31[0] * e31[0] + e31[1] * e31[1] + e31[2] * e31[2]);
    slength[e * 3] = sqrt(e12[0] * e12[0] + e12[1] * e12[1] + e12[2] * e12[2]);
    double nx = e12[1] * e31[2] - e12[2] * e31[1];
    double ny = e12[2] * e31[0] - e12[0] * e31[2];
    double nz = e12[0] * e31[1] - e12[1] * e31[0];
    const double A = 0.5 * sqrt(nx * nx + ny * ny + nz * nz);
    farea[e] = A;
    const double t = thick[e];
    const double mtot = dens[e] * A * t;
    const double mt = mtot / 3, mr = mtot / 3 * (t * t) / 12;
    const int32_t* mc = mcode + e * 18;
    for (int r = 0; r < 18; ++r)
        if (mc[r]) sm[mc[r] - 1] += ((r % 6) >= 3) ? mr : mt;
}

// Keff = K + cnum * diag(M) / cden on the free slots of the joint-diagonal blocks
// (solve.c:199-207)
__global__ void k_add_mass_diag(double* __restrict__ K, const int32_t* __restrict__ diagblk,
                                const int32_t* __restrict__ eq, const double* __restrict__ sm,
                                double cnum, double cden, long nslot, int bs) {
    long s = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (s >= nslot) return;
    const int32_t q = eq[s];
    if (q < 0) return;
    const long j = s / bs;
    const int k = int(s % bs);
    K[(long)diagblk[j] * bs * bs + k * bs + k] += cnum * sm[q] / cden;
}

__global__ void k_nm_vec(NewmarkVec op, NewmarkArgs A, long n) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) nm_apply(op, A, i);
}

__global__ void k_linear_nm(LinearNewmarkVec op, LinearNewmarkArgs A, long n) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (op == LNM_REFF) {
        const double meff = A.sm[i] * ((1 - A.am) * (A.u[i] * A.a0 + A.v[i] * A.a2 + A.a[i] * A.a3) -
                                       A.am * A.a[i]) / (1 - A.af);
        const double* p = A.P + i * A.ld;
        double r = (A.k == 0) ? p[A.k] + meff : p[A.k] + A.af / (1 - A.af) * p[A.k - 1] + meff;
        if (A.Ku) r -= A.af / (1 - A.af) * A.Ku[i];
        A.r[i] = r;
    } else {
        const double an = A.a0 * (A.unew[i] - A.u[i]) - A.a2 * A.v[i] - A.a3 * A.a[i];
        A.anew[i] = an;
        A.vnew[i] = A.v[i] + A.a6 * A.a[i] + A.a7 * an;
    }
}

}  // namespace

void launch_mass(Dev& d, double* sm, int type) {
    const double* xb = d.x[G_BASE];
    if (type == 0) {
        for (size_t c = 0; c + 1 < d.col_ptr_tr.size(); ++c) {
            int n = d.col_ptr_tr[c + 1] - d.col_ptr_tr[c];
            if (!n) continue;
            hipLaunchKernelGGL(k_mass_truss, dim3(nblk(n)), dim3(TPB), 0, d.s,
                               d.col_el_tr + d.col_ptr_tr[c], n, d.minc, d.mcode, xb, d.dens,
                               d.carea, d.llen, sm);
        }
    } else if (type == 1 && !d.ntr && !d.nsh && d.bs == 6) {
        hipLaunchKernelGGL(k_mass_frame_members, dim3(nblk(d.nfr)), dim3(TPB), 0, d.s, d.nfr, d.minc, xb, d.dens,
                           d.carea, d.llen, d.fh, d.ofs);
        hipLaunchKernelGGL(k_mass_frame_gather, dim3(nblk(d.nslot)), dim3(TPB), 0, d.s, d.jnt_ptr_fr, d.jnt_item_fr,
                           d.eq_of_slot, d.nslot, d.fh, sm);
    } else if (type == 1) {
        for (size_t c = 0; c + 1 < d.col_ptr_fr.size(); ++c) {
            int n = d.col_ptr_fr[c + 1] - d.col_ptr_fr[c];
            if (!n) continue;
            hipLaunchKernelGGL(k_mass_frame, dim3(nblk(n)), dim3(TPB), 0, d.s,
                               d.col_el_fr + d.col_ptr_fr[c], n, d.minc, d.mcode, xb, d.dens,
                               d.carea, d.llen, sm, d.ofs);
        }
    } else {
        for (size_t c = 0; c + 1 < d.col_ptr_sh.size(); ++c) {
            int n = d.col_ptr_sh[c + 1] - d.col_ptr_sh[c];
            if (!n) continue;
            hipLaunchKernelGGL(k_mass_shell, dim3(nblk(n)), dim3(TPB), 0, d.s,
                               d.col_el_sh + d.col_ptr_sh[c], n, d.minc, d.mcode, xb, d.dens,
                               d.thick, d.farea, d.slength, sm);
        }
    }
    CB_CHECK(hipGetLastError());
}

void add_mass_diag(Dev& d, const double* sm, double cnum, double cden) {
    hipLaunchKernelGGL(k_add_mass_diag, dim3(nblk(d.nslot)), dim3(TPB), 0, d.s, d.K, d.diagblk,
                       d.eq_of_slot, sm, cnum, cden, d.nslot, d.bs);
    CB_CHECK(hipGetLastError());
}

void linear_newmark_vec(Dev& d, LinearNewmarkVec op, const LinearNewmarkArgs& a) {
    if (!d.neq) return;
    hipLaunchKernelGGL(k_linear_nm, dim3(nblk(d.neq)), dim3(TPB), 0, d.s, op, a, d.neq);
    CB_CHECK(hipGetLastError());
}

void newmark_vec(Dev& d, NewmarkVec op, const NewmarkArgs& a) {
    if (!d.neq) return;
    hipLaunchKernelGGL(k_nm_vec, dim3(nblk(d.neq)), dim3(TPB), 0, d.s, op, a, d.neq);
    CB_CHECK(hipGetLastError());
}

}  // namespace cb
