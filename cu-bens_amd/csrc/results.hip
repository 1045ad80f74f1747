// Results of a converged state: support reactions, and the averaged element
// forces of the reference's results3-5.txt.
//
// Support reactions.
// The reference computes none.  They are the internal force vector its
// element routines accumulate into f, T_i^T ef_i summed over the elements
// (truss.c:349-372, frame.c:1273-1309, shell.c:1728-1785), read at the slots
// of the constrained DOFs, where forces_* skip the scatter (mcode == 0).  Every
// element's end forces are self-equilibrated, so at an equilibrium state the
// reactions balance the applied load: sum(reactions) + lambda sum(q) = 0 per
// translational direction, up to the residual the convergence test allowed.
//
// One thread per element, one launch per colour class (elements of a class
// share no joint), so the joint sums are atomic-free and deterministic.
#include "dev.h"

namespace cb {
namespace {

constexpr int TPB = 256;
inline int nblk(long n) { return int((n + TPB - 1) / TPB); }

// truss.c:349-372: f -= ef_end * c at both ends.
__global__ void k_react_truss(const int32_t* __restrict__ els, int n, const int32_t* __restrict__ minc,
                              const double* __restrict__ c1, const double* __restrict__ c2,
                              const double* __restrict__ c3, const double* __restrict__ ef,
                              double* __restrict__ rx, double sgn) {
    const int it = blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= n) return;
    const long e = els[it];
    const double c[3] = {c1[e], c2[e], c3[e]};
    for (int a = 0; a < 2; ++a) {
        const long j = minc[e * 2 + a] - 1;
        const double v = sgn * ef[e * 2 + a];
        for (int k = 0; k < 3; ++k) rx[j * 7 + k] -= v * c[k];
    }
}

// Frame end forces in the member triad, 14 per member: per end the force
// triple, the moment triple and the warping bimoment; T is block diagonal in
// the triad with 1 on the warping entries (frame.c:286-295).  A member end
// offset o carries the end force F to the joint with the moment o x F added
// (G^T, frame.c:297-324); ofs is null when no member has offsets.
__global__ void k_react_frame(const int32_t* __restrict__ els, int n, const int32_t* __restrict__ minc,
                              const double* __restrict__ c1, const double* __restrict__ c2,
                              const double* __restrict__ c3, const double* __restrict__ ef,
                              double* __restrict__ rx, long ntr, const double* __restrict__ ofs) {
    const int it = blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= n) return;
    const long e = els[it];
    const long t = ntr + e * 3;
    const double R[3][3] = {{c1[t], c1[t + 1], c1[t + 2]},
                            {c2[t], c2[t + 1], c2[t + 2]},
                            {c3[t], c3[t + 1], c3[t + 2]}};
    const double* fe = ef + ntr * 2 + e * 14;
    for (int a = 0; a < 2; ++a) {
        const long j = minc[ntr * 2 + e * 2 + a] - 1;
        double F[2][3];
        for (int h = 0; h < 2; ++h)
            for (int m = 0; m < 3; ++m) {
                const int b = 7 * a + 3 * h;
                F[h][m] = R[0][m] * fe[b] + R[1][m] * fe[b + 1] + R[2][m] * fe[b + 2];
            }
        if (ofs) {
            const double* o = ofs + e * 6 + 3 * a;
            F[1][0] += o[1] * F[0][2] - o[2] * F[0][1];
            F[1][1] += o[2] * F[0][0] - o[0] * F[0][2];
            F[1][2] += o[0] * F[0][1] - o[1] * F[0][0];
        }
        for (int h = 0; h < 2; ++h)
            for (int m = 0; m < 3; ++m) rx[j * 7 + 3 * h + m] += F[h][m];
        rx[j * 7 + 6] += fe[7 * a + 6];
    }
}

// Shell forces in the element triad, 6 per vertex (force and moment triples),
// T block diagonal in the triad (shell.c:1775-1785).
__global__ void k_react_shell(const int32_t* __restrict__ els, int n, const int32_t* __restrict__ minc,
                              const double* __restrict__ c1, const double* __restrict__ c2,
                              const double* __restrict__ c3, const double* __restrict__ ef,
                              double* __restrict__ rx, long ioff, long toff, long eoff) {
    const int it = blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= n) return;
    const long e = els[it];
    const long t = toff + e * 3;
    const double R[3][3] = {{c1[t], c1[t + 1], c1[t + 2]},
                            {c2[t], c2[t + 1], c2[t + 2]},
                            {c3[t], c3[t + 1], c3[t + 2]}};
    const double* fe = ef + eoff + e * 18;
    for (int a = 0; a < 3; ++a) {
        const long j = minc[ioff + e * 3 + a] - 1;
        for (int h = 0; h < 2; ++h)
            for (int m = 0; m < 3; ++m) {
                const int b = 6 * a + 3 * h;
                rx[j * 7 + 3 * h + m] += R[0][m] * fe[b] + R[1][m] * fe[b + 1] + R[2][m] * fe[b + 2];
            }
    }
}

// Averaged element forces as output() writes them (misc.c:419-449): per truss
// (|N_1| + |N_2|)/2, per frame member the 7 entries (|f_j| + |f_{7+j}|)/2, per
// shell the 6 entries (|f_j| + |f_{6+j}| + |f_{12+j}|)/3; out [NE_TR + 7 NE_FR + 6 NE_SH].
__global__ void k_avg_forces(const double* __restrict__ ef, double* __restrict__ out, long ntr, long nfr,
                             long nsh) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const long n1 = ntr, n2 = n1 + 7 * nfr, n3 = n2 + 6 * nsh;
    if (i >= n3) return;
    if (i < n1) {
        out[i] = (fabs(ef[i * 2]) + fabs(ef[i * 2 + 1])) / 2;
    } else if (i < n2) {
        const long e = (i - n1) / 7, j = (i - n1) % 7;
        const double* f = ef + ntr * 2 + e * 14;
        out[i] = (fabs(f[j]) + fabs(f[7 + j])) / 2;
    } else {
        const long e = (i - n2) / 6, j = (i - n2) % 6;
        const double* f = ef + ntr * 2 + nfr * 14 + e * 18;
        out[i] = (fabs(f[j]) + fabs(f[6 + j]) + fabs(f[12 + j])) / 3;
    }
}

// keep the constrained slots (jcode == 0), clear the others
__global__ void k_react_mask(const int32_t* __restrict__ jcode7, double* __restrict__ rx, long n) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n && jcode7[i] != 0) rx[i] = 0.0;
}

}  // namespace

void launch_reactions(Dev& d, int g, double* rx) {
    const long ns = d.nj * 7;
    CB_CHECK(hipMemsetAsync(rx, 0, sizeof(double) * size_t(ns), d.s));
    for (size_t c = 0; c + 1 < d.col_ptr_tr.size(); ++c) {
        const int n = d.col_ptr_tr[c + 1] - d.col_ptr_tr[c];
        if (n)
            hipLaunchKernelGGL(k_react_truss, dim3(nblk(n)), dim3(TPB), 0, d.s, d.col_el_tr + d.col_ptr_tr[c], n,
                               d.minc, d.c1[g], d.c2[g], d.c3[g], d.ef[g], rx,
                               // the linear truss end forces carry the opposite sign
                               // (truss.c:300-305): EA/L0 (d1 - d2) at end 1
                               d.anaflag == 1 ? -1.0 : 1.0);
    }
    for (size_t c = 0; c + 1 < d.col_ptr_fr.size(); ++c) {
        const int n = d.col_ptr_fr[c + 1] - d.col_ptr_fr[c];
        if (n)
            hipLaunchKernelGGL(k_react_frame, dim3(nblk(n)), dim3(TPB), 0, d.s, d.col_el_fr + d.col_ptr_fr[c], n,
                               d.minc, d.c1[g], d.c2[g], d.c3[g], d.ef[g], rx, d.ntr, d.ofs);
    }
    const long io = d.ntr * 2 + d.nfr * 2, to = d.ntr + d.nfr * 3, eo = d.ntr * 2 + d.nfr * 14;
    for (size_t c = 0; c + 1 < d.col_ptr_sh.size(); ++c) {
        const int n = d.col_ptr_sh[c + 1] - d.col_ptr_sh[c];
        if (n)
            hipLaunchKernelGGL(k_react_shell, dim3(nblk(n)), dim3(TPB), 0, d.s, d.col_el_sh + d.col_ptr_sh[c], n,
                               d.minc, d.c1[g], d.c2[g], d.c3[g], d.ef[g], rx, io, to, eo);
    }
    hipLaunchKernelGGL(k_react_mask, dim3(nblk(ns)), dim3(TPB), 0, d.s, d.jcode, rx, ns);
    CB_CHECK(hipGetLastError());
}

void launch_avg_forces(Dev& d, int g, double* out) {
    const long n = d.ntr + 7 * d.nfr + 6 * d.nsh;
    if (n == 0) return;
    hipLaunchKernelGGL(k_avg_forces, dim3(nblk(n)), dim3(TPB), 0, d.s, d.ef[g], out, d.ntr, d.nfr, d.nsh);
    CB_CHECK(hipGetLastError());
}

}  // namespace cb
