// Device state of one model and the launchers of the hot-path kernels.
//
// HBM layout:
//  * per-element state keeps the reference's flat layouts (three values per
//    element for triads, 18 per shell for element forces, ...) so the state can
//    be exchanged with the reference-shaped host arrays by plain copies; each
//    thread of an element kernel reads its own contiguous 24-144 byte records;
//  * the global tangent is joint-block CSR (BSR) with bs x bs row-major blocks,
//    both triangles stored, over bs DOF slots per joint; slots of fixed DOFs
//    carry an identity row/column so the block system is nonsingular;
//  * driver vectors (d, f, r, ...) live in the reference's equation order
//    (NEQ), the solver vectors in slot order (NJ*bs);
//  * the band factor is stored as dense T x T tiles (tile rows of w+1 tiles).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "model.h"

namespace cb {

#define CB_CHECK(x)                                                                       \
    do {                                                                                  \
        hipError_t e_ = (x);                                                              \
        if (e_ != hipSuccess) throw ::cb::DeviceError(std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct DeviceError {
    std::string msg;
    explicit DeviceError(std::string m) : msg(std::move(m)) {}
};

// Generations of per-element state, as in the reference: base (last converged),
// _i (current iterate), _ip (previous iterate).
enum Gen { G_BASE = 0, G_I = 1, G_IP = 2 };

// Outcome of a frame force pass under ANAFLAG 3: the first member whose pass
// returned (NE_FR when none did), its code (1 yield overshoot, 2 unloading)
// and the regula-falsi fraction of a code 1.
struct PlasticStatus {
    int32_t first;
    int32_t code;
    double tau;
};

struct Dev {
    uint64_t uid = 0;   // unique per dev_create: caches keyed on a model use this, never its address
    hipStream_t s = nullptr;
    long nj = 0, ntr = 0, nfr = 0, nsh = 0, neq = 0, nslot = 0, nnzb = 0;
    int bs = 6;
    int anaflag = 2;
    // topology (int32)
    int32_t *minc = nullptr, *mcode = nullptr, *jcode = nullptr, *eq_of_slot = nullptr;
    int32_t *rowptr = nullptr, *colidx = nullptr, *diagblk = nullptr;
    int32_t *blk_tr = nullptr, *blk_sh = nullptr, *col_el_tr = nullptr, *col_el_sh = nullptr;
    int32_t *blk_fr = nullptr, *col_el_fr = nullptr;
    int32_t *blkrow = nullptr, *gat_ptr_fr = nullptr, *gat_item_fr = nullptr;   // frame tangent gather
    int32_t *jnt_ptr_fr = nullptr, *jnt_item_fr = nullptr;                       // frame force gather
    double* fh = nullptr;   // frame end forces in global axes, 12 per member
    std::vector<int32_t> col_ptr_tr, col_ptr_sh, col_ptr_fr;
    // properties; llen, farea and slength are the working copies the element
    // kernels read (the Newmark mass routines overwrite them)
    double *emod = nullptr, *carea = nullptr, *llen = nullptr, *dens = nullptr;
    double *gmod = nullptr, *istrong = nullptr, *iweak = nullptr, *ipolar = nullptr, *iwarp = nullptr;
    double *auxpt = nullptr;
    // member-end offsets [NE_FR * 6] in global axes (zero for a member without
    // them); null when no member has offsets
    double *ofs = nullptr;
    // member end bending releases [NE_FR], the four flags as a mask (bit a:
    // flag a); null when no member has releases
    int32_t *relmask = nullptr;
    double *nu = nullptr, *thick = nullptr, *xlocal = nullptr, *farea = nullptr, *slength = nullptr;
    // ANAFLAG 3: yield stresses [NE_TR + NE_FR + NE_SH], plastic moduli
    // [NE_FR], yldflag (2 per member: 0 elastic, 1 yielded, 2 unloading), and
    // the scratch of the frame force pass: each member's ef_i (14) and
    // code / flags (4), its regula-falsi fraction, the pass's outcome
    double *yield = nullptr, *zstrong = nullptr, *zweak = nullptr;
    int32_t* yld = nullptr;
    double *pl_efc = nullptr, *pl_tau = nullptr;
    int32_t* pl_code = nullptr;
    PlasticStatus* pl_st = nullptr;
    PlasticStatus* pl_host = nullptr;   // pinned
    // state
    double *x[3] = {nullptr, nullptr, nullptr};  // [G_BASE]=x, [G_I]=x_temp, [G_IP]=x_ip
    double *c1[3] = {}, *c2[3] = {}, *c3[3] = {};
    double *ef[3] = {}, *defllen[3] = {}, *deffarea[3] = {}, *defslen[3] = {};
    // tangent values (FP64 and optional FP32 copy)
    double* K = nullptr;
    float* K32 = nullptr;
    // reduction scratch
    double* red = nullptr;
    double* red_host = nullptr;   // pinned
    int red_blocks = 0;
};

// ---- setup ---------------------------------------------------------------
void dev_create(Dev& d, const Model& m);
void dev_destroy(Dev& d);
void dev_reset_state(Dev& d, const Model& m);   // all generations <- model geometry

template <class T> T* dalloc(size_t n) {
    T* p = nullptr;
    CB_CHECK(hipMalloc(&p, sizeof(T) * (n ? n : 1)));
    return p;
}
template <class T> void h2d(T* dst, const T* src, size_t n, hipStream_t s) {
    if (n) CB_CHECK(hipMemcpyAsync(dst, src, sizeof(T) * n, hipMemcpyHostToDevice, s));
}
template <class T> void d2h(T* dst, const T* src, size_t n, hipStream_t s) {
    if (n) CB_CHECK(hipMemcpyAsync(dst, src, sizeof(T) * n, hipMemcpyDeviceToHost, s));
}
template <class T> void d2d(T* dst, const T* src, size_t n, hipStream_t s) {
    if (n) CB_CHECK(hipMemcpyAsync(dst, src, sizeof(T) * n, hipMemcpyDeviceToDevice, s));
}

// ---- element kernels (kernels.hip) -----------------------------------------
// K <- tangent at generation g (coords x[gx], triads c[g], forces ef[g], ...).
void launch_tangent(Dev& d, int gx, int g, bool want_fp32);
void launch_tangent_fp32(Dev& d, int gx, int g);   // the tangent in K32 only
// f += internal forces; reads ef_ip (clears its membrane entries, shell.c:1773),
// writes ef_i; triads c_ip (previous) and c_i (current), coordinates x[gx].
void launch_forces(Dev& d, int gx, double* f, const double* dd);
// coordinates x[gx] += dd (x_ip <- old x[gx]); element lengths/areas/triads -> _i.
void launch_updatc(Dev& d, int gx, const double* dd);
// copy generations of per-element state: dst <- src (ef optional)
void copy_gen(Dev& d, int dst, int src, bool with_ef, bool with_x);

// frame members (frame.hip): tangent of generation g into K, internal forces
// (ef_ip -> ef_i, triads _ip and _i), member lengths/triads after updatc
void launch_frame_tangent(Dev& d, int g);
// frame-only models: K <- tangent of generation g, one thread per joint block
// summing its members' contributions in colour order, fixed slots as identity
void launch_frame_tangent_gather(Dev& d, int g);
// frame-only models: the same tangent written straight into band tiles of
// tile width 64 (free entries on or below the diagonal; the tiles must be zero
// elsewhere), w tiles below the diagonal; *bad = 2 for an entry outside
// (fp32: the member tangents in FP32, rounded once into the FP64 band)
void launch_frame_tangent_band(Dev& d, int g, double* tiles, int w, int* bad, bool fp32 = false);
void launch_frame_forces(Dev& d, double* f, const double* dd);
void launch_frame_update(Dev& d, int gx);
// ANAFLAG 3: the code of the last frame force pass (0, 1 or 2, as forces_fr
// returns it) and the regula-falsi fraction of a code 1 (waits for the stream)
int frame_plastic_result(Dev& d, double* tau);
// ANAFLAG 3: flags 2 (unloading) back to 0 at a commit
void launch_unloaded_elastic(Dev& d);

// ---- results of a converged state (results.hip) --------------------------------
// rx[NJ*7] <- the element forces T^T ef of generation g summed at every
// constrained joint slot (jcode == 0), zero at the free slots.
void launch_reactions(Dev& d, int g, double* rx);
// out[NE_TR + 7 NE_FR + 6 NE_SH] <- the averaged element forces of generation g
// that the reference writes to results3-5.txt (misc.c:419-449).
void launch_avg_forces(Dev& d, int g, double* out);

// ---- nonlinear Newmark (newmark.hip) -------------------------------------------
// sm += lumped masses of element type (0 truss, 1 frame, 2 shell) at the
// converged coordinates; overwrites the working lengths / areas.
void launch_mass(Dev& d, double* sm, int type);
// K(slot diagonal) += cnum * sm[eq] / cden
void add_mass_diag(Dev& d, const double* sm, double cnum, double cden);

enum NewmarkVec { NM_PREDICT0 = 0, NM_PREDICT, NM_CORRECT, NM_REFF, NM_UPDATE, NM_DYN };
struct NewmarkArgs {
    const double* P = nullptr;   // load histories [NEQ][ld]
    long ld = 0, k = 0;
    double sub_dt = 1, lpf = 1, am = 0, af = 0, a0 = 0, a2 = 0, a3 = 0, a6 = 0, a7 = 0;
    int flip = 0;
    double *qtot = nullptr, *r = nullptr, *ft = nullptr, *fip = nullptr, *dt = nullptr, *dd = nullptr;
    double *sm = nullptr, *ac = nullptr, *vc = nullptr, *out = nullptr;
    const double *um = nullptr, *vm = nullptr, *amv = nullptr;
};
void newmark_vec(Dev& d, NewmarkVec op, const NewmarkArgs& a);

// linear Newmark time step (solve.c:328-423):
//   LNM_REFF:   r = P_k [+ af/(1-af) P_{k-1}, k > 0]
//                   + sm ((1-am)(u a0 + v a2 + a a3) - am a)/(1-af) [- af/(1-af) Ku, Ku set]
//   LNM_UPDATE: a_new = a0 (u_new - u) - a2 v - a3 a,  v_new = v + a6 a + a7 a_new
enum LinearNewmarkVec { LNM_REFF = 0, LNM_UPDATE = 1 };
struct LinearNewmarkArgs {
    const double* P = nullptr;   // load histories [NEQ][ld]
    long ld = 0, k = 0;
    double am = 0, af = 0, a0 = 0, a2 = 0, a3 = 0, a6 = 0, a7 = 0;
    const double *sm = nullptr, *u = nullptr, *v = nullptr, *a = nullptr, *Ku = nullptr;
    double* r = nullptr;
    const double* unew = nullptr;
    double *vnew = nullptr, *anew = nullptr;
};
void linear_newmark_vec(Dev& d, LinearNewmarkVec op, const LinearNewmarkArgs& a);

// ---- vectors -----------------------------------------------------------------
// Fused dot products, deterministic two-level reduction: at most DOTS_MAX
// products per pass, partial sums of at most DOTS_BLOCKS workgroups each.
constexpr int DOTS_MAX = 8;
constexpr int DOTS_BLOCKS = 512;
constexpr long DOTS_PARTS = long(DOTS_MAX) * DOTS_BLOCKS;
// out[k] (device) = sum_i a_k[i]*b_k[i], k < n <= DOTS_MAX, over device
// vectors of length len, queued on stream s; part: DOTS_PARTS device doubles.
void launch_dots(hipStream_t s, int n, const double* const* a, const double* const* b, long len, double* part,
                 double* out);
// the same on the model's stream, results to the host
void multi_dot(Dev& d, int n, const double* const* a, const double* const* b, long len, double* out);
void vfill(Dev& d, double* y, double v, long n);
void vcopy(Dev& d, double* y, const double* x, long n);
// y = a*x + b*y
void vaxpby(Dev& d, double* y, double a, const double* x, double b, long n);
// r = lam*q - f
void vresidual(Dev& d, double* r, double lam, const double* q, const double* f, long n);
// neq <-> slot layouts
void to_slots(Dev& d, double* ext, const double* v);
void from_slots(Dev& d, double* v, const double* ext);

// ---- sparse block operator -----------------------------------------------------
void bsr_spmv(Dev& d, const double* x, double* y);                 // FP64 values
void bsr_spmv32(Dev& d, const double* x, double* y);               // FP32 values, FP64 vectors

}  // namespace cb
