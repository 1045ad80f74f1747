// Learned preconditioner: a message-passing network on the joint graph of the
// tangent emits a factorized block-sparse approximate inverse
//     M = S G^T G S,      S = diag(K)^{-1/2} (free slots), G block lower-triangular
// over the joint-block pattern of K.  The diagonal blocks of G start from the
// exact block-Jacobi factor C_i (inverse Cholesky factor of the scaled 6x6
// diagonal block, so C_i^T C_i = K~_ii^{-1}) and the network adds corrections:
//     G_ii = C_i + a * (network diagonal output),  G_ij = a * (network edge output), i > j.
// With the network's output at zero M is exactly block-Jacobi.
//
// Network (widths: node/edge inputs 48, hidden 64, outputs 48 of which 36 used):
//   H0 = relu(Xn Wn + bn)                 E0 = relu(Xe We + be)
//   for l = 1..3:
//     M_l = relu([H[i], H[j], E] Wm_l + bm_l)  per directed edge (i <- j)
//     E  += M_l;  A_l[i] = sum_j M_l(i<-j)
//     H  += relu([H, A_l] Wu_l + bu_l)
//   Do = [H[i], H[j], E] Wo + bo (edges),  Dd = H Wd + bd (joints)
// All GEMMs run on MFMA in BF16 with FP32 accumulation; master weights, Adam
// state and the preconditioner blocks are FP32; the Krylov vectors are FP64.
#pragma once

#include <cstdint>
#include <vector>

#include "dev.h"

namespace cb {

typedef uint16_t bf16_t;

class Learned {
public:
    static constexpr int H = 64, FIN = 48, FOUT = 48, NL = 3;
    static constexpr float ALPHA = 0.1f;

    explicit Learned(unsigned seed);
    ~Learned();
    long nparams() const { return np_; }
    void get_params(float* host) const;
    void set_params(const float* host);
    void get_grad(float* host) const;

    // Builds the edge graph of d's pattern (once per device model) and sizes
    // the activations.
    void bind(Dev& d);
    // Features from the tangent in d.K, network forward, preconditioner blocks.
    void setup(Dev& d);
    // z = M v on slot-layout FP64 vectors.
    void apply(Dev& d, const double* v, double* z);
    // Loss (1/B) sum_b of a norm of the error t_b = M~ r~_b - u~_b in the
    // scaled space (K~ = S K S, r~ = S r, u~ = u / S, M~ = G^T G): objective 1
    // (default) t.K~t / r~.u~ (energy norm), 0 |K~ t|^2 / |r~|^2, 2
    // |t|^2 / |u~|^2; and its gradient w.r.t. all parameters (device, FP32),
    // for slot-layout pairs (r_b, u_b) with K u_b = r_b, at the current
    // setup().  Returns the loss.
    double loss_grad(Dev& d, const double* const* r, const double* const* u, int B);
    // Adam on the gradient of the last loss_grad.
    void adam_step(float lr);
    // Host copies of the preconditioner blocks (diag [nv*36], edge [ne*36]).
    void get_blocks(float* gd, float* go) const;
    long nv() const { return nv_; }
    long ne() const { return ne_; }
    long steps() const { return t_; }
    void set_objective(int k) { objective_ = k; }
    int objective() const { return objective_; }

private:
    uint64_t owner_ = 0;   // uid of the model the graph was built for
    hipStream_t s_ = nullptr;
    long nv_ = 0, ne_ = 0, nslot_ = 0;
    long np_ = 0;
    long t_ = 0;
    int objective_ = 1;
    // parameters (FP32 master), gradient, Adam moments, BF16 copies
    float *p_ = nullptr, *g_ = nullptr, *m1_ = nullptr, *m2_ = nullptr;
    bf16_t *wt_ = nullptr, *wn_ = nullptr;   // per weight: W^T [N][K] and W [K][N]
    float* partial_ = nullptr;
    int nblk_dw_ = 0;
    // graph
    int32_t *erow_ = nullptr, *ecol_ = nullptr, *eblk_ = nullptr, *eptr_ = nullptr, *erev_ = nullptr;
    // features and preconditioner
    bf16_t *xn_ = nullptr, *xe_ = nullptr;
    float *scale_ = nullptr, *cblk_ = nullptr, *gd_ = nullptr, *go_ = nullptr, *mask_ = nullptr;
    // activations: H (fp32+bf16) per layer, E per layer, M, U, A per MP layer
    float *hf_[NL + 1] = {}, *ef_[NL + 1] = {}, *mf_[NL + 1] = {}, *uf_[NL + 1] = {};
    bf16_t *hb_[NL + 1] = {}, *eb_[NL + 1] = {}, *ab_[NL + 1] = {};
    float *dof_ = nullptr, *ddf_ = nullptr;
    // backward scratch
    float *dh_ = nullptr, *de_ = nullptr, *dtmp_e_ = nullptr, *dtmp_n_ = nullptr, *dga_ = nullptr;
    bf16_t *dyb_e_ = nullptr, *dyb_n_ = nullptr;
    float *dgd_ = nullptr, *dgo_ = nullptr;
    // vector scratch (slot layout, FP64)
    double *v1_ = nullptr, *v2_ = nullptr, *v3_ = nullptr, *v4_ = nullptr, *v5_ = nullptr, *v6_ = nullptr;
    double* red_ = nullptr;
    double* red_host_ = nullptr;

    struct Wdesc { long off; int K, N; long boff; };
    std::vector<Wdesc> w_;
    void release_graph();
    void refresh_bf16();
    // out = mask S K S v (the scaled tangent on slot vectors)
    void ktilde(Dev& d, const double* v, double* out);
};

}  // namespace cb
