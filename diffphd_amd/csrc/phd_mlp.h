// Learned warm start of the PD iteration: a per-vertex MLP
// 13 -> 128 -> 128 -> 128 -> 3 (GELU) that predicts the converged step's
// offset from the x + h v guess. BF16 weights on MFMA with FP32 accumulation;
// FP32 master weights and Adam state. No counterpart in the reference (its
// initial iterate is q_t, deformable_projective_dynamics_forward_gpu.cpp:1069).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "phd_solver.h"

namespace phd {

constexpr int kMlpIn = 13;      // features per vertex
constexpr int kMlpInPad = 16;   // padded to one MFMA k-step
constexpr int kMlpHidden = 128;
constexpr int kMlpOut = 3;
constexpr int kMlpOutPad = 16;

// Flat FP32 parameter layout (W stored [in][out]); padded rows of W1 and
// padded columns of W4 / b4 stay zero.
struct MlpLayout {
    static constexpr int W1 = 0;
    static constexpr int b1 = W1 + kMlpInPad * kMlpHidden;
    static constexpr int W2 = b1 + kMlpHidden;
    static constexpr int b2 = W2 + kMlpHidden * kMlpHidden;
    static constexpr int W3 = b2 + kMlpHidden;
    static constexpr int b3 = W3 + kMlpHidden * kMlpHidden;
    static constexpr int W4 = b3 + kMlpHidden;
    static constexpr int b4 = W4 + kMlpHidden * kMlpOutPad;
    static constexpr int total = b4 + kMlpOutPad;
};

class WarmStartNet {
public:
    bool ready() const { return ready_; }
    void init(uint64_t seed, hipStream_t s);
    void set_params(const float* p, hipStream_t s);
    void get_params(float* p, hipStream_t s);
    // out (n x 3) = net(feat), feat: n x 16 standardized features (device)
    void forward(hipStream_t s, int n, const float* feat, float* out);
    // full-batch MSE loss and its gradient (device grad of MlpLayout::total)
    double loss_grad(hipStream_t s, int n, const float* feat, const float* target, float* grad);
    // `steps` Adam steps on consecutive minibatches of a seeded permutation
    void train(hipStream_t s, int n, const float* feat, const float* target, int steps, int batch, float lr,
               uint64_t seed, float* losses);

    std::vector<float> fmean = std::vector<float>(kMlpIn, 0.0f), fstd = std::vector<float>(kMlpIn, 1.0f);
    float tscale = 1.0f;

private:
    void pack(hipStream_t s);
    void grad_pass(hipStream_t s, int nrows, const float* feat, const float* target, const int* idx, float loss_scale);
    bool ready_ = false;
    int step_ = 0;
    DevBuf<float> master_, m_, v_, grad_, partial_, loss_part_;
    DevBuf<uint16_t> wf_, wb_;  // BF16: forward layout [out][in], backward layout [in][out]
    DevBuf<int> idx_;
};

// Raw per-vertex features (n x 16, columns 13..15 zero):
//  0-2 (q - X) / L; 3-5 Jacobi step r_i / A_ii / L of the PD residual
//  r = b(q + h v) - A (q + h v) at the x + h v guess; 6-8 h v / L; 9 mean
//  log10 element stiffness; 10 contact flag; 11 obstacle distance / L
//  (clamped to [-4, 4]; 4 off the contact set); 12 h^2 |g| / L.
void launch_mlp_features(hipStream_t s, int n, const double* q, const double* v, const double* X, const double* res,
                         const float* adiag, const float* logk, const float* contact, const float* dist, double h, double L,
                         double gacc, const float* fmean, const float* fstd, int standardize, float* feat);

// Initial iterates: g0 = q + h v, g1 = g0 + tscale * net (Dirichlet vertices fixed).
void launch_warm_guess(hipStream_t s, int nv, const double* q, const double* v, double h, const float* net, float tscale,
                       const unsigned char* dmask, const double* dval, double* g0, double* g1);

// MFMA lane-layout probe: D (16 x 16, row-major) = A (16 x 16) * B (16 x 16),
// inputs rounded to BF16. Test entry point.
void mlp_mfma_probe(const float* A, const float* B, float* D);

}  // namespace phd
