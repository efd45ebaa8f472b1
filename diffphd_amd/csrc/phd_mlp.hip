// Learned warm start: BF16 MFMA kernels for the per-vertex MLP (inference and
// a fused forward + backward pass), gradient reduction, Adam, and the
// feature kernel. One wavefront carries 16 vertices through all layers with
// v_mfma_f32_16x16x16_bf16; activations stay in the wave's LDS slice.
//
// MFMA 16x16x16 BF16 lane layout (lane l, r = 0..3):
//   A (16 x 16, rows i, cols k):  A[l % 16][4 (l / 16) + r]
//   B (16 x 16, rows k, cols n):  B[4 (l / 16) + r][l % 16]
//   D (16 x 16, rows i, cols n):  D[4 (l / 16) + r][l % 16]
// so a D tile is, element for element, a B fragment of the next product.
#include "phd_mlp.h"

#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <stdexcept>

namespace phd {
namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int H = kMlpHidden;
constexpr int LDX = kMlpInPad + 8;  // LDS row strides (BF16 elements), 8-byte aligned rows
constexpr int LDH = H + 8;
constexpr int WPB = 4;              // waves per block
// BF16 weight copies, per layer: forward [out][in], backward [in][out]
constexpr int WO1 = 0, WO2 = WO1 + kMlpInPad * H, WO3 = WO2 + H * H, WO4 = WO3 + H * H, WTOT = WO4 + H * kMlpOutPad;

__host__ __device__ inline uint16_t f2bf(float f) {  // round to nearest even
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ inline float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ inline float gelu_d(float x) {
    return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x);
}

__device__ inline f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }

__device__ inline void wave_sync() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); }

void check(bool ok, const char* msg) {
    if (!ok) throw std::runtime_error(msg);
}

// A fragment from a row-major [16][ld] BF16 tile at column offset k0
__device__ inline bf16x4 lds_a(const uint16_t* T, int ld, int k0, int lane) {
    return *reinterpret_cast<const bf16x4*>(T + (lane & 15) * ld + k0 + 4 * (lane >> 4));
}
// A fragment of the transpose: A'[i][k] = T[k][c0 + i]
__device__ inline bf16x4 lds_at(const uint16_t* T, int ld, int c0, int lane) {
    bf16x4 r;
#pragma unroll
    for (int j = 0; j < 4; ++j) r[j] = (short)T[(4 * (lane >> 4) + j) * ld + c0 + (lane & 15)];
    return r;
}
// B fragment from a global BF16 matrix stored [n][K] (k contiguous)
__device__ inline bf16x4 gl_b(const uint16_t* W, int K, int n0, int k0, int lane) {
    return *reinterpret_cast<const bf16x4*>(W + (size_t)(n0 + (lane & 15)) * K + k0 + 4 * (lane >> 4));
}
__device__ inline bf16x4 d_as_b(const f32x4& d) {
    bf16x4 r;
#pragma unroll
    for (int j = 0; j < 4; ++j) r[j] = (short)f2bf(d[j]);
    return r;
}
__device__ inline void store_d(uint16_t* T, int ld, int c0, int lane, const f32x4& d) {
#pragma unroll
    for (int r = 0; r < 4; ++r) T[(4 * (lane >> 4) + r) * ld + c0 + (lane & 15)] = f2bf(d[r]);
}
__device__ inline float col_sum(const f32x4& d) {  // sum over the 16 rows of column lane % 16
    float s = d[0] + d[1] + d[2] + d[3];
    s += __shfl_xor(s, 16, 64);
    s += __shfl_xor(s, 32, 64);
    return s;
}

// One hidden layer forward: z = A W + b (A from LDS), H = gelu(z) into LDS.
template <int K>
__device__ inline void layer_fwd(const uint16_t* A, int lda, const uint16_t* Wf, const float* b, uint16_t* Hout,
                                 f32x4 (&z)[H / 16], int lane) {
#pragma unroll
    for (int nt = 0; nt < H / 16; ++nt) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < K / 16; ++ks) acc = mfma(lds_a(A, lda, 16 * ks, lane), gl_b(Wf, K, 16 * nt, 16 * ks, lane), acc);
        const float bb = b[16 * nt + (lane & 15)];
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[r] += bb;
        z[nt] = acc;
        f32x4 h;
#pragma unroll
        for (int r = 0; r < 4; ++r) h[r] = gelu(acc[r]);
        store_d(Hout, LDH, 16 * nt, lane, h);
    }
}

// Backward through a hidden layer given dz (D layout, 8 tiles): writes dW, db
// into P and returns dA = dz W^T (D layout) for the layer below (if wanted).
template <bool WANT_DA>
__device__ inline void layer_bwd(const uint16_t* Ain, int lda, int kin, const f32x4 (&dz)[H / 16], const uint16_t* Wb,
                                 uint16_t* G, float* P, int offW, int offb, f32x4 (&dA)[H / 16], int lane) {
#pragma unroll
    for (int bt = 0; bt < H / 16; ++bt) {
        const float s = col_sum(dz[bt]);
        if (lane < 16) P[offb + 16 * bt + lane] = s;
    }
    for (int a = 0; a < kin / 16; ++a) {
        const bf16x4 at = lds_at(Ain, lda, 16 * a, lane);
#pragma unroll
        for (int bt = 0; bt < H / 16; ++bt) {
            const f32x4 acc = mfma(at, d_as_b(dz[bt]), f32x4{0.f, 0.f, 0.f, 0.f});
#pragma unroll
            for (int r = 0; r < 4; ++r) P[offW + (16 * a + 4 * (lane >> 4) + r) * H + 16 * bt + (lane & 15)] = acc[r];
        }
    }
    if (!WANT_DA) return;
#pragma unroll
    for (int bt = 0; bt < H / 16; ++bt) store_d(G, LDH, 16 * bt, lane, dz[bt]);
    wave_sync();
#pragma unroll
    for (int it = 0; it < H / 16; ++it) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < H / 16; ++ks) acc = mfma(lds_a(G, LDH, 16 * ks, lane), gl_b(Wb, H, 16 * it, 16 * ks, lane), acc);
        dA[it] = acc;
    }
    wave_sync();
}

// Forward (TRAIN = false: outputs) or forward + backward (TRAIN = true: per-wave
// gradient partials and loss) of 16 rows per wavefront. Row r of the pass is
// row idx[r] of feat / target (idx == nullptr: identity).
template <bool TRAIN>
__global__ void __launch_bounds__(256) k_mlp_pass(int nrows, const float* __restrict__ feat, const float* __restrict__ target,
                                                  const int* __restrict__ idx, const uint16_t* __restrict__ wf,
                                                  const uint16_t* __restrict__ wb, const float* __restrict__ master,
                                                  float* __restrict__ out, float* __restrict__ partial,
                                                  float* __restrict__ loss_part, float loss_scale) {
    __shared__ __align__(16) uint16_t sX[WPB][16 * LDX];
    __shared__ __align__(16) uint16_t sH[WPB][3][16 * LDH];
    __shared__ __align__(16) uint16_t sG[WPB][16 * LDH];
    const int lane = threadIdx.x % 64, wv = threadIdx.x / 64;
    const int w = blockIdx.x * WPB + wv;
    const int v0 = 16 * w;
    if (v0 >= nrows) return;
    uint16_t* X0 = sX[wv];
    uint16_t* H1 = sH[wv][0];
    uint16_t* H2 = sH[wv][1];
    uint16_t* H3 = sH[wv][2];
    uint16_t* G = sG[wv];
    {  // features: 16 rows x 16 columns, 4 per lane
        const int row = lane / 4, c = 4 * (lane % 4);
        const int gr = v0 + row;
        float4 f = make_float4(0.f, 0.f, 0.f, 0.f);
        if (gr < nrows) f = *reinterpret_cast<const float4*>(feat + (size_t)(idx ? idx[gr] : gr) * kMlpInPad + c);
        X0[row * LDX + c] = f2bf(f.x);
        X0[row * LDX + c + 1] = f2bf(f.y);
        X0[row * LDX + c + 2] = f2bf(f.z);
        X0[row * LDX + c + 3] = f2bf(f.w);
    }
    wave_sync();
    f32x4 z1[H / 16], z2[H / 16], z3[H / 16];
    layer_fwd<kMlpInPad>(X0, LDX, wf + WO1, master + MlpLayout::b1, H1, z1, lane);
    wave_sync();
    layer_fwd<H>(H1, LDH, wf + WO2, master + MlpLayout::b2, H2, z2, lane);
    wave_sync();
    layer_fwd<H>(H2, LDH, wf + WO3, master + MlpLayout::b3, H3, z3, lane);
    wave_sync();
    f32x4 o = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < H / 16; ++ks) o = mfma(lds_a(H3, LDH, 16 * ks, lane), gl_b(wf + WO4, H, 0, 16 * ks, lane), o);
    const int col = lane & 15;
    const float b4 = master[MlpLayout::b4 + col];
#pragma unroll
    for (int r = 0; r < 4; ++r) o[r] += b4;
    if (!TRAIN) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int gr = v0 + 4 * (lane >> 4) + r;
            if (col < kMlpOut && gr < nrows) out[(size_t)gr * kMlpOut + col] = o[r];
        }
        return;
    }
    // ---- backward: loss = loss_scale * sum (o - t)^2 over valid rows, 3 outputs
    float* P = partial + (size_t)w * MlpLayout::total;
    f32x4 d4;
    float lsum = 0.f;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int gr = v0 + 4 * (lane >> 4) + r;
        float d = 0.f;
        if (col < kMlpOut && gr < nrows) {
            const float e = o[r] - target[(size_t)(idx ? idx[gr] : gr) * kMlpOut + col];
            lsum += e * e;
            d = 2.f * loss_scale * e;
        }
        d4[r] = d;
    }
#pragma unroll
    for (int off = 32; off >= 1; off /= 2) lsum += __shfl_xor(lsum, off, 64);
    if (lane == 0) loss_part[w] = lsum * loss_scale;
    // layer 4: db4, dW4 = H3^T d4, dH3 = d4 W4^T
    {
        const float s = col_sum(d4);
        if (lane < 16) P[MlpLayout::b4 + lane] = s;
#pragma unroll
        for (int a = 0; a < H / 16; ++a) {
            const f32x4 acc = mfma(lds_at(H3, LDH, 16 * a, lane), d_as_b(d4), f32x4{0.f, 0.f, 0.f, 0.f});
#pragma unroll
            for (int r = 0; r < 4; ++r) P[MlpLayout::W4 + (16 * a + 4 * (lane >> 4) + r) * kMlpOutPad + col] = acc[r];
        }
        store_d(G, LDH, 0, lane, d4);
        wave_sync();
        f32x4 dh[H / 16];
#pragma unroll
        for (int it = 0; it < H / 16; ++it)
            dh[it] = mfma(lds_a(G, LDH, 0, lane), gl_b(wb + WO4, kMlpOutPad, 16 * it, 0, lane), f32x4{0.f, 0.f, 0.f, 0.f});
        wave_sync();
#pragma unroll
        for (int it = 0; it < H / 16; ++it)
#pragma unroll
            for (int r = 0; r < 4; ++r) z3[it][r] = dh[it][r] * gelu_d(z3[it][r]);  // z3 now holds dz3
    }
    f32x4 dA[H / 16];
    layer_bwd<true>(H2, LDH, H, z3, wb + WO3, G, P, MlpLayout::W3, MlpLayout::b3, dA, lane);
#pragma unroll
    for (int it = 0; it < H / 16; ++it)
#pragma unroll
        for (int r = 0; r < 4; ++r) z2[it][r] = dA[it][r] * gelu_d(z2[it][r]);
    layer_bwd<true>(H1, LDH, H, z2, wb + WO2, G, P, MlpLayout::W2, MlpLayout::b2, dA, lane);
#pragma unroll
    for (int it = 0; it < H / 16; ++it)
#pragma unroll
        for (int r = 0; r < 4; ++r) z1[it][r] = dA[it][r] * gelu_d(z1[it][r]);
    layer_bwd<false>(X0, LDX, kMlpInPad, z1, nullptr, G, P, MlpLayout::W1, MlpLayout::b1, dA, lane);
}

__global__ void k_mlp_reduce(int nw, const float* __restrict__ partial, float* __restrict__ grad) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= MlpLayout::total) return;
    float s = 0.f;
    for (int w = 0; w < nw; ++w) s += partial[(size_t)w * MlpLayout::total + p];
    grad[p] = s;
}

__global__ void k_mlp_adam(float* __restrict__ master, float* __restrict__ m, float* __restrict__ v,
                           const float* __restrict__ g, float lr, float c1, float c2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= MlpLayout::total) return;
    const float gp = g[p];
    const float mp = 0.9f * m[p] + 0.1f * gp;
    const float vp = 0.999f * v[p] + 0.001f * gp * gp;
    m[p] = mp;
    v[p] = vp;
    master[p] -= lr * (mp / c1) / (sqrtf(vp / c2) + 1e-8f);
}

// BF16 copies of the four weight matrices in both layouts
__global__ void k_mlp_pack(const float* __restrict__ master, uint16_t* __restrict__ wf, uint16_t* __restrict__ wb) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    int off, in, out, nin, nout, wo;
    if (t < WO2) { off = MlpLayout::W1; nin = kMlpInPad; nout = H; wo = WO1; in = t / H; out = t % H; }
    else if (t < WO3) { off = MlpLayout::W2; nin = H; nout = H; wo = WO2; in = (t - WO2) / H; out = (t - WO2) % H; }
    else if (t < WO4) { off = MlpLayout::W3; nin = H; nout = H; wo = WO3; in = (t - WO3) / H; out = (t - WO3) % H; }
    else if (t < WTOT) { off = MlpLayout::W4; nin = H; nout = kMlpOutPad; wo = WO4; in = (t - WO4) / kMlpOutPad; out = (t - WO4) % kMlpOutPad; }
    else return;
    const uint16_t b = f2bf(master[off + in * nout + out]);
    wf[wo + out * nin + in] = b;
    wb[wo + in * nout + out] = b;
}

__global__ void k_mlp_features(int n, const double* __restrict__ q, const double* __restrict__ v, const double* __restrict__ X,
                               const double* __restrict__ res, const float* __restrict__ adiag,
                               const float* __restrict__ logk, const float* __restrict__ contact,
                               const float* __restrict__ dist, double h, double L, double gacc,
                               const float* __restrict__ fmean, const float* __restrict__ fstd, int standardize,
                               float* __restrict__ feat) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float f[kMlpInPad];
    const double iL = 1.0 / L;
    for (int d = 0; d < 3; ++d) {
        const double qd = q[3 * i + d], vd = v[3 * i + d], xd = X[3 * i + d];
        f[d] = (float)((qd - xd) * iL);
        f[3 + d] = (float)(res[3 * i + d] / (double)adiag[i] * iL);
        f[6 + d] = (float)(h * vd * iL);
    }
    f[9] = logk[i];
    f[10] = contact ? contact[i] : 0.f;
    f[11] = dist ? dist[i] : 4.f;
    f[12] = (float)(h * h * gacc * iL);
    for (int c = kMlpIn; c < kMlpInPad; ++c) f[c] = 0.f;
    if (standardize)
        for (int c = 0; c < kMlpIn; ++c) f[c] = (f[c] - fmean[c]) / fstd[c];
    for (int c = 0; c < kMlpInPad; ++c) feat[(size_t)i * kMlpInPad + c] = f[c];
}

// g0 = q + h v, g1 = g0 + tscale * net (Dirichlet vertices keep their values)
__global__ void k_warm_guess(int nv, const double* __restrict__ q, const double* __restrict__ v, double h,
                             const float* __restrict__ net, float tscale, const unsigned char* __restrict__ dmask,
                             const double* __restrict__ dval, double* __restrict__ g0, double* __restrict__ g1) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nv) return;
    const bool fixed = dmask && dmask[i];
    for (int d = 0; d < 3; ++d) {
        const int k = 3 * i + d;
        const double a = fixed ? dval[k] : q[k] + h * v[k];
        g0[k] = a;
        g1[k] = fixed ? a : a + (net ? (double)tscale * (double)net[k] : 0.0);
    }
}

__global__ void k_mfma_probe(const float* __restrict__ A, const float* __restrict__ B, float* __restrict__ D) {
    const int l = threadIdx.x;
    bf16x4 a, b;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        a[j] = (short)f2bf(A[(l & 15) * 16 + 4 * (l >> 4) + j]);
        b[j] = (short)f2bf(B[(4 * (l >> 4) + j) * 16 + (l & 15)]);
    }
    const f32x4 d = mfma(a, b, f32x4{0.f, 0.f, 0.f, 0.f});
#pragma unroll
    for (int r = 0; r < 4; ++r) D[(4 * (l >> 4) + r) * 16 + (l & 15)] = d[r];
}

}  // namespace

void WarmStartNet::init(uint64_t seed, hipStream_t s) {
    std::vector<float> p(MlpLayout::total, 0.0f);
    std::mt19937_64 rng(seed);
    auto glorot = [&](int off, int nin_real, int nin, int nout_real, int nout, float gain) {
        std::uniform_real_distribution<float> U(-1.f, 1.f);
        const float a = gain * std::sqrt(6.0f / (float)(nin_real + nout_real));
        for (int i = 0; i < nin_real; ++i)
            for (int o = 0; o < nout_real; ++o) p[off + i * nout + o] = a * U(rng);
        (void)nin;
    };
    glorot(MlpLayout::W1, kMlpIn, kMlpInPad, H, H, 1.f);
    glorot(MlpLayout::W2, H, H, H, H, 1.f);
    glorot(MlpLayout::W3, H, H, H, H, 1.f);
    glorot(MlpLayout::W4, H, H, kMlpOut, kMlpOutPad, 0.1f);
    master_.upload(p);
    m_.alloc(MlpLayout::total);
    v_.alloc(MlpLayout::total);
    hip_check(hipMemsetAsync(m_.p, 0, sizeof(float) * MlpLayout::total, s), "adam m");
    hip_check(hipMemsetAsync(v_.p, 0, sizeof(float) * MlpLayout::total, s), "adam v");
    grad_.alloc(MlpLayout::total);
    wf_.alloc(WTOT);
    wb_.alloc(WTOT);
    step_ = 0;
    pack(s);
    ready_ = true;
}

void WarmStartNet::pack(hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_pack, dim3((WTOT + 255) / 256), dim3(256), 0, s, master_.p, wf_.p, wb_.p);
}

void WarmStartNet::set_params(const float* p, hipStream_t s) {
    check(ready_, "warm-start network not initialized");
    hip_check(hipMemcpyAsync(master_.p, p, sizeof(float) * MlpLayout::total, hipMemcpyHostToDevice, s), "params");
    pack(s);
    hip_check(hipStreamSynchronize(s), "params");
}

void WarmStartNet::get_params(float* p, hipStream_t s) {
    check(ready_, "warm-start network not initialized");
    hip_check(hipMemcpyAsync(p, master_.p, sizeof(float) * MlpLayout::total, hipMemcpyDeviceToHost, s), "params");
    hip_check(hipStreamSynchronize(s), "params");
}

void WarmStartNet::forward(hipStream_t s, int n, const float* feat, float* out) {
    check(ready_, "warm-start network not initialized");
    if (n <= 0) return;
    const int nw = (n + 15) / 16;
    hipLaunchKernelGGL(k_mlp_pass<false>, dim3((nw + WPB - 1) / WPB), dim3(64 * WPB), 0, s, n, feat, nullptr, nullptr,
                       wf_.p, wb_.p, master_.p, out, nullptr, nullptr, 0.f);
}

void WarmStartNet::grad_pass(hipStream_t s, int nrows, const float* feat, const float* target, const int* idx, float loss_scale) {
    const int nw = (nrows + 15) / 16;
    if (partial_.n < (size_t)nw * MlpLayout::total) partial_.alloc((size_t)nw * MlpLayout::total);
    if (loss_part_.n < (size_t)nw) loss_part_.alloc(nw);
    hipLaunchKernelGGL(k_mlp_pass<true>, dim3((nw + WPB - 1) / WPB), dim3(64 * WPB), 0, s, nrows, feat, target, idx, wf_.p,
                       wb_.p, master_.p, nullptr, partial_.p, loss_part_.p, loss_scale);
    hipLaunchKernelGGL(k_mlp_reduce, dim3((MlpLayout::total + 255) / 256), dim3(256), 0, s, nw, partial_.p, grad_.p);
}

double WarmStartNet::loss_grad(hipStream_t s, int n, const float* feat, const float* target, float* grad) {
    check(ready_, "warm-start network not initialized");
    check(n > 0, "empty batch");
    grad_pass(s, n, feat, target, nullptr, 1.0f / (3.0f * (float)n));
    const int nw = (n + 15) / 16;
    std::vector<float> lp(nw);
    hip_check(hipMemcpyAsync(grad, grad_.p, sizeof(float) * MlpLayout::total, hipMemcpyDeviceToDevice, s), "grad");
    hip_check(hipMemcpyAsync(lp.data(), loss_part_.p, sizeof(float) * nw, hipMemcpyDeviceToHost, s), "loss");
    hip_check(hipStreamSynchronize(s), "loss_grad");
    double L = 0;
    for (float x : lp) L += x;
    return L;
}

void WarmStartNet::train(hipStream_t s, int n, const float* feat, const float* target, int steps, int batch, float lr,
                         uint64_t seed, float* losses) {
    check(ready_, "warm-start network not initialized");
    check(n > 0 && batch > 0 && steps >= 0, "invalid training sizes");
    batch = std::min(batch, n);
    std::mt19937_64 rng(seed);
    std::vector<int> perm(n);
    std::iota(perm.begin(), perm.end(), 0);
    int pos = n;
    const int nw = (batch + 15) / 16;
    std::vector<float> lp(nw);
    for (int st = 0; st < steps; ++st) {
        if (pos + batch > n) {
            std::shuffle(perm.begin(), perm.end(), rng);
            idx_.upload(perm);
            pos = 0;
        }
        grad_pass(s, batch, feat, target, idx_.p + pos, 1.0f / (3.0f * (float)batch));
        pos += batch;
        ++step_;
        const float c1 = 1.0f - std::pow(0.9f, (float)step_), c2 = 1.0f - std::pow(0.999f, (float)step_);
        hipLaunchKernelGGL(k_mlp_adam, dim3((MlpLayout::total + 255) / 256), dim3(256), 0, s, master_.p, m_.p, v_.p, grad_.p,
                           lr, c1, c2);
        pack(s);
        if (losses) {
            hip_check(hipMemcpyAsync(lp.data(), loss_part_.p, sizeof(float) * nw, hipMemcpyDeviceToHost, s), "loss");
            hip_check(hipStreamSynchronize(s), "train step");
            double L = 0;
            for (float x : lp) L += x;
            losses[st] = (float)L;
        }
    }
    hip_check(hipStreamSynchronize(s), "train");
}

void launch_mlp_features(hipStream_t s, int n, const double* q, const double* v, const double* X, const double* res,
                         const float* adiag, const float* logk, const float* contact, const float* dist, double h, double L,
                         double gacc, const float* fmean, const float* fstd, int standardize, float* feat) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_mlp_features, dim3((n + 127) / 128), dim3(128), 0, s, n, q, v, X, res, adiag, logk, contact, dist,
                       h, L, gacc, fmean, fstd, standardize, feat);
}

void launch_warm_guess(hipStream_t s, int nv, const double* q, const double* v, double h, const float* net, float tscale,
                       const unsigned char* dmask, const double* dval, double* g0, double* g1) {
    if (nv <= 0) return;
    hipLaunchKernelGGL(k_warm_guess, dim3((nv + 127) / 128), dim3(128), 0, s, nv, q, v, h, net, tscale, dmask, dval, g0, g1);
}

void mlp_mfma_probe(const float* A, const float* B, float* D) {
    DevBuf<float> a, b, d;
    a.upload(A, 256);
    b.upload(B, 256);
    d.alloc(256);
    hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, 0, a.p, b.p, d.p);
    hip_check(hipDeviceSynchronize(), "mfma probe");
    d.download(D, 256);
}

}  // namespace phd
