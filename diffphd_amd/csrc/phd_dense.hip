// Dense SPD solve for the contact system: blocked right-looking Cholesky with
// the right-hand side as an extra block row, then blocked backward substitution.
#include "phd_dense.h"

#include <cstring>

namespace phd {
namespace {

constexpr int CB = kCholBlock;

template <class TV>
__global__ void __launch_bounds__(256) k_chol_assemble(int nJ, int np, int64_t lda, const int* __restrict__ J,
                                                       const double* __restrict__ H, int K, const double* __restrict__ nrm,
                                                       const double* __restrict__ om, const double* __restrict__ E,
                                                       const double* __restrict__ rhs, TV* __restrict__ M) {
    const int64_t rows = (int64_t)np + CB;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= rows * np) return;
    const int r = (int)(idx % rows), c = (int)(idx / rows);
    if (r < c) return;
    double v;
    if (r >= np) {
        v = (r == np && c < nJ) ? rhs[c] : 0.0;
    } else if (r >= nJ || c >= nJ) {
        v = (r == c) ? 1.0 : 0.0;
    } else {
        const int i = J[r], j = J[c];
        const double nd = nrm[3 * i] * nrm[3 * j] + nrm[3 * i + 1] * nrm[3 * j + 1] + nrm[3 * i + 2] * nrm[3 * j + 2];
        // H is symmetric: read along its row j so that consecutive r stay close
        v = om[i] * nd * H[(int64_t)j * K + i] * om[j];
        if (r == c) v += E[i];
    }
    M[r + (int64_t)c * lda] = (TV)v;
}

__device__ __forceinline__ float lane_bcast(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double lane_bcast(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, l), hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

template <class TV> struct V16;
template <> struct V16<float> { using type = float4; static constexpr int N = 4; };
template <> struct V16<double> { using type = double2; static constexpr int N = 2; };

constexpr int kColStride = 2 * CB + 4;  // one published pivot column, see k_chol_panel

// Steps [j, j + 16) of the panel chain. Window z[m] = column j + m of the
// lane's row; columns past 63 never matter, and from step j on only
// m <= 63 - j can reach one, so this quarter of the steps stops at m = MB.
template <class TV, int MB>
__device__ __forceinline__ void panel_quarter(int& j, TV (&z)[CB], TV* colS, bool diag_wave, int r, bool& bad_any,
                                              TV* out, int64_t out_ld) {
    using VT = typename V16<TV>::type;
    constexpr int VN = V16<TV>::N;
    for (int jj = 0; jj < 16; ++jj, ++j) {
        TV* cb = colS + (j & 1) * kColStride;
        TV coef = TV(0);
        if (diag_wave) {
            const TV p = lane_bcast(z[0], j);
            const bool bad = !(p > TV(0));
            bad_any |= bad;
            const TV d = bad ? TV(1) : sqrt(p);
            const TV inv = TV(1) / d;
            coef = (r > j) ? z[0] * inv : (r == j ? d : TV(0));
            // L(r, j) at CB + r - j: L(j + m, j) lands at CB + m for every step
            cb[CB + r - j] = coef;
            if (r == CB - 1) {  // the two slots past the column, stale from step j - 2
                cb[2 * CB - j] = TV(0);
                cb[2 * CB + 1 - j] = TV(0);
            }
        }
        __syncthreads();
        if (!diag_wave) {
            coef = z[0] * (TV(1) / cb[CB]);
            out[r + (int64_t)j * out_ld] = coef;
        }
#pragma unroll
        for (int q = 0; q < (MB + 1) / 16; ++q) {
            TV c[16];
#pragma unroll
            for (int e = 0; e < 16 / VN; ++e) {
                const VT t = *reinterpret_cast<const VT*>(cb + CB + 16 * q + VN * e);
                memcpy(c + VN * e, &t, sizeof(VT));
            }
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                const int m = 16 * q + e;
                if (m >= 1) z[m - 1] = z[m] - coef * c[e];
            }
        }
    }
}

// Panel step k, fused: workgroup w of T - k + 1, two waves. Wave 0 factors the
// diagonal block (every workgroup repeats it): lane r holds row r as a window
// that shifts one column per step, the pivot comes from lane j by readlane,
// and the pivot column L(., j) is published through LDS (two alternating
// buffers, so one barrier per step). Wave 1 solves X L_kk^T = B in lockstep,
// B = block row k + 1 + w of M for w < T - k (w = T - k - 1 is the
// right-hand-side block) or the identity for w = T - k, which gives
// Xinv_k = L_kk^-T (column-major 64 x 64) for the backward substitution.
template <class TV>
__global__ void __launch_bounds__(128) k_chol_panel(int k, int T, int64_t lda, TV* __restrict__ M, TV* __restrict__ Xinv,
                                                    int* __restrict__ fail) {
    __shared__ __align__(16) TV colS[2 * kColStride];
    const int wave = threadIdx.x / 64, r = threadIdx.x % 64;
    const int w = blockIdx.x;
    const bool diag_wave = wave == 0;
    const int64_t c0 = (int64_t)k * CB;
    TV z[CB];
    TV* out;
    int64_t out_ld;
    if (diag_wave) {
#pragma unroll
        for (int c = 0; c < CB; ++c) z[c] = M[c0 + r + (c0 + c) * lda];
        out = nullptr;
        out_ld = 0;
    } else if (w < T - k) {
        const int64_t i0 = (int64_t)(k + 1 + w) * CB;
#pragma unroll
        for (int c = 0; c < CB; ++c) z[c] = M[i0 + r + (c0 + c) * lda];
        out = M + i0 + c0 * lda;
        out_ld = lda;
    } else {
#pragma unroll
        for (int c = 0; c < CB; ++c) z[c] = (c == r) ? TV(1) : TV(0);
        out = Xinv + (int64_t)k * CB * CB;
        out_ld = CB;
    }
    for (int e = threadIdx.x; e < 2 * kColStride; e += 128) colS[e] = TV(0);
    __syncthreads();
    bool bad_any = false;
    int j = 0;
    panel_quarter<TV, 63>(j, z, colS, diag_wave, r, bad_any, out, out_ld);
    panel_quarter<TV, 47>(j, z, colS, diag_wave, r, bad_any, out, out_ld);
    panel_quarter<TV, 31>(j, z, colS, diag_wave, r, bad_any, out, out_ld);
    panel_quarter<TV, 15>(j, z, colS, diag_wave, r, bad_any, out, out_ld);
    if (diag_wave && bad_any && w == 0) fail[0] = 1;
}

// Trailing update of step k: A_ij -= L_ik L_jk^T for k < j <= i <= T. Thread
// (tx, ty) owns rows tx + 16q and columns ty + 16p of the 64 x 64 tile.
template <class TV>
__global__ void __launch_bounds__(256) k_chol_update(int k, int64_t lda, TV* __restrict__ M) {
    const int j = k + 1 + blockIdx.x;
    const int i = k + 1 + blockIdx.y;
    if (i < j) return;
    __shared__ TV As[CB][CB];
    __shared__ TV Bs[CB][CB];
    const int t = threadIdx.x;
    const int64_t c0 = (int64_t)k * CB;
    {
        TV ta[CB * CB / 256], tb[CB * CB / 256];
#pragma unroll
        for (int u = 0; u < CB * CB / 256; ++u) {
            const int e = t + 256 * u, row = e % CB, l = e / CB;
            ta[u] = M[(int64_t)i * CB + row + (c0 + l) * lda];
            tb[u] = M[(int64_t)j * CB + row + (c0 + l) * lda];
        }
#pragma unroll
        for (int u = 0; u < CB * CB / 256; ++u) {
            const int e = t + 256 * u, row = e % CB, l = e / CB;
            As[l][row] = ta[u];
            Bs[l][row] = tb[u];
        }
    }
    __syncthreads();
    const int tx = t % 16, ty = t / 16;
    TV acc[4][4];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int p = 0; p < 4; ++p) acc[q][p] = 0;
    for (int l = 0; l < CB; ++l) {
        TV a[4], b[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) { a[q] = As[l][tx + 16 * q]; b[q] = Bs[l][ty + 16 * q]; }
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int p = 0; p < 4; ++p) acc[q][p] += a[q] * b[p];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            TV* dst = M + (int64_t)i * CB + tx + 16 * q + ((int64_t)j * CB + ty + 16 * p) * lda;
            *dst -= acc[q][p];
        }
}

// y = row np of the layout (L^-1 rhs after the factorization).
template <class TV>
__global__ void k_chol_gety(int np, int64_t lda, const TV* __restrict__ M, TV* __restrict__ y) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < np) y[c] = M[np + (int64_t)c * lda];
}

// Backward substitution step b with Xinv_b = L_bb^-T (column-major): every
// workgroup forms x_b = Xinv_b y_b; workgroup b stores it, workgroup a < b
// updates y_a -= L_ba^T x_b.
template <class TV>
__global__ void __launch_bounds__(64) k_chol_bsolve(int b, int64_t lda, const TV* __restrict__ M,
                                                    const TV* __restrict__ Xinv, TV* __restrict__ y, TV* __restrict__ x) {
    __shared__ TV Ls[CB][CB + 1];  // Ls[l][cc] = L(64 b + l, 64 a + cc)
    __shared__ TV ys[CB], xs[CB];
    const int r = threadIdx.x;
    const int a = blockIdx.x;
    const int64_t o = (int64_t)b * CB, oa = (int64_t)a * CB;
    const TV* Xb = Xinv + (int64_t)b * CB * CB;
    TV xcol[CB];
#pragma unroll
    for (int c = 0; c < CB; ++c) xcol[c] = Xb[c * CB + r];
    if (a < b) {
        TV lv[CB];
#pragma unroll
        for (int cc = 0; cc < CB; ++cc) lv[cc] = M[o + r + (oa + cc) * lda];
#pragma unroll
        for (int cc = 0; cc < CB; ++cc) Ls[r][cc] = lv[cc];
    }
    ys[r] = y[o + r];
    __syncthreads();
    TV xb = 0;
#pragma unroll
    for (int c = 0; c < CB; ++c) xb += xcol[c] * ys[c];
    if (a == b) {
        x[o + r] = xb;
        return;
    }
    xs[r] = xb;
    __syncthreads();
    TV acc = 0;
#pragma unroll
    for (int l = 0; l < CB; ++l) acc += Ls[l][r] * xs[l];
    y[oa + r] -= acc;
}

__global__ void k_clear_flag(int* f) {
    if (threadIdx.x == 0) f[0] = 0;
}

}  // namespace

template <class TV>
void launch_chol_assemble(hipStream_t s, int nJ, const int* J, const double* H, int K, const double* nrm,
                          const double* om, const double* E, const double* rhs, TV* M) {
    const int np = chol_padded(nJ);
    const int64_t lda = chol_lda(nJ);
    const int64_t total = lda * np;
    hipLaunchKernelGGL(k_chol_assemble<TV>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, nJ, np, lda, J, H, K,
                       nrm, om, E, rhs, M);
}

template <class TV>
void launch_chol_solve(hipStream_t s, int n, TV* M, TV* work, TV* x, int* fail) {
    const int np = chol_padded(n);
    const int64_t lda = chol_lda(n);
    const int T = np / CB;
    TV* y = work;
    TV* Xinv = work + np;
    for (int k = 0; k < T; ++k) {
        hipLaunchKernelGGL(k_chol_panel<TV>, dim3(T - k + 1), dim3(128), 0, s, k, T, lda, M, Xinv, fail);
        if (k + 1 < T) hipLaunchKernelGGL(k_chol_update<TV>, dim3(T - k - 1, T - k), dim3(256), 0, s, k, lda, M);
    }
    hipLaunchKernelGGL(k_chol_gety<TV>, dim3((np + 255) / 256), dim3(256), 0, s, np, lda, (const TV*)M, y);
    for (int b = T - 1; b >= 0; --b)
        hipLaunchKernelGGL(k_chol_bsolve<TV>, dim3(b + 1), dim3(64), 0, s, b, lda, (const TV*)M, (const TV*)Xinv, y, x);
}

void launch_clear_flag(hipStream_t s, int* flag) { hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(64), 0, s, flag); }

template void launch_chol_assemble<float>(hipStream_t, int, const int*, const double*, int, const double*, const double*,
                                          const double*, const double*, float*);
template void launch_chol_assemble<double>(hipStream_t, int, const int*, const double*, int, const double*, const double*,
                                           const double*, const double*, double*);
template void launch_chol_solve<float>(hipStream_t, int, float*, float*, float*, int*);
template void launch_chol_solve<double>(hipStream_t, int, double*, double*, double*, int*);

}  // namespace phd
