// Dense SPD solve for the contact system: blocked right-looking Cholesky with
// the right-hand side as an extra block row, then blocked backward substitution.
#include "phd_dense.h"

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

// Panel step k. Workgroup w handles row block i = k + w (i == T is the
// right-hand-side block). Every workgroup factors the diagonal block itself
// (left-looking, lane = row), then solves its block X L_kk^T = A_ik.
template <class TV>
__global__ void __launch_bounds__(64) k_chol_panel(int k, int T, int64_t lda, TV* __restrict__ M, int* __restrict__ fail) {
    __shared__ TV D[CB][CB + 1];
    __shared__ TV X[CB][CB + 1];
    const int r = threadIdx.x;
    const int i = k + blockIdx.x;
    const int64_t c0 = (int64_t)k * CB;
    TV* Mk = M + c0 * lda;
    for (int c = 0; c < CB; ++c) D[r][c] = Mk[c0 + r + (int64_t)c * lda];
    if (i != k)
        for (int c = 0; c < CB; ++c) X[r][c] = Mk[(int64_t)i * CB + r + (int64_t)c * lda];
    __syncthreads();
    for (int j = 0; j < CB; ++j) {
        if (r >= j) {
            TV s0 = D[r][j], s1 = 0, s2 = 0, s3 = 0;
            int l = 0;
            for (; l + 3 < j; l += 4) {
                s0 -= D[r][l] * D[j][l];
                s1 -= D[r][l + 1] * D[j][l + 1];
                s2 -= D[r][l + 2] * D[j][l + 2];
                s3 -= D[r][l + 3] * D[j][l + 3];
            }
            for (; l < j; ++l) s0 -= D[r][l] * D[j][l];
            D[r][j] = (s0 + s1) + (s2 + s3);
        }
        __syncthreads();
        const TV p = D[j][j];
        const bool bad = !(p > TV(0));
        const TV d = bad ? TV(1) : sqrt(p);
        if (bad && r == j && blockIdx.x == 0) fail[0] = 1;
        __syncthreads();
        if (r > j) D[r][j] = D[r][j] / d;
        else if (r == j) D[j][j] = d;
        __syncthreads();
    }
    if (i == k) {
        for (int c = 0; c <= r; ++c) Mk[c0 + r + (int64_t)c * lda] = D[r][c];
        return;
    }
    for (int j = 0; j < CB; ++j) {
        TV s0 = X[r][j], s1 = 0, s2 = 0, s3 = 0;
        int l = 0;
        for (; l + 3 < j; l += 4) {
            s0 -= X[r][l] * D[j][l];
            s1 -= X[r][l + 1] * D[j][l + 1];
            s2 -= X[r][l + 2] * D[j][l + 2];
            s3 -= X[r][l + 3] * D[j][l + 3];
        }
        for (; l < j; ++l) s0 -= X[r][l] * D[j][l];
        X[r][j] = ((s0 + s1) + (s2 + s3)) / D[j][j];
    }
    for (int c = 0; c < CB; ++c) Mk[(int64_t)i * CB + r + (int64_t)c * lda] = X[r][c];
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
    for (int e = t; e < CB * CB; e += 256) {
        const int row = e % CB, l = e / CB;
        As[l][row] = M[(int64_t)i * CB + row + (c0 + l) * lda];
        Bs[l][row] = M[(int64_t)j * CB + row + (c0 + l) * lda];
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

// Backward substitution step b: x_b = L_bb^-T y_b (every workgroup), then
// workgroup c < b updates y_c -= L_bc^T x_b; workgroup b stores x_b.
template <class TV>
__global__ void __launch_bounds__(64) k_chol_bsolve(int b, int64_t lda, const TV* __restrict__ M, TV* __restrict__ y,
                                                    TV* __restrict__ x) {
    __shared__ TV D[CB][CB + 1];
    __shared__ TV xs[CB];
    const int r = threadIdx.x;
    const int c = blockIdx.x;
    const int64_t o = (int64_t)b * CB;
    for (int cc = 0; cc < CB; ++cc) D[r][cc] = M[o + r + (o + cc) * lda];
    __syncthreads();
    TV s = y[o + r];
    for (int j = CB - 1; j >= 0; --j) {
        if (r == j) xs[j] = s / D[j][j];
        __syncthreads();
        const TV xj = xs[j];
        if (r < j) s -= D[j][r] * xj;
    }
    if (c == b) {
        x[o + r] = xs[r];
        return;
    }
    __syncthreads();
    const int64_t oc = (int64_t)c * CB;
    for (int cc = 0; cc < CB; ++cc) D[r][cc] = M[o + r + (oc + cc) * lda];  // D[l][cc] = L_bc(l, cc)
    __syncthreads();
    TV acc = 0;
    for (int l = 0; l < CB; ++l) acc += D[l][r] * xs[l];
    y[oc + r] -= acc;
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
void launch_chol_solve(hipStream_t s, int n, TV* M, TV* y, TV* x, int* fail) {
    const int np = chol_padded(n);
    const int64_t lda = chol_lda(n);
    const int T = np / CB;
    for (int k = 0; k < T; ++k) {
        hipLaunchKernelGGL(k_chol_panel<TV>, dim3(T - k + 1), dim3(64), 0, s, k, T, lda, M, fail);
        if (k + 1 < T) hipLaunchKernelGGL(k_chol_update<TV>, dim3(T - k - 1, T - k), dim3(256), 0, s, k, lda, M);
    }
    hipLaunchKernelGGL(k_chol_gety<TV>, dim3((np + 255) / 256), dim3(256), 0, s, np, lda, (const TV*)M, y);
    for (int b = T - 1; b >= 0; --b) hipLaunchKernelGGL(k_chol_bsolve<TV>, dim3(b + 1), dim3(64), 0, s, b, lda, (const TV*)M, y, x);
}

void launch_clear_flag(hipStream_t s, int* flag) { hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(64), 0, s, flag); }

template void launch_chol_assemble<float>(hipStream_t, int, const int*, const double*, int, const double*, const double*,
                                          const double*, const double*, float*);
template void launch_chol_assemble<double>(hipStream_t, int, const int*, const double*, int, const double*, const double*,
                                           const double*, const double*, double*);
template void launch_chol_solve<float>(hipStream_t, int, float*, float*, float*, int*);
template void launch_chol_solve<double>(hipStream_t, int, double*, double*, double*, int*);

}  // namespace phd
