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

// Panel step k, part 1: one wave factors the diagonal block, lane r holding
// row r in registers as a window that shifts by one column per step (x[m] =
// A(r, j + m) at step j), so the current column is always x[0] and the
// update reads the published pivot column at static offsets from a moving
// base. Results are collected in LDS and written once at the end, so the
// per-step barriers wait on LDS only. Writes L_kk into M and a column-major
// copy plus the reciprocal pivots into Ld.
template <class TV>
__global__ void __launch_bounds__(64) k_chol_diag(int k, int64_t lda, TV* __restrict__ M, TV* __restrict__ Ld,
                                                  int* __restrict__ fail) {
    __shared__ TV col[2 * CB];          // L(l, j) of the current step, zero-padded
    __shared__ TV Lo[CB * CB + CB];     // Lo[j * CB + l] = L(l, j); reciprocal pivots at the end
    const int r = threadIdx.x;
    const int64_t c0 = (int64_t)k * CB;
    TV* Mk = M + c0 * lda;
    TV x[CB];
#pragma unroll
    for (int c = 0; c < CB; ++c) x[c] = Mk[c0 + r + (int64_t)c * lda];
    col[CB + r] = TV(0);
    bool bad_any = false;
    for (int j = 0; j < CB; ++j) {
        const TV p = __shfl(x[0], j);
        const bool bad = !(p > TV(0));
        bad_any |= bad;
        const TV d = bad ? TV(1) : sqrt(p);
        const TV inv = TV(1) / d;
        const TV lrj = (r > j) ? x[0] * inv : (r == j ? d : TV(0));
        col[r] = lrj;
        Lo[j * CB + r] = lrj;
        if (r == j) Lo[CB * CB + j] = inv;
        // one wave: its LDS operations complete in program order, so a
        // wavefront-scope fence (compiler ordering only) replaces the barrier
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        const TV* cj = col + j;
#pragma unroll
        for (int m = 1; m < CB; ++m) x[m - 1] = x[m] - lrj * cj[m];
        x[CB - 1] = TV(0);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
    if (bad_any && r == 0) fail[0] = 1;
#pragma unroll
    for (int e = 0; e < CB + 1; ++e) Ld[e * CB + r] = Lo[e * CB + r];
#pragma unroll
    for (int c = 0; c < CB; ++c)
        if (c <= r) Mk[c0 + r + (int64_t)c * lda] = Lo[c * CB + r];
}

// Panel step k, part 2: workgroup w solves X L_kk^T = A_ik for row block
// i = k + 1 + w (i == T is the right-hand-side block), lane r = row r held in
// a shifting register window as above, L_kk staged in LDS (read-only here).
template <class TV>
__global__ void __launch_bounds__(64) k_chol_trsm(int k, int64_t lda, TV* __restrict__ M, const TV* __restrict__ Ld) {
    __shared__ TV Lc[CB * CB + 2 * CB];  // Lc[j * CB + l] = L(l, j); reciprocal pivots at CB*CB; zero tail
    const int r = threadIdx.x;
    const int i = k + 1 + blockIdx.x;
    const int64_t c0 = (int64_t)k * CB;
    TV* Mk = M + c0 * lda;
    // all loads in flight at once (a rolled loop pays the memory latency per load)
#pragma unroll
    for (int e = 0; e < CB + 1; ++e) Lc[e * CB + r] = Ld[e * CB + r];
    TV y[CB];
#pragma unroll
    for (int c = 0; c < CB; ++c) y[c] = Mk[(int64_t)i * CB + r + (int64_t)c * lda];
    __shared__ TV rp[CB];
    __syncthreads();
    rp[r] = Lc[CB * CB + r];
    Lc[CB * CB + r] = TV(0);  // the window below reads past column 63: keep it zero
    Lc[CB * CB + CB + r] = TV(0);
    __syncthreads();
    for (int j = 0; j < CB; ++j) {
        const TV yj = y[0] * rp[j];
        Mk[(int64_t)i * CB + r + (int64_t)j * lda] = yj;
        const TV* lj = Lc + j * (CB + 1);  // lj[m] = L(j + m, j)
#pragma unroll
        for (int m = 1; m < CB; ++m) y[m - 1] = y[m] - yj * lj[m];
        y[CB - 1] = TV(0);
    }
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

// Backward substitution step b: x_b = L_bb^-T y_b (every workgroup, lane r
// holding y_r, shuffles instead of barriers), then workgroup c < b updates
// y_c -= L_bc^T x_b; workgroup b stores x_b.
template <class TV>
__global__ void __launch_bounds__(64) k_chol_bsolve(int b, int64_t lda, const TV* __restrict__ M, TV* __restrict__ y,
                                                    TV* __restrict__ x) {
    __shared__ TV Lr[CB][CB + 1];  // Lr[row][col] of the loaded block
    const int r = threadIdx.x;
    const int c = blockIdx.x;
    const int64_t o = (int64_t)b * CB;
#pragma unroll
    for (int cc = 0; cc < CB; ++cc) Lr[r][cc] = M[o + r + (o + cc) * lda];
    __syncthreads();
    TV s = y[o + r];
    TV xv[CB];
    TV mine = 0;
#pragma unroll
    for (int j = CB - 1; j >= 0; --j) {
        const TV xj = __shfl(s, j) / Lr[j][j];
        xv[j] = xj;
        if (r == j) mine = xj;
        if (r < j) s -= Lr[j][r] * xj;
    }
    if (c == b) {
        x[o + r] = mine;
        return;
    }
    __syncthreads();
    const int64_t oc = (int64_t)c * CB;
#pragma unroll
    for (int cc = 0; cc < CB; ++cc) Lr[r][cc] = M[o + r + (oc + cc) * lda];  // Lr[l][cc] = L_bc(l, cc)
    __syncthreads();
    TV acc = 0;
#pragma unroll
    for (int l = 0; l < CB; ++l) acc += Lr[l][r] * xv[l];
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
void launch_chol_solve(hipStream_t s, int n, TV* M, TV* work, TV* x, int* fail) {
    const int np = chol_padded(n);
    const int64_t lda = chol_lda(n);
    const int T = np / CB;
    TV* y = work;
    TV* Ld = work + np;
    for (int k = 0; k < T; ++k) {
        hipLaunchKernelGGL(k_chol_diag<TV>, dim3(1), dim3(64), 0, s, k, lda, M, Ld, fail);
        hipLaunchKernelGGL(k_chol_trsm<TV>, dim3(T - k), dim3(64), 0, s, k, lda, M, (const TV*)Ld);
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
