// Dense SPD solve for the contact system: blocked right-looking Cholesky with
// the right-hand side as an extra block row, then blocked backward substitution.
#include "phd_dense.h"

#include <cstring>
#include <type_traits>

namespace phd {
namespace {

constexpr int CB = kCholBlock;

template <class TV, class TH>
__global__ void __launch_bounds__(256) k_chol_assemble(int nJ, int np, int64_t lda, const int* __restrict__ J,
                                                       const TH* __restrict__ H, int K, const double* __restrict__ nrm,
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
        const double nd = nrm ? nrm[3 * i] * nrm[3 * j] + nrm[3 * i + 1] * nrm[3 * j + 1] + nrm[3 * i + 2] * nrm[3 * j + 2]
                              : 1.0;
        // H is symmetric: read along its row j so that consecutive r stay close
        v = om[i] * nd * (double)H[(int64_t)j * K + i] * om[j];
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

// 1 / sqrt(p) for the pivots: the hardware reciprocal square root (1 ulp in
// FP32); in FP64 refined by two Newton steps y <- y (3 - p y^2) / 2.
__device__ __forceinline__ float pivot_rsqrt(float p) { return __builtin_amdgcn_rsqf(p); }
__device__ __forceinline__ double pivot_rsqrt(double p) {
    double y = __builtin_amdgcn_rsq(p);
    const double h = 0.5 * p;
    y = y * (1.5 - h * y * y);
    y = y * (1.5 - h * y * y);
    return y;
}

template <class TV> struct V16;
template <> struct V16<float> { using type = float4; static constexpr int N = 4; };
template <> struct V16<double> { using type = double2; static constexpr int N = 2; };

constexpr int kColStride = 2 * CB + 4;  // one published pivot column, see k_chol_step
constexpr int kTSB = CB + 4;             // row stride of the published L columns (16-byte aligned rows)

// Pivot of step j from the window's column j (wave 0): publishes the column
// into buffer j & 1, L(r, j) at CB + r - j for r > j so that L(j + m, j) sits
// at CB + m for every j, and 1 / L(j, j) at CB (the slot of the diagonal,
// which no update reads). Returns L(r, j) for r > j, 1 / L(j, j) on lane j and
// zero above it (lanes r <= j only ever touch their own upper triangle).
template <class TV>
__device__ __forceinline__ TV panel_pivot(int j, TV zj, TV* colS, int r, bool& bad_any) {
    TV* cb = colS + (j & 1) * kColStride;
    const TV p = lane_bcast(zj, j);
    const bool bad = !(p > TV(0));
    bad_any |= bad;
    const TV inv = bad ? TV(1) : pivot_rsqrt(p);
    const TV coef = (r > j) ? zj * inv : (r == j ? inv : TV(0));
    cb[CB + r - j] = coef;
    if (r == CB - 1) {  // the two slots past the column, stale from step j - 2
        cb[2 * CB - j] = TV(0);
        cb[2 * CB + 1 - j] = TV(0);
    }
    return coef;
}

template <class TV>
__device__ __forceinline__ void read16(const TV* src, TV (&c)[16]) {
    using VT = typename V16<TV>::type;
    constexpr int VN = V16<TV>::N;
#pragma unroll
    for (int e = 0; e < 16 / VN; ++e) {
        const VT t = reinterpret_cast<const VT*>(src)[e];
        memcpy(c + VN * e, &t, sizeof(VT));
    }
}

// Steps [j, j + 16) of the panel chain. Window z[m] = column j + m of the
// lane's row; columns past 63 never matter, and from step j on only
// m <= 63 - j can reach one, so this quarter of the steps stops at m = MB.
// Role 0 (the diagonal wave) takes L(j + 1, j) from lane j + 1 by readlane,
// updates its column j + 1 and publishes the pivot of step j + 1 before the
// rest of step j's update, so the pivot chain never waits on LDS. Role 1
// solves its block row against the published columns; role 2 only keeps the
// barriers. The published column is read in chunks of 16, each chunk's reads
// issued before the previous chunk's update so the LDS latency overlaps the
// FMAs.
template <class TV, int MB>
__device__ __forceinline__ void panel_quarter(int& j, TV (&z)[CB], TV* colS, int role, int r, bool& bad_any, TV& coef,
                                              TV* out, int64_t out_ld) {
    constexpr int NQ = (MB + 1) / 16;
    for (int jj = 0; jj < 16; ++jj, ++j) {
        __syncthreads();
        if (role == 2) continue;
        const bool diag_wave = role == 0;
        const TV* cb = colS + (j & 1) * kColStride + CB;
        TV c[2][16];
        read16<TV>(cb, c[0]);
        if (NQ > 1) read16<TV>(cb + 16, c[1]);
        const bool chain = diag_wave && (j + 1 < CB);
        TV z0 = TV(0), coef_next = TV(0);
        if (chain) {
            z0 = z[1] - coef * lane_bcast(coef, j + 1);
            coef_next = panel_pivot<TV>(j + 1, z0, colS, r, bad_any);
        }
        if (!diag_wave) {
            coef = z[0] * c[0][0];
            out[r + (int64_t)j * out_ld] = coef;
        }
        z[0] = chain ? z0 : z[1] - coef * c[0][1];
#pragma unroll
        for (int m = 2; m < 16; ++m) z[m - 1] = z[m] - coef * c[0][m];
#pragma unroll
        for (int q = 1; q < NQ; ++q) {
            if (q + 1 < NQ) read16<TV>(cb + 16 * (q + 1), c[(q + 1) & 1]);
#pragma unroll
            for (int e = 0; e < 16; ++e) z[16 * q + e - 1] = z[16 * q + e] - coef * c[q & 1][e];
        }
        if (diag_wave) coef = coef_next;
    }
}

// ---- The panel in 16-column sub-blocks ------------------------------------
// The diagonal wave (lane r holds row r of the diagonal block in z) factors
// sub-block SB's 16 columns with readlane pivots and rank-1 updates restricted
// to those columns, so a step waits on register broadcasts only; L(r, j)
// replaces z[j] (0 above the diagonal) and inv[jj] = 1 / L(j, j).
template <class TV, int SB>
__device__ __forceinline__ void diag_chain(TV (&z)[CB], int r, bool& bad_any, TV& inv_mine) {
#pragma unroll
    for (int jj = 0; jj < 16; ++jj) {
        const int j = 16 * SB + jj;
        const TV p = lane_bcast(z[j], j);
        const bool bad = !(p > TV(0));
        bad_any |= bad;
        const TV iv = bad ? TV(1) : pivot_rsqrt(p);
        if (r == j) inv_mine = iv;
        const TV l = (r > j) ? z[j] * iv : (r == j ? p * iv : TV(0));
        z[j] = l;
#pragma unroll
        for (int m = j + 1; m < 16 * SB + 16; ++m) z[m] -= l * lane_bcast(l, m);
    }
}

// Sub-block SB of L through LDS: Lt[j][m] = L(m, j), invp[j] = 1 / L(j, j)
// (lane j holds its own pivot's inverse).
template <class TV, int SB>
__device__ __forceinline__ void diag_publish(const TV (&z)[CB], int r, TV inv_mine, TV (*Lt)[kTSB], TV* invp) {
#pragma unroll
    for (int jj = 0; jj < 16; ++jj) Lt[16 * SB + jj][r] = z[16 * SB + jj];
    if (r >= 16 * SB && r < 16 * SB + 16) invp[r] = inv_mine;
}

__device__ __forceinline__ float vget(const float4& t, int e) { return e == 0 ? t.x : e == 1 ? t.y : e == 2 ? t.z : t.w; }
__device__ __forceinline__ double vget(const double2& t, int e) { return e == 0 ? t.x : t.y; }

// Rank-16 update of the columns right of sub-block SB: z[m] -= sum_j z[j] L(m, j)
// over the sub-block's j (z[j] = L(r, j) in the diagonal wave, X(r, j) in the
// block-row wave); L(m, j) read as broadcast 16-byte vectors. FP64 only: the
// FP32 panel runs these updates on the matrix cores (diag_trail_mfma,
// row_trail_mfma), which keeps LDS latency off the panel's critical path.
template <class TV, int SB>
__device__ __forceinline__ void panel_trail(TV (&z)[CB], const TV (*Lt)[kTSB]) {
    using VT = typename V16<TV>::type;
    constexpr int VN = V16<TV>::N;
    constexpr int M0 = 16 * (SB + 1);
#pragma unroll
    for (int jj = 0; jj < 16; ++jj) {
        const TV x = z[16 * SB + jj];
#pragma unroll
        for (int m = M0; m < CB; m += VN) {
            const VT t = *reinterpret_cast<const VT*>(&Lt[16 * SB + jj][m]);
#pragma unroll
            for (int e = 0; e < VN; ++e) z[m + e] -= x * vget(t, e);
        }
    }
}

// The block-row wave's forward substitution on sub-block SB: X(r, j) =
// (B(r, j) - sum_{i<j} X(r, i) L(j, i)) / L(j, j), right-looking within the
// sub-block; X is written to out (column-major, out_ld).
template <class TV, int SB>
__device__ __forceinline__ void row_solve(TV (&z)[CB], const TV (*Lt)[kTSB], const TV* invp, int r, TV* out,
                                          int64_t out_ld) {
    using VT = typename V16<TV>::type;
    constexpr int VN = V16<TV>::N, NV = 16 / VN, j0 = 16 * SB;
    // the sub-block's pivot inverses and L row j + 1 are read while row j is
    // applied, so each step waits on registers, not on LDS
    VT ip[NV], cur[NV], nxt[NV];
#pragma unroll
    for (int e = 0; e < NV; ++e) {
        ip[e] = reinterpret_cast<const VT*>(invp + j0)[e];
        cur[e] = reinterpret_cast<const VT*>(&Lt[j0][j0])[e];
    }
#pragma unroll
    for (int jj = 0; jj < 16; ++jj) {
        const int j = j0 + jj;
        if (jj + 1 < 16) {
#pragma unroll
            for (int e = 0; e < NV; ++e) nxt[e] = reinterpret_cast<const VT*>(&Lt[j + 1][j0])[e];
        }
        const TV x = z[j] * vget(ip[jj / VN], jj % VN);
        z[j] = x;
        out[r + (int64_t)j * out_ld] = x;
#pragma unroll
        for (int m = jj + 1; m < 16; ++m) z[j0 + m] -= x * vget(cur[m / VN], m % VN);
#pragma unroll
        for (int e = 0; e < NV; ++e) cur[e] = nxt[e];
    }
}

// S[l][row] = M(r0 + row, c0 + l) for the 64 x 64 block at (r0, c0); 256
// threads, all loads in flight before the LDS stores.
constexpr int kTS = CB + 4;  // LDS tile row stride (16-byte aligned rows)

template <class TV>
__device__ __forceinline__ void stage_block(const TV* __restrict__ M, int64_t lda, int64_t r0, int64_t c0,
                                            TV (*S)[kTS]) {
    const int t = threadIdx.x;
    TV v[CB * CB / 256];
#pragma unroll
    for (int u = 0; u < CB * CB / 256; ++u) {
        const int e = t + 256 * u;
        v[u] = M[r0 + e % CB + (c0 + e / CB) * lda];
    }
#pragma unroll
    for (int u = 0; u < CB * CB / 256; ++u) {
        const int e = t + 256 * u;
        S[e / CB][e % CB] = v[u];
    }
}

// acc[p] = sum_l A[l][tx] B[l][16 ty + p]: thread (tx, ty) of 256 owns row tx
// and columns 16 ty .. 16 ty + 15 of a rank-64 product tile (a wave holds one
// column group, so B is read as broadcast 16-byte vectors).
template <class TV>
__device__ __forceinline__ void tile_product(const TV (*A)[kTS], const TV (*B)[kTS], int tx, int ty, TV (&acc)[16]) {
#pragma unroll
    for (int p = 0; p < 16; ++p) acc[p] = 0;
#pragma unroll 4
    for (int l = 0; l < CB; ++l) {
        const TV a = A[l][tx];
        TV b[16];
        read16<TV>(&B[l][16 * ty], b);
#pragma unroll
        for (int p = 0; p < 16; ++p) acc[p] += a * b[p];
    }
}

// FP32 rank-64 tile product on the matrix cores: C(row, col) = sum_l
// A[l][row] B[l][col], wave w owning rows 32 (w / 2) .. and columns
// 32 (w % 2) .. of the 64 x 64 tile. v_mfma_f32_32x32x2f32 is fed the
// transposed product D = C^T block (operand A lane i + 32 h = B[l][32 wj + i],
// operand B lane j + 32 h = A[l][32 wi + j], l = 2 s + h at step s), so that
// result register v of lane L holds C(32 wi + (L % 32), 32 wj + mfma_col(v, L)):
// the 32 lanes of a half-wave hold 32 consecutive rows of one column of the
// column-major matrix.
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ int mfma_col(int v, int lane) { return (v & 3) + 4 * (lane >> 5) + 8 * (v >> 2); }

__device__ __forceinline__ f32x16 tile_product_mfma(const float (*A)[kTS], const float (*B)[kTS], int wave, int lane) {
    const int wi = wave >> 1, wj = wave & 1, i = lane & 31, h = lane >> 5;
    // two accumulation chains (steps st and st + CB / 4) so that consecutive
    // MFMAs do not wait on each other's results
    f32x16 acc0, acc1;
#pragma unroll
    for (int v = 0; v < 16; ++v) acc0[v] = acc1[v] = 0.f;
#pragma unroll 4
    for (int st = 0; st < CB / 4; ++st) {
        const int l0 = 2 * st + h, l1 = CB / 2 + 2 * st + h;
        acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(B[l0][32 * wj + i], A[l0][32 * wi + i], acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(B[l1][32 * wj + i], A[l1][32 * wi + i], acc1, 0, 0, 0);
    }
    return acc0 + acc1;
}

// Exchange of the upper half-wave of x with the lower half-wave of y
// (v_permlane32_swap_b32): x <- [x.lo, y.lo], y <- [x.hi, y.hi].
__device__ __forceinline__ void swap_halves(float& x, float& y) {
    const auto p = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(y), false, false);
    x = __uint_as_float(p[0]);
    y = __uint_as_float(p[1]);
}

// z[m] -= C(r, m) for the columns m >= M0 of column tile TJ (m = 32 TJ ..
// 32 TJ + 31), given the tile's two row halves in the MFMA result layout (lo:
// rows 0..31, hi: rows 32..63; register v of lane L holds row 32 t + L % 32,
// column 32 TJ + mfma_col(v, L)). One half-wave swap per register pair puts
// row r in lane r: afterwards x holds column 32 TJ + c and y column 32 TJ + c + 4.
template <int TJ, int M0>
__device__ __forceinline__ void trail_apply(float (&z)[CB], const f32x16& lo, const f32x16& hi) {
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        const int c = 32 * TJ + (v & 3) + 8 * (v >> 2);
        if (c + 4 < M0) continue;
        float x = lo[v], y = hi[v];
        swap_halves(x, y);
        if (c >= M0) z[c] -= x;
        z[c + 4] -= y;
    }
}

// The panel's rank-16 updates on the matrix cores, C(r, m) = sum_j P(r, j)
// L(m, j) over sub-block SB's columns j for the columns m right of it, as
// 32 x 32 tiles of v_mfma_f32_32x32x2f32 (step s covers j = 16 SB + 2 s + h,
// h the half-wave). Operand lane i + 32 h of a row tile t is P(32 t + i, j):
// the half-wave swap of the register pair (z[j0 + 2 s], z[j0 + 2 s + 1]) gives
// both row tiles of step s at once.
// Diagonal wave: P = L (its own rows), only the lower part m <= r matters, so
// the column tile 32..63 needs only the row tile 32..63.
template <int SB>
__device__ __forceinline__ void diag_trail_mfma(float (&z)[CB]) {
    constexpr int j0 = 16 * SB, M0 = 16 * (SB + 1);
    float p0[8], p1[8];
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        p0[s] = z[j0 + 2 * s];
        p1[s] = z[j0 + 2 * s + 1];
        swap_halves(p0[s], p1[s]);
    }
    f32x16 zero;
#pragma unroll
    for (int v = 0; v < 16; ++v) zero[v] = 0.f;
    if constexpr (M0 < 32) {
        f32x16 a0 = zero, a1 = zero;
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            a0 = __builtin_amdgcn_mfma_f32_32x32x2f32(p0[s], p0[s], a0, 0, 0, 0);
            a1 = __builtin_amdgcn_mfma_f32_32x32x2f32(p0[s], p1[s], a1, 0, 0, 0);
        }
        trail_apply<0, M0>(z, a0, a1);
    }
    f32x16 a1 = zero;
#pragma unroll
    for (int s = 0; s < 8; ++s) a1 = __builtin_amdgcn_mfma_f32_32x32x2f32(p1[s], p1[s], a1, 0, 0, 0);
    trail_apply<1, M0>(z, zero, a1);
}

// Block-row wave: P = X (its rows of the solved block row), L(m, j) read from
// the published columns, Lt[j][m] = L(m, j), one 4-byte read per lane and step.
template <int SB>
__device__ __forceinline__ void row_trail_mfma(float (&z)[CB], const float (*Lt)[kTSB], int lane) {
    constexpr int j0 = 16 * SB, M0 = 16 * (SB + 1);
    const int i = lane & 31, h = lane >> 5;
    float p0[8], p1[8];
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        p0[s] = z[j0 + 2 * s];
        p1[s] = z[j0 + 2 * s + 1];
        swap_halves(p0[s], p1[s]);
    }
    f32x16 zero;
#pragma unroll
    for (int v = 0; v < 16; ++v) zero[v] = 0.f;
#pragma unroll
    for (int tj = M0 / 32; tj < 2; ++tj) {
        float l[8];
#pragma unroll
        for (int s = 0; s < 8; ++s) l[s] = Lt[j0 + 2 * s + h][32 * tj + i];
        f32x16 a0 = zero, a1 = zero;
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            a0 = __builtin_amdgcn_mfma_f32_32x32x2f32(l[s], p0[s], a0, 0, 0, 0);
            a1 = __builtin_amdgcn_mfma_f32_32x32x2f32(l[s], p1[s], a1, 0, 0, 0);
        }
        if (tj == 0) trail_apply<0, M0>(z, a0, a1);
        else trail_apply<1, M0>(z, a0, a1);
    }
}

// The FP32 panel's trailing updates on the matrix cores, FP64's on the VALU.
template <class TV, int SB>
__device__ __forceinline__ void diag_trail(TV (&z)[CB], const TV (*Lt)[kTSB]) {
    if constexpr (std::is_same<TV, float>::value) diag_trail_mfma<SB>(z);
    else panel_trail<TV, SB>(z, Lt);
}
template <class TV, int SB>
__device__ __forceinline__ void row_trail(TV (&z)[CB], const TV (*Lt)[kTSB], int lane) {
    if constexpr (std::is_same<TV, float>::value) row_trail_mfma<SB>(z, Lt, lane);
    else panel_trail<TV, SB>(z, Lt);
}

// Rank-64 update of step kp on block (ib, jb): A_ij -= L_{i,kp} L_{j,kp}^T
// (256 threads; the staging tiles are S0 and S1).
template <class TV>
__device__ __forceinline__ void update_block(TV* __restrict__ M, int64_t lda, int kp, int ib, int jb, TV (*S0)[kTS],
                                             TV (*S1)[kTS]) {
    const int t = threadIdx.x, tx = t % 64, ty = t / 64;
    if constexpr (std::is_same<TV, float>::value) {
        // the target block is read together with the staging loads, so its
        // latency overlaps theirs instead of following the product
        const int lane = t % 64, wave = t / 64;
        float* dst = M + (int64_t)ib * CB + 32 * (wave >> 1) + (lane & 31) + ((int64_t)jb * CB + 32 * (wave & 1)) * lda;
        float old[16];
#pragma unroll
        for (int v = 0; v < 16; ++v) old[v] = dst[mfma_col(v, lane) * lda];
        stage_block<TV>(M, lda, (int64_t)ib * CB, (int64_t)kp * CB, S0);
        stage_block<TV>(M, lda, (int64_t)jb * CB, (int64_t)kp * CB, S1);
        __syncthreads();
        const f32x16 acc = tile_product_mfma(S0, S1, wave, lane);
#pragma unroll
        for (int v = 0; v < 16; ++v) dst[mfma_col(v, lane) * lda] = old[v] - acc[v];
    } else {
        stage_block<TV>(M, lda, (int64_t)ib * CB, (int64_t)kp * CB, S0);
        stage_block<TV>(M, lda, (int64_t)jb * CB, (int64_t)kp * CB, S1);
        __syncthreads();
        TV acc[16];
        tile_product<TV>(S0, S1, tx, ty, acc);
        TV* dst = M + (int64_t)ib * CB + tx + ((int64_t)jb * CB + 16 * ty) * lda;
#pragma unroll
        for (int p = 0; p < 16; ++p) dst[p * lda] -= acc[p];
    }
}

// Step k of the factorization, one launch of 256-thread workgroups in two parts
// that touch disjoint blocks, so the panel of step k runs next to the trailing
// update of step k - 1:
//  * workgroups w < T - k + 1 (panel of block column k): the rank-64 update of
//    step k - 1 is applied here, in LDS, to the two blocks the workgroup
//    reads: the diagonal block (k, k) (every workgroup repeats it) and block
//    (k + 1 + w, k); w = T - k - 1 is the right-hand-side block row and
//    w = T - k the identity. Then wave 0 factors the diagonal block and
//    wave 1 solves X L_kk^T = B, writing L(k + 1 + w, k) into M, or
//    Xinv_k = L_kk^-T (column-major 64 x 64) for the identity. FP32 works in
//    16-column sub-blocks: wave 0 runs each sub-block's pivot chain on
//    readlane broadcasts and publishes its columns through LDS, wave 1
//    solves against them one sub-block behind, and both apply the rank-16
//    update of the columns to the right as MFMA tiles. FP64 publishes one
//    pivot column per barrier (two alternating LDS buffers);
//  * the remaining workgroups apply the update of step k - 1 to the blocks
//    (i, j) with k < j <= i <= T.
// Four workgroups per CU (the FP32 MFMA accumulators stay in VGPRs, so the
// update workgroups of a step run at four waves per SIMD).
template <class TV>
__global__ void __launch_bounds__(256, 4) k_chol_step(int k, int T, int64_t lda, TV* __restrict__ M,
                                                   TV* __restrict__ Xinv, int* __restrict__ fail) {
    __shared__ __align__(16) TV S0[CB][kTS];
    __shared__ __align__(16) TV S1[CB][kTS];
    __shared__ __align__(16) TV colS[2 * kColStride];
    const int t = threadIdx.x, tx = t % 64, ty = t / 64;
    const int NP = T - k + 1;
    const int kp = k - 1;
    if ((int)blockIdx.x >= NP) {
        int u = (int)blockIdx.x - NP, jb = k + 1, cnt = T - jb + 1;
        while (u >= cnt) { u -= cnt; ++jb; cnt = T - jb + 1; }
        update_block<TV>(M, lda, kp, jb + u, jb, S0, S1);
        return;
    }
    const int w = blockIdx.x;
    const int wave = t / 64, r = t % 64;
    const int role = wave < 2 ? wave : 2;
    const bool has_row = w < T - k;          // block row k + 1 + w of M (else the identity)
    const int64_t c0 = (int64_t)k * CB, i0 = (int64_t)(k + 1 + w) * CB;
    TV z[CB];
    using VT = typename V16<TV>::type;
    constexpr int VN = V16<TV>::N;
    if (k > 0) {
        // FP32: the two target blocks are read together with the staging loads
        const int row = 32 * (wave >> 1) + (r & 31);
        float aD[16], aR[16];
        if constexpr (std::is_same<TV, float>::value) {
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int64_t off = row + (c0 + 32 * (wave & 1) + mfma_col(v, r)) * lda;
                aD[v] = M[c0 + off];
                aR[v] = has_row ? M[i0 + off] : 0.f;
            }
        }
        stage_block<TV>(M, lda, c0, (int64_t)kp * CB, S0);
        if (has_row) stage_block<TV>(M, lda, i0, (int64_t)kp * CB, S1);
        __syncthreads();
        if constexpr (std::is_same<TV, float>::value) {
            const f32x16 accD = tile_product_mfma(S0, S0, wave, r);
            f32x16 accR;
            if (has_row) accR = tile_product_mfma(S1, S0, wave, r);
            __syncthreads();
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int col = 32 * (wave & 1) + mfma_col(v, r);
                S0[row][col] = aD[v] - accD[v];
                if (has_row) S1[row][col] = aR[v] - accR[v];
            }
        } else {
            TV accD[16], accR[16];
            tile_product<TV>(S0, S0, tx, ty, accD);
            if (has_row) tile_product<TV>(S1, S0, tx, ty, accR);
            TV aD[16], aR[16];
#pragma unroll
            for (int p = 0; p < 16; ++p) {
                const int64_t off = tx + (c0 + 16 * ty + p) * lda;
                aD[p] = M[c0 + off];
                aR[p] = has_row ? M[i0 + off] : TV(0);
            }
            __syncthreads();
#pragma unroll
            for (int p = 0; p < 16; ++p) {
                S0[tx][16 * ty + p] = aD[p] - accD[p];
                if (has_row) S1[tx][16 * ty + p] = aR[p] - accR[p];
            }
        }
        __syncthreads();
        if (role == 0) {
#pragma unroll
            for (int q = 0; q < CB / VN; ++q) {
                const VT t4 = reinterpret_cast<const VT*>(&S0[r][0])[q];
#pragma unroll
                for (int e = 0; e < VN; ++e) z[VN * q + e] = vget(t4, e);
            }
        } else if (role == 1) {
#pragma unroll
            for (int q = 0; q < CB / VN; ++q) {
                const VT t4 = has_row ? reinterpret_cast<const VT*>(&S1[r][0])[q] : VT{};
#pragma unroll
                for (int e = 0; e < VN; ++e) z[VN * q + e] = has_row ? vget(t4, e) : (VN * q + e == r ? TV(1) : TV(0));
            }
        }
    } else if (role == 0) {
#pragma unroll
        for (int c = 0; c < CB; ++c) z[c] = M[c0 + r + (c0 + c) * lda];
    } else if (role == 1) {
#pragma unroll
        for (int c = 0; c < CB; ++c) z[c] = has_row ? M[i0 + r + (c0 + c) * lda] : (c == r ? TV(1) : TV(0));
    }
    TV* out = nullptr;
    int64_t out_ld = 0;
    if (role == 1) {
        out = has_row ? M + i0 + c0 * lda : Xinv + (int64_t)k * CB * CB;
        out_ld = has_row ? lda : CB;
    }
    bool bad_any = false;
    if constexpr (std::is_same<TV, double>::value) {
        // FP64: one pivot per barrier (the sub-block form below needs twice the
        // FP32 registers for its broadcast operands and spills)
        for (int e = t; e < 2 * kColStride; e += 256) colS[e] = TV(0);
        __syncthreads();
        TV coef = TV(0);
        if (role == 0) coef = panel_pivot<TV>(0, z[0], colS, r, bad_any);
        int j = 0;
        panel_quarter<TV, 63>(j, z, colS, role, r, bad_any, coef, out, out_ld);
        panel_quarter<TV, 47>(j, z, colS, role, r, bad_any, coef, out, out_ld);
        panel_quarter<TV, 31>(j, z, colS, role, r, bad_any, coef, out, out_ld);
        panel_quarter<TV, 15>(j, z, colS, role, r, bad_any, coef, out, out_ld);
        if (role == 0 && bad_any && w == 0) fail[0] = 1;
        return;
    }
    __syncthreads();  // every wave holds its rows: S0 now carries the published L columns
    TV (*Lt)[kTSB] = reinterpret_cast<TV (*)[kTSB]>(&S0[0][0]);
    TV* invp = colS;
    TV inv = TV(1);
    // the diagonal wave runs one sub-block ahead of the block-row wave
    if (role == 0) { diag_chain<TV, 0>(z, r, bad_any, inv); diag_publish<TV, 0>(z, r, inv, Lt, invp); }
    __syncthreads();
    if (role == 0) {
        diag_trail<TV, 0>(z, Lt);
        diag_chain<TV, 1>(z, r, bad_any, inv);
        diag_publish<TV, 1>(z, r, inv, Lt, invp);
    } else if (role == 1) {
        row_solve<TV, 0>(z, Lt, invp, r, out, out_ld);
        row_trail<TV, 0>(z, Lt, r);
    }
    __syncthreads();
    if (role == 0) {
        diag_trail<TV, 1>(z, Lt);
        diag_chain<TV, 2>(z, r, bad_any, inv);
        diag_publish<TV, 2>(z, r, inv, Lt, invp);
    } else if (role == 1) {
        row_solve<TV, 1>(z, Lt, invp, r, out, out_ld);
        row_trail<TV, 1>(z, Lt, r);
    }
    __syncthreads();
    if (role == 0) {
        diag_trail<TV, 2>(z, Lt);
        diag_chain<TV, 3>(z, r, bad_any, inv);
        diag_publish<TV, 3>(z, r, inv, Lt, invp);
    } else if (role == 1) {
        row_solve<TV, 2>(z, Lt, invp, r, out, out_ld);
        row_trail<TV, 2>(z, Lt, r);
    }
    __syncthreads();
    if (role == 1) row_solve<TV, 3>(z, Lt, invp, r, out, out_ld);
    if (role == 0 && bad_any && w == 0) fail[0] = 1;
}

// y = row np of the layout (L^-1 rhs after the factorization).
template <class TV>
__global__ void k_chol_gety(int np, int64_t lda, const TV* __restrict__ M, TV* __restrict__ y) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < np) y[c] = M[np + (int64_t)c * lda];
}

// Backward substitution over the group of blocks b0 .. b0 + 3 (those below T),
// with Xinv_b = L_bb^-T (column-major). Wave w owns block b0 + w. Every
// workgroup solves the group: from the top block down, x_i = Xinv_i y_i, then
// y_w -= L_{i,w}^T x_i for the group's blocks w < i. Workgroup b0 stores the
// group's x; workgroup a < b0 updates y_a -= sum_i L_{b0+i,a}^T x_i. Four
// blocks per launch cut the number of dependent launches by four.
constexpr int kBsGroup = 4;

template <class TV>
__global__ void __launch_bounds__(256) k_chol_bsolve(int b0, int T, int64_t lda, const TV* __restrict__ M,
                                                     const TV* __restrict__ Xinv, TV* __restrict__ y, TV* __restrict__ x) {
    __shared__ TV ys[kBsGroup][CB];
    __shared__ TV xs[kBsGroup][CB];
    __shared__ TV part[kBsGroup][CB];
    const int t = threadIdx.x, w = t / CB, r = t % CB;
    const int a = blockIdx.x;
    const int G = min(kBsGroup, T - b0);
    const int64_t bw = (int64_t)b0 + w;
    const bool valid = w < G;
    TV X[CB];  // X[c] = Xinv_bw(r, c)
    if (valid) {
        const TV* Xb = Xinv + bw * CB * CB;
#pragma unroll
        for (int c = 0; c < CB; ++c) X[c] = Xb[c * CB + r];
        ys[w][r] = y[bw * CB + r];
    }
    __syncthreads();
    for (int i = G - 1; i >= 0; --i) {
        TV Lv[CB];  // Lv[l] = L(64 (b0 + i) + l, 64 (b0 + w) + r)
        if (w < i) {
            const TV* Lc = M + ((int64_t)b0 + i) * CB + (bw * CB + r) * lda;
#pragma unroll
            for (int l = 0; l < CB; ++l) Lv[l] = Lc[l];
        }
        if (w == i) {
            TV s = 0;
#pragma unroll
            for (int c = 0; c < CB; ++c) s += X[c] * ys[i][c];
            xs[i][r] = s;
        }
        __syncthreads();
        if (w < i) {
            TV s = 0;
#pragma unroll
            for (int l = 0; l < CB; ++l) s += Lv[l] * xs[i][l];
            ys[w][r] -= s;
        }
        __syncthreads();
    }
    if (a == b0) {
        if (valid) x[bw * CB + r] = xs[w][r];
        return;
    }
    TV s = 0;
    if (valid) {
        const TV* Lc = M + bw * CB + ((int64_t)a * CB + r) * lda;
        TV U[CB];
#pragma unroll
        for (int l = 0; l < CB; ++l) U[l] = Lc[l];
#pragma unroll
        for (int l = 0; l < CB; ++l) s += U[l] * xs[w][l];
    }
    part[w][r] = s;
    __syncthreads();
    if (w == 0) y[(int64_t)a * CB + r] -= (part[0][r] + part[1][r]) + (part[2][r] + part[3][r]);
}

__global__ void k_clear_flag(int* f) {
    if (threadIdx.x == 0) f[0] = 0;
}

}  // namespace

template <class TV, class TH>
void launch_chol_assemble(hipStream_t s, int nJ, const int* J, const TH* H, int K, const double* nrm,
                          const double* om, const double* E, const double* rhs, TV* M) {
    const int np = chol_padded(nJ);
    const int64_t lda = chol_lda(nJ);
    const int64_t total = lda * np;
    hipLaunchKernelGGL((k_chol_assemble<TV, TH>), dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, nJ, np, lda, J, H,
                       K, nrm, om, E, rhs, M);
}

template <class TV>
void launch_chol_solve(hipStream_t s, int n, TV* M, TV* work, TV* x, int* fail) {
    const int np = chol_padded(n);
    const int64_t lda = chol_lda(n);
    const int T = np / CB;
    TV* y = work;
    TV* Xinv = work + np;
    for (int k = 0; k < T; ++k) {
        // update blocks of step k - 1 left to the second part: columns k + 1 .. T - 1, rows j .. T
        const int nu = k == 0 ? 0 : (T - k) * (T - k + 1) / 2 - 1;
        hipLaunchKernelGGL(k_chol_step<TV>, dim3(T - k + 1 + nu), dim3(256), 0, s, k, T, lda, M, Xinv, fail);
    }
    hipLaunchKernelGGL(k_chol_gety<TV>, dim3((np + 255) / 256), dim3(256), 0, s, np, lda, (const TV*)M, y);
    for (int b0 = kBsGroup * ((T - 1) / kBsGroup); b0 >= 0; b0 -= kBsGroup)
        hipLaunchKernelGGL(k_chol_bsolve<TV>, dim3(b0 + 1), dim3(256), 0, s, b0, T, lda, (const TV*)M, (const TV*)Xinv, y,
                           x);
}

void launch_clear_flag(hipStream_t s, int* flag) { hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(64), 0, s, flag); }

template void launch_chol_assemble<float, double>(hipStream_t, int, const int*, const double*, int, const double*,
                                                  const double*, const double*, const double*, float*);
template void launch_chol_assemble<float, float>(hipStream_t, int, const int*, const float*, int, const double*,
                                                 const double*, const double*, const double*, float*);
template void launch_chol_assemble<double, double>(hipStream_t, int, const int*, const double*, int, const double*,
                                                   const double*, const double*, const double*, double*);
template void launch_chol_solve<float>(hipStream_t, int, float*, float*, float*, int*);
template void launch_chol_solve<double>(hipStream_t, int, double*, double*, double*, int*);

}  // namespace phd
