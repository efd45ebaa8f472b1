// HIP kernels (gfx950) of the PD forward/backward time step.
//
// Data layout in HBM:
//  * positions / forces: FP64 AoS, 3 per vertex, original vertex order;
//  * solver vectors of the factor: SoA per axis (w[d * n + j]) in postorder;
//  * factor values V (FP32 or FP64): row k of S = L^-1 holds the contiguous
//    column range [fd(k), k] at V[base[k] + j], base[k] a multiple of 64, so
//    both products stream V with 256/512-byte coalesced wavefront accesses and
//    no column indices at all.
#include <hip/hip_runtime.h>

#include "phd_kernels.h"

#include <cstring>
#include <stdexcept>
#include "phd_math.h"

namespace phd {

namespace {

constexpr int WAVE = 64;

template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = WAVE / 2; off > 0; off >>= 1) v += __shfl_down(v, off, WAVE);
    return v;
}

inline int blocks_for(int64_t n, int bs) { return (int)((n + bs - 1) / bs); }

// -----------------------------------------------------------------------------
// S = L^-1 expansion: one workgroup (4 waves) per tile of 64 consecutive
// postorder columns; lane l carries column 64t + l of L^-1. The path (the
// tile's rows, then the ancestors of its last column) is walked in chain
// segments: runs of at most 64 consecutive columns i -> i+1 = parent(i) whose
// patterns nest (pattern(i) = {i} + pattern(i+1)), so every column of the
// segment shares the rows R below its last column. Per segment:
//   1. its 64-column block of X is staged in LDS and wave 0 solves it against
//      the segment's dense diagonal block (writing the S entries);
//   2. the four waves update X on R, four rows at a time, with the product of
//      L(R, segment) and the LDS block accumulated in registers: one global
//      read-modify-write per (row, column) per segment.
// -----------------------------------------------------------------------------
constexpr int EXP_SEG = 64;

__global__ void __launch_bounds__(256) k_expand_s(int n, int t_begin, const int64_t* __restrict__ Lp,
                                                  const int* __restrict__ Li, const double* __restrict__ Lx,
                                                  const int* __restrict__ parent, const int* __restrict__ depth,
                                                  const int* __restrict__ fd, const int64_t* __restrict__ base,
                                                  const unsigned char* __restrict__ chain, double* __restrict__ ws,
                                                  int64_t ws_stride, double* __restrict__ V64, float* __restrict__ V32) {
    __shared__ double xs[EXP_SEG][WAVE];
    __shared__ int64_t cpos[EXP_SEG];
    const int t = t_begin + blockIdx.x;
    const int l = threadIdx.x % WAVE;
    const int wv = threadIdx.x / WAVE;
    const int g0 = WAVE * t;
    const int glast = min(n - 1, g0 + WAVE - 1);
    const int ng = glast - g0 + 1;
    const int dl = depth[glast];
    const int nrows = ng + dl;
    double* X = ws + (int64_t)blockIdx.x * ws_stride;
    for (int r = wv; r < nrows; r += 4) X[(int64_t)r * WAVE + l] = (r == l && l < ng) ? 1.0 : 0.0;
    __syncthreads();
    const int j = g0 + l;
    // row of node k in X: tile rows first, then the ancestors by increasing height
    auto lk_of = [&](int k) { return k <= glast ? k - g0 : ng + dl - depth[k] - 1; };
    int a = g0;
    while (a != -1) {
        int b = a;
        while (b - a + 1 < EXP_SEG && chain[b]) ++b;
        const int w = b - a + 1;
        const int la = lk_of(a);  // segment rows are consecutive in X
        for (int s2 = wv; s2 < w; s2 += 4) xs[s2][l] = X[(int64_t)(la + s2) * WAVE + l];
        if (threadIdx.x < w) cpos[threadIdx.x] = Lp[a + threadIdx.x] + (w - (int)threadIdx.x);
        __syncthreads();
        if (wv == 0) {
            for (int s2 = 0; s2 < w; ++s2) {
                const int i = a + s2;
                const int64_t p0 = Lp[i];
                const double x = xs[s2][l] / Lx[p0];
                xs[s2][l] = x;
                if (l < ng && j >= fd[i] && j <= i) {
                    const int64_t at = base[i] + j;
                    V64[at] = x;
                    V32[at] = (float)x;
                }
                for (int u = s2 + 1; u < w; ++u) xs[u][l] -= Lx[p0 + (u - s2)] * x;
            }
        }
        __syncthreads();
        const int64_t pb = Lp[b] + 1;
        const int nr = (int)(Lp[b + 1] - pb);
        for (int r0 = 4 * wv; r0 < nr; r0 += 16) {
            double acc[4] = {0.0, 0.0, 0.0, 0.0};
            const int m = min(4, nr - r0);
            for (int s2 = 0; s2 < w; ++s2) {
                const double xv = xs[s2][l];
                const int64_t c = cpos[s2] + r0;
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (q < m) acc[q] += Lx[c + q] * xv;
            }
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (q < m) {
                    double* xk = X + (int64_t)lk_of(Li[pb + r0 + q]) * WAVE + l;
                    *xk -= acc[q];
                }
        }
        __syncthreads();
        a = (b < glast) ? b + 1 : parent[b];
    }
}

// 16-byte vector of a value type: 4 floats or 2 doubles per lane load.
template <class TV> struct Vec16;
template <> struct Vec16<float> { using type = float4; static constexpr int N = 4; };
template <> struct Vec16<double> { using type = double2; static constexpr int N = 2; };

// -----------------------------------------------------------------------------
// y = S w by (group of 64 consecutive rows, chunk of CW columns) units: the
// workgroup stages w on the chunk in LDS once, then its four waves stream the
// rows of the group that cross the chunk (16-byte vectors, the whole chunk in
// flight per row) against it. Row k's dot product on the chunk goes to its
// slot rbase[k] + chunk - fd(k) / CW; k_spmv_rows_slots sums the slots. CW is
// one of kSpmvChunks (picked by the autotuner; the plan is built for it).
// -----------------------------------------------------------------------------
template <class TV, int CW>
__global__ void __launch_bounds__(256) k_spmv_rows_grp(int nunits, const int* __restrict__ gu_group,
                                                       const int* __restrict__ gu_chunk, const int* __restrict__ fd,
                                                       const int64_t* __restrict__ base, const int* __restrict__ rbase,
                                                       const TV* __restrict__ V, const TV* __restrict__ w, int n,
                                                       TV* __restrict__ pslot) {
    using V16 = typename Vec16<TV>::type;
    constexpr int VN = Vec16<TV>::N;
    constexpr int U = CW / (VN * WAVE);
    __shared__ __align__(16) TV ws[3][CW];
    const int u = blockIdx.x;
    if (u >= nunits) return;
    const int g = gu_group[u], ch = gu_chunk[u];
    const int c0 = ch * CW, c1 = c0 + CW - 1;
    const int ldw = spmv_ldw(n);
    // columns >= n stage as zeros whatever the padding of w holds: the masked
    // lanes below still multiply 0 by the staged value, and 0 * NaN is NaN
    for (int e = (int)threadIdx.x * VN; e < CW; e += 256 * VN)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            V16 t;
            if (c0 + e + VN <= n) {
                t = *reinterpret_cast<const V16*>(w + (int64_t)d * ldw + c0 + e);
            } else {
                TV tt[VN];
#pragma unroll
                for (int i = 0; i < VN; ++i) tt[i] = (c0 + e + i < n) ? w[(int64_t)d * ldw + c0 + e + i] : TV(0);
                memcpy(&t, tt, sizeof(t));
            }
            *reinterpret_cast<V16*>(&ws[d][e]) = t;
        }
    __syncthreads();
    const int lane = threadIdx.x % WAVE, wv = threadIdx.x / WAVE;
    // the metadata of this wave's 16 rows (64 g + wv + 4 l), one row per lane,
    // loaded together so that each row waits on one memory round trip
    int mf = 0, mrb = 0;
    int64_t mb = 0;
    {
        const int kk = 64 * g + wv + 4 * lane;
        if (lane < 16 && kk < n) { mf = fd[kk]; mb = base[kk]; mrb = rbase[kk]; }
    }
    // row l's values are loaded one row ahead of its FMAs and reduction
    const int k0 = 64 * g + wv;
    const int nr = k0 >= n ? 0 : min(16, (n - k0 + 3) / 4);
    auto bounds = [&](int l, int& f, int& lo, int& hi) {
        f = __builtin_amdgcn_readlane(mf, l);
        lo = max(f, c0);
        hi = min(k0 + 4 * l, c1);
    };
    auto load = [&](int l, TV (&v)[U][VN]) {
        int f, lo, hi;
        bounds(l, f, lo, hi);
        if (lo > hi) return;
        const int64_t bk = ((int64_t)__builtin_amdgcn_readlane((int)(mb >> 32), l) << 32) |
                           (uint32_t)__builtin_amdgcn_readlane((int)mb, l);
        const TV* vr = V + bk;
        const int cs = (lo & ~(VN - 1)) + VN * lane;
#pragma unroll
        for (int q = 0; q < U; ++q) {
            const int c = cs + q * VN * WAVE;
            V16 t;
            if (c <= hi) t = *reinterpret_cast<const V16*>(vr + c);
            else memset(&t, 0, sizeof(t));
            memcpy(v[q], &t, sizeof(t));
        }
    };
    auto finish = [&](int l, const TV (&v)[U][VN]) {
        int f, lo, hi;
        bounds(l, f, lo, hi);
        if (lo > hi) return;
        const int cs = (lo & ~(VN - 1)) + VN * lane;
        TV a0 = 0, a1 = 0, a2 = 0;
#pragma unroll
        for (int q = 0; q < U; ++q) {
            const int c = cs + q * VN * WAVE;
            if (c <= hi) {
                TV x0[VN], x1[VN], x2[VN];
                memcpy(x0, &ws[0][c - c0], sizeof(V16));
                memcpy(x1, &ws[1][c - c0], sizeof(V16));
                memcpy(x2, &ws[2][c - c0], sizeof(V16));
#pragma unroll
                for (int e = 0; e < VN; ++e) {
                    const TV vv = (c + e >= lo && c + e <= hi) ? v[q][e] : TV(0);
                    a0 += vv * x0[e];
                    a1 += vv * x1[e];
                    a2 += vv * x2[e];
                }
            }
        }
        a0 = wave_sum(a0);
        a1 = wave_sum(a1);
        a2 = wave_sum(a2);
        if (lane == 0) {
            TV* p = pslot + 3 * (int64_t)(__builtin_amdgcn_readlane(mrb, l) + ch - f / CW);
            p[0] = a0;
            p[1] = a1;
            p[2] = a2;
        }
    };
    TV va[U][VN], vb[U][VN];
    if (nr > 0) load(0, va);
#pragma unroll 1
    for (int l = 0; l < nr; l += 2) {
        if (l + 1 < nr) load(l + 1, vb);
        finish(l, va);
        if (l + 1 >= nr) break;
        if (l + 2 < nr) load(l + 2, va);
        finish(l + 1, vb);
    }
}

template <class TV>
__global__ void __launch_bounds__(256) k_spmv_rows_slots(int n, const int* __restrict__ rbase,
                                                         const TV* __restrict__ pslot, TV* __restrict__ y) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    // rows near the top of the etree cross up to ~140 chunks: eight independent
    // partial sums keep eight loads in flight per lane
    TV a[8][3] = {};
    const int s0 = rbase[k], s1 = rbase[k + 1];
    int s = s0;
    for (; s + 7 < s1; s += 8)
#pragma unroll
        for (int u = 0; u < 8; ++u)
#pragma unroll
            for (int d = 0; d < 3; ++d) a[u][d] += pslot[3 * (int64_t)(s + u) + d];
    for (; s < s1; ++s)
#pragma unroll
        for (int d = 0; d < 3; ++d) a[0][d] += pslot[3 * (int64_t)s + d];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const TV t = ((a[0][d] + a[1][d]) + (a[2][d] + a[3][d])) + ((a[4][d] + a[5][d]) + (a[6][d] + a[7][d]));
        y[(int64_t)d * n + k] = t;
    }
}

// -----------------------------------------------------------------------------
// x = S^T y: one wavefront per (tile, item range) unit. A row of the tile is
// 64 columns = 64 / VN lanes of 16-byte vectors, so one load instruction
// covers VN rows (lane sub-group `sub`); four such loads are in flight. The
// item metadata of 64 rows is loaded cooperatively (one row per lane) and
// broadcast with cross-lane reads. The sub-group partial sums are reduced at
// the end.
// -----------------------------------------------------------------------------
template <class TV>
__global__ void __launch_bounds__(256) k_spmv_cols(int nunits, const int* __restrict__ cu_tile,
                                                   const int* __restrict__ cu_i0, const int* __restrict__ cu_i1,
                                                   const int* __restrict__ items, const int* __restrict__ fd,
                                                   const int64_t* __restrict__ base, const TV* __restrict__ V,
                                                   const TV* __restrict__ y, int n, TV* __restrict__ xpart) {
    using V16 = typename Vec16<TV>::type;
    constexpr int VN = Vec16<TV>::N;
    constexpr int LPR = WAVE / VN;  // lanes per row
    constexpr int U = 4;
    const int u = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    const int lane = threadIdx.x % WAVE;
    if (u >= nunits) return;
    const int t = cu_tile[u];
    const int sub = lane / LPR, g = lane % LPR;
    const int jb = 64 * t + VN * g;
    const int i0 = cu_i0[u], i1 = cu_i1[u];
    TV acc0[VN], acc1[VN], acc2[VN];
#pragma unroll
    for (int e = 0; e < VN; ++e) acc0[e] = acc1[e] = acc2[e] = 0;
    for (int it0 = i0; it0 < i1; it0 += WAVE) {
        const int my = it0 + lane;
        const int cnt = min(WAVE, i1 - it0);
        int mk = -1, mf = 0x7fffffff;
        int mblo = 0, mbhi = 0;
        TV my0 = 0, my1 = 0, my2 = 0;
        if (my < i1) {
            mk = items[my];
            mf = fd[mk];
            const int64_t b = base[mk];
            mblo = (int)(b & 0xffffffff);
            mbhi = (int)(b >> 32);
            my0 = y[mk];
            my1 = y[n + mk];
            my2 = y[2 * n + mk];
        }
        for (int q = 0; q < cnt; q += VN * U) {
            TV v[U][VN];
            TV ys0[U], ys1[U], ys2[U];
            int kk[U], ff[U];
#pragma unroll
            for (int s = 0; s < U; ++s) {
                const int src = q + s * VN + sub;
                kk[s] = __shfl(mk, src, WAVE);
                ff[s] = __shfl(mf, src, WAVE);
                const int64_t bb = ((int64_t)__shfl(mbhi, src, WAVE) << 32) | (uint32_t)__shfl(mblo, src, WAVE);
                ys0[s] = __shfl(my0, src, WAVE);
                ys1[s] = __shfl(my1, src, WAVE);
                ys2[s] = __shfl(my2, src, WAVE);
                V16 tv;
                if (src < cnt) tv = *reinterpret_cast<const V16*>(V + bb + jb);
                else memset(&tv, 0, sizeof(tv));
                memcpy(v[s], &tv, sizeof(tv));
                if (src >= cnt) kk[s] = -1;
            }
#pragma unroll
            for (int s = 0; s < U; ++s)
#pragma unroll
                for (int e = 0; e < VN; ++e) {
                    const int j = jb + e;
                    const TV vv = (j >= ff[s] && j <= kk[s]) ? v[s][e] : TV(0);
                    acc0[e] += vv * ys0[s];
                    acc1[e] += vv * ys1[s];
                    acc2[e] += vv * ys2[s];
                }
        }
    }
#pragma unroll
    for (int off = LPR; off < WAVE; off *= 2)
#pragma unroll
        for (int e = 0; e < VN; ++e) {
            acc0[e] += __shfl_xor(acc0[e], off, WAVE);
            acc1[e] += __shfl_xor(acc1[e], off, WAVE);
            acc2[e] += __shfl_xor(acc2[e], off, WAVE);
        }
    if (sub == 0) {
        TV* xp = xpart + (int64_t)u * 3 * WAVE;
#pragma unroll
        for (int e = 0; e < VN; ++e) {
            xp[VN * g + e] = acc0[e];
            xp[WAVE + VN * g + e] = acc1[e];
            xp[2 * WAVE + VN * g + e] = acc2[e];
        }
    }
}

template <class TV>
__global__ void k_spmv_cols_fin(int n, const int* __restrict__ tile_unit_ptr, const TV* __restrict__ xpart,
                                const int* __restrict__ perm, const double* __restrict__ add,
                                const unsigned char* __restrict__ dmask, const double* __restrict__ dval,
                                double* __restrict__ out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int t = j / WAVE, lane = j % WAVE;
    double a0 = 0, a1 = 0, a2 = 0;
    for (int u = tile_unit_ptr[t]; u < tile_unit_ptr[t + 1]; ++u) {
        const TV* xp = xpart + (int64_t)u * 3 * WAVE;
        a0 += (double)xp[lane];
        a1 += (double)xp[WAVE + lane];
        a2 += (double)xp[2 * WAVE + lane];
    }
    const int v = perm[j];
    if (dmask && dmask[v]) {
        out[3 * v] = dval[3 * v];
        out[3 * v + 1] = dval[3 * v + 1];
        out[3 * v + 2] = dval[3 * v + 2];
        return;
    }
    if (add) { a0 += add[3 * v]; a1 += add[3 * v + 1]; a2 += add[3 * v + 2]; }
    out[3 * v] = a0;
    out[3 * v + 1] = a1;
    out[3 * v + 2] = a2;
}

template <class TV>
__global__ void k_perm_in(int n, const int* __restrict__ perm, const double* __restrict__ r,
                          const unsigned char* __restrict__ dmask, TV* __restrict__ w) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int v = perm[j];
    const bool z = dmask && dmask[v];
    const int ldw = spmv_ldw(n);
    w[j] = z ? TV(0) : (TV)r[3 * v];
    w[ldw + j] = z ? TV(0) : (TV)r[3 * v + 1];
    w[2 * ldw + j] = z ? TV(0) : (TV)r[3 * v + 2];
}

// -----------------------------------------------------------------------------
// Local step (ProjectiveDynamicsLocalStep, deformable_projective_dynamics_forward.cpp:186-292):
// one thread per element; per sample F = x_e G_s, one SVD shared by every
// projection, contribution sum_energies w A_s^T vec(p(F)).
// -----------------------------------------------------------------------------
template <int EDIM, class TC>
__global__ void __launch_bounds__(256) k_local_step(DevMaterials m, const double* __restrict__ q,
                                                    double* __restrict__ elem_out) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m.ne) return;
    double x[EDIM][3];
#pragma unroll
    for (int k = 0; k < EDIM; ++k) {
        const int v = m.elems[k * m.ne + e];
        x[k][0] = q[3 * v]; x[k][1] = q[3 * v + 1]; x[k][2] = q[3 * v + 2];
    }
    double acc[EDIM][3];
#pragma unroll
    for (int k = 0; k < EDIM; ++k) acc[k][0] = acc[k][1] = acc[k][2] = 0.0;
    const double wc = m.wc ? m.wc[e] : 0.0;
    const double wv = m.wv ? m.wv[e] : 0.0;
    const double wm = m.mw ? m.mw[e] : 0.0;
    double mf[3] = {0.0, 0.0, 0.0}, ma = 0.0;
    if (wm != 0.0) { mf[0] = m.mfib[e]; mf[1] = m.mfib[m.ne + e]; mf[2] = m.mfib[2 * m.ne + e]; ma = m.mact[e]; }
    const double* Ge = m.G + (int64_t)e * m.g_stride;
    for (int s = 0; s < m.samples; ++s) {
        const double* G = Ge + s * EDIM * 3;
        double F[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) F[i] = 0.0;
#pragma unroll
        for (int k = 0; k < EDIM; ++k)
#pragma unroll
            for (int jj = 0; jj < 3; ++jj) {
                const double g = G[k * 3 + jj];
                F[0 + 3 * jj] += x[k][0] * g;
                F[1 + 3 * jj] += x[k][1] * g;
                F[2 + 3 * jj] += x[k][2] * g;
            }
        // F in FP64 (positions minus the sum of their gradients cancel to
        // O(1)); the SVD and the projections in TC
        TC Fc[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) Fc[i] = (TC)F[i];
        TC U[9], sg[3], V[9];
        svd3(Fc, U, sg, V);
        double P[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) P[i] = 0.0;
        if (wc != 0.0) {
            const TC one[3] = {TC(1), TC(1), TC(1)};
            TC R[9];
            udv(U, one, V, R);
#pragma unroll
            for (int i = 0; i < 9; ++i) P[i] += wc * (double)R[i];
        }
        if (wv != 0.0) {
            TC D[3], sd[3], Pv[9];
            volume_D(sg, D);
            sd[0] = sg[0] + D[0]; sd[1] = sg[1] + D[1]; sd[2] = sg[2] + D[2];
            udv(U, sd, V, Pv);
#pragma unroll
            for (int i = 0; i < 9; ++i) P[i] += wv * (double)Pv[i];
        }
        for (int h = 0; h < m.n_nh; ++h) {
            const double wn = m.wn[h][e];
            if (wn == 0.0) continue;
            TC ss[3], Pn[9];
            nh_prox((TC)m.nh_mu[h], (TC)m.nh_la[h], (TC)m.nh_k[h], sg, ss);
            udv(U, ss, V, Pn);
#pragma unroll
            for (int i = 0; i < 9; ++i) P[i] += wn * (double)Pn[i];
        }
        if (wm != 0.0) {
            // muscle (ProjectToManifold, pd_muscle_energy.cpp:118-125): Bp = a F m / |F m|,
            // contribution w M^T Bp = w Bp m^T (FP64 in both precisions)
            double Fm[3];
#pragma unroll
            for (int i = 0; i < 3; ++i) Fm[i] = F[i] * mf[0] + F[i + 3] * mf[1] + F[i + 6] * mf[2];
            const double l = sqrt(Fm[0] * Fm[0] + Fm[1] * Fm[1] + Fm[2] * Fm[2]);
            const double r = l <= DBL_EPSILON ? 0.0 : wm * ma / l;
#pragma unroll
            for (int j = 0; j < 3; ++j)
#pragma unroll
                for (int i = 0; i < 3; ++i) P[i + 3 * j] += r * Fm[i] * mf[j];
        }
#pragma unroll
        for (int k = 0; k < EDIM; ++k)
#pragma unroll
            for (int i = 0; i < 3; ++i)
                acc[k][i] += P[i] * G[k * 3] + P[i + 3] * G[k * 3 + 1] + P[i + 6] * G[k * 3 + 2];
    }
    double* o = elem_out + (int64_t)e * EDIM * 3;
#pragma unroll
    for (int k = 0; k < EDIM; ++k) { o[3 * k] = acc[k][0]; o[3 * k + 1] = acc[k][1]; o[3 * k + 2] = acc[k][2]; }
}

__global__ void k_gather_rhs(int nv, const int* __restrict__ inc_ptr, const int* __restrict__ inc_idx,
                             const double* __restrict__ elem_out, const double* __restrict__ c_dir,
                             const double* __restrict__ q_tilde, double inv_h2m, double* __restrict__ pd_rhs,
                             double* __restrict__ b) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    double s0 = 0, s1 = 0, s2 = 0;
    for (int p = inc_ptr[v]; p < inc_ptr[v + 1]; ++p) {
        const int64_t i = inc_idx[p];
        s0 += elem_out[3 * i];
        s1 += elem_out[3 * i + 1];
        s2 += elem_out[3 * i + 2];
    }
    s0 -= c_dir[3 * v]; s1 -= c_dir[3 * v + 1]; s2 -= c_dir[3 * v + 2];
    pd_rhs[3 * v] = s0; pd_rhs[3 * v + 1] = s1; pd_rhs[3 * v + 2] = s2;
    b[3 * v] = inv_h2m * q_tilde[3 * v] + s0;
    b[3 * v + 1] = inv_h2m * q_tilde[3 * v + 1] + s1;
    b[3 * v + 2] = inv_h2m * q_tilde[3 * v + 2] + s2;
}

__global__ void k_gather_vertex(int nv, const int* __restrict__ inc_ptr, const int* __restrict__ inc_idx,
                                const double* __restrict__ elem_out, double* __restrict__ out) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    double s0 = 0, s1 = 0, s2 = 0;
    for (int p = inc_ptr[v]; p < inc_ptr[v + 1]; ++p) {
        const int64_t i = inc_idx[p];
        s0 += elem_out[3 * i];
        s1 += elem_out[3 * i + 1];
        s2 += elem_out[3 * i + 2];
    }
    out[3 * v] = s0; out[3 * v + 1] = s1; out[3 * v + 2] = s2;
}

// r = b - A q (three right-hand sides, AoS), A in ELL form: entry j of row v
// at j * nv + v (padding: column v, value 0), so consecutive lanes read
// consecutive addresses; entries are added in row order.
__global__ void __launch_bounds__(256) k_residual(int nv, int width, const int* __restrict__ Ai, const double* __restrict__ Ax,
                                                  const double* __restrict__ b, const double* __restrict__ q,
                                                  const unsigned char* __restrict__ dmask,
                                                  const unsigned char* __restrict__ pmask, double* __restrict__ r) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    if (dmask && dmask[v]) { r[3 * v] = r[3 * v + 1] = r[3 * v + 2] = 0.0; return; }
    double s0 = b[3 * v], s1 = b[3 * v + 1], s2 = b[3 * v + 2];
    if (pmask) {
        // per-DOF Dirichlet: axis d of the matrix drops the couplings to the
        // DOFs fixed on axis d, and their rows are identities (residual 0)
        for (int j = 0; j < width; ++j) {
            const int c = Ai[(int64_t)j * nv + v];
            const double a = Ax[(int64_t)j * nv + v];
            if (!pmask[3 * c]) s0 -= a * q[3 * c];
            if (!pmask[3 * c + 1]) s1 -= a * q[3 * c + 1];
            if (!pmask[3 * c + 2]) s2 -= a * q[3 * c + 2];
        }
        r[3 * v] = pmask[3 * v] ? 0.0 : s0;
        r[3 * v + 1] = pmask[3 * v + 1] ? 0.0 : s1;
        r[3 * v + 2] = pmask[3 * v + 2] ? 0.0 : s2;
        return;
    }
    for (int j = 0; j < width; ++j) {
        const int c = Ai[(int64_t)j * nv + v];
        const double a = Ax[(int64_t)j * nv + v];
        s0 -= a * q[3 * c];
        s1 -= a * q[3 * c + 1];
        s2 -= a * q[3 * c + 2];
    }
    r[3 * v] = s0; r[3 * v + 1] = s1; r[3 * v + 2] = s2;
}

// -----------------------------------------------------------------------------
// Per-DOF Dirichlet conditions on the shared factor. A DOF fixed on its own
// (its vertex keeps a free DOF) stays in the shared matrix A; axis d's system
// A_ff x_f = r_f (F_d: the DOFs fixed on axis d) is the bordered system
// [A E; E^T 0][y; mu] = [r; 0]: y = A^-1 r - Z G_d^-1 (A^-1 r)_F with
// Z = A^-1 E (columns of A^-1 at the partially fixed vertices, one per vertex
// for every axis: A is the same scalar matrix on each) and G_d = Z_F.
// k_pdir_coef: c_d[p] = (G_d^-1 (y_F - t_F))[i] for the vertex p that is entry i
// of F_d (0 elsewhere); k_pdir_apply: y -= Z c per axis, then y = o on F_d.
// -----------------------------------------------------------------------------
__global__ void k_pdir_coef(int np, PdirAxes ax, const int* __restrict__ map, const int* __restrict__ fv,
                            const double* __restrict__ ginv, const double* __restrict__ y, const double* __restrict__ t,
                            double* __restrict__ coef) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int d = blockIdx.y;
    if (p >= np) return;
    const int i = map[d * np + p];
    double s = 0.0;
    if (i >= 0) {
        const int nf = ax.off[d + 1] - ax.off[d];
        const double* g = ginv + ax.goff[d] + (int64_t)i * nf;
        for (int j = 0; j < nf; ++j) {
            const int v = fv[ax.off[d] + j];
            s += g[j] * (y[3 * v + d] - (t ? t[3 * v + d] : 0.0));
        }
    }
    coef[d * np + p] = s;
}

__global__ void __launch_bounds__(256) k_pdir_apply(int nv, int np, const double* __restrict__ Z,
                                                    const double* __restrict__ coef, const unsigned char* __restrict__ pmask,
                                                    const double* __restrict__ o, double* __restrict__ y) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int p = 0; p < np; ++p) {
        const double z = Z[(int64_t)p * nv + v];
        s0 += z * coef[p];
        s1 += z * coef[np + p];
        s2 += z * coef[2 * np + p];
    }
    const double s[3] = {s0, s1, s2};
#pragma unroll
    for (int d = 0; d < 3; ++d) y[3 * v + d] = pmask[3 * v + d] ? (o ? o[3 * v + d] : 0.0) : y[3 * v + d] - s[d];
}

// out[q np + p] = Z[q nv + pv[p]]: the rows of the columns at the partial vertices
__global__ void k_pdir_gather(int np, int nv, const int* __restrict__ pv, const double* __restrict__ Z,
                              double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)np * np) return;
    const int q = (int)(i / np), p = (int)(i % np);
    out[i] = Z[(int64_t)q * nv + pv[p]];
}

// Z[(p0 + a) nv + v] = x[3 v + a] for a < na: axis slot a of one A^-1 product
// carries the column of vertex p0 + a
__global__ void k_pdir_columns(int nv, int p0, int na, const double* __restrict__ x, double* __restrict__ Z) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    for (int a = 0; a < na; ++a) Z[(int64_t)(p0 + a) * nv + v] = x[3 * v + a];
}

// -----------------------------------------------------------------------------
// Reductions: block partial sums of up to 32 dot products, then one wavefront
// per product adds the block partials in a fixed order (lane b sums blocks
// b, b + 64, ..., then a butterfly), so results do not depend on scheduling.
// -----------------------------------------------------------------------------
constexpr int DOT_BLOCKS = 512;
constexpr int DOT_THREADS = 256;

__global__ void __launch_bounds__(DOT_THREADS) k_dots(int n, DotArgs a, double* __restrict__ partial) {
    __shared__ double sh[DOT_THREADS / WAVE][32];
    double acc[32];
#pragma unroll
    for (int i = 0; i < 32; ++i) acc[i] = 0.0;
    for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < n; idx += gridDim.x * blockDim.x) {
#pragma unroll
        for (int i = 0; i < 32; ++i)
            if (i < a.m) acc[i] += a.a[i][idx] * a.b[i][idx];
    }
    const int lane = threadIdx.x % WAVE, wid = threadIdx.x / WAVE;
#pragma unroll
    for (int i = 0; i < 32; ++i) {
        if (i < a.m) {
            const double s = wave_sum(acc[i]);
            if (lane == 0) sh[wid][i] = s;
        }
    }
    __syncthreads();
    if (threadIdx.x < a.m) {
        double s = 0;
        for (int w = 0; w < DOT_THREADS / WAVE; ++w) s += sh[w][threadIdx.x];
        partial[blockIdx.x * 32 + threadIdx.x] = s;
    }
}

__global__ void __launch_bounds__(WAVE) k_dots_fin(const double* __restrict__ partial, double* __restrict__ out) {
    const int i = blockIdx.x, lane = threadIdx.x;
    double s = 0;
#pragma unroll
    for (int b = lane; b < DOT_BLOCKS; b += WAVE) s += partial[b * 32 + i];
    s = wave_sum(s);
    if (lane == 0) out[i] = s;
}

__global__ void k_axpbypcz(int n, double a, const double* __restrict__ x, double b, const double* __restrict__ y, double c,
                           const double* __restrict__ z, double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double r = a * x[i];
    if (y) r += b * y[i];
    if (z) r += c * z[i];
    out[i] = r;
}

// Anderson history push: with a previous iterate, dQ = x - x_last and dG = g - g_last;
// then x_last = x and g_last = g (each thread reads its entries before it writes them).
__global__ void k_aa_push(int n, const double* __restrict__ x, const double* __restrict__ g, double* __restrict__ x_last,
                          double* __restrict__ g_last, double* __restrict__ dQ, double* __restrict__ dG) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double xi = x[i], gi = g[i];
    if (dQ) {
        dQ[i] = xi - x_last[i];
        dG[i] = gi - g_last[i];
    }
    x_last[i] = xi;
    g_last[i] = gi;
}

struct AAPtrs {
    const double* dQ[8];
    const double* dG[8];
    double gamma[8];
};

__global__ void k_aa_combine(int n, const double* __restrict__ q_prev, const double* __restrict__ g, int m, AAPtrs p,
                             double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double c = q_prev[i] + g[i];
    for (int a = 0; a < m; ++a) c -= (p.dQ[a][i] + p.dG[a][i]) * p.gamma[a];
    out[i] = c;
}

__global__ void k_set_dirichlet(int nd, const int* __restrict__ ddof, const double* __restrict__ dval, double* __restrict__ q) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nd) q[ddof[i]] = dval[i];
}
__global__ void k_zero_dirichlet(int nd, const int* __restrict__ ddof, double* __restrict__ q) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nd) q[ddof[i]] = 0.0;
}
__global__ void k_nan_to_zero(int n, double* __restrict__ x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && !isfinite(x[i])) x[i] = 0.0;
}
// x *= clip / |x| when |x| = sqrt(*norm2) is finite and above clip (the norm
// comes from a device reduction, so no host round trip is needed).
__global__ void k_clip_scale(int n, double clip, const double* __restrict__ norm2, double* __restrict__ x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double nn = sqrt(*norm2);
    if (isfinite(nn) && nn > clip) x[i] *= clip / nn;
}
__global__ void k_scale(int n, double a, double* __restrict__ x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] *= a;
}

// -----------------------------------------------------------------------------
// Contact: obstacle geometry (planar/spherical as in friction/*.cpp, plus the
// heightfield ground and the two-obstacle press).
// -----------------------------------------------------------------------------
__device__ void hf_eval(const DevBoundary& b, double x, double y, double& h, double& hx, double& hy) {
    h = b.z0; hx = 0; hy = 0;
    for (int i = 0; i < b.nwaves; ++i) {
        const double a = b.waves[4 * i], kx = b.waves[4 * i + 1], ky = b.waves[4 * i + 2], ph = b.waves[4 * i + 3];
        const double arg = kx * x + ky * y + ph;
        double sn, cs;
        sincos(arg, &sn, &cs);
        h += a * sn;
        hx += a * kx * cs;
        hy += a * ky * cs;
    }
}

__device__ void boundary_eval(const DevBoundary& b, const double* q, double& dist, double* nn) {
    if (b.type == 1) {
        dist = q[0] * b.n[0] + q[1] * b.n[1] + q[2] * b.n[2] + b.off;
        nn[0] = b.n[0]; nn[1] = b.n[1]; nn[2] = b.n[2];
        return;
    }
    if (b.type == 2) {
        const double d0 = q[0] - b.c[0], d1 = q[1] - b.c[1], d2 = q[2] - b.c[2];
        const double rho = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
        dist = rho - b.r;
        nn[0] = d0 / rho; nn[1] = d1 / rho; nn[2] = d2 / rho;
        return;
    }
    double h, hx, hy;
    hf_eval(b, q[0], q[1], h, hx, hy);
    const double s = sqrt(1 + hx * hx + hy * hy);
    const double dg = (q[2] - h) / s;
    if (b.type == 4) {
        const double dp = b.plane_z - q[2];
        if (dp < dg) { dist = dp; nn[0] = 0; nn[1] = 0; nn[2] = -1; return; }
    }
    dist = dg;
    nn[0] = -hx / s; nn[1] = -hy / s; nn[2] = 1 / s;
}

__global__ void k_contact_frames(int K, DevBoundary b, const int* __restrict__ nodes, const double* __restrict__ q,
                                 double* __restrict__ nrm, double* __restrict__ t1, double* __restrict__ t2) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= K) return;
    const int v = nodes[c];
    double d, nn[3];
    boundary_eval(b, q + 3 * v, d, nn);
    // GetLocalFrame: t1 = argmax_i |n x e_i| normalized, t2 = n x t1 normalized
    double ux[3] = {0, 0, 0}, best = 0;
    for (int i = 0; i < 3; ++i) {
        double e[3] = {0, 0, 0};
        e[i] = 1;
        const double x0 = nn[1] * e[2] - nn[2] * e[1], x1 = nn[2] * e[0] - nn[0] * e[2], x2 = nn[0] * e[1] - nn[1] * e[0];
        const double sq = x0 * x0 + x1 * x1 + x2 * x2;
        if (sq > best) { best = sq; ux[0] = x0; ux[1] = x1; ux[2] = x2; }
    }
    const double un = sqrt(best);
    const double a0 = ux[0] / un, a1 = ux[1] / un, a2 = ux[2] / un;
    double y0 = nn[1] * a2 - nn[2] * a1, y1 = nn[2] * a0 - nn[0] * a2, y2 = nn[0] * a1 - nn[1] * a0;
    const double yn = sqrt(y0 * y0 + y1 * y1 + y2 * y2);
    nrm[3 * c] = nn[0]; nrm[3 * c + 1] = nn[1]; nrm[3 * c + 2] = nn[2];
    t1[3 * c] = a0; t1[3 * c + 1] = a1; t1[3 * c + 2] = a2;
    t2[3 * c] = y0 / yn; t2[3 * c + 1] = y1 / yn; t2[3 * c + 2] = y2 / yn;
}

// Mesh boundary (MeshFrictionalBoundary, mesh_frictional_boundary.cpp): one
// wave per candidate scans the opposite triangle set (lanes stride the faces)
// for the closest point at the current positions q; ties go to the lowest face
// index, as the reference's sequential scan keeps the first minimum.
// FRAMES: frame-start normal (p - closest) / |p - closest| (the face normal
// when p is on the surface), Gram-Schmidt tangents (forward_gpu.cpp:496-509),
// and the candidate filter |p - closest| - radius <= 5 radius (:460-478).
// DISTANCE: (p - closest) . refn - radius (GetSignedDistanceWithReferenceNormal).
template <bool FRAMES>
__global__ void __launch_bounds__(256) k_mesh_query(int K, DevBoundary b, const int* __restrict__ nodes,
                                                    const double* __restrict__ q, const double* __restrict__ refn,
                                                    double* __restrict__ out0, double* __restrict__ t1,
                                                    double* __restrict__ t2, unsigned char* __restrict__ enabled) {
    const int c = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    const int lane = threadIdx.x % WAVE;
    if (c >= K) return;
    const int v = nodes[c];
    const bool right = v < b.split;
    const int f0 = right ? b.nfl : 0, nf = right ? b.nfr : b.nfl;
    const double p[3] = {q[3 * v], q[3 * v + 1], q[3 * v + 2]};
    double best = INFINITY;
    int bf = -1;
    for (int f = lane; f < nf; f += WAVE) {
        const int* fv = b.faces + 3 * (int64_t)(f0 + f);
        double cl[3];
        closest_on_triangle(p, q + 3 * fv[0], q + 3 * fv[1], q + 3 * fv[2], cl);
        const double d2 = (p[0] - cl[0]) * (p[0] - cl[0]) + (p[1] - cl[1]) * (p[1] - cl[1]) + (p[2] - cl[2]) * (p[2] - cl[2]);
        if (d2 < best) { best = d2; bf = f; }
    }
    // lexicographic minimum of (distance^2, face) over the wave
#pragma unroll
    for (int off = WAVE / 2; off > 0; off >>= 1) {
        const double ob = __shfl_xor(best, off, WAVE);
        const int of = __shfl_xor(bf, off, WAVE);
        if (of >= 0 && (bf < 0 || ob < best || (ob == best && of < bf))) { best = ob; bf = of; }
    }
    if (lane != 0) return;
    if (bf < 0) {  // no opposite faces: infinitely far, never a candidate
        if (FRAMES) {
            out0[3 * c] = 0; out0[3 * c + 1] = 0; out0[3 * c + 2] = 1;
            t1[3 * c] = 1; t1[3 * c + 1] = 0; t1[3 * c + 2] = 0;
            t2[3 * c] = 0; t2[3 * c + 1] = 1; t2[3 * c + 2] = 0;
            enabled[c] = 0;
        } else {
            out0[c] = INFINITY;
        }
        return;
    }
    const int* fv = b.faces + 3 * (int64_t)(f0 + bf);
    const double *A = q + 3 * fv[0], *B = q + 3 * fv[1], *C = q + 3 * fv[2];
    double cl[3];
    closest_on_triangle(p, A, B, C, cl);
    const double dp[3] = {p[0] - cl[0], p[1] - cl[1], p[2] - cl[2]};
    if (!FRAMES) {
        out0[c] = dp[0] * refn[3 * c] + dp[1] * refn[3 * c + 1] + dp[2] * refn[3 * c + 2] - b.radius;
        return;
    }
    const double dn = sqrt(dp[0] * dp[0] + dp[1] * dp[1] + dp[2] * dp[2]);
    double n[3];
    if (dn > 1e-12) {
        n[0] = dp[0] / dn; n[1] = dp[1] / dn; n[2] = dp[2] / dn;
    } else {
        const double e1[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]}, e2[3] = {C[0] - A[0], C[1] - A[1], C[2] - A[2]};
        double fn[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        const double fl = sqrt(fn[0] * fn[0] + fn[1] * fn[1] + fn[2] * fn[2]);
        if (fl > DBL_EPSILON) { n[0] = fn[0] / fl; n[1] = fn[1] / fl; n[2] = fn[2] / fl; }
        else { n[0] = 0; n[1] = 0; n[2] = 1; }
    }
    // t1 from the axis least parallel to n (first minimum), t2 = n x t1
    int ax = 0;
    for (int d = 1; d < 3; ++d) if (fabs(n[d]) < fabs(n[ax])) ax = d;
    double u[3] = {-n[ax] * n[0], -n[ax] * n[1], -n[ax] * n[2]};
    u[ax] += 1.0;
    const double ul = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    for (int d = 0; d < 3; ++d) u[d] /= ul;
    double w[3] = {n[1] * u[2] - n[2] * u[1], n[2] * u[0] - n[0] * u[2], n[0] * u[1] - n[1] * u[0]};
    const double wl = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    for (int d = 0; d < 3; ++d) {
        out0[3 * c + d] = n[d];
        t1[3 * c + d] = u[d];
        t2[3 * c + d] = w[d] / wl;
    }
    enabled[c] = (dn - b.radius > 5.0 * b.radius) ? 0 : 1;
}

__global__ void k_contact_distance(int K, DevBoundary b, const int* __restrict__ nodes, const double* __restrict__ q,
                                   double* __restrict__ delta) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= K) return;
    double d, nn[3];
    boundary_eval(b, q + 3 * nodes[c], d, nn);
    delta[c] = d;
}

// out_c = (S_C^T y)_c over the etree path of candidate c. Paths run up to
// the etree depth (~6,000 rows), so each path is cut into pieces of at most
// kContactPiece entries: one wavefront per piece (four index -> value load
// chains in flight per lane), then one thread per candidate adds its pieces.
template <class TV>
__global__ void k_contact_gather(int np, const int* __restrict__ pc_b, const int* __restrict__ pc_e,
                                 const int* __restrict__ cp_row, const TV* __restrict__ cp_val, const TV* __restrict__ y,
                                 int n, double* __restrict__ part) {
    const int pc = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    const int lane = threadIdx.x % WAVE;
    if (pc >= np) return;
    double a0 = 0, a1 = 0, a2 = 0;
    const int end = pc_e[pc];
    int p = pc_b[pc] + lane;
    for (; p + 3 * WAVE < end; p += 4 * WAVE) {
        int k[4];
        TV v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) { k[u] = cp_row[p + u * WAVE]; v[u] = cp_val[p + u * WAVE]; }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const double vv = (double)v[u];
            a0 += vv * (double)y[k[u]];
            a1 += vv * (double)y[n + k[u]];
            a2 += vv * (double)y[2 * n + k[u]];
        }
    }
    for (; p < end; p += WAVE) {
        const int k = cp_row[p];
        const double v = (double)cp_val[p];
        a0 += v * (double)y[k];
        a1 += v * (double)y[n + k];
        a2 += v * (double)y[2 * n + k];
    }
    a0 = wave_sum(a0); a1 = wave_sum(a1); a2 = wave_sum(a2);
    if (lane == 0) { part[3 * pc] = a0; part[3 * pc + 1] = a1; part[3 * pc + 2] = a2; }
}

__global__ void k_contact_gather_fin(int K, const int* __restrict__ piece_ptr, const double* __restrict__ part,
                                     double* __restrict__ out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= K) return;
    double a0 = 0, a1 = 0, a2 = 0;
    for (int q = piece_ptr[c]; q < piece_ptr[c + 1]; ++q) { a0 += part[3 * q]; a1 += part[3 * q + 1]; a2 += part[3 * q + 2]; }
    out[3 * c] = a0; out[3 * c + 1] = a1; out[3 * c + 2] = a2;
}

// y_k += (S_C f)_k for the rows crossed by candidate paths; rows near the
// root cross every candidate, so rows are cut into pieces as above.
template <class TV>
__global__ void k_contact_scatter(int np, const int* __restrict__ pc_b, const int* __restrict__ pc_e,
                                  const int* __restrict__ rc_c, const TV* __restrict__ rc_val, const double* __restrict__ f,
                                  double* __restrict__ part) {
    const int pc = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    const int lane = threadIdx.x % WAVE;
    if (pc >= np) return;
    double a0 = 0, a1 = 0, a2 = 0;
    const int end = pc_e[pc];
    int p = pc_b[pc] + lane;
    for (; p + 3 * WAVE < end; p += 4 * WAVE) {
        int cc[4];
        TV v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) { cc[u] = rc_c[p + u * WAVE]; v[u] = rc_val[p + u * WAVE]; }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const double vv = (double)v[u];
            a0 += vv * f[3 * cc[u]];
            a1 += vv * f[3 * cc[u] + 1];
            a2 += vv * f[3 * cc[u] + 2];
        }
    }
    for (; p < end; p += WAVE) {
        const int c = rc_c[p];
        const double v = (double)rc_val[p];
        a0 += v * f[3 * c];
        a1 += v * f[3 * c + 1];
        a2 += v * f[3 * c + 2];
    }
    a0 = wave_sum(a0); a1 = wave_sum(a1); a2 = wave_sum(a2);
    if (lane == 0) { part[3 * pc] = a0; part[3 * pc + 1] = a1; part[3 * pc + 2] = a2; }
}

template <class TV>
__global__ void k_contact_scatter_fin(int nrows, const int* __restrict__ piece_ptr, const int* __restrict__ rc_row,
                                      const double* __restrict__ part, int n, TV* __restrict__ y) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    double a0 = 0, a1 = 0, a2 = 0;
    for (int q = piece_ptr[r]; q < piece_ptr[r + 1]; ++q) { a0 += part[3 * q]; a1 += part[3 * q + 1]; a2 += part[3 * q + 2]; }
    const int k = rc_row[r];
    y[k] += (TV)a0;
    y[n + k] += (TV)a1;
    y[2 * n + k] += (TV)a2;
}

__global__ void k_add_node_force(int K, const int* __restrict__ nodes, const double* __restrict__ f, double* __restrict__ r) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= K) return;
    const int v = nodes[c];
    r[3 * v] += f[3 * c];
    r[3 * v + 1] += f[3 * c + 1];
    r[3 * v + 2] += f[3 * c + 2];
}

template <class TM>
__global__ void k_assemble_msys(int K, const double* __restrict__ H, const double* __restrict__ nrm,
                                const double* __restrict__ om, const double* __restrict__ E, TM* __restrict__ M) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)K * K) return;
    const int i = (int)(idx / K), j = (int)(idx % K);
    const double nd = nrm[3 * i] * nrm[3 * j] + nrm[3 * i + 1] * nrm[3 * j + 1] + nrm[3 * i + 2] * nrm[3 * j + 2];
    double v = om[i] * nd * H[idx] * om[j];
    if (i == j) v += E[i];
    M[idx] = (TM)v;
}

__global__ void k_h_apply(int K, const double* __restrict__ H, const double* __restrict__ f, double* __restrict__ out) {
    const int c = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    const int lane = threadIdx.x % WAVE;
    if (c >= K) return;
    const double* Hr = H + (int64_t)c * K;
    double a0 = 0, a1 = 0, a2 = 0;
    for (int j = lane; j < K; j += WAVE) {
        const double h = Hr[j];
        a0 += h * f[3 * j];
        a1 += h * f[3 * j + 1];
        a2 += h * f[3 * j + 2];
    }
    a0 = wave_sum(a0); a1 = wave_sum(a1); a2 = wave_sum(a2);
    if (lane == 0) { out[3 * c] = a0; out[3 * c + 1] = a1; out[3 * c + 2] = a2; }
}

// -----------------------------------------------------------------------------
// Backward: per-sample 9x9 proximal-map Jacobians (SetupProjectiveDynamicsLocalStepDifferential,
// deformable_projective_dynamics_backward.cpp:10-80) and their matrix-free application
// (ApplyProjectiveDynamicsLocalStepDifferential, :82-139).
// -----------------------------------------------------------------------------
struct KnPtrs { const double* kn[4]; };

// Spectral filter of a sample's proximal-map block (backward option
// alg_phd_bwd_kpd_filter) on one eigenvalue mu of the block scaled by the
// sample's weight in A (eta = 1 - mu is the eigenvalue of the sample's share of
// the Hessian A - K): mu > 1 moves to 1 + theta (mu - 1) (theta = 0 clamps eta
// at 0, -1 reflects it to |eta|: the filters of ProjectedNewtonMatrix,
// deformable_newton_forward.cpp:408-433 with eps = 0), then mu is raised to lo
// (-rho in the trust region, which keeps A^-1 K's spectrum inside (-rho, 1)).
__device__ __forceinline__ double kpd_filter_eig(double theta, double lo, double mu) {
    return fmax(mu > 1.0 ? 1.0 + theta * (mu - 1.0) : mu, lo);
}

template <int EDIM, class TW>
__global__ void __launch_bounds__(64) k_kpd_build(DevMaterials m, const double* __restrict__ kc, const double* __restrict__ kv,
                                                  KnPtrs kn, const double* __restrict__ q, int filter, double theta,
                                                  double lo, int* __restrict__ nfilt, TW* __restrict__ W) {
    // one thread per (sample, element), elements fastest: the writes of W
    // (element-minor: entry j of sample s at (81 s + j) ne + e) are coalesced
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (int64_t)m.ne * m.samples) return;
    const int e = (int)(gid % m.ne), s = (int)(gid / m.ne);
    const double* G = m.G + (int64_t)e * m.g_stride + s * EDIM * 3;
    double F[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) F[i] = 0.0;
    for (int k = 0; k < EDIM; ++k) {
        const int v = m.elems[k * m.ne + e];
#pragma unroll
        for (int jj = 0; jj < 3; ++jj) {
            const double g = G[k * 3 + jj];
            F[0 + 3 * jj] += q[3 * v] * g;
            F[1 + 3 * jj] += q[3 * v + 1] * g;
            F[2 + 3 * jj] += q[3 * v + 2] * g;
        }
    }
    double U[9], sg[3], V[9];
    svd3(F, U, sg, V);
    const double wc = kc ? kc[e] : 0.0, wv = kv ? kv[e] : 0.0;
    // muscle block w (a / l)(I - n n^T) dF m  m^T with n = F m / l (ProjectToManifoldDifferential,
    // pd_muscle_energy.cpp:127-165; backward.cpp:54-80), added outside the spectral filter
    double mc = 0.0, mn[3] = {0.0, 0.0, 0.0}, mf[3] = {0.0, 0.0, 0.0};
    if (m.mw && m.mw[e] != 0.0) {
        mf[0] = m.mfib[e]; mf[1] = m.mfib[m.ne + e]; mf[2] = m.mfib[2 * m.ne + e];
        double Fm[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) Fm[i] = F[i] * mf[0] + F[i + 3] * mf[1] + F[i + 6] * mf[2];
        const double l = sqrt(Fm[0] * Fm[0] + Fm[1] * Fm[1] + Fm[2] * Fm[2]);
        if (l > DBL_EPSILON) {
            mc = m.mw[e] * m.mact[e] / l;
            mn[0] = Fm[0] / l; mn[1] = Fm[1] / l; mn[2] = Fm[2] / l;
        }
    }
    // column c = p + 3 q: entry i + 3 j is mc (d_ip - n_i n_p) m_q m_j
    auto muscle_col = [&](int c, double* col) {
        if (mc == 0.0) return;
        const int p = c % 3, qq = c / 3;
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int i = 0; i < 3; ++i) col[i + 3 * j] += mc * ((i == p ? 1.0 : 0.0) - mn[i] * mn[p]) * mf[qq] * mf[j];
    };
    double R[9], Sp[9], D[3], ssn[4][3];
    const double one[3] = {1.0, 1.0, 1.0};
    udv(U, one, V, R);
    udv(V, sg, V, Sp);
    if (wv != 0.0) volume_D(sg, D);
    for (int h = 0; h < m.n_nh; ++h)
        if (kn.kn[h] && kn.kn[h][e] != 0.0) nh_prox(m.nh_mu[h], m.nh_la[h], m.nh_k[h], sg, ssn[h]);
    // col = sum_energies k_e dBp/dF [dF]
    auto op = [&](const double* dF, double* col) {
#pragma unroll
        for (int i = 0; i < 9; ++i) col[i] = 0.0;
        if (wc != 0.0) {
            double dR[9];
            dR_from_dF(R, Sp, dF, dR);
#pragma unroll
            for (int i = 0; i < 9; ++i) col[i] += wc * dR[i];
        }
        if (wv != 0.0) {
            double dP[9];
            volume_diff(U, sg, V, D, dF, dP);
#pragma unroll
            for (int i = 0; i < 9; ++i) col[i] += wv * dP[i];
        }
        for (int h = 0; h < m.n_nh; ++h) {
            if (!kn.kn[h] || kn.kn[h][e] == 0.0) continue;
            double dP[9];
            nh_diff(m.nh_mu[h], m.nh_la[h], m.nh_k[h], U, sg, V, ssn[h], dF, dP);
            const double w = kn.kn[h][e];
#pragma unroll
            for (int i = 0; i < 9; ++i) col[i] += w * dP[i];
        }
    };
    TW* out = W + (int64_t)s * 81 * m.ne + e;
    if (filter != 0) {
        // The proximal maps are isotropic, so in the basis E_kl = u_k v_l^T of
        // the SVD of F the block splits into a 3x3 stretch block (E_00, E_11,
        // E_22) and three 2x2 blocks (E_kl, E_lk), k < l: nine applications of
        // the differential give all of them, and their eigenvalues are those of
        // the 9x9 block.
        double aw = 0.0;
        if (m.wc) aw += m.wc[e];
        if (m.wv) aw += m.wv[e];
        for (int h = 0; h < m.n_nh; ++h) if (m.wn[h]) aw += m.wn[h][e];
        if (aw > 0.0) {
            double Bs[9], P[3][3];  // stretch block (column-major); pair blocks a, b, c of [[a, b], [b, c]]
            double dF[9], col[9], t[9];
            for (int i = 0; i < 3; ++i) {  // stretch: input E_ii
#pragma unroll
                for (int a = 0; a < 3; ++a)
#pragma unroll
                    for (int b = 0; b < 3; ++b) dF[a + 3 * b] = U[a + 3 * i] * V[b + 3 * i];
                op(dF, col);
                double ct[9];
                mat_tmul(U, col, ct);  // U^T col
                mat_mul(ct, V, t);     // U^T col V
#pragma unroll
                for (int k = 0; k < 3; ++k) Bs[k + 3 * i] = t[4 * k];
            }
            const int pk[3] = {0, 0, 1}, pl[3] = {1, 2, 2};
            for (int p = 0; p < 3; ++p) {
                const int k = pk[p], l = pl[p];
                double x[2][2];  // x[out][in]: out/in 0 = E_kl, 1 = E_lk
                for (int in = 0; in < 2; ++in) {
                    const int ki = in == 0 ? k : l, li = in == 0 ? l : k;
#pragma unroll
                    for (int a = 0; a < 3; ++a)
#pragma unroll
                        for (int b = 0; b < 3; ++b) dF[a + 3 * b] = U[a + 3 * ki] * V[b + 3 * li];
                    op(dF, col);
                    double ct[9];
                    mat_tmul(U, col, ct);
                    mat_mul(ct, V, t);
                    x[0][in] = t[k + 3 * l];
                    x[1][in] = t[l + 3 * k];
                }
                P[p][0] = x[0][0] / aw;
                P[p][1] = 0.5 * (x[0][1] + x[1][0]) / aw;
                P[p][2] = x[1][1] / aw;
            }
            double S3[9], Z[9];
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) S3[i + 3 * j] = 0.5 * (Bs[i + 3 * j] + Bs[j + 3 * i]) / aw;
            sym3_eigen(S3, Z);
            double shift = 0.0;
            bool changed = false;
            double ls[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const double mu = S3[4 * k];
                ls[k] = kpd_filter_eig(theta, lo, mu);
                changed |= ls[k] != mu;
                shift = fmax(shift, fabs(ls[k] - mu));
            }
            double pl0[3], pl1[3], pcs[3], psn[3];
            for (int p = 0; p < 3; ++p) {
                double l0, l1;
                sym2_eigen(P[p][0], P[p][1], P[p][2], l0, l1, pcs[p], psn[p]);
                pl0[p] = kpd_filter_eig(theta, lo, l0);
                pl1[p] = kpd_filter_eig(theta, lo, l1);
                changed |= (pl0[p] != l0) | (pl1[p] != l1);
                shift = fmax(shift, fmax(fabs(pl0[p] - l0), fabs(pl1[p] - l1)));
            }
            if (changed) {
                if (shift > 1e-9 && nfilt) atomicAdd(nfilt, 1);
                // filtered blocks B' (times aw), then column c of the 9x9 block is
                // U B'(U^T E_c V) V^T with E_c the unit matrix of entry c
                double Bf[9];
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 3; ++j)
                        Bf[i + 3 * j] = aw * (Z[i] * ls[0] * Z[j] + Z[i + 3] * ls[1] * Z[j + 3] + Z[i + 6] * ls[2] * Z[j + 6]);
                double Pf[3][3];
                for (int p = 0; p < 3; ++p) {
                    const double c = pcs[p], sn = psn[p];
                    // eigenvectors (c, -sn) for l0 and (sn, c) for l1
                    Pf[p][0] = aw * (c * c * pl0[p] + sn * sn * pl1[p]);
                    Pf[p][1] = aw * (-c * sn * pl0[p] + sn * c * pl1[p]);
                    Pf[p][2] = aw * (sn * sn * pl0[p] + c * c * pl1[p]);
                }
                for (int c = 0; c < 9; ++c) {
                    const int ac = c % 3, bc = c / 3;
                    double T[9], S[9];
#pragma unroll
                    for (int k = 0; k < 3; ++k)
#pragma unroll
                        for (int l = 0; l < 3; ++l) T[k + 3 * l] = U[ac + 3 * k] * V[bc + 3 * l];
#pragma unroll
                    for (int k = 0; k < 3; ++k) S[4 * k] = Bf[k] * T[0] + Bf[k + 3] * T[4] + Bf[k + 6] * T[8];
                    for (int p = 0; p < 3; ++p) {
                        const int k = pk[p], l = pl[p];
                        const double tkl = T[k + 3 * l], tlk = T[l + 3 * k];
                        S[k + 3 * l] = Pf[p][0] * tkl + Pf[p][1] * tlk;
                        S[l + 3 * k] = Pf[p][1] * tkl + Pf[p][2] * tlk;
                    }
                    double US[9], cl[9];
                    mat_mul(U, S, US);
#pragma unroll
                    for (int i = 0; i < 3; ++i)
#pragma unroll
                        for (int j = 0; j < 3; ++j)
                            cl[i + 3 * j] = US[i] * V[j] + US[i + 3] * V[j + 3] + US[i + 6] * V[j + 6];
                    muscle_col(c, cl);
#pragma unroll
                    for (int i = 0; i < 9; ++i) out[(int64_t)(i + 9 * c) * m.ne] = (TW)cl[i];
                }
                return;
            }
        }
    }
    for (int c = 0; c < 9; ++c) {
        double dF[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) dF[i] = (i == c) ? 1.0 : 0.0;
        double col[9];
        op(dF, col);
        muscle_col(c, col);
#pragma unroll
        for (int i = 0; i < 9; ++i) out[(int64_t)(i + 9 * c) * m.ne] = (TW)col[i];
    }
}

template <int EDIM, class TW>
__global__ void __launch_bounds__(128) k_kpd_apply(DevMaterials m, const TW* __restrict__ W, const double* __restrict__ dq,
                                                   double* __restrict__ elem_out) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m.ne) return;
    double x[EDIM][3], acc[EDIM][3];
#pragma unroll
    for (int k = 0; k < EDIM; ++k) {
        const int v = m.elems[k * m.ne + e];
        x[k][0] = dq[3 * v]; x[k][1] = dq[3 * v + 1]; x[k][2] = dq[3 * v + 2];
        acc[k][0] = acc[k][1] = acc[k][2] = 0.0;
    }
    const double* Ge = m.G + (int64_t)e * m.g_stride;
    for (int s = 0; s < m.samples; ++s) {
        const double* G = Ge + s * EDIM * 3;
        double dF[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) dF[i] = 0.0;
#pragma unroll
        for (int k = 0; k < EDIM; ++k)
#pragma unroll
            for (int jj = 0; jj < 3; ++jj) {
                const double g = G[k * 3 + jj];
                dF[0 + 3 * jj] += x[k][0] * g;
                dF[1 + 3 * jj] += x[k][1] * g;
                dF[2 + 3 * jj] += x[k][2] * g;
            }
        const TW* Ws = W + (int64_t)s * 81 * m.ne + e;  // element-minor: coalesced across the wave
        double y[9];
#pragma unroll
        for (int r = 0; r < 9; ++r) {
            double a = 0;
#pragma unroll
            for (int c = 0; c < 9; ++c) a += (double)Ws[(int64_t)(r + 9 * c) * m.ne] * dF[c];
            y[r] = a;
        }
#pragma unroll
        for (int k = 0; k < EDIM; ++k)
#pragma unroll
            for (int i = 0; i < 3; ++i) acc[k][i] += y[i] * G[k * 3] + y[i + 3] * G[k * 3 + 1] + y[i + 6] * G[k * 3 + 2];
    }
    double* o = elem_out + (int64_t)e * EDIM * 3;
#pragma unroll
    for (int k = 0; k < EDIM; ++k) { o[3 * k] = acc[k][0]; o[3 * k + 1] = acc[k][1]; o[3 * k + 2] = acc[k][2]; }
}

// dl/d(mat_w) per merged type (PdEnergyForceDifferential require_dw,
// deformable_pd_energy.cpp:368-403 with StressTensor/stiffness = F - p for the
// PD energies and the true Neo-Hookean P / (2 mu_mean) for pd_neohookean).
template <int EDIM>
__global__ void __launch_bounds__(128) k_dmatw(DevMaterials m, const double* __restrict__ q, const double* __restrict__ z,
                                               double vs, double* __restrict__ g_cor, double* __restrict__ g_vol,
                                               double* __restrict__ g_nh) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m.ne) return;
    double x[EDIM][3], zz[EDIM][3];
#pragma unroll
    for (int k = 0; k < EDIM; ++k) {
        const int v = m.elems[k * m.ne + e];
        x[k][0] = q[3 * v]; x[k][1] = q[3 * v + 1]; x[k][2] = q[3 * v + 2];
        zz[k][0] = z[3 * v]; zz[k][1] = z[3 * v + 1]; zz[k][2] = z[3 * v + 2];
    }
    double sc = 0, sv = 0, sn[4] = {0, 0, 0, 0};
    const double* Ge = m.G + (int64_t)e * m.g_stride;
    for (int s = 0; s < m.samples; ++s) {
        const double* G = Ge + s * EDIM * 3;
        double F[9], ZG[9];  // ZG = sum_k z_k G_k^T: z_e^T A^T vec(P) = <ZG, P>
#pragma unroll
        for (int i = 0; i < 9; ++i) { F[i] = 0.0; ZG[i] = 0.0; }
#pragma unroll
        for (int k = 0; k < EDIM; ++k)
#pragma unroll
            for (int jj = 0; jj < 3; ++jj) {
                const double g = G[k * 3 + jj];
#pragma unroll
                for (int i = 0; i < 3; ++i) { F[i + 3 * jj] += x[k][i] * g; ZG[i + 3 * jj] += zz[k][i] * g; }
            }
        double U[9], sg[3], V[9];
        svd3(F, U, sg, V);
        if (g_cor) {
            const double one[3] = {1, 1, 1};
            double R[9];
            udv(U, one, V, R);
            double a = 0;
#pragma unroll
            for (int i = 0; i < 9; ++i) a += ZG[i] * (F[i] - R[i]);
            sc += a;
        }
        if (g_vol) {
            double D[3], sd[3], P[9];
            volume_D(sg, D);
            sd[0] = sg[0] + D[0]; sd[1] = sg[1] + D[1]; sd[2] = sg[2] + D[2];
            udv(U, sd, V, P);
            double a = 0;
#pragma unroll
            for (int i = 0; i < 9; ++i) a += ZG[i] * (F[i] - P[i]);
            sv += a;
        }
        if (g_nh && m.n_nh > 0) {
            const double J = F[0] * (F[4] * F[8] - F[7] * F[5]) - F[3] * (F[1] * F[8] - F[7] * F[2]) +
                             F[6] * (F[1] * F[5] - F[4] * F[2]);
            const double lj = log(fmax(J, 1e-12));
            // F^-T = cof(F) / det
            double FiT[9];
            FiT[0] = (F[4] * F[8] - F[5] * F[7]) / J;
            FiT[1] = (F[5] * F[6] - F[3] * F[8]) / J;
            FiT[2] = (F[3] * F[7] - F[4] * F[6]) / J;
            FiT[3] = (F[2] * F[7] - F[1] * F[8]) / J;
            FiT[4] = (F[0] * F[8] - F[2] * F[6]) / J;
            FiT[5] = (F[1] * F[6] - F[0] * F[7]) / J;
            FiT[6] = (F[1] * F[5] - F[2] * F[4]) / J;
            FiT[7] = (F[2] * F[3] - F[0] * F[5]) / J;
            FiT[8] = (F[0] * F[4] - F[1] * F[3]) / J;
            for (int h = 0; h < m.n_nh; ++h) {
                double a = 0;
#pragma unroll
                for (int i = 0; i < 9; ++i) a += ZG[i] * (m.nh_mu[h] * (F[i] - FiT[i]) + m.nh_la[h] * lj * FiT[i]);
                sn[h] += a / m.nh_k[h];
            }
        }
    }
    if (g_cor) g_cor[e] = -vs * sc;
    if (g_vol) g_vol[e] = -vs * sv;
    if (g_nh)
        for (int h = 0; h < m.n_nh; ++h) g_nh[(int64_t)h * m.ne + e] = -vs * sn[h];
}

// Gathers factor values V[base[row] + col] (setup of the contact paths).
__global__ void k_gather_vals(int64_t m, const int* __restrict__ rows, const int* __restrict__ cols,
                              const int64_t* __restrict__ base, const double* __restrict__ V64, double* __restrict__ o64,
                              float* __restrict__ o32) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const double v = V64[base[rows[i]] + cols[i]];
    o64[i] = v;
    o32[i] = (float)v;
}
// Dense S_C[ridx][c] = value (row-major, K columns) for the contact Gram H = S_C^T S_C.
__global__ void k_scatter_dense(int64_t m, const int* __restrict__ ridx, const int* __restrict__ cidx,
                                const double* __restrict__ val, int K, double* __restrict__ D) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    D[(int64_t)ridx[i] * K + cidx[i]] = val[i];
}
__global__ void k_symmetrize_lower(int K, double* __restrict__ H) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)K * K) return;
    const int i = (int)(idx / K), j = (int)(idx % K);
    // column-major lower triangle (i >= j in column-major = H[j*K + i]) mirrored to the upper part
    if (i < j) H[(int64_t)j * K + i] = H[(int64_t)i * K + j];
}

__global__ void k_gather_nodes(int K, const int* __restrict__ nodes, const double* __restrict__ q, double* __restrict__ out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= K) return;
    const int v = nodes[c];
    out[3 * c] = q[3 * v];
    out[3 * c + 1] = q[3 * v + 1];
    out[3 * c + 2] = q[3 * v + 2];
}
// G[a][b] = (d_a . d_b) H[c_a][c_b]: Gram of node-local directions under A^-1
// (computed in FP64, stored as TG).
template <class TG>
__global__ void k_assemble_g(int nb, const int* __restrict__ cidx, const double* __restrict__ dirs,
                             const double* __restrict__ H, int K, TG* __restrict__ G) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)nb * nb) return;
    const int a = (int)(idx / nb), b = (int)(idx % nb);
    const double dd = dirs[3 * a] * dirs[3 * b] + dirs[3 * a + 1] * dirs[3 * b + 1] + dirs[3 * a + 2] * dirs[3 * b + 2];
    G[idx] = (TG)(dd * H[(int64_t)cidx[a] * K + cidx[b]]);
}

// ub[a] = d_a . u[c_a]: the candidate vectors projected on the basis rows,
// J_basis u (backward_gpu.cpp:998-1024 projects the same way, on the host).
__global__ void k_basis_project(int nb, const int* __restrict__ cidx, const double* __restrict__ dirs,
                                const double* __restrict__ u, const double* __restrict__ u2, double* __restrict__ ub) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nb) return;
    const int c = cidx[a];
    double x = u[3 * c], y = u[3 * c + 1], z = u[3 * c + 2];
    if (u2) { x += u2[3 * c]; y += u2[3 * c + 1]; z += u2[3 * c + 2]; }
    ub[a] = dirs[3 * a] * x + dirs[3 * a + 1] * y + dirs[3 * a + 2] * z;
}

// F[3c + d] = -sum_a g[a] d_a[d] over the basis rows a of candidate c, in
// increasing a (clist holds them grouped by candidate, cptr the ranges):
// -J_basis^T g, the correction of backward_gpu.cpp:924.
__global__ void k_basis_expand(int K, const int* __restrict__ cptr, const int* __restrict__ clist,
                               const double* __restrict__ dirs, const double* __restrict__ g, double* __restrict__ F) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= K) return;
    double f0 = 0.0, f1 = 0.0, f2 = 0.0;
    for (int p = cptr[c]; p < cptr[c + 1]; ++p) {
        const int a = clist[p];
        f0 -= g[a] * dirs[3 * a];
        f1 -= g[a] * dirs[3 * a + 1];
        f2 -= g[a] * dirs[3 * a + 2];
    }
    F[3 * c] = f0;
    F[3 * c + 1] = f1;
    F[3 * c + 2] = f2;
}

// out[3a + d] = sum_b H[rows[a]][cols[b]] f[3 cols[b] + d]: the candidate block
// of A^-1 applied to impulses supported on `cols`. A workgroup of kHWaves
// wavefronts takes kHRows rows: wave w walks the columns b = w WAVE + lane,
// stepping kHWaves WAVE, and loads each column index and f_c once for all
// kHRows rows; kHCols columns are in flight per step, with the next step's
// indices loaded one step ahead. The waves' partial sums are added in wave
// order through LDS, so the result does not depend on scheduling.
constexpr int kHRows = 4;
constexpr int kHWaves = 4;
constexpr int kHCols = 4;
template <class TH>
__device__ __forceinline__ void h_apply_rows(int a0, int nr, const int* __restrict__ rows, int nc,
                                             const int* __restrict__ cols, const TH* __restrict__ H, int K,
                                             const double* __restrict__ f, double* __restrict__ out) {
    __shared__ double part[kHWaves][kHRows * 3];
    const int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE;
    constexpr int stride = kHWaves * WAVE;
    const TH* Hr[kHRows];
#pragma unroll
    for (int r = 0; r < kHRows; ++r) Hr[r] = H + (int64_t)rows[min(a0 + r, nr - 1)] * K;  // rows past nr: not written
    double acc[kHRows][3];
#pragma unroll
    for (int r = 0; r < kHRows; ++r) acc[r][0] = acc[r][1] = acc[r][2] = 0.0;
    int b = w * WAVE + lane;
    int cn[kHCols];  // the next step's column indices
#pragma unroll
    for (int u = 0; u < kHCols; ++u) cn[u] = b + u * stride < nc ? cols[b + u * stride] : 0;
    for (; b + (kHCols - 1) * stride < nc; b += kHCols * stride) {
        int c[kHCols];
#pragma unroll
        for (int u = 0; u < kHCols; ++u) {
            c[u] = cn[u];
            const int bn = b + (kHCols + u) * stride;
            cn[u] = bn < nc ? cols[bn] : 0;
        }
        double fv[kHCols][3];
        TH h[kHCols][kHRows];
#pragma unroll
        for (int u = 0; u < kHCols; ++u) {
            fv[u][0] = f[3 * c[u]]; fv[u][1] = f[3 * c[u] + 1]; fv[u][2] = f[3 * c[u] + 2];
#pragma unroll
            for (int r = 0; r < kHRows; ++r) h[u][r] = Hr[r][c[u]];
        }
        __builtin_amdgcn_sched_barrier(0);  // every index, H and f load is in flight before the first product
#pragma unroll
        for (int u = 0; u < kHCols; ++u)
#pragma unroll
            for (int r = 0; r < kHRows; ++r) {
                const double hv = (double)h[u][r];
                acc[r][0] += hv * fv[u][0]; acc[r][1] += hv * fv[u][1]; acc[r][2] += hv * fv[u][2];
            }
    }
    for (; b < nc; b += stride) {
        const int c0 = cols[b];
        const double f00 = f[3 * c0], f01 = f[3 * c0 + 1], f02 = f[3 * c0 + 2];
#pragma unroll
        for (int r = 0; r < kHRows; ++r) {
            const double hv = (double)Hr[r][c0];
            acc[r][0] += hv * f00; acc[r][1] += hv * f01; acc[r][2] += hv * f02;
        }
    }
#pragma unroll
    for (int r = 0; r < kHRows; ++r)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double v = wave_sum(acc[r][d]);
            if (lane == 0) part[w][3 * r + d] = v;
        }
    __syncthreads();
    if (threadIdx.x < kHRows * 3 && a0 + (int)threadIdx.x / 3 < nr) {
        double v = part[0][threadIdx.x];
#pragma unroll
        for (int q = 1; q < kHWaves; ++q) v += part[q][threadIdx.x];
        out[3 * a0 + threadIdx.x] = v;
    }
}
template <class TH>
__global__ __launch_bounds__(kHWaves * WAVE) void k_h_apply_idx(int nr, const int* __restrict__ rows, int nc,
                                                                const int* __restrict__ cols, const TH* __restrict__ H,
                                                                int K, const double* __restrict__ f,
                                                                double* __restrict__ out) {
    const int a0 = blockIdx.x * kHRows;
    if (a0 >= nr) return;  // uniform over the workgroup
    h_apply_rows<TH>(a0, nr, rows, nc, cols, H, K, f, out);
}
// Compressed contact system on the index set J: M[a][b] = om_i (n_i . n_j) H[i][j] om_j + (a == b) E_i.
template <class TM>
__global__ void k_assemble_msys_idx(int nJ, const int* __restrict__ J, const double* __restrict__ H, int K,
                                    const double* __restrict__ nrm, const double* __restrict__ om,
                                    const double* __restrict__ E, TM* __restrict__ M) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)nJ * nJ) return;
    const int a = (int)(idx / nJ), b = (int)(idx % nJ);
    const int i = J[a], j = J[b];
    const double nd = nrm[3 * i] * nrm[3 * j] + nrm[3 * i + 1] * nrm[3 * j + 1] + nrm[3 * i + 2] * nrm[3 * j + 2];
    double v = om[i] * nd * H[(int64_t)i * K + j] * om[j];
    if (a == b) v += E[i];
    M[idx] = (TM)v;
}

// Plain streaming copy (calibration of the HBM byte counters).
// Four 16-byte loads in flight per lane, one pass over the data.
__global__ void __launch_bounds__(256) k_copy(int64_t n, const float4* __restrict__ a, float4* __restrict__ b) {
    const int64_t i0 = (int64_t)blockIdx.x * blockDim.x * 4 + threadIdx.x;
    float4 v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int64_t i = i0 + (int64_t)q * blockDim.x;
        if (i < n) v[q] = a[i];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int64_t i = i0 + (int64_t)q * blockDim.x;
        if (i < n) b[i] = v[q];
    }
}

}  // namespace

// ============================================================================
// launchers
// ============================================================================
void launch_expand_s(hipStream_t s, int n, int t_begin, int t_end, const int64_t* Lp, const int* Li, const double* Lx,
                     const int* parent, const int* depth, const int* fd, const int64_t* base, const unsigned char* chain,
                     double* ws, int64_t ws_stride, double* V64, float* V32) {
    if (t_end <= t_begin) return;
    hipLaunchKernelGGL(k_expand_s, dim3(t_end - t_begin), dim3(256), 0, s, n, t_begin, Lp, Li, Lx, parent, depth, fd, base,
                       chain, ws, ws_stride, V64, V32);
}

template <class TV>
void launch_spmv_rows_grp(hipStream_t s, int nunits, const int* gu_group, const int* gu_chunk, const int* fd,
                          const int64_t* base, const int* rbase, const TV* V, const TV* w, int n, TV* pslot, int chunk_w) {
    if (nunits == 0) return;
    switch (chunk_w) {
    case 512:
        hipLaunchKernelGGL((k_spmv_rows_grp<TV, 512>), dim3(nunits), dim3(256), 0, s, nunits, gu_group, gu_chunk, fd, base,
                           rbase, V, w, n, pslot);
        break;
    case 1024:
        hipLaunchKernelGGL((k_spmv_rows_grp<TV, 1024>), dim3(nunits), dim3(256), 0, s, nunits, gu_group, gu_chunk, fd, base,
                           rbase, V, w, n, pslot);
        break;
    case 2048:
        hipLaunchKernelGGL((k_spmv_rows_grp<TV, 2048>), dim3(nunits), dim3(256), 0, s, nunits, gu_group, gu_chunk, fd, base,
                           rbase, V, w, n, pslot);
        break;
    default:
        throw std::invalid_argument("row-group chunk width must be 512, 1024 or 2048");
    }
}
template <class TV>
void launch_spmv_rows_slots(hipStream_t s, int n, const int* rbase, const TV* pslot, TV* y) {
    hipLaunchKernelGGL(k_spmv_rows_slots<TV>, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, rbase, pslot, y);
}
template <class TV>
void launch_spmv_cols(hipStream_t s, int nunits, const int* cu_tile, const int* cu_i0, const int* cu_i1, const int* items,
                      const int* fd, const int64_t* base, const TV* V, const TV* y, int n, TV* xpart) {
    if (nunits == 0) return;
    hipLaunchKernelGGL(k_spmv_cols<TV>, dim3(blocks_for(nunits, 4)), dim3(256), 0, s, nunits, cu_tile, cu_i0, cu_i1,
                       items, fd, base, V, y, n, xpart);
}
template <class TV>
void launch_spmv_cols_fin(hipStream_t s, int n, const int* tile_unit_ptr, const TV* xpart, const int* perm,
                          const double* add, const unsigned char* dmask, const double* dval, double* out) {
    hipLaunchKernelGGL(k_spmv_cols_fin<TV>, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, tile_unit_ptr, xpart, perm, add,
                       dmask, dval, out);
}
template <class TV>
void launch_perm_in(hipStream_t s, int n, const int* perm, const double* r, const unsigned char* dmask, TV* w) {
    hipLaunchKernelGGL(k_perm_in<TV>, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, perm, r, dmask, w);
}

#define PHD_INST_TV(TV)                                                                                                  \
    template void launch_spmv_rows_grp<TV>(hipStream_t, int, const int*, const int*, const int*, const int64_t*,       \
                                           const int*, const TV*, const TV*, int, TV*, int);                           \
    template void launch_spmv_rows_slots<TV>(hipStream_t, int, const int*, const TV*, TV*);                          \
    template void launch_spmv_cols<TV>(hipStream_t, int, const int*, const int*, const int*, const int*, const int*,    \
                                       const int64_t*, const TV*, const TV*, int, TV*);                                  \
    template void launch_spmv_cols_fin<TV>(hipStream_t, int, const int*, const TV*, const int*, const double*,          \
                                           const unsigned char*, const double*, double*);                                \
    template void launch_perm_in<TV>(hipStream_t, int, const int*, const double*, const unsigned char*, TV*);           \
    template void launch_contact_gather<TV>(hipStream_t, const ContactPieces&, const int*, const TV*, const TV*, int,      \
                                            double*, double*);                                                         \
    template void launch_contact_scatter<TV>(hipStream_t, const ContactPieces&, const int*, const int*, const TV*,          \
                                             const double*, double*, int, TV*);

void launch_local_step(hipStream_t s, const DevMaterials& m, const double* q, double* elem_out, int block, bool fp32) {
    if (block != 64 && block != 128 && block != 256) throw std::invalid_argument("local-step block must be 64, 128 or 256");
    const dim3 g(blocks_for(m.ne, block)), b(block);
    if (m.edim == 8) {
        if (fp32) hipLaunchKernelGGL((k_local_step<8, float>), g, b, 0, s, m, q, elem_out);
        else hipLaunchKernelGGL((k_local_step<8, double>), g, b, 0, s, m, q, elem_out);
    } else {
        if (fp32) hipLaunchKernelGGL((k_local_step<4, float>), g, b, 0, s, m, q, elem_out);
        else hipLaunchKernelGGL((k_local_step<4, double>), g, b, 0, s, m, q, elem_out);
    }
}
void launch_gather_rhs(hipStream_t s, int nv, const int* inc_ptr, const int* inc_idx, const double* elem_out,
                       const double* c_dir, const double* q_tilde, double inv_h2m, double* pd_rhs, double* b) {
    hipLaunchKernelGGL(k_gather_rhs, dim3(blocks_for(nv, 256)), dim3(256), 0, s, nv, inc_ptr, inc_idx, elem_out, c_dir,
                       q_tilde, inv_h2m, pd_rhs, b);
}
void launch_gather_vertex(hipStream_t s, int nv, const int* inc_ptr, const int* inc_idx, const double* elem_out, double* out) {
    hipLaunchKernelGGL(k_gather_vertex, dim3(blocks_for(nv, 256)), dim3(256), 0, s, nv, inc_ptr, inc_idx, elem_out, out);
}
void launch_pdir_coef(hipStream_t s, int np, const PdirAxes& ax, const int* map, const int* fv, const double* ginv,
                      const double* y, const double* t, double* coef) {
    if (np == 0) return;
    hipLaunchKernelGGL(k_pdir_coef, dim3(blocks_for(np, 256), 3), dim3(256), 0, s, np, ax, map, fv, ginv, y, t, coef);
}
void launch_pdir_apply(hipStream_t s, int nv, int np, const double* Z, const double* coef, const unsigned char* pmask,
                       const double* o, double* y) {
    if (np == 0) return;
    hipLaunchKernelGGL(k_pdir_apply, dim3(blocks_for(nv, 256)), dim3(256), 0, s, nv, np, Z, coef, pmask, o, y);
}
void launch_pdir_gather(hipStream_t s, int np, int nv, const int* pv, const double* Z, double* out) {
    const int64_t m = (int64_t)np * np;
    hipLaunchKernelGGL(k_pdir_gather, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, np, nv, pv, Z, out);
}
void launch_pdir_columns(hipStream_t s, int nv, int p0, int na, const double* x, double* Z) {
    hipLaunchKernelGGL(k_pdir_columns, dim3(blocks_for(nv, 256)), dim3(256), 0, s, nv, p0, na, x, Z);
}
void launch_residual(hipStream_t s, int nv, int width, const int* Ai, const double* Ax, const double* b, const double* q,
                     const unsigned char* dmask, double* r, const unsigned char* pmask) {
    hipLaunchKernelGGL(k_residual, dim3(blocks_for(nv, 256)), dim3(256), 0, s, nv, width, Ai, Ax, b, q, dmask, pmask, r);
}
int dots_partial_size() { return DOT_BLOCKS * 32; }
void launch_dots(hipStream_t s, int n, const DotArgs& args, double* partial, double* out) {
    hipLaunchKernelGGL(k_dots, dim3(DOT_BLOCKS), dim3(DOT_THREADS), 0, s, n, args, partial);
    hipLaunchKernelGGL(k_dots_fin, dim3(args.m), dim3(WAVE), 0, s, (const double*)partial, out);
}
void launch_axpbypcz(hipStream_t s, int n, double a, const double* x, double b, const double* y, double c,
                     const double* z, double* out) {
    hipLaunchKernelGGL(k_axpbypcz, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, a, x, b, y, c, z, out);
}
void launch_aa_push(hipStream_t s, int n, const double* x, const double* g, double* x_last, double* g_last, double* dQ,
                    double* dG) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_aa_push, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, x, g, x_last, g_last, dQ, dG);
}
void launch_aa_combine(hipStream_t s, int n, const double* q_prev, const double* g, int m, const double* const* dQ,
                       const double* const* dG, const double* gamma, double* out) {
    AAPtrs p;
    for (int a = 0; a < m && a < 8; ++a) { p.dQ[a] = dQ[a]; p.dG[a] = dG[a]; p.gamma[a] = gamma[a]; }
    hipLaunchKernelGGL(k_aa_combine, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, q_prev, g, m, p, out);
}
void launch_set_dirichlet(hipStream_t s, int nd, const int* ddof, const double* dval, double* q) {
    if (nd == 0) return;
    hipLaunchKernelGGL(k_set_dirichlet, dim3(blocks_for(nd, 256)), dim3(256), 0, s, nd, ddof, dval, q);
}
void launch_zero_dirichlet(hipStream_t s, int nd, const int* ddof, double* q) {
    if (nd == 0) return;
    hipLaunchKernelGGL(k_zero_dirichlet, dim3(blocks_for(nd, 256)), dim3(256), 0, s, nd, ddof, q);
}
void launch_nan_to_zero(hipStream_t s, int n, double* x) {
    hipLaunchKernelGGL(k_nan_to_zero, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, x);
}
void launch_clip_scale(hipStream_t s, int n, double clip, const double* norm2, double* x) {
    hipLaunchKernelGGL(k_clip_scale, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, clip, norm2, x);
}
void launch_scale(hipStream_t s, int n, double a, double* x) {
    hipLaunchKernelGGL(k_scale, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, a, x);
}
void launch_contact_frames(hipStream_t s, int K, const DevBoundary& b, const int* nodes, const double* q, double* nrm,
                           double* t1, double* t2, unsigned char* enabled_out) {
    if (K == 0) return;
    if (b.type == 5) {
        if (!enabled_out) throw std::invalid_argument("mesh boundary frames need the candidate filter output");
        hipLaunchKernelGGL(k_mesh_query<true>, dim3(blocks_for(K, 4)), dim3(256), 0, s, K, b, nodes, q, nullptr, nrm, t1,
                           t2, enabled_out);
        return;
    }
    hipLaunchKernelGGL(k_contact_frames, dim3(blocks_for(K, 128)), dim3(128), 0, s, K, b, nodes, q, nrm, t1, t2);
}
void launch_contact_distance(hipStream_t s, int K, const DevBoundary& b, const int* nodes, const double* q, double* delta,
                             const double* refn) {
    if (K == 0) return;
    if (b.type == 5) {
        if (!refn) throw std::invalid_argument("mesh boundary distances need the frame-start normals");
        hipLaunchKernelGGL(k_mesh_query<false>, dim3(blocks_for(K, 4)), dim3(256), 0, s, K, b, nodes, q, refn, delta,
                           nullptr, nullptr, nullptr);
        return;
    }
    hipLaunchKernelGGL(k_contact_distance, dim3(blocks_for(K, 128)), dim3(128), 0, s, K, b, nodes, q, delta);
}
template <class TV>
void launch_contact_gather(hipStream_t s, const ContactPieces& pc, const int* cp_row, const TV* cp_val, const TV* y, int n,
                           double* part, double* out) {
    if (pc.nlists == 0) return;
    hipLaunchKernelGGL(k_contact_gather<TV>, dim3(blocks_for(pc.npieces, 4)), dim3(256), 0, s, pc.npieces, pc.b, pc.e, cp_row,
                       cp_val, y, n, part);
    hipLaunchKernelGGL(k_contact_gather_fin, dim3(blocks_for(pc.nlists, 256)), dim3(256), 0, s, pc.nlists, pc.ptr,
                       (const double*)part, out);
}
template <class TV>
void launch_contact_scatter(hipStream_t s, const ContactPieces& pc, const int* rc_row, const int* rc_c, const TV* rc_val,
                            const double* f, double* part, int n, TV* y) {
    if (pc.nlists == 0) return;
    hipLaunchKernelGGL(k_contact_scatter<TV>, dim3(blocks_for(pc.npieces, 4)), dim3(256), 0, s, pc.npieces, pc.b, pc.e, rc_c,
                       rc_val, f, part);
    hipLaunchKernelGGL(k_contact_scatter_fin<TV>, dim3(blocks_for(pc.nlists, 256)), dim3(256), 0, s, pc.nlists, pc.ptr,
                       rc_row, (const double*)part, n, y);
}
PHD_INST_TV(float)
PHD_INST_TV(double)

void launch_gather_vals(hipStream_t s, int64_t m, const int* rows, const int* cols, const int64_t* base,
                        const double* V64, double* o64, float* o32) {
    if (m == 0) return;
    hipLaunchKernelGGL(k_gather_vals, dim3(blocks_for(m, 256)), dim3(256), 0, s, m, rows, cols, base, V64, o64, o32);
}
void launch_scatter_dense(hipStream_t s, int64_t m, const int* ridx, const int* cidx, const double* val, int K, double* D) {
    if (m == 0) return;
    hipLaunchKernelGGL(k_scatter_dense, dim3(blocks_for(m, 256)), dim3(256), 0, s, m, ridx, cidx, val, K, D);
}
void launch_symmetrize_lower(hipStream_t s, int K, double* H) {
    if (K == 0) return;
    hipLaunchKernelGGL(k_symmetrize_lower, dim3(blocks_for((int64_t)K * K, 256)), dim3(256), 0, s, K, H);
}
void launch_gather_nodes(hipStream_t s, int K, const int* nodes, const double* q, double* out) {
    if (K == 0) return;
    hipLaunchKernelGGL(k_gather_nodes, dim3(blocks_for(K, 128)), dim3(128), 0, s, K, nodes, q, out);
}
namespace {
// Gs (ld x ld, column-major, zero past n) = the full symmetric matrix whose
// lower triangle G (n x n, column-major) holds; 32 x 32 tiles through LDS so
// that the mirrored reads stay coalesced.
template <class TG, class TV>
__global__ void __launch_bounds__(256) k_sym_full(int n, int64_t ld, const TG* __restrict__ G, TV* __restrict__ Gs) {
    __shared__ TG t[32][33];  // t[q][i] = G(32 sr + i, 32 sc + q), lower entries only
    const int tr = blockIdx.x, tc = blockIdx.y;
    const int lx = threadIdx.x % 32, ly = threadIdx.x / 32;
    const int sr = max(tr, tc), sc = min(tr, tc);
    for (int q = ly; q < 32; q += 8) {
        const int r = 32 * sr + lx, c = 32 * sc + q;
        t[q][lx] = (r < n && c < n && r >= c) ? G[r + (int64_t)c * n] : TG(0);
    }
    __syncthreads();
    for (int q = ly; q < 32; q += 8) {
        const int r = 32 * tr + lx, c = 32 * tc + q;
        const TG v = (tr > tc || (tr == tc && lx >= q)) ? t[q][lx] : t[lx][q];
        Gs[r + (int64_t)c * ld] = (TV)v;
    }
}

// y = Gs x for a symmetric Gs (ld x ld, zero past n): y_i = (column i) . x,
// four columns per wavefront so that each x load serves four columns; 16-byte
// loads of Gs, FP64 accumulation.
template <class TV>
__global__ void __launch_bounds__(256) k_gemv_sym(int n, int64_t ld, const TV* __restrict__ Gs,
                                                  const double* __restrict__ x, double* __restrict__ y) {
    using V16 = typename Vec16<TV>::type;
    constexpr int VN = Vec16<TV>::N;
    const int i0 = 4 * (blockIdx.x * 4 + threadIdx.x / WAVE);
    const int lane = threadIdx.x % WAVE;
    if (i0 >= n) return;
    const int nc = min(4, n - i0);
    double acc[4] = {0, 0, 0, 0};
    for (int e0 = VN * lane; e0 < n; e0 += VN * WAVE) {
        double xv[VN];
#pragma unroll
        for (int e = 0; e < VN; ++e) xv[e] = (e0 + e < n) ? x[e0 + e] : 0.0;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            if (c < nc) {
                const V16 g = *reinterpret_cast<const V16*>(Gs + (int64_t)(i0 + c) * ld + e0);
                TV gv[VN];
                memcpy(gv, &g, sizeof(g));
#pragma unroll
                for (int e = 0; e < VN; ++e) acc[c] += (double)gv[e] * xv[e];
            }
        }
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const double t = wave_sum(acc[c]);
        if (lane == 0 && c < nc) y[i0 + c] = t;
    }
}
}  // namespace

template <class TG, class TV>
void launch_sym_full(hipStream_t s, int n, int64_t ld, const TG* G, TV* Gs) {
    if (n == 0) return;
    const unsigned t = (unsigned)(ld / 32);
    hipLaunchKernelGGL((k_sym_full<TG, TV>), dim3(t, t), dim3(256), 0, s, n, ld, G, Gs);
}
template <class TV>
void launch_gemv_sym(hipStream_t s, int n, int64_t ld, const TV* Gs, const double* x, double* y) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_gemv_sym<TV>, dim3((n + 15) / 16), dim3(256), 0, s, n, ld, Gs, x, y);
}
template void launch_sym_full<double, float>(hipStream_t, int, int64_t, const double*, float*);
template void launch_sym_full<double, double>(hipStream_t, int, int64_t, const double*, double*);
template void launch_sym_full<float, float>(hipStream_t, int, int64_t, const float*, float*);
template void launch_gemv_sym<float>(hipStream_t, int, int64_t, const float*, const double*, double*);
template void launch_gemv_sym<double>(hipStream_t, int, int64_t, const double*, const double*, double*);

template <class TG>
void launch_assemble_g(hipStream_t s, int nb, const int* cidx, const double* dirs, const double* H, int K, TG* G) {
    if (nb == 0) return;
    hipLaunchKernelGGL((k_assemble_g<TG>), dim3(blocks_for((int64_t)nb * nb, 256)), dim3(256), 0, s, nb, cidx, dirs, H, K, G);
}
template void launch_assemble_g<double>(hipStream_t, int, const int*, const double*, const double*, int, double*);
template void launch_assemble_g<float>(hipStream_t, int, const int*, const double*, const double*, int, float*);
void launch_basis_project(hipStream_t s, int nb, const int* cidx, const double* dirs, const double* u, const double* u2,
                          double* ub) {
    if (nb == 0) return;
    hipLaunchKernelGGL(k_basis_project, dim3(blocks_for(nb, 256)), dim3(256), 0, s, nb, cidx, dirs, u, u2, ub);
}
void launch_basis_expand(hipStream_t s, int K, const int* cptr, const int* clist, const double* dirs, const double* g,
                         double* F) {
    if (K == 0) return;
    hipLaunchKernelGGL(k_basis_expand, dim3(blocks_for(K, 256)), dim3(256), 0, s, K, cptr, clist, dirs, g, F);
}
template <class TH>
void launch_h_apply_idx(hipStream_t s, int nr, const int* rows, int nc, const int* cols, const TH* H, int K,
                        const double* f, double* out) {
    if (nr == 0) return;
    hipLaunchKernelGGL(k_h_apply_idx<TH>, dim3(blocks_for(nr, kHRows)), dim3(kHWaves * WAVE), 0, s, nr, rows, nc, cols, H, K,
                       f, out);
}
template void launch_h_apply_idx<float>(hipStream_t, int, const int*, int, const int*, const float*, int, const double*,
                                        double*);
template void launch_h_apply_idx<double>(hipStream_t, int, const int*, int, const int*, const double*, int, const double*,
                                         double*);
__global__ void k_to_float(int64_t n, const double* __restrict__ a, float* __restrict__ b) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}
void launch_to_float(hipStream_t s, int64_t n, const double* a, float* b) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_to_float, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, a, b);
}
template <class TM>
void launch_assemble_msys_idx(hipStream_t s, int nJ, const int* J, const double* H, int K, const double* nrm,
                              const double* om, const double* E, TM* M) {
    if (nJ == 0) return;
    hipLaunchKernelGGL(k_assemble_msys_idx<TM>, dim3(blocks_for((int64_t)nJ * nJ, 256)), dim3(256), 0, s, nJ, J, H, K,
                       nrm, om, E, M);
}
template void launch_assemble_msys_idx<float>(hipStream_t, int, const int*, const double*, int, const double*,
                                              const double*, const double*, float*);
template void launch_assemble_msys_idx<double>(hipStream_t, int, const int*, const double*, int, const double*,
                                               const double*, const double*, double*);
void launch_copy(hipStream_t s, int64_t n16, const void* a, void* b) {
    hipLaunchKernelGGL(k_copy, dim3((unsigned)((n16 + 1023) / 1024)), dim3(256), 0, s, n16, (const float4*)a, (float4*)b);
}
void launch_add_node_force(hipStream_t s, int K, const int* nodes, const double* f, double* r) {
    if (K == 0) return;
    hipLaunchKernelGGL(k_add_node_force, dim3(blocks_for(K, 128)), dim3(128), 0, s, K, nodes, f, r);
}
template <class TM>
void launch_assemble_msys(hipStream_t s, int K, const double* H, const double* nrm, const double* om, const double* E, TM* M) {
    if (K == 0) return;
    hipLaunchKernelGGL(k_assemble_msys<TM>, dim3(blocks_for((int64_t)K * K, 256)), dim3(256), 0, s, K, H, nrm, om, E, M);
}
template void launch_assemble_msys<float>(hipStream_t, int, const double*, const double*, const double*, const double*, float*);
template void launch_assemble_msys<double>(hipStream_t, int, const double*, const double*, const double*, const double*, double*);
void launch_h_apply(hipStream_t s, int K, const double* H, const double* f, double* out) {
    if (K == 0) return;
    hipLaunchKernelGGL(k_h_apply, dim3(blocks_for(K, 4)), dim3(256), 0, s, K, H, f, out);
}

template <class TW>
void launch_kpd_build(hipStream_t s, const DevMaterials& m, const double* kc, const double* kv, const double* const* kn,
                      const double* q, int filter, double theta, double lo, int* nfilt, TW* W) {
    KnPtrs p;
    for (int h = 0; h < 4; ++h) p.kn[h] = (h < m.n_nh) ? kn[h] : nullptr;
    const int64_t total = (int64_t)m.ne * m.samples;
    if (m.edim == 8)
        hipLaunchKernelGGL((k_kpd_build<8, TW>), dim3(blocks_for(total, 64)), dim3(64), 0, s, m, kc, kv, p, q, filter, theta,
                           lo, nfilt, W);
    else
        hipLaunchKernelGGL((k_kpd_build<4, TW>), dim3(blocks_for(total, 64)), dim3(64), 0, s, m, kc, kv, p, q, filter, theta,
                           lo, nfilt, W);
}
template <class TW>
void launch_kpd_apply(hipStream_t s, const DevMaterials& m, const TW* W, const double* dq, double* elem_out) {
    if (m.edim == 8)
        hipLaunchKernelGGL((k_kpd_apply<8, TW>), dim3(blocks_for(m.ne, 128)), dim3(128), 0, s, m, W, dq, elem_out);
    else
        hipLaunchKernelGGL((k_kpd_apply<4, TW>), dim3(blocks_for(m.ne, 128)), dim3(128), 0, s, m, W, dq, elem_out);
}
template void launch_kpd_build<float>(hipStream_t, const DevMaterials&, const double*, const double*, const double* const*,
                                       const double*, int, double, double, int*, float*);
template void launch_kpd_build<double>(hipStream_t, const DevMaterials&, const double*, const double*, const double* const*,
                                       const double*, int, double, double, int*, double*);
template void launch_kpd_apply<float>(hipStream_t, const DevMaterials&, const float*, const double*, double*);
template void launch_kpd_apply<double>(hipStream_t, const DevMaterials&, const double*, const double*, double*);

// dl/da and dl/d(act_w) per muscle element (ActuationForceDifferential,
// deformable_actuation.cpp:120-177, contracted with z): with Fm = F m,
// Fzm = (sum_k z_k G_k^T) m and l = |Fm|, per sample
//   z^T df/da = (w / l) Fm . Fzm,  z^T df/dw_act = -(vol/s)(1 - a / l) Fm . Fzm
// (df_k/da = -(vol/s) dP/da G_k with P = k (1 - a/l) F m m^T).
template <int EDIM>
__global__ void __launch_bounds__(128) k_dact(DevMaterials m, const double* __restrict__ q, const double* __restrict__ z,
                                              double vs, double* __restrict__ g_da, double* __restrict__ g_dw) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m.ne) return;
    const double wm = m.mw[e];
    double sa = 0.0, sw = 0.0;
    if (wm != 0.0) {
        const double mf[3] = {m.mfib[e], m.mfib[m.ne + e], m.mfib[2 * m.ne + e]};
        const double a = m.mact[e];
        double x[EDIM][3], zz[EDIM][3];
#pragma unroll
        for (int k = 0; k < EDIM; ++k) {
            const int v = m.elems[k * m.ne + e];
            x[k][0] = q[3 * v]; x[k][1] = q[3 * v + 1]; x[k][2] = q[3 * v + 2];
            zz[k][0] = z[3 * v]; zz[k][1] = z[3 * v + 1]; zz[k][2] = z[3 * v + 2];
        }
        const double* Ge = m.G + (int64_t)e * m.g_stride;
        for (int s = 0; s < m.samples; ++s) {
            const double* G = Ge + s * EDIM * 3;
            double Fm[3] = {0.0, 0.0, 0.0}, Zm[3] = {0.0, 0.0, 0.0};
#pragma unroll
            for (int k = 0; k < EDIM; ++k) {
                const double gm = G[k * 3] * mf[0] + G[k * 3 + 1] * mf[1] + G[k * 3 + 2] * mf[2];
#pragma unroll
                for (int i = 0; i < 3; ++i) { Fm[i] += x[k][i] * gm; Zm[i] += zz[k][i] * gm; }
            }
            const double l = sqrt(Fm[0] * Fm[0] + Fm[1] * Fm[1] + Fm[2] * Fm[2]);
            if (l <= DBL_EPSILON) continue;
            const double t = Fm[0] * Zm[0] + Fm[1] * Zm[1] + Fm[2] * Zm[2];
            sa += wm / l * t;
            sw -= vs * (1.0 - a / l) * t;
        }
    }
    g_da[e] = sa;
    g_dw[e] = sw;
}

void launch_dact(hipStream_t s, const DevMaterials& m, const double* q, const double* z, double vol_over_s, double* g_da,
                 double* g_dw) {
    if (m.edim == 8)
        hipLaunchKernelGGL(k_dact<8>, dim3(blocks_for(m.ne, 128)), dim3(128), 0, s, m, q, z, vol_over_s, g_da, g_dw);
    else
        hipLaunchKernelGGL(k_dact<4>, dim3(blocks_for(m.ne, 128)), dim3(128), 0, s, m, q, z, vol_over_s, g_da, g_dw);
}

// -----------------------------------------------------------------------------
// The DiffPD baseline forward's energy and force (ComputePdEnergy / PdEnergyForce,
// deformable_pd_energy.cpp:156-240; ActuationEnergy / ActuationForce,
// deformable_actuation.cpp:28-77): per element
//   E_e = sum_s kc/2 |F - R|^2 + kv/2 |F - P_vol|^2 + mw/2 (|F m| - a)^2
//   f_e = -sum_s A_s^T vec(kc (F - R) + kv (F - P_vol) + mw (1 - a/l) F m m^T)
// with kc, kv the energies' stiffness() times vol/s (the mean for per-element
// parameters, as the reference's EnergyDensity / StressTensor read it).
// -----------------------------------------------------------------------------
template <int EDIM>
__global__ void __launch_bounds__(128) k_pd_energy_force(DevMaterials m, const double* __restrict__ kc,
                                                         const double* __restrict__ kv, const double* __restrict__ q,
                                                         double* __restrict__ energy, double* __restrict__ elem_out) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m.ne) return;
    double x[EDIM][3], acc[EDIM][3];
#pragma unroll
    for (int k = 0; k < EDIM; ++k) {
        const int v = m.elems[k * m.ne + e];
        x[k][0] = q[3 * v]; x[k][1] = q[3 * v + 1]; x[k][2] = q[3 * v + 2];
        acc[k][0] = acc[k][1] = acc[k][2] = 0.0;
    }
    const double wc = kc ? kc[e] : 0.0, wv = kv ? kv[e] : 0.0;
    const double wm = m.mw ? m.mw[e] : 0.0;
    double mf[3] = {0.0, 0.0, 0.0}, ma = 0.0;
    if (wm != 0.0) { mf[0] = m.mfib[e]; mf[1] = m.mfib[m.ne + e]; mf[2] = m.mfib[2 * m.ne + e]; ma = m.mact[e]; }
    double en = 0.0;
    const double* Ge = m.G + (int64_t)e * m.g_stride;
    for (int s = 0; s < m.samples; ++s) {
        const double* G = Ge + s * EDIM * 3;
        double F[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) F[i] = 0.0;
#pragma unroll
        for (int k = 0; k < EDIM; ++k)
#pragma unroll
            for (int jj = 0; jj < 3; ++jj) {
                const double g = G[k * 3 + jj];
                F[0 + 3 * jj] += x[k][0] * g;
                F[1 + 3 * jj] += x[k][1] * g;
                F[2 + 3 * jj] += x[k][2] * g;
            }
        double P[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) P[i] = 0.0;
        if (wc != 0.0 || wv != 0.0) {
            double U[9], sg[3], V[9];
            svd3(F, U, sg, V);
            if (wc != 0.0) {
                const double one[3] = {1.0, 1.0, 1.0};
                double R[9], n2 = 0.0;
                udv(U, one, V, R);
#pragma unroll
                for (int i = 0; i < 9; ++i) { const double d = F[i] - R[i]; n2 += d * d; P[i] += wc * d; }
                en += 0.5 * wc * n2;
            }
            if (wv != 0.0) {
                double D[3], sd[3], Pv[9], n2 = 0.0;
                volume_D(sg, D);
                sd[0] = sg[0] + D[0]; sd[1] = sg[1] + D[1]; sd[2] = sg[2] + D[2];
                udv(U, sd, V, Pv);
#pragma unroll
                for (int i = 0; i < 9; ++i) { const double d = F[i] - Pv[i]; n2 += d * d; P[i] += wv * d; }
                en += 0.5 * wv * n2;
            }
        }
        if (wm != 0.0) {
            double Fm[3];
#pragma unroll
            for (int i = 0; i < 3; ++i) Fm[i] = F[i] * mf[0] + F[i + 3] * mf[1] + F[i + 6] * mf[2];
            const double l = sqrt(Fm[0] * Fm[0] + Fm[1] * Fm[1] + Fm[2] * Fm[2]);
            en += 0.5 * wm * (l - ma) * (l - ma);
            if (l > DBL_EPSILON) {
                const double r = wm * (1.0 - ma / l);
#pragma unroll
                for (int j = 0; j < 3; ++j)
#pragma unroll
                    for (int i = 0; i < 3; ++i) P[i + 3 * j] += r * Fm[i] * mf[j];
            }
        }
#pragma unroll
        for (int k = 0; k < EDIM; ++k)
#pragma unroll
            for (int i = 0; i < 3; ++i)
                acc[k][i] -= P[i] * G[k * 3] + P[i + 3] * G[k * 3 + 1] + P[i + 6] * G[k * 3 + 2];
    }
    energy[e] = en;
    double* o = elem_out + (int64_t)e * EDIM * 3;
#pragma unroll
    for (int k = 0; k < EDIM; ++k) { o[3 * k] = acc[k][0]; o[3 * k + 1] = acc[k][1]; o[3 * k + 2] = acc[k][2]; }
}

void launch_pd_energy_force(hipStream_t s, const DevMaterials& m, const double* kc, const double* kv, const double* q,
                            double* energy, double* elem_out) {
    if (m.edim == 8)
        hipLaunchKernelGGL(k_pd_energy_force<8>, dim3(blocks_for(m.ne, 128)), dim3(128), 0, s, m, kc, kv, q, energy, elem_out);
    else
        hipLaunchKernelGGL(k_pd_energy_force<4>, dim3(blocks_for(m.ne, 128)), dim3(128), 0, s, m, kc, kv, q, energy, elem_out);
}

// PdLhsSolve's frozen nodes (forward.cpp:690-936) on the shared factor: with
// Z[c] = A^-1 e_c (one column per frozen node, the same on the three axes),
// x += Z mu per axis (mu: 3 values per frozen node), then x = src on the frozen
// nodes.
__global__ void __launch_bounds__(256) k_frozen_apply(int nv, int nc, const double* __restrict__ Z,
                                                      const int* __restrict__ zcol, const double* __restrict__ mu,
                                                      double* __restrict__ x) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int c = 0; c < nc; ++c) {
        const double z = Z[(int64_t)zcol[c] * nv + v];
        s0 += z * mu[3 * c];
        s1 += z * mu[3 * c + 1];
        s2 += z * mu[3 * c + 2];
    }
    x[3 * v] += s0;
    x[3 * v + 1] += s1;
    x[3 * v + 2] += s2;
}
__global__ void k_copy_nodes(int nc, const int* __restrict__ nodes, const double* __restrict__ src, double* __restrict__ dst) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const int v = nodes[c];
    dst[3 * v] = src[3 * v];
    dst[3 * v + 1] = src[3 * v + 1];
    dst[3 * v + 2] = src[3 * v + 2];
}
// G[i * nc + j] = Z[zcol[j]][nodes[i]]: the frozen block of A^-1
__global__ void k_gather_z(int nc, int nv, const int* __restrict__ nodes, const int* __restrict__ zcol,
                           const double* __restrict__ Z, double* __restrict__ G) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nc * nc) return;
    const int i = t / nc, j = t % nc;
    G[t] = Z[(int64_t)zcol[j] * nv + nodes[i]];
}
void launch_gather_z(hipStream_t s, int nc, int nv, const int* nodes, const int* zcol, const double* Z, double* G) {
    if (nc > 0) hipLaunchKernelGGL(k_gather_z, dim3(blocks_for((int64_t)nc * nc, 256)), dim3(256), 0, s, nc, nv, nodes, zcol, Z, G);
}
void launch_frozen_apply(hipStream_t s, int nv, int nc, const double* Z, const int* zcol, const double* mu, double* x) {
    hipLaunchKernelGGL(k_frozen_apply, dim3(blocks_for(nv, 256)), dim3(256), 0, s, nv, nc, Z, zcol, mu, x);
}
void launch_copy_nodes(hipStream_t s, int nc, const int* nodes, const double* src, double* dst) {
    if (nc > 0) hipLaunchKernelGGL(k_copy_nodes, dim3(blocks_for(nc, 128)), dim3(128), 0, s, nc, nodes, src, dst);
}

void launch_dmatw(hipStream_t s, const DevMaterials& m, const double* q, const double* z, double vol_over_s, double* g_cor,
                  double* g_vol, double* g_nh) {
    if (m.edim == 8)
        hipLaunchKernelGGL(k_dmatw<8>, dim3(blocks_for(m.ne, 128)), dim3(128), 0, s, m, q, z, vol_over_s, g_cor, g_vol, g_nh);
    else
        hipLaunchKernelGGL(k_dmatw<4>, dim3(blocks_for(m.ne, 128)), dim3(128), 0, s, m, q, z, vol_over_s, g_cor, g_vol, g_nh);
}

// -----------------------------------------------------------------------------
// The per-candidate part of the NCP iteration on the device
// (ForwardProjectiveDynamicsAlgPhd, forward_gpu.cpp:1149-1302): the smoothed
// Signorini weights, the index sets of the compressed contact system, the
// multiplier update and the Coulomb step. Only the set sizes and the dense
// solve's pivot flag reach the host.
// -----------------------------------------------------------------------------
namespace {

// omega, E, h (:1160-1197) per candidate, the current impulses' support flags
// and F_g = omega lambda n + lt1 t1 + lt2 t2 (the impulses g carries); uC, the
// candidates' rows of S^T S r, becomes (A^-1 b)_C + H F_prev by adding q_C.
__global__ void k_ncp_sets(int K, const unsigned char* __restrict__ enabled, const double* __restrict__ delta,
                           const double* __restrict__ lambda,
                           const double* __restrict__ rprec, const double* __restrict__ nrm, const double* __restrict__ t1,
                           const double* __restrict__ t2, const double* __restrict__ qC, double* __restrict__ uC,
                           const double* __restrict__ F, const double* __restrict__ lt1, const double* __restrict__ lt2,
                           int do_fric, double* __restrict__ om, double* __restrict__ E, double* __restrict__ hvec,
                           double* __restrict__ Fg, unsigned char* __restrict__ flags) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= K) return;
    if (enabled && !enabled[c]) {  // outside the replayed active set: no impulse, lambda stays 0
        om[c] = 0.0;
        E[c] = 1.0;
        hvec[c] = 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d) { uC[3 * c + d] += qC[3 * c + d]; Fg[3 * c + d] = 0.0; }
        flags[c] = 0;
        return;
    }
    const double dj = delta[c], lj = lambda[c], rj = rprec[c];
    double den = sqrt(dj * dj + rj * rj * lj * lj);
    if (den < 1e-14) den = 1e-14;
    const double o = 1.0 - dj / den;
    om[c] = o;
    E[c] = (1.0 - rj * lj / den) * rj;
    const double phi = dj + rj * lj - den;
    double hq = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double q = qC[3 * c + d];
        uC[3 * c + d] += q;
        hq += nrm[3 * c + d] * q;
    }
    hvec[c] = -phi + o * hq;
    const double l1 = lt1[c], l2 = lt2[c];
    const bool fr = do_fric && (l1 != 0.0 || l2 != 0.0);
    const bool inG = o * lj != 0.0 || fr;
#pragma unroll
    for (int d = 0; d < 3; ++d)
        Fg[3 * c + d] = inG ? o * lj * nrm[3 * c + d] + (fr ? l1 * t1[3 * c + d] + l2 * t2[3 * c + d] : 0.0) : 0.0;
    const bool inF = F[3 * c] != 0.0 || F[3 * c + 1] != 0.0 || F[3 * c + 2] != 0.0;
    flags[c] = (unsigned char)((o != 0.0 ? kNcpJ : 0) | (inF ? kNcpF : 0) | (inG ? kNcpG : 0) |
                               (o != 0.0 || inF ? kNcpP : 0));
}

// Lists (in increasing candidate order) of the candidates whose flags contain
// mask m.m[i] (all bits), i < nl; counts[i] = list sizes; pos[c] = position of
// c in list 0, -1 when absent. One workgroup; each thread owns a contiguous run
// of candidates, counts its members of every list in one pass over the flags,
// and the nl counts are scanned together (one pair of barriers for all lists).
struct Masks { int m[4]; };
__global__ void __launch_bounds__(1024) k_ncp_compact(int K, const unsigned char* __restrict__ flags, int nl, Masks mk,
                                                      int* __restrict__ lists, int* __restrict__ counts,
                                                      int* __restrict__ pos) {
    __shared__ int sh[4][16];
    const int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE, nw = blockDim.x / WAVE;
    const int per = (K + blockDim.x - 1) / blockDim.x;
    const int c0 = min(K, (int)threadIdx.x * per), c1 = min(K, c0 + per);
    int cnt[4] = {0, 0, 0, 0};
    for (int c = c0; c < c1; ++c) {
        const int f = flags[c];
#pragma unroll
        for (int i = 0; i < 4; ++i) cnt[i] += i < nl && (f & mk.m[i]) == mk.m[i];
    }
    // inclusive scan of the four counts within the wavefront, then over the wavefront totals
    int x[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        x[i] = cnt[i];
#pragma unroll
        for (int o = 1; o < WAVE; o <<= 1) {
            const int y = __shfl_up(x[i], o, WAVE);
            if (lane >= o) x[i] += y;
        }
        if (lane == WAVE - 1) sh[i][w] = x[i];
    }
    __syncthreads();
    if (w < 4) {  // wavefront i scans the totals of list i
        int t = lane < nw ? sh[w][lane] : 0;
#pragma unroll
        for (int o = 1; o < WAVE; o <<= 1) {
            const int y = __shfl_up(t, o, WAVE);
            if (lane >= o) t += y;
        }
        if (lane < nw) sh[w][lane] = t;
    }
    __syncthreads();
    int at[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) at[i] = (w > 0 ? sh[i][w - 1] : 0) + x[i] - cnt[i];
    for (int c = c0; c < c1; ++c) {
        const int f = flags[c];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if (i >= nl) break;
            const bool in = (f & mk.m[i]) == mk.m[i];
            if (in) lists[(int64_t)i * K + at[i]] = c;
            if (i == 0) pos[c] = in ? at[0] : -1;
            at[i] += in;
        }
    }
    if (threadIdx.x < nl) counts[threadIdx.x] = sh[threadIdx.x][nw - 1];
}

// For a < nP (c = P[a]): uC_c -= (H F_prev)_a; for c in J (a' = pos_J[c]) the
// right-hand side of the compressed system, h_c - omega_c n_c . (uC_c + (H F_g)_a').
__global__ void k_ncp_rhs(int nP, const int* __restrict__ P, const double* __restrict__ hprev, const int* __restrict__ posJ,
                          const double* __restrict__ hg, const double* __restrict__ hvec, const double* __restrict__ om,
                          const double* __restrict__ nrm, double* __restrict__ uC, double* __restrict__ rJ) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nP) return;
    const int c = P[a];
    double u[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        u[d] = uC[3 * c + d] - hprev[3 * a + d];
        uC[3 * c + d] = u[d];
    }
    const int j = posJ[c];
    if (j < 0) return;
    double val = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) val += nrm[3 * c + d] * (u[d] + hg[3 * j + d]);
    rJ[j] = hvec[c] - om[c] * val;
}

// lambda_c <- clamp(lambda_c + dl_c, 0, 1e8): dl from the compressed solve on J,
// h_c / E_c for the decoupled rest (:1241-1253).
template <class TV>
__global__ void k_ncp_lambda(int K, const int* __restrict__ posJ, const TV* __restrict__ x, const double* __restrict__ hvec,
                             const double* __restrict__ E, const double* __restrict__ om, double lam_max,
                             double* __restrict__ lambda) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= K) return;
    const double dl = om[c] != 0.0 ? (double)x[posJ[c]] : hvec[c] / E[c];
    lambda[c] = fmin(fmax(lambda[c] + dl, 0.0), lam_max);
}

// Friction sets: lambda > 0 (kNcpL) and omega lambda != 0 (kNcpN, carrying the
// normal impulse F_n = omega lambda n).
__global__ void k_fric_sets(int K, const double* __restrict__ lambda, const double* __restrict__ om,
                            const double* __restrict__ nrm, double* __restrict__ Fn, unsigned char* __restrict__ flags) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= K) return;
    const double l = lambda[c], ol = om[c] * l;
#pragma unroll
    for (int d = 0; d < 3; ++d) Fn[3 * c + d] = ol != 0.0 ? ol * nrm[3 * c + d] : 0.0;
    flags[c] = (unsigned char)((l > 0.0 ? kNcpL : 0) | (ol != 0.0 ? kNcpN : 0));
}

// The NCP-integrated Coulomb step (:1255-1302) for lambda_c > 0: the unconstrained
// step of the contact vertex is uC_c + (H F_n)_c - q_c(frame start); its
// tangential part u_t gives the stick impulse W_tt lt = -u_t (W_tt = H_cc times
// the tangents' Gram matrix), kept if inside the cone mu lambda, else the slip
// impulse -mu lambda u_t / |u_t|. Zero for lambda_c <= 0.
__global__ void k_fric_coulomb(int K, const double* __restrict__ lambda, const int* __restrict__ posL,
                               const double* __restrict__ hn, const double* __restrict__ uC,
                               const double* __restrict__ qframe, const double* __restrict__ nrm,
                               const double* __restrict__ t1, const double* __restrict__ t2, const double* __restrict__ Hdiag,
                               double fmu, double* __restrict__ lt1, double* __restrict__ lt2) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= K) return;
    const double lam = lambda[c];
    if (lam <= 0.0) { lt1[c] = 0.0; lt2[c] = 0.0; return; }
    const int a = posL[c];
    double step[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) step[d] = uC[3 * c + d] + hn[3 * a + d] - qframe[3 * c + d];
    const double* n = nrm + 3 * c;
    const double* e1 = t1 + 3 * c;
    const double* e2 = t2 + 3 * c;
    const double sn = step[0] * n[0] + step[1] * n[1] + step[2] * n[2];
    double ut[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) ut[d] = step[d] - sn * n[d];
    const double u1 = ut[0] * e1[0] + ut[1] * e1[1] + ut[2] * e1[2];
    const double u2 = ut[0] * e2[0] + ut[1] * e2[1] + ut[2] * e2[2];
    const double hc = Hdiag[c];
    const double w11 = hc * (e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
    const double w12 = hc * (e1[0] * e2[0] + e1[1] * e2[1] + e1[2] * e2[2]);
    const double w22 = hc * (e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2]);
    const double det = w11 * w22 - w12 * w12;
    double s1 = 0.0, s2 = 0.0;
    if (fabs(det) > 1e-20) { s1 = (w22 * -u1 - w12 * -u2) / det; s2 = (-w12 * -u1 + w11 * -u2) / det; }
    const double cone = fmu * lam;
    if (sqrt(s1 * s1 + s2 * s2) <= cone + 1e-14) { lt1[c] = s1; lt2[c] = s2; return; }
    const double um = sqrt(u1 * u1 + u2 * u2);
    if (um > 1e-14) { lt1[c] = -cone * u1 / um; lt2[c] = -cone * u2 / um; }
    else { lt1[c] = 0.0; lt2[c] = 0.0; }
}

// New impulses F = omega lambda n + lt1 t1 + lt2 t2 and their change dF (the S
// pass carries only the change).
__global__ void k_ncp_impulses(int K, const double* __restrict__ om, const double* __restrict__ lambda,
                               const double* __restrict__ nrm, const double* __restrict__ lt1, const double* __restrict__ lt2,
                               const double* __restrict__ t1, const double* __restrict__ t2, int do_fric,
                               double* __restrict__ F, double* __restrict__ dF) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= K) return;
    const double ol = om[c] * lambda[c];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double f = ol * nrm[3 * c + d] + (do_fric ? lt1[c] * t1[3 * c + d] + lt2[c] * t2[3 * c + d] : 0.0);
        dF[3 * c + d] = f - F[3 * c + d];
        F[3 * c + d] = f;
    }
}

// k_h_apply_idx with the row and column counts read on the device (grid sized
// for the largest possible row list).
template <class TH>
__global__ __launch_bounds__(kHWaves * WAVE) void k_h_apply_idx_dev(const int* __restrict__ nr_dev,
                                                                    const int* __restrict__ rows,
                                                                    const int* __restrict__ nc_dev,
                                                                    const int* __restrict__ cols,
                                                                    const TH* __restrict__ H, int K,
                                                                    const double* __restrict__ f,
                                                                    double* __restrict__ out) {
    const int nr = *nr_dev;
    const int a0 = blockIdx.x * kHRows;
    if (a0 >= nr) return;  // uniform over the workgroup
    h_apply_rows<TH>(a0, nr, rows, *nc_dev, cols, H, K, f, out);
}

}  // namespace

void launch_ncp_sets(hipStream_t s, int K, const unsigned char* enabled, const double* delta, const double* lambda,
                     const double* rprec, const double* nrm, const double* t1, const double* t2, const double* qC, double* uC,
                     const double* F, const double* lt1, const double* lt2, int do_fric, double* om, double* E, double* hvec,
                     double* Fg, unsigned char* flags) {
    if (K == 0) return;
    hipLaunchKernelGGL(k_ncp_sets, dim3(blocks_for(K, 128)), dim3(128), 0, s, K, enabled, delta, lambda, rprec, nrm, t1, t2,
                       qC, uC, F, lt1, lt2, do_fric, om, E, hvec, Fg, flags);
}
void launch_ncp_compact(hipStream_t s, int K, const unsigned char* flags, int nl, const int* masks, int* lists, int* counts,
                        int* pos) {
    Masks mk = {{0, 0, 0, 0}};
    nl = nl < 4 ? nl : 4;
    for (int i = 0; i < nl; ++i) mk.m[i] = masks[i];
    hipLaunchKernelGGL(k_ncp_compact, dim3(1), dim3(1024), 0, s, K, flags, nl, mk, lists, counts, pos);
}
void launch_ncp_rhs(hipStream_t s, int nP, const int* P, const double* hprev, const int* posJ, const double* hg,
                    const double* hvec, const double* om, const double* nrm, double* uC, double* rJ) {
    if (nP == 0) return;
    hipLaunchKernelGGL(k_ncp_rhs, dim3(blocks_for(nP, 128)), dim3(128), 0, s, nP, P, hprev, posJ, hg, hvec, om, nrm, uC, rJ);
}
template <class TV>
void launch_ncp_lambda(hipStream_t s, int K, const int* posJ, const TV* x, const double* hvec, const double* E,
                       const double* om, double lam_max, double* lambda) {
    if (K == 0) return;
    hipLaunchKernelGGL((k_ncp_lambda<TV>), dim3(blocks_for(K, 128)), dim3(128), 0, s, K, posJ, x, hvec, E, om, lam_max,
                       lambda);
}
template void launch_ncp_lambda<float>(hipStream_t, int, const int*, const float*, const double*, const double*,
                                       const double*, double, double*);
template void launch_ncp_lambda<double>(hipStream_t, int, const int*, const double*, const double*, const double*,
                                        const double*, double, double*);
void launch_fric_sets(hipStream_t s, int K, const double* lambda, const double* om, const double* nrm, double* Fn,
                      unsigned char* flags) {
    if (K == 0) return;
    hipLaunchKernelGGL(k_fric_sets, dim3(blocks_for(K, 128)), dim3(128), 0, s, K, lambda, om, nrm, Fn, flags);
}
void launch_fric_coulomb(hipStream_t s, int K, const double* lambda, const int* posL, const double* hn, const double* uC,
                         const double* qframe, const double* nrm, const double* t1, const double* t2, const double* Hdiag,
                         double fmu, double* lt1, double* lt2) {
    if (K == 0) return;
    hipLaunchKernelGGL(k_fric_coulomb, dim3(blocks_for(K, 128)), dim3(128), 0, s, K, lambda, posL, hn, uC, qframe, nrm, t1,
                       t2, Hdiag, fmu, lt1, lt2);
}
void launch_ncp_impulses(hipStream_t s, int K, const double* om, const double* lambda, const double* nrm, const double* lt1,
                         const double* lt2, const double* t1, const double* t2, int do_fric, double* F, double* dF) {
    if (K == 0) return;
    hipLaunchKernelGGL(k_ncp_impulses, dim3(blocks_for(K, 128)), dim3(128), 0, s, K, om, lambda, nrm, lt1, lt2, t1, t2,
                       do_fric, F, dF);
}
template <class TH>
void launch_h_apply_idx_dev(hipStream_t s, int max_rows, const int* nr_dev, const int* rows, const int* nc_dev,
                            const int* cols, const TH* H, int K, const double* f, double* out) {
    if (max_rows == 0) return;
    hipLaunchKernelGGL(k_h_apply_idx_dev<TH>, dim3(blocks_for(max_rows, kHRows)), dim3(kHWaves * WAVE), 0, s, nr_dev, rows,
                       nc_dev, cols, H, K, f, out);
}
template void launch_h_apply_idx_dev<float>(hipStream_t, int, const int*, const int*, const int*, const int*,
                                            const float*, int, const double*, double*);
template void launch_h_apply_idx_dev<double>(hipStream_t, int, const int*, const int*, const int*, const int*,
                                             const double*, int, const double*, double*);

}  // namespace phd
