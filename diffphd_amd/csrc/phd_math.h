// Element-level math shared by the HIP kernels: 3x3 SVD, polar factor, the
// corotated / volume / Neo-Hookean proximal projections and their
// differentials. Templated on the arithmetic type; every routine is
// register-resident (no local arrays indexed by run-time values).
//
// Conventions follow the reference (Eigen::JacobiSVD at
// cpp/core/src/pd_energy/deformation_gradient_auxiliary_data.cpp:6): singular
// values sorted descending and non-negative, U and V orthogonal (U may be a
// reflection when det F < 0). Matrices are column-major: M[i + 3j] = M(i, j).
#pragma once
#include <hip/hip_runtime.h>
#include <cfloat>

namespace phd {

template <class T> struct Eps;
template <> struct Eps<double> { static constexpr double v = DBL_EPSILON; };
template <> struct Eps<float> { static constexpr float v = FLT_EPSILON; };

template <class T>
__device__ __forceinline__ void mat_mul(const T* A, const T* B, T* C) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) C[i + 3 * j] = A[i] * B[3 * j] + A[i + 3] * B[1 + 3 * j] + A[i + 6] * B[2 + 3 * j];
}
// C = A^T B
template <class T>
__device__ __forceinline__ void mat_tmul(const T* A, const T* B, T* C) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            C[i + 3 * j] = A[3 * i] * B[3 * j] + A[1 + 3 * i] * B[1 + 3 * j] + A[2 + 3 * i] * B[2 + 3 * j];
}
// C = U diag(d) V^T
template <class T>
__device__ __forceinline__ void udv(const T* U, const T* d, const T* V, T* C) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            C[i + 3 * j] = U[i] * d[0] * V[j] + U[i + 3] * d[1] * V[j + 3] + U[i + 6] * d[2] * V[j + 6];
}

template <class T>
__device__ __forceinline__ void jacobi_rotate(T* S, T* V, int p, int q) {
    const T spq = S[p + 3 * q];
    const T tau = (S[q + 3 * q] - S[p + 3 * p]) / (T(2) * spq);
    const T t = (tau >= T(0) ? T(1) : T(-1)) / (fabs(tau) + sqrt(T(1) + tau * tau));
    const T c = T(1) / sqrt(T(1) + t * t), s = c * t;
    // S <- J^T S J with J the Givens rotation in plane (p, q)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const T skp = S[k + 3 * p], skq = S[k + 3 * q];
        S[k + 3 * p] = c * skp - s * skq;
        S[k + 3 * q] = s * skp + c * skq;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const T spk = S[p + 3 * k], sqk = S[q + 3 * k];
        S[p + 3 * k] = c * spk - s * sqk;
        S[q + 3 * k] = s * spk + c * sqk;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const T vkp = V[k + 3 * p], vkq = V[k + 3 * q];
        V[k + 3 * p] = c * vkp - s * vkq;
        V[k + 3 * q] = s * vkp + c * vkq;
    }
}

// Symmetric 3x3 eigen-decomposition by cyclic Jacobi rotations: S (column-
// major, symmetric) is overwritten by its diagonal form (eigenvalues on the
// diagonal), the eigenvectors accumulate in the columns of Z.
template <class T>
__device__ void sym3_eigen(T* S, T* Z) {
#pragma unroll
    for (int i = 0; i < 9; ++i) Z[i] = (i % 4 == 0) ? T(1) : T(0);
    for (int sweep = 0; sweep < 10; ++sweep) {
        const T off = fabs(S[3]) + fabs(S[6]) + fabs(S[7]);
        const T dia = fabs(S[0]) + fabs(S[4]) + fabs(S[8]);
        if (!(off > T(1e-30) * dia)) break;
        if (S[3] != T(0)) jacobi_rotate(S, Z, 0, 1);
        if (S[6] != T(0)) jacobi_rotate(S, Z, 0, 2);
        if (S[7] != T(0)) jacobi_rotate(S, Z, 1, 2);
    }
}

// Symmetric 2x2 [[a, b], [b, c]]: eigenvalues l0, l1 with eigenvectors
// (cs, -sn) and (sn, cs) (one Jacobi rotation).
template <class T>
__device__ __forceinline__ void sym2_eigen(T a, T b, T c, T& l0, T& l1, T& cs, T& sn) {
    if (b == T(0)) { l0 = a; l1 = c; cs = T(1); sn = T(0); return; }
    const T tau = (c - a) / (T(2) * b);
    const T t = (tau >= T(0) ? T(1) : T(-1)) / (fabs(tau) + sqrt(T(1) + tau * tau));
    cs = T(1) / sqrt(T(1) + t * t);
    sn = cs * t;
    l0 = a - t * b;
    l1 = c + t * b;
}

// SVD F = U diag(sig) V^T: cyclic Jacobi eigen-analysis of F^T F with exact
// Givens rotations (the structure of McAdams et al. 2011, "Computing the SVD
// of 3x3 matrices with minimal branching and elementary floating point
// operations", with exact instead of approximate rotations so that FP64 keeps
// full precision), descending sort, then U from a Gram-Schmidt QR of F V whose
// diagonal gives the singular values.
template <class T>
__device__ void svd3(const T* F, T* U, T* sig, T* V) {
    T S[9];
    mat_tmul(F, F, S);
#pragma unroll
    for (int i = 0; i < 9; ++i) V[i] = T(0);
    V[0] = V[4] = V[8] = T(1);
    const T tol = Eps<T>::v * Eps<T>::v;
    for (int sweep = 0; sweep < 12; ++sweep) {
        const T off = S[3] * S[3] + S[6] * S[6] + S[7] * S[7];
        const T dia = S[0] * S[0] + S[4] * S[4] + S[8] * S[8];
        if (off <= tol * dia) break;
        if (S[3] != T(0)) jacobi_rotate(S, V, 0, 1);
        if (S[6] != T(0)) jacobi_rotate(S, V, 0, 2);
        if (S[7] != T(0)) jacobi_rotate(S, V, 1, 2);
    }
    // sort eigenvalues descending (swap V columns)
    T l0 = S[0], l1 = S[4], l2 = S[8];
    auto swapc = [&](int a, int b) {
#pragma unroll
        for (int k = 0; k < 3; ++k) { const T t = V[k + 3 * a]; V[k + 3 * a] = V[k + 3 * b]; V[k + 3 * b] = t; }
    };
    if (l0 < l1) { T t = l0; l0 = l1; l1 = t; swapc(0, 1); }
    if (l0 < l2) { T t = l0; l0 = l2; l2 = t; swapc(0, 2); }
    if (l1 < l2) { T t = l1; l1 = l2; l2 = t; swapc(1, 2); }
    T B[9];
    mat_mul(F, V, B);
    const T tiny = T(1e-30);
    // column 0
    T n0 = sqrt(B[0] * B[0] + B[1] * B[1] + B[2] * B[2]);
    if (n0 > tiny) { U[0] = B[0] / n0; U[1] = B[1] / n0; U[2] = B[2] / n0; }
    else { U[0] = T(1); U[1] = T(0); U[2] = T(0); n0 = T(0); }
    // column 1
    const T d1 = U[0] * B[3] + U[1] * B[4] + U[2] * B[5];
    T c1[3] = {B[3] - d1 * U[0], B[4] - d1 * U[1], B[5] - d1 * U[2]};
    T n1 = sqrt(c1[0] * c1[0] + c1[1] * c1[1] + c1[2] * c1[2]);
    if (n1 > tiny && n0 > T(0)) { U[3] = c1[0] / n1; U[4] = c1[1] / n1; U[5] = c1[2] / n1; }
    else {
        // any unit vector orthogonal to U0: cross with the least aligned axis
        int ax = 0;
        if (fabs(U[1]) < fabs(U[ax])) ax = 1;
        if (fabs(U[2]) < fabs(U[ax])) ax = 2;
        T e[3] = {T(0), T(0), T(0)};
        e[ax] = T(1);
        T x[3] = {U[1] * e[2] - U[2] * e[1], U[2] * e[0] - U[0] * e[2], U[0] * e[1] - U[1] * e[0]};
        const T xn = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
        U[3] = x[0] / xn; U[4] = x[1] / xn; U[5] = x[2] / xn;
        n1 = T(0);
    }
    // column 2 = U0 x U1, signed singular value from the projection of F v2
    U[6] = U[1] * U[5] - U[2] * U[4];
    U[7] = U[2] * U[3] - U[0] * U[5];
    U[8] = U[0] * U[4] - U[1] * U[3];
    T n2 = U[6] * B[6] + U[7] * B[7] + U[8] * B[8];
    if (n2 < T(0)) { n2 = -n2; U[6] = -U[6]; U[7] = -U[7]; U[8] = -U[8]; }
    sig[0] = n0; sig[1] = n1; sig[2] = n2;
}

// dRFromdF, cpp/core/src/common/geometry.cpp:85-113 (R, S of the polar
// decomposition F = R S, S = V Sig V^T).
template <class T>
__device__ void dR_from_dF(const T* R, const T* S, const T* dF, T* out) {
    T RtdF[9];
    mat_tmul(R, dF, RtdF);
    // lhs = R^T dF - dF^T R = RtdF - RtdF^T
    const T b0 = RtdF[0 + 3 * 1] - RtdF[1 + 3 * 0];
    const T b1 = RtdF[0 + 3 * 2] - RtdF[2 + 3 * 0];
    const T b2 = RtdF[1 + 3 * 2] - RtdF[2 + 3 * 1];
    const T a00 = S[0] + S[4], a11 = S[0] + S[8], a22 = S[4] + S[8];
    const T a01 = S[1 + 3 * 2], a02 = -S[0 + 3 * 2], a12 = S[0 + 3 * 1];
    // solve the symmetric 3x3 system by Cramer's rule (A.inverse() in the reference)
    const T c00 = a11 * a22 - a12 * a12, c01 = a02 * a12 - a01 * a22, c02 = a01 * a12 - a02 * a11;
    const T c11 = a00 * a22 - a02 * a02, c12 = a01 * a02 - a00 * a12, c22 = a00 * a11 - a01 * a01;
    const T det = a00 * c00 + a01 * c01 + a02 * c02;
    const T x = (c00 * b0 + c01 * b1 + c02 * b2) / det;
    const T y = (c01 * b0 + c11 * b1 + c12 * b2) / det;
    const T z = (c02 * b0 + c12 * b1 + c22 * b2) / det;
    T W[9] = {T(0), -x, -y, x, T(0), -z, y, z, T(0)};  // column-major: W(0,1)=x, W(0,2)=y, W(1,2)=z
    mat_mul(R, W, out);
}

// dSvd, cpp/core/src/common/geometry.cpp:159-186 (rotation terms dropped when
// sig_i - sig_j <= 10 eps).
template <class T>
__device__ void dsvd3(const T* U, const T* sig, const T* V, const T* dF, T* dU, T* dsig, T* dV) {
    T tmp[9], dP[9];
    mat_tmul(U, dF, tmp);
    mat_mul(tmp, V, dP);
    dsig[0] = dP[0]; dsig[1] = dP[4]; dsig[2] = dP[8];
    const T eps = T(10) * Eps<T>::v;
    T wU[9], wV[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) { wU[i] = T(0); wV[i] = T(0); }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = i + 1; j < 3; ++j) {
            if (sig[i] - sig[j] > eps) {
                const T sij = T(1) / (sig[j] * sig[j] - sig[i] * sig[i]);
                const T pij = dP[i + 3 * j], pji = dP[j + 3 * i];
                wU[i + 3 * j] = sij * (pij * sig[j] + sig[i] * pji);
                wU[j + 3 * i] = -sij * (pji * sig[i] + sig[j] * pij);
                wV[i + 3 * j] = sij * (sig[i] * pij + pji * sig[j]);
                wV[j + 3 * i] = -sij * (sig[j] * pji + pij * sig[i]);
            }
        }
    mat_mul(U, wU, dU);
    mat_mul(V, wV, dV);
}

// VolumePdElementEnergy::ProjectToManifold iteration,
// cpp/core/src/pd_energy/volume_pd_element_energy.cpp:21-59: D = sig* - sig.
template <class T>
__device__ void volume_D(const T* sig, T* D) {
    const T eps = Eps<T>::v;
    const T p = cbrt(sig[0] * sig[1] * sig[2]);
#pragma unroll
    for (int i = 0; i < 3; ++i) D[i] = sig[i] / p - sig[i];
    for (int it = 0; it < 50; ++it) {
        const T s0 = sig[0] + D[0], s1 = sig[1] + D[1], s2 = sig[2] + D[2];
        const T C = s0 * s1 * s2 - T(1);
        const T g0 = s1 * s2, g1 = s0 * s2, g2 = s0 * s1;
        const T f = (g0 * D[0] + g1 * D[1] + g2 * D[2] - C) / (g0 * g0 + g1 * g1 + g2 * g2);
        const T n0 = f * g0, n1 = f * g1, n2 = f * g2;
        const T diff = fmax(fabs(n0 - D[0]), fmax(fabs(n1 - D[1]), fabs(n2 - D[2])));
        if (diff <= eps) break;
        D[0] = n0; D[1] = n1; D[2] = n2;
    }
}

// Symmetric 3x3 solve by LDL^T without pivoting (row-major H).
template <class T>
__device__ void ldlt3(const T* H, const T* b, T* x) {
    const T d0 = H[0];
    const T l10 = H[3] / d0, l20 = H[6] / d0;
    const T d1 = H[4] - l10 * l10 * d0;
    const T l21 = (H[7] - l20 * l10 * d0) / d1;
    const T d2 = H[8] - l20 * l20 * d0 - l21 * l21 * d1;
    T y0 = b[0], y1 = b[1] - l10 * y0, y2 = b[2] - l20 * y0 - l21 * y1;
    y0 /= d0; y1 /= d1; y2 /= d2;
    x[2] = y2;
    x[1] = y1 - l21 * x[2];
    x[0] = y0 - l10 * x[1] - l20 * x[2];
}

template <class T>
__device__ void nh_hessian(T mu, T la, T k, const T* s, T* H) {
    const T L = log(fmax(s[0] * s[1] * s[2], T(1e-12)));
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            H[3 * i + j] = (i == j) ? mu + (mu + la - la * L) / (s[i] * s[i]) + k : la / (s[i] * s[j]);
}

// NeoHookeanPdElementEnergy::SolveProximal,
// cpp/core/src/pd_energy/neohookean_pd_element_energy.cpp:58-103.
template <class T>
__device__ void nh_prox(T mu, T la, T k, const T* sF, T* s) {
    const T smin = T(1e-4), tol = T(1e-10);
#pragma unroll
    for (int i = 0; i < 3; ++i) s[i] = fmax(sF[i], smin);
    for (int it = 0; it < 32; ++it) {
        const T L = log(fmax(s[0] * s[1] * s[2], T(1e-12)));
        T g[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) g[i] = mu * s[i] - mu / s[i] + la * L / s[i] + k * (s[i] - sF[i]);
        if (sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]) < tol) break;
        T H[9];
        nh_hessian(mu, la, k, s, H);
        H[0] += T(1e-9); H[4] += T(1e-9); H[8] += T(1e-9);
        T d[3];
        ldlt3(H, g, d);
        T alpha = T(1);
        T sn[3] = {s[0] - d[0], s[1] - d[1], s[2] - d[2]};
        for (int ls = 0; ls < 20; ++ls) {
            if (fmin(sn[0], fmin(sn[1], sn[2])) > smin * T(0.5)) break;
            alpha *= T(0.5);
#pragma unroll
            for (int i = 0; i < 3; ++i) sn[i] = s[i] - alpha * d[i];
        }
        const T step = fmax(fabs(sn[0] - s[0]), fmax(fabs(sn[1] - s[1]), fabs(sn[2] - s[2])));
#pragma unroll
        for (int i = 0; i < 3; ++i) s[i] = fmax(sn[i], smin);
        if (step < tol) break;
    }
}

// Volume projection differential, volume_pd_element_energy.cpp:61-155.
template <class T>
__device__ void volume_diff(const T* U, const T* sig, const T* V, const T* D, const T* dF, T* out) {
    T g[3] = {(sig[1] + D[1]) * (sig[2] + D[2]), (sig[0] + D[0]) * (sig[2] + D[2]), (sig[0] + D[0]) * (sig[1] + D[1])};
    // H(i,j) = prod over k != i,j of (sig_k + D_k), zero diagonal
    const T h01 = sig[2] + D[2], h02 = sig[1] + D[1], h12 = sig[0] + D[0];
    T dU[9], dV[9], ds[3];
    dsvd3(U, sig, V, dF, dU, ds, dV);
    const T Dn = sqrt(D[0] * D[0] + D[1] * D[1] + D[2] * D[2]);
    const T gn = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
    const T la_abs = Dn / gn;
    T mneg = T(0), mpos = T(0);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        mneg += (D[i] - la_abs * g[i]) * (D[i] - la_abs * g[i]);
        mpos += (D[i] + la_abs * g[i]) * (D[i] + la_abs * g[i]);
    }
    const T la = (mneg < mpos) ? -la_abs : la_abs;
    // [I + la H, g; g^T, 0] [dD; dla] = [-la H ds; -g.ds], solved by elimination with partial pivoting
    T M[4][5];
    const T Hm[3][3] = {{T(0), h01, h02}, {h01, T(0), h12}, {h02, h12, T(0)}};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
#pragma unroll
        for (int j = 0; j < 3; ++j) M[i][j] = (i == j ? T(1) : T(0)) + la * Hm[i][j];
        M[i][3] = g[i];
        M[i][4] = -la * (Hm[i][0] * ds[0] + Hm[i][1] * ds[1] + Hm[i][2] * ds[2]);
    }
    M[3][0] = g[0]; M[3][1] = g[1]; M[3][2] = g[2]; M[3][3] = T(0);
    M[3][4] = -(g[0] * ds[0] + g[1] * ds[1] + g[2] * ds[2]);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        int p = c;
#pragma unroll
        for (int r = c + 1; r < 4; ++r) if (fabs(M[r][c]) > fabs(M[p][c])) p = r;
        if (p != c) {
#pragma unroll
            for (int j = 0; j < 5; ++j) { const T t = M[c][j]; M[c][j] = M[p][j]; M[p][j] = t; }
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            if (r == c) continue;
            const T f = M[r][c] / M[c][c];
#pragma unroll
            for (int j = 0; j < 5; ++j) M[r][j] -= f * M[c][j];
        }
    }
    const T dD[3] = {M[0][4] / M[0][0], M[1][4] / M[1][1], M[2][4] / M[2][2]};
    T t1[9], t2[9], t3[9];
    udv(dU, D, V, t1);
    udv(U, dD, V, t2);
    udv(U, D, dV, t3);
#pragma unroll
    for (int i = 0; i < 9; ++i) out[i] = dF[i] + t1[i] + t2[i] + t3[i];
}

// Neo-Hookean proximal-map differential, neohookean_pd_element_energy.cpp:128-162.
template <class T>
__device__ void nh_diff(T mu, T la, T k, const T* U, const T* sig, const T* V, const T* ss, const T* dF, T* out) {
    T dU[9], dV[9], dsF[3];
    dsvd3(U, sig, V, dF, dU, dsF, dV);
    T H[9];
    nh_hessian(mu, la, k, ss, H);
    H[0] += T(1e-9); H[4] += T(1e-9); H[8] += T(1e-9);
    const T rhs[3] = {k * dsF[0], k * dsF[1], k * dsF[2]};
    T dss[3];
    ldlt3(H, rhs, dss);
    T t1[9], t2[9], t3[9];
    udv(dU, ss, V, t1);
    udv(U, dss, V, t2);
    udv(U, ss, dV, t3);
#pragma unroll
    for (int i = 0; i < 9; ++i) out[i] = t1[i] + t2[i] + t3[i];
}


// Closest point of triangle (a, b, c) to p by its Voronoi regions (Ericson,
// Real-Time Collision Detection, 5.1.5): vertex regions, then edge regions,
// then the face, each decided by signs of dot products of the edge vectors
// with p's offsets (the region order and tests of the reference's
// MeshFrictionalBoundary, so that ties resolve alike).
__host__ __device__ inline void closest_on_triangle(const double* p, const double* a, const double* b, const double* c,
                                                    double* out, double* wt = nullptr) {
    double ab[3], ac[3], ap[3], bp[3], cp[3];
    for (int i = 0; i < 3; ++i) {
        ab[i] = b[i] - a[i]; ac[i] = c[i] - a[i];
        ap[i] = p[i] - a[i]; bp[i] = p[i] - b[i]; cp[i] = p[i] - c[i];
    }
    auto dot = [](const double* x, const double* y) { return x[0] * y[0] + x[1] * y[1] + x[2] * y[2]; };
    auto weights = [&](double wa, double wb, double wc) { if (wt) { wt[0] = wa; wt[1] = wb; wt[2] = wc; } };
    const double d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0 && d2 <= 0) { for (int i = 0; i < 3; ++i) out[i] = a[i]; weights(1, 0, 0); return; }
    const double d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0 && d4 <= d3) { for (int i = 0; i < 3; ++i) out[i] = b[i]; weights(0, 1, 0); return; }
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) {
        const double t = d1 / (d1 - d3);
        for (int i = 0; i < 3; ++i) out[i] = a[i] + t * ab[i];
        weights(1 - t, t, 0);
        return;
    }
    const double d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0 && d5 <= d6) { for (int i = 0; i < 3; ++i) out[i] = c[i]; weights(0, 0, 1); return; }
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) {
        const double t = d2 / (d2 - d6);
        for (int i = 0; i < 3; ++i) out[i] = a[i] + t * ac[i];
        weights(1 - t, 0, t);
        return;
    }
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
        const double t = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        for (int i = 0; i < 3; ++i) out[i] = b[i] + t * (c[i] - b[i]);
        weights(0, 1 - t, t);
        return;
    }
    const double den = 1 / (va + vb + vc);
    const double v = vb * den, w = vc * den;
    for (int i = 0; i < 3; ++i) out[i] = a[i] + ab[i] * v + ac[i] * w;
    weights(1 - v - w, v, w);
}

}  // namespace phd
