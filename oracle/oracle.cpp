// =============================================================================
// oracle.cpp -- FP64 CPU restatement of DiffPhD's `pd_eigen_alg_phd` projective
// dynamics (PD) forward/backward time step.
//
// TEST INFRASTRUCTURE ONLY. This file is the correctness oracle: it is compiled
// into oracle/_build/liboracle.so and loaded solely by tests/, by
// __graft_entry__.smoke() and by bench.py's cpu_baseline leg, always as the
// checker and never as the product path.
//
// Every function cites the reference file:line it restates (paths relative to
// the reference repository root). The arithmetic is plain FP64 with no external
// libraries: own 3x3 one-sided Jacobi SVD, own dense LDL^T/Cholesky and an own
// up-looking sparse LDL^T under a recursive-coordinate-bisection nested
// dissection ordering (the reference uses METIS NodeND; the ordering changes
// only round-off because S^T S = A^-1 for any symmetric permutation).
// =============================================================================
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <deque>
#include <limits>
#include <map>
#include <numeric>
#include <set>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

namespace orc {

using Vec = std::vector<double>;

static void check(bool ok, const std::string& msg) {
    if (!ok) throw std::runtime_error(msg);
}

// ---------------------------------------------------------------------------
// Small dense helpers (column-major 3x3 like Eigen: M[i + 3*j] = M(i, j)).
// ---------------------------------------------------------------------------
struct M3 {
    double a[9];
    double& operator()(int i, int j) { return a[i + 3 * j]; }
    double operator()(int i, int j) const { return a[i + 3 * j]; }
    static M3 zero() { M3 m; for (double& x : m.a) x = 0; return m; }
    static M3 eye() { M3 m = zero(); m(0, 0) = m(1, 1) = m(2, 2) = 1; return m; }
};
static M3 mul(const M3& A, const M3& B) {
    M3 C = M3::zero();
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0;
            for (int k = 0; k < 3; ++k) s += A(i, k) * B(k, j);
            C(i, j) = s;
        }
    return C;
}
static M3 tr(const M3& A) {
    M3 B;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) B(i, j) = A(j, i);
    return B;
}
static M3 add(const M3& A, const M3& B) { M3 C; for (int i = 0; i < 9; ++i) C.a[i] = A.a[i] + B.a[i]; return C; }
static M3 scale(const M3& A, double s) { M3 C; for (int i = 0; i < 9; ++i) C.a[i] = A.a[i] * s; return C; }
static M3 diag_scale(const M3& U, const double* d, const M3& V) {  // U diag(d) V^T
    M3 C = M3::zero();
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0;
            for (int k = 0; k < 3; ++k) s += U(i, k) * d[k] * V(j, k);
            C(i, j) = s;
        }
    return C;
}
static double det3(const M3& A) {
    return A(0, 0) * (A(1, 1) * A(2, 2) - A(1, 2) * A(2, 1))
         - A(0, 1) * (A(1, 0) * A(2, 2) - A(1, 2) * A(2, 0))
         + A(0, 2) * (A(1, 0) * A(2, 1) - A(1, 1) * A(2, 0));
}
static M3 inv3(const M3& A) {
    const double d = det3(A);
    M3 B;
    B(0, 0) = (A(1, 1) * A(2, 2) - A(1, 2) * A(2, 1)) / d;
    B(0, 1) = (A(0, 2) * A(2, 1) - A(0, 1) * A(2, 2)) / d;
    B(0, 2) = (A(0, 1) * A(1, 2) - A(0, 2) * A(1, 1)) / d;
    B(1, 0) = (A(1, 2) * A(2, 0) - A(1, 0) * A(2, 2)) / d;
    B(1, 1) = (A(0, 0) * A(2, 2) - A(0, 2) * A(2, 0)) / d;
    B(1, 2) = (A(0, 2) * A(1, 0) - A(0, 0) * A(1, 2)) / d;
    B(2, 0) = (A(1, 0) * A(2, 1) - A(1, 1) * A(2, 0)) / d;
    B(2, 1) = (A(0, 1) * A(2, 0) - A(0, 0) * A(2, 1)) / d;
    B(2, 2) = (A(0, 0) * A(1, 1) - A(0, 1) * A(1, 0)) / d;
    return B;
}

// Dense symmetric positive (semi)definite solve by LDL^T without pivoting on a
// row-major n x n matrix. Stands in for Eigen's `.ldlt().solve()` on the small
// SPD systems of the path (M_sys, AA Gram, G_T, NH prox Hessian); for an SPD
// matrix every stable LDL^T gives the same solution up to round-off.
static Vec dense_ldlt_solve(std::vector<double> M, int n, const Vec& b) {
    Vec d(n, 0.0);
    for (int j = 0; j < n; ++j) {
        double djj = M[j * n + j];
        for (int k = 0; k < j; ++k) djj -= M[j * n + k] * M[j * n + k] * d[k];
        d[j] = djj;
        for (int i = j + 1; i < n; ++i) {
            double s = M[i * n + j];
            for (int k = 0; k < j; ++k) s -= M[i * n + k] * M[j * n + k] * d[k];
            M[i * n + j] = (djj != 0.0) ? s / djj : 0.0;
        }
    }
    Vec x = b;
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < i; ++k) x[i] -= M[i * n + k] * x[k];
    for (int i = 0; i < n; ++i) x[i] = (d[i] != 0.0) ? x[i] / d[i] : 0.0;
    for (int i = n - 1; i >= 0; --i)
        for (int k = i + 1; k < n; ++k) x[i] -= M[k * n + i] * x[k];
    return x;
}

static double norm(const Vec& v) { double s = 0; for (double x : v) s += x * x; return std::sqrt(s); }
static double dot(const Vec& a, const Vec& b) { double s = 0; for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i]; return s; }

// ---------------------------------------------------------------------------
// 3x3 SVD by one-sided Jacobi: F V = U diag(sig), sig sorted descending and
// non-negative (the conventions of Eigen::JacobiSVD used at
// cpp/core/src/pd_energy/deformation_gradient_auxiliary_data.cpp:6 and
// cpp/core/src/common/geometry.cpp:25,123).
// ---------------------------------------------------------------------------
void svd3(const M3& F, M3& U, double sig[3], M3& V) {
    M3 B = F;
    V = M3::eye();
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0, scale_sq = 0;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                double alpha = 0, beta = 0, gamma = 0;
                for (int i = 0; i < 3; ++i) {
                    alpha += B(i, p) * B(i, p);
                    beta += B(i, q) * B(i, q);
                    gamma += B(i, p) * B(i, q);
                }
                scale_sq = std::max(scale_sq, alpha * beta);
                if (std::fabs(gamma) <= 1e-300) continue;
                off = std::max(off, gamma * gamma / std::max(alpha * beta, 1e-300));
                const double zeta = (beta - alpha) / (2 * gamma);
                const double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1 + zeta * zeta));
                const double c = 1 / std::sqrt(1 + t * t), s = c * t;
                for (int i = 0; i < 3; ++i) {
                    const double bp = B(i, p), bq = B(i, q);
                    B(i, p) = c * bp - s * bq;
                    B(i, q) = s * bp + c * bq;
                    const double vp = V(i, p), vq = V(i, q);
                    V(i, p) = c * vp - s * vq;
                    V(i, q) = s * vp + c * vq;
                }
            }
        (void)scale_sq;
        if (off < 1e-32) break;
    }
    double nrm[3];
    for (int j = 0; j < 3; ++j) nrm[j] = std::sqrt(B(0, j) * B(0, j) + B(1, j) * B(1, j) + B(2, j) * B(2, j));
    int ord[3] = {0, 1, 2};
    std::sort(ord, ord + 3, [&](int a, int b) { return nrm[a] > nrm[b]; });
    M3 Bs, Vs;
    for (int j = 0; j < 3; ++j) {
        sig[j] = nrm[ord[j]];
        for (int i = 0; i < 3; ++i) { Bs(i, j) = B(i, ord[j]); Vs(i, j) = V(i, ord[j]); }
    }
    V = Vs;
    U = M3::zero();
    // Columns with a vanishing singular value are completed to an orthonormal
    // basis (sig is sorted, so only trailing columns can vanish).
    const double tiny = 1e-300;
    int valid = 0;
    for (int j = 0; j < 3; ++j) {
        if (!(sig[j] > tiny)) break;
        for (int i = 0; i < 3; ++i) U(i, j) = Bs(i, j) / sig[j];
        ++valid;
    }
    if (valid == 0) { U = M3::eye(); return; }
    if (valid == 1) {
        // any unit vector orthogonal to U0: cross with the least-aligned axis
        int ax = 0;
        for (int d = 1; d < 3; ++d) if (std::fabs(U(d, 0)) < std::fabs(U(ax, 0))) ax = d;
        double e[3] = {0, 0, 0};
        e[ax] = 1;
        double c[3] = {U(1, 0) * e[2] - U(2, 0) * e[1], U(2, 0) * e[0] - U(0, 0) * e[2], U(0, 0) * e[1] - U(1, 0) * e[0]};
        const double cn = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
        for (int i = 0; i < 3; ++i) U(i, 1) = c[i] / cn;
        valid = 2;
    }
    if (valid == 2) {
        U(0, 2) = U(1, 0) * U(2, 1) - U(2, 0) * U(1, 1);
        U(1, 2) = U(2, 0) * U(0, 1) - U(0, 0) * U(2, 1);
        U(2, 2) = U(0, 0) * U(1, 1) - U(1, 0) * U(0, 1);
    }
}

// PolarDecomposition, cpp/core/src/common/geometry.cpp:24-31: R = U V^T, S = V Sig V^T.
void polar3(const M3& F, M3& R, M3& S) {
    M3 U, V;
    double s[3];
    svd3(F, U, s, V);
    R = mul(U, tr(V));
    S = diag_scale(V, s, V);
}

// dRFromdF, cpp/core/src/common/geometry.cpp:85-113.
M3 dR_from_dF(const M3& R, const M3& S, const M3& dF) {
    const M3 lhs = add(mul(tr(R), dF), scale(mul(tr(dF), R), -1.0));
    M3 A = M3::zero();
    A(0, 0) = S(0, 0) + S(1, 1);
    A(1, 1) = S(0, 0) + S(2, 2);
    A(2, 2) = S(1, 1) + S(2, 2);
    A(0, 1) = A(1, 0) = S(1, 2);
    A(0, 2) = A(2, 0) = -S(0, 2);
    A(1, 2) = A(2, 1) = S(0, 1);
    const M3 Ai = inv3(A);
    const double b[3] = {lhs(0, 1), lhs(0, 2), lhs(1, 2)};
    double xyz[3];
    for (int i = 0; i < 3; ++i) xyz[i] = Ai(i, 0) * b[0] + Ai(i, 1) * b[1] + Ai(i, 2) * b[2];
    M3 W = M3::zero();
    W(0, 1) = xyz[0]; W(0, 2) = xyz[1];
    W(1, 0) = -xyz[0]; W(1, 2) = xyz[2];
    W(2, 0) = -xyz[1]; W(2, 1) = -xyz[2];
    return mul(R, W);
}

// dSvd, cpp/core/src/common/geometry.cpp:159-186 (rotation terms dropped when
// sig_i - sig_j <= 10 * DBL_EPSILON).
void dsvd3(const M3& U, const double sig[3], const M3& V, const M3& dF, M3& dU, double dsig[3], M3& dV) {
    const M3 dP = mul(mul(tr(U), dF), V);
    for (int i = 0; i < 3; ++i) dsig[i] = dP(i, i);
    const double eps = 10 * std::numeric_limits<double>::epsilon();
    M3 Sij = M3::zero();
    for (int i = 0; i < 3; ++i)
        for (int j = i + 1; j < 3; ++j)
            if (sig[i] - sig[j] > eps) {
                Sij(i, j) = 1.0 / (sig[j] * sig[j] - sig[i] * sig[i]);
                Sij(j, i) = -Sij(i, j);
            }
    M3 dwU, dwV;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            dwU(i, j) = Sij(i, j) * (dP(i, j) * sig[j] + sig[i] * dP(j, i));
            dwV(i, j) = Sij(i, j) * (sig[i] * dP(i, j) + dP(j, i) * sig[j]);
        }
    dU = mul(U, dwU);
    dV = mul(V, dwV);
}

// ---------------------------------------------------------------------------
// PD element energies (cpp/core/src/pd_energy/*.cpp).
// ---------------------------------------------------------------------------
enum EnergyType { kCorotated = 0, kVolume = 1, kNeoHookean = 2 };

struct Energy {
    int type = kCorotated;
    double stiffness = 0;              // stiffness_(), mean when per-element
    std::vector<double> elem_stiff;    // element_stiffnesses_ (empty: homogeneous)
    double mu = 0, la = 0;             // NH mean Lame parameters
    std::vector<char> member;          // element in this energy's index set
    double elem_w(int e) const { return elem_stiff.empty() ? stiffness : elem_stiff[e]; }
};

// VolumePdElementEnergy::ProjectToManifold, volume_pd_element_energy.cpp:21-59.
static void volume_project_D(const double sig[3], double D[3]) {
    const double eps = std::numeric_limits<double>::epsilon();
    check(std::min(sig[0], std::min(sig[1], sig[2])) >= eps, "Singular F.");
    const double p = std::pow(sig[0] * sig[1] * sig[2], 1.0 / 3.0);
    for (int i = 0; i < 3; ++i) D[i] = sig[i] / p - sig[i];
    for (int it = 0; it < 50; ++it) {
        const double C = (sig[0] + D[0]) * (sig[1] + D[1]) * (sig[2] + D[2]) - 1;
        double g[3];
        for (int j = 0; j < 3; ++j) {
            g[j] = 1;
            for (int k = 0; k < 3; ++k) if (k != j) g[j] *= (sig[k] + D[k]);
        }
        const double gD = g[0] * D[0] + g[1] * D[1] + g[2] * D[2];
        const double gg = g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
        double Dn[3], diff = 0;
        for (int j = 0; j < 3; ++j) { Dn[j] = (gD - C) / gg * g[j]; diff = std::max(diff, std::fabs(Dn[j] - D[j])); }
        if (diff <= eps) break;
        for (int j = 0; j < 3; ++j) D[j] = Dn[j];
    }
}

// NeoHookeanPdElementEnergy::HessianProximal, neohookean_pd_element_energy.cpp:37-56.
static void nh_hessian(const Energy& en, const double s[3], double H[9]) {
    const double k = en.stiffness;
    const double L = std::log(std::max(s[0] * s[1] * s[2], 1e-12));
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            H[i * 3 + j] = (i == j) ? en.mu + (en.mu + en.la - en.la * L) / (s[i] * s[i]) + k
                                    : en.la / (s[i] * s[j]);
}

// NeoHookeanPdElementEnergy::SolveProximal, neohookean_pd_element_energy.cpp:58-103.
void nh_solve_prox(const Energy& en, const double sF[3], double s[3]) {
    const double k = en.stiffness, smin = 1e-4, tol = 1e-10;
    for (int i = 0; i < 3; ++i) s[i] = std::max(sF[i], smin);
    for (int it = 0; it < 32; ++it) {
        const double L = std::log(std::max(s[0] * s[1] * s[2], 1e-12));
        Vec g(3);
        for (int i = 0; i < 3; ++i)
            g[i] = en.mu * s[i] - en.mu / s[i] + en.la * L / s[i] + k * (s[i] - sF[i]);
        if (norm(g) < tol) break;
        double H[9];
        nh_hessian(en, s, H);
        for (int i = 0; i < 3; ++i) H[i * 4] += 1e-9;
        const Vec delta = dense_ldlt_solve(std::vector<double>(H, H + 9), 3, g);
        double alpha = 1, sn[3];
        for (int i = 0; i < 3; ++i) sn[i] = s[i] - alpha * delta[i];
        for (int ls = 0; ls < 20; ++ls) {
            const double mn = std::min(sn[0], std::min(sn[1], sn[2]));
            if (mn > smin * 0.5) break;
            alpha *= 0.5;
            for (int i = 0; i < 3; ++i) sn[i] = s[i] - alpha * delta[i];
        }
        double step = 0;
        for (int i = 0; i < 3; ++i) step = std::max(step, std::fabs(sn[i] - s[i]));
        for (int i = 0; i < 3; ++i) s[i] = std::max(sn[i], smin);
        if (step < tol) break;
    }
}

// F auxiliary data (deformation_gradient_auxiliary_data.cpp:3-12).
struct FAux {
    M3 F, U, V, R, S;
    double sig[3];
    void init(const M3& F_) {
        F = F_;
        svd3(F, U, sig, V);
        R = mul(U, tr(V));
        S = diag_scale(V, sig, V);
    }
};

// ProjectToManifold(F_auxiliary): corotated_pd_element_energy.cpp:21-24,
// volume_pd_element_energy.cpp:21-59, neohookean_pd_element_energy.cpp:112-119.
// Symmetric eigen-decomposition of an n x n matrix (n <= 9, row-major A,
// overwritten) by cyclic Jacobi rotations: eigenvalues in w, eigenvectors in
// the columns of Z (row-major). Textbook two-sided Jacobi (Golub & Van Loan,
// Matrix Computations, sec. 8.5), run until the off-diagonal mass is zero to
// round-off.
void jacobi_eigen(int n, double* A, double* w, double* Z) {
    for (int i = 0; i < n * n; ++i) Z[i] = 0;
    for (int i = 0; i < n; ++i) Z[i * n + i] = 1;
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0, tot = 0;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                tot += A[i * n + j] * A[i * n + j];
                if (i != j) off += A[i * n + j] * A[i * n + j];
            }
        if (off <= 1e-32 * tot || off == 0) break;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = A[p * n + q];
                if (apq == 0) continue;
                const double theta = (A[q * n + q] - A[p * n + p]) / (2 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1));
                const double c = 1 / std::sqrt(t * t + 1), sn = t * c;
                for (int k = 0; k < n; ++k) {  // A <- A J (columns p, q)
                    const double akp = A[k * n + p], akq = A[k * n + q];
                    A[k * n + p] = c * akp - sn * akq;
                    A[k * n + q] = sn * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {  // A <- J^T A (rows p, q)
                    const double apk = A[p * n + k], aqk = A[q * n + k];
                    A[p * n + k] = c * apk - sn * aqk;
                    A[q * n + k] = sn * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {  // Z <- Z J
                    const double zkp = Z[k * n + p], zkq = Z[k * n + q];
                    Z[k * n + p] = c * zkp - sn * zkq;
                    Z[k * n + q] = sn * zkp + c * zkq;
                }
            }
    }
    for (int i = 0; i < n; ++i) w[i] = A[i * n + i];
}

// Spectral filter of one sample's weighted proximal-map Jacobian M (9 x 9,
// column-major, M = sum_energies k_e dBp/dF) relative to the sample's weight
// in A, aw = sum_energies w_e vol / s. With mu an eigenvalue of sym(M) / aw
// (eta = 1 - mu is the eigenvalue of the sample's share of the Hessian A - K):
//   mu > 1:  mu <- 1 + theta (mu - 1)   (theta = 1 leaves the implicit-function
//            operator as it is, 0 clamps eta at 0, -1 reflects eta to |eta|:
//            the clamp (eps = 0) and abs filters of ProjectedNewtonMatrix,
//            deformable_newton_forward.cpp:408-433, applied per sample block)
//   then mu <- max(mu, lo)               (lo = -rho bounds the backbone operator
//            A^-1 K from below; -inf for the plain clamp / abs filters).
// With theta <= 0 and lo = -rho the block's spectrum lies in [-rho, 1], so
// |x^T K x| <= x^T (A - M/h^2) x and A^-1 K is a contraction.
// M is left untouched (bit for bit) when no eigenvalue moves, as the reference
// re-assembles only when an eigenvalue was modified (:431-433). Returns the
// largest eigenvalue shift (0: unchanged).
double filter_prox_jacobian(double* M, double aw, double theta, double lo) {
    if (!(aw > 0)) return 0;
    double A[81], w[9], Z[81];
    for (int i = 0; i < 9; ++i)
        for (int j = 0; j < 9; ++j) A[i * 9 + j] = 0.5 * (M[i + 9 * j] + M[j + 9 * i]) / aw;
    jacobi_eigen(9, A, w, Z);
    double shift = 0;
    bool changed = false;
    for (int k = 0; k < 9; ++k) {
        double mu = w[k] > 1 ? 1 + theta * (w[k] - 1) : w[k];
        mu = std::max(mu, lo);
        if (mu != w[k]) { changed = true; shift = std::max(shift, std::fabs(mu - w[k])); w[k] = mu; }
    }
    if (!changed) return 0;
    for (int i = 0; i < 9; ++i)
        for (int j = 0; j < 9; ++j) {
            double acc = 0;
            for (int k = 0; k < 9; ++k) acc += Z[i * 9 + k] * w[k] * Z[j * 9 + k];
            M[i + 9 * j] = aw * acc;
        }
    return shift > 0 ? shift : 1e-300;
}

// (theta, lo) of a filter mode: 1 clamp, 2 abs, 3 trust region at stage
// `stage` (0: the unfiltered operator, 1: clamp with lo = -rho, 2: abs with lo = -rho).
static void filter_params(int mode, int stage, double rho, double& theta, double& lo, bool& active) {
    active = true;
    lo = -std::numeric_limits<double>::infinity();
    if (mode == 1) { theta = 0; return; }
    if (mode == 2) { theta = -1; return; }
    if (mode == 3 && stage > 0) { theta = stage == 1 ? 0.0 : -1.0; lo = -rho; return; }
    active = false;
    theta = 1;
}

M3 project(const Energy& en, const FAux& a) {
    if (en.type == kCorotated) return a.R;
    if (en.type == kVolume) {
        double D[3], sd[3];
        volume_project_D(a.sig, D);
        for (int i = 0; i < 3; ++i) sd[i] = a.sig[i] + D[i];
        return diag_scale(a.U, sd, a.V);
    }
    double ss[3];
    nh_solve_prox(en, a.sig, ss);
    return diag_scale(a.U, ss, a.V);
}

// ProjectToManifoldDifferential(F_auxiliary, projection, dF):
// corotated_pd_element_energy.cpp:27-31, volume_pd_element_energy.cpp:61-155,
// neohookean_pd_element_energy.cpp:128-162.
M3 project_diff(const Energy& en, const FAux& a, const M3& dF) {
    if (en.type == kCorotated) return dR_from_dF(a.R, a.S, dF);
    const double* sig = a.sig;
    if (en.type == kVolume) {
        double D[3];
        volume_project_D(sig, D);
        double g[3];
        for (int j = 0; j < 3; ++j) { g[j] = 1; for (int k = 0; k < 3; ++k) if (k != j) g[j] *= (sig[k] + D[k]); }
        double H[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                if (i == j) { H[i * 3 + j] = 0; continue; }
                double h = 1;
                for (int k = 0; k < 3; ++k) if (k != i && k != j) h *= (sig[k] + D[k]);
                H[i * 3 + j] = h;
            }
        M3 dU, dV;
        double dsig[3];
        dsvd3(a.U, sig, a.V, dF, dU, dsig, dV);
        const double Dn = std::sqrt(D[0] * D[0] + D[1] * D[1] + D[2] * D[2]);
        const double gn = std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
        const double la_abs = Dn / gn;
        double mneg = 0, mpos = 0;
        for (int i = 0; i < 3; ++i) {
            mneg += (D[i] - la_abs * g[i]) * (D[i] - la_abs * g[i]);
            mpos += (D[i] + la_abs * g[i]) * (D[i] + la_abs * g[i]);
        }
        const double la = (mneg < mpos) ? -la_abs : la_abs;
        // [I + la H, g; g^T, 0] [dD; dla] = [-la H dsig; -g.dsig]
        double A[16] = {0};
        for (int i = 0; i < 3; ++i) {
            for (int j = 0; j < 3; ++j) A[i * 4 + j] = (i == j ? 1.0 : 0.0) + la * H[i * 3 + j];
            A[i * 4 + 3] = g[i];
            A[3 * 4 + i] = g[i];
        }
        double rhs[4];
        for (int i = 0; i < 3; ++i) {
            double s = 0;
            for (int j = 0; j < 3; ++j) s += H[i * 3 + j] * dsig[j];
            rhs[i] = -la * s;
        }
        rhs[3] = -(g[0] * dsig[0] + g[1] * dsig[1] + g[2] * dsig[2]);
        // Gaussian elimination with partial pivoting (stands in for A.inverse() * b).
        double Mx[4][5];
        for (int i = 0; i < 4; ++i) { for (int j = 0; j < 4; ++j) Mx[i][j] = A[i * 4 + j]; Mx[i][4] = rhs[i]; }
        for (int c = 0; c < 4; ++c) {
            int p = c;
            for (int r = c + 1; r < 4; ++r) if (std::fabs(Mx[r][c]) > std::fabs(Mx[p][c])) p = r;
            for (int j = 0; j < 5; ++j) std::swap(Mx[c][j], Mx[p][j]);
            for (int r = 0; r < 4; ++r) {
                if (r == c) continue;
                const double f = Mx[r][c] / Mx[c][c];
                for (int j = c; j < 5; ++j) Mx[r][j] -= f * Mx[c][j];
            }
        }
        double dD[3];
        for (int i = 0; i < 3; ++i) dD[i] = Mx[i][4] / Mx[i][i];
        M3 out = dF;
        out = add(out, diag_scale(dU, D, a.V));
        out = add(out, diag_scale(a.U, dD, a.V));
        out = add(out, diag_scale(a.U, D, dV));
        return out;
    }
    // Neo-Hookean proximal map.
    double ss[3];
    nh_solve_prox(en, sig, ss);
    M3 dU, dV;
    double dsF[3];
    dsvd3(a.U, sig, a.V, dF, dU, dsF, dV);
    double H[9];
    nh_hessian(en, ss, H);
    for (int i = 0; i < 3; ++i) H[i * 4] += 1e-9;
    Vec rhs(3);
    for (int i = 0; i < 3; ++i) rhs[i] = en.stiffness * dsF[i];
    const Vec dss = dense_ldlt_solve(std::vector<double>(H, H + 9), 3, rhs);
    const double dsd[3] = {dss[0], dss[1], dss[2]};
    M3 out = diag_scale(dU, ss, a.V);
    out = add(out, diag_scale(a.U, dsd, a.V));
    out = add(out, diag_scale(a.U, ss, dV));
    return out;
}

// StressTensor(F_auxiliary, projection): pd_element_energy.cpp:46-50 for the
// PD energies (stiffness_ * (F - p)); for pd_neohookean the true Neo-Hookean
// PK1 with the mean Lame parameters, neohookean_pd_element_energy.cpp:188-197.
M3 stress(const Energy& en, const FAux& a, const M3& p) {
    if (en.type != kNeoHookean) return scale(add(a.F, scale(p, -1.0)), en.stiffness);
    const M3& F = a.F;
    const double J = det3(F);
    const double lj = std::log(std::max(J, 1e-12));
    const M3 FinvT = tr(inv3(F));
    return add(scale(add(F, scale(FinvT, -1.0)), en.mu), scale(FinvT, en.la * lj));
}

// ---------------------------------------------------------------------------
// Frictional boundaries (friction/*.cpp, deformable_friction.cpp:7-89) plus
// the heightfield ground and the two-obstacle press used by the contact-rich
// press scene (no reference counterpart: the reference has one planar or
// spherical boundary per body).
// ---------------------------------------------------------------------------
enum BoundaryType { kNone = 0, kPlanar = 1, kSpherical = 2, kHeightfield = 3, kPress = 4, kMesh = 5 };

// Closest point of triangle (a, b, c) to p (mesh_frictional_boundary.cpp:19-63,
// the Voronoi-region method of Ericson, Real-Time Collision Detection 5.1.5):
// vertex regions a, b, the edge ab, vertex c, the edge ac, the edge bc, then
// the face interior by barycentric weights.
void tri_closest(const double* p, const double* a, const double* b, const double* c, double* out, double* wt = nullptr) {
    double ab[3], ac[3], ap[3], bp[3], cp[3];
    for (int i = 0; i < 3; ++i) {
        ab[i] = b[i] - a[i]; ac[i] = c[i] - a[i];
        ap[i] = p[i] - a[i]; bp[i] = p[i] - b[i]; cp[i] = p[i] - c[i];
    }
    auto dot = [](const double* x, const double* y) { return x[0] * y[0] + x[1] * y[1] + x[2] * y[2]; };
    auto wts = [&](double wa, double wb, double wc) { if (wt) { wt[0] = wa; wt[1] = wb; wt[2] = wc; } };
    const double d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0 && d2 <= 0) { for (int i = 0; i < 3; ++i) out[i] = a[i]; wts(1, 0, 0); return; }
    const double d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0 && d4 <= d3) { for (int i = 0; i < 3; ++i) out[i] = b[i]; wts(0, 1, 0); return; }
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) { const double t = d1 / (d1 - d3); for (int i = 0; i < 3; ++i) out[i] = a[i] + t * ab[i]; wts(1 - t, t, 0); return; }
    const double d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0 && d5 <= d6) { for (int i = 0; i < 3; ++i) out[i] = c[i]; wts(0, 0, 1); return; }
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) { const double t = d2 / (d2 - d6); for (int i = 0; i < 3; ++i) out[i] = a[i] + t * ac[i]; wts(1 - t, 0, t); return; }
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
        const double t = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        for (int i = 0; i < 3; ++i) out[i] = b[i] + t * (c[i] - b[i]);
        wts(0, 1 - t, t);
        return;
    }
    const double den = 1 / (va + vb + vc);
    for (int i = 0; i < 3; ++i) out[i] = a[i] + ab[i] * (vb * den) + ac[i] * (vc * den);
    wts(1 - vb * den - vc * den, vb * den, vc * den);
}

struct Boundary {
    int type = kNone;
    double n[3] = {0, 0, 1}, off = 0;     // planar
    double c[3] = {0, 0, 0}, r = 0;       // spherical
    double z0 = 0;                         // heightfield base
    std::vector<double> waves;             // (amp, kx, ky, phase) quadruples
    double plane_z = 0;                    // press plane height (normal -z)
    // mesh (MeshFrictionalBoundary): the body's own left / right triangle sets
    double radius = 0;
    int split = 0, nfl = 0, nfr = 0;
    std::vector<int> faces;

    // Closest point to p on the set opposite to `node` (right set for node <
    // split, else left) at positions X, and that face's unit normal; false for
    // an empty set (mesh_frictional_boundary.cpp:126-160, 176-212).
    bool mesh_closest(const double* p, const Vec& X, int node, double cl[3], double fn[3]) const {
        const bool right = node < split;
        const int f0 = right ? nfl : 0, nf = right ? nfr : nfl;
        double best = std::numeric_limits<double>::infinity();
        int bf = -1;
        for (int f = 0; f < nf; ++f) {
            const int* fv = &faces[3 * (size_t)(f0 + f)];
            double c[3];
            tri_closest(p, &X[3 * (size_t)fv[0]], &X[3 * (size_t)fv[1]], &X[3 * (size_t)fv[2]], c);
            const double d2 = (p[0] - c[0]) * (p[0] - c[0]) + (p[1] - c[1]) * (p[1] - c[1]) + (p[2] - c[2]) * (p[2] - c[2]);
            if (d2 < best) { best = d2; bf = f; for (int i = 0; i < 3; ++i) cl[i] = c[i]; }
        }
        if (bf < 0) return false;
        const int* fv = &faces[3 * (size_t)(f0 + bf)];
        const double *A = &X[3 * (size_t)fv[0]], *B = &X[3 * (size_t)fv[1]], *C = &X[3 * (size_t)fv[2]];
        const double e1[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]}, e2[3] = {C[0] - A[0], C[1] - A[1], C[2] - A[2]};
        fn[0] = e1[1] * e2[2] - e1[2] * e2[1]; fn[1] = e1[2] * e2[0] - e1[0] * e2[2]; fn[2] = e1[0] * e2[1] - e1[1] * e2[0];
        const double fl = std::sqrt(fn[0] * fn[0] + fn[1] * fn[1] + fn[2] * fn[2]);
        if (fl > std::numeric_limits<double>::epsilon()) for (int i = 0; i < 3; ++i) fn[i] /= fl;
        else { fn[0] = 0; fn[1] = 0; fn[2] = 1; }
        return true;
    }
    // GetClosestPointAndNormal (:176-212) and the Gram-Schmidt tangents of the
    // alg_phd forward (forward_gpu.cpp:496-509); returns the unsigned distance
    // minus the radius (GetDistance), +inf for an empty opposite set.
    double mesh_frame(const double* p, const Vec& X, int node, double t1[3], double t2[3], double nn[3]) const {
        double cl[3], fn[3];
        if (!mesh_closest(p, X, node, cl, fn)) {
            nn[0] = 0; nn[1] = 0; nn[2] = 1; t1[0] = 1; t1[1] = 0; t1[2] = 0; t2[0] = 0; t2[1] = 1; t2[2] = 0;
            return std::numeric_limits<double>::infinity();
        }
        const double dp[3] = {p[0] - cl[0], p[1] - cl[1], p[2] - cl[2]};
        const double dn = std::sqrt(dp[0] * dp[0] + dp[1] * dp[1] + dp[2] * dp[2]);
        for (int i = 0; i < 3; ++i) nn[i] = dn > 1e-12 ? dp[i] / dn : fn[i];
        int ax = 0;
        for (int d = 1; d < 3; ++d) if (std::fabs(nn[d]) < std::fabs(nn[ax])) ax = d;
        double e[3] = {0, 0, 0};
        e[ax] = 1;
        const double ne = nn[ax];
        double u[3] = {e[0] - ne * nn[0], e[1] - ne * nn[1], e[2] - ne * nn[2]};
        const double ul = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
        for (int i = 0; i < 3; ++i) t1[i] = u[i] / ul;
        double w[3] = {nn[1] * t1[2] - nn[2] * t1[1], nn[2] * t1[0] - nn[0] * t1[2], nn[0] * t1[1] - nn[1] * t1[0]};
        const double wl = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
        for (int i = 0; i < 3; ++i) t2[i] = w[i] / wl;
        return dn - radius;
    }
    // GetSignedDistanceWithReferenceNormal (:216-255)
    double mesh_signed(const double* p, const Vec& X, int node, const double* ref) const {
        double cl[3], fn[3];
        if (!mesh_closest(p, X, node, cl, fn)) return std::numeric_limits<double>::infinity();
        return (p[0] - cl[0]) * ref[0] + (p[1] - cl[1]) * ref[1] + (p[2] - cl[2]) * ref[2] - radius;
    }

    void hf(double x, double y, double& h, double& hx, double& hy) const {
        h = z0; hx = 0; hy = 0;
        for (size_t i = 0; i + 3 < waves.size(); i += 4) {
            const double a = waves[i], kx = waves[i + 1], ky = waves[i + 2], ph = waves[i + 3];
            const double arg = kx * x + ky * y + ph;
            h += a * std::sin(arg);
            hx += a * kx * std::cos(arg);
            hy += a * ky * std::cos(arg);
        }
    }
    // Which obstacle governs a point (press only): 0 ground, 1 plane.
    int press_side(const double* q, double& d_ground, double& d_plane) const {
        double h, hx, hy;
        hf(q[0], q[1], h, hx, hy);
        d_ground = (q[2] - h) / std::sqrt(1 + hx * hx + hy * hy);
        d_plane = plane_z - q[2];
        return (d_plane < d_ground) ? 1 : 0;
    }
    double distance(const double* q) const {
        if (type == kPlanar) return q[0] * n[0] + q[1] * n[1] + q[2] * n[2] + off;
        if (type == kSpherical) {
            const double d0 = q[0] - c[0], d1 = q[1] - c[1], d2 = q[2] - c[2];
            return std::sqrt(d0 * d0 + d1 * d1 + d2 * d2) - r;
        }
        if (type == kHeightfield) {
            double h, hx, hy;
            hf(q[0], q[1], h, hx, hy);
            return (q[2] - h) / std::sqrt(1 + hx * hx + hy * hy);
        }
        double dg, dp;
        return press_side(q, dg, dp) ? dp : dg;
    }
    void normal(const double* q, double out[3]) const {
        if (type == kPlanar) { for (int i = 0; i < 3; ++i) out[i] = n[i]; return; }
        if (type == kSpherical) {
            double d[3] = {q[0] - c[0], q[1] - c[1], q[2] - c[2]};
            const double rho = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            check(rho > std::numeric_limits<double>::epsilon(), "SphericalFrictionalBoundary::GetNormal singular at sphere center.");
            for (int i = 0; i < 3; ++i) out[i] = d[i] / rho;
            return;
        }
        double dg = 0, dp = 0;
        if (type == kPress && press_side(q, dg, dp)) { out[0] = 0; out[1] = 0; out[2] = -1; return; }
        double h, hx, hy;
        hf(q[0], q[1], h, hx, hy);
        const double s = std::sqrt(1 + hx * hx + hy * hy);
        out[0] = -hx / s; out[1] = -hy / s; out[2] = 1 / s;
    }
    // GetLocalFrame (planar_frictional_boundary.cpp:25-40,
    // spherical_frictional_boundary.cpp:25-42): columns t1, t2, n.
    void frame(const double* q, double t1[3], double t2[3], double nn[3]) const {
        normal(q, nn);
        double ux[3] = {0, 0, 0}, best = 0;
        for (int i = 0; i < 3; ++i) {
            double e[3] = {0, 0, 0};
            e[i] = 1;
            const double x[3] = {nn[1] * e[2] - nn[2] * e[1], nn[2] * e[0] - nn[0] * e[2], nn[0] * e[1] - nn[1] * e[0]};
            const double sq = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
            if (sq > best) { best = sq; for (int k = 0; k < 3; ++k) ux[k] = x[k]; }
        }
        const double un = std::sqrt(best);
        for (int k = 0; k < 3; ++k) t1[k] = ux[k] / un;
        double uy[3] = {nn[1] * t1[2] - nn[2] * t1[1], nn[2] * t1[0] - nn[0] * t1[2], nn[0] * t1[1] - nn[1] * t1[0]};
        const double yn = std::sqrt(uy[0] * uy[0] + uy[1] * uy[1] + uy[2] * uy[2]);
        for (int k = 0; k < 3; ++k) t2[k] = uy[k] / yn;
    }
};

// Per-contact Coulomb step (forward_gpu.cpp:1262-1302): the stick impulse
// solves W_tt lt = -u_t (W_tt = the contact's 2x2 tangential Delassus block,
// row-major, u_t the tangential displacement of the unconstrained step); it is
// kept inside the cone |lt| <= mu lambda, otherwise the impulse slides,
// lt = -mu lambda u_t / |u_t|.
void coulomb_step(const double* Wt, double u1, double u2, double fmu, double lam, double& lt1, double& lt2) {
    const double det = Wt[0] * Wt[3] - Wt[1] * Wt[2];
    double s1 = 0, s2 = 0;
    if (std::fabs(det) > 1e-20) {
        s1 = (Wt[3] * -u1 - Wt[1] * -u2) / det;
        s2 = (-Wt[2] * -u1 + Wt[0] * -u2) / det;
    }
    const double cone = fmu * lam;
    if (std::sqrt(s1 * s1 + s2 * s2) <= cone + 1e-14) { lt1 = s1; lt2 = s2; return; }
    const double um = std::sqrt(u1 * u1 + u2 * u2);
    if (um > 1e-14) { lt1 = -cone * u1 / um; lt2 = -cone * u2 / um; }
    else lt1 = lt2 = 0.0;
}

// ---------------------------------------------------------------------------
// Sparse LDL^T (restates the CSparse-style up-looking factorization of
// RealSim/src/Scomponent/tools/math/SparseLDLT.cpp:433-518) on a symmetric
// matrix given in full CSR, after a symmetric permutation `perm` (new -> old).
// ---------------------------------------------------------------------------
struct SparseLDLT {
    int n = 0;
    std::vector<int> perm, iperm;      // perm[new] = old
    std::vector<int> Lp, Li;           // CSC of unit lower L (strict)
    std::vector<double> Lx, D;

    void factor(int n_, const std::vector<int>& Ap, const std::vector<int>& Ai, const std::vector<double>& Ax,
                const std::vector<int>& perm_) {
        n = n_;
        perm = perm_;
        iperm.assign(n, 0);
        for (int i = 0; i < n; ++i) iperm[perm[i]] = i;
        // Permuted upper-triangular pattern by columns: C = P A P^T, entries (i,k), i<=k, stored by column k.
        std::vector<std::vector<std::pair<int, double>>> col(n);
        for (int r = 0; r < n; ++r)
            for (int p = Ap[r]; p < Ap[r + 1]; ++p) {
                const int c = Ai[p];
                const int i = iperm[r], k = iperm[c];
                if (i <= k) col[k].push_back({i, Ax[p]});
            }
        // Elimination tree and column counts (LDL symbolic, T. Davis 2005).
        std::vector<int> parent(n, -1), flag(n, -1), Lnz(n, 0);
        for (int k = 0; k < n; ++k) {
            flag[k] = k;
            for (auto& e : col[k]) {
                int i = e.first;
                for (; flag[i] != k; i = parent[i]) {
                    if (parent[i] == -1) parent[i] = k;
                    Lnz[i]++;
                    flag[i] = k;
                }
            }
        }
        Lp.assign(n + 1, 0);
        for (int k = 0; k < n; ++k) Lp[k + 1] = Lp[k] + Lnz[k];
        Li.assign(Lp[n], 0);
        Lx.assign(Lp[n], 0.0);
        D.assign(n, 0.0);
        // Numeric up-looking factorization (row k of L from a sparse triangular solve).
        std::vector<double> Y(n, 0.0);
        std::vector<int> pattern(n), fill(n, 0);
        for (int k = 0; k < n; ++k) {
            Y[k] = 0;
            int top = n;
            flag[k] = k;
            for (auto& e : col[k]) {
                int i = e.first;
                Y[i] += e.second;
                int len = 0;
                for (; flag[i] != k; i = parent[i]) { pattern[len++] = i; flag[i] = k; }
                while (len > 0) pattern[--top] = pattern[--len];
            }
            D[k] = Y[k];
            Y[k] = 0;
            for (; top < n; ++top) {
                const int i = pattern[top];
                const double yi = Y[i];
                Y[i] = 0;
                int p;
                for (p = Lp[i]; p < Lp[i] + fill[i]; ++p) Y[Li[p]] -= Lx[p] * yi;
                const double lki = yi / D[i];
                D[k] -= lki * yi;
                Li[p] = k;
                Lx[p] = lki;
                fill[i]++;
            }
            check(D[k] != 0.0, "LDLT zero pivot.");
        }
    }
    // x = A^-1 b.
    void solve(const double* b, double* x) const {
        Vec y(n);
        for (int i = 0; i < n; ++i) y[i] = b[perm[i]];
        for (int j = 0; j < n; ++j)
            for (int p = Lp[j]; p < Lp[j + 1]; ++p) y[Li[p]] -= Lx[p] * y[j];
        for (int j = 0; j < n; ++j) y[j] /= D[j];
        for (int j = n - 1; j >= 0; --j)
            for (int p = Lp[j]; p < Lp[j + 1]; ++p) y[j] -= Lx[p] * y[Li[p]];
        for (int i = 0; i < n; ++i) x[perm[i]] = y[i];
    }
};

// Nested dissection by recursive coordinate bisection: split at the median of
// the widest axis, separator = nodes of the lower half adjacent to the upper
// half; order (lower, upper, separator).
static void rcb_order(const std::vector<int>& nodes, const Vec& X, const std::vector<std::vector<int>>& adj,
                      std::vector<int>& out, std::vector<int>& side) {
    if (nodes.size() <= 32) { for (int v : nodes) out.push_back(v); return; }
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int v : nodes)
        for (int d = 0; d < 3; ++d) { lo[d] = std::min(lo[d], X[3 * v + d]); hi[d] = std::max(hi[d], X[3 * v + d]); }
    int ax = 0;
    for (int d = 1; d < 3; ++d) if (hi[d] - lo[d] > hi[ax] - lo[ax]) ax = d;
    std::vector<double> key;
    for (int v : nodes) key.push_back(X[3 * v + ax]);
    std::vector<double> sk = key;
    std::nth_element(sk.begin(), sk.begin() + sk.size() / 2, sk.end());
    const double med = sk[sk.size() / 2];
    // side: 1 lower, 2 upper (marks valid only for this call's nodes)
    for (size_t i = 0; i < nodes.size(); ++i) side[nodes[i]] = (key[i] < med) ? 1 : 2;
    std::vector<int> L, U, Sep;
    for (int v : nodes) {
        if (side[v] == 2) { U.push_back(v); continue; }
        bool touches = false;
        for (int w : adj[v]) if (side[w] == 2) { touches = true; break; }
        (touches ? Sep : L).push_back(v);
    }
    for (int v : nodes) side[v] = 0;
    if (L.empty() || U.empty()) { for (int v : nodes) out.push_back(v); return; }
    rcb_order(L, X, adj, out, side);
    rcb_order(U, X, adj, out, side);
    for (int v : Sep) out.push_back(v);
}

// ---------------------------------------------------------------------------
// The deformable body: mesh, FE samples, energies, boundary conditions, state.
// ---------------------------------------------------------------------------
struct AlgPhdReplayCache {  // cpp/core/include/fem/deformable.h:336-355 (fields used by the backward)
    std::vector<int> contact_candidates;
    Vec q_tilde, q_sol;
    Vec lambda, omega;
    std::vector<std::array<double, 3>> normals;
};

struct Body {
    int edim = 8;                 // 8 hex, 4 tet
    int nv = 0, ne = 0, dofs = 0;
    Vec X;                        // 3*nv rest positions
    std::vector<int> elems;       // edim*ne
    double density = 0, element_volume = 0;
    int samples = 8;
    std::vector<double> grad;     // per (element, sample): edim x 3 row-major grad_undeformed_sample_weight
    std::vector<Energy> energies;
    // PdMuscleEnergy groups (pd_muscle_energies_, deformable_actuation.cpp:5-26):
    // stiffness, unit fiber direction m, elements; the actuation vector `act`
    // holds one level per (group, element) in group order (act_dofs_)
    struct Muscle {
        double stiffness = 0;
        double m[3] = {0, 0, 0};
        std::vector<int> elems;
    };
    std::vector<Muscle> muscles;
    int act_dofs = 0;
    Vec act;
    Vec last_dl_da, last_dl_dact_w;
    std::map<int, double> dirichlet;
    std::vector<double> gravity;  // empty or 3
    // PlanarContactStateForce (planar_contact_state_force.cpp:3-14): unit normal,
    // offset / |normal|, exponent p, stiffness kn, friction clamp kf, mu; each
    // state force's parameters sit at its offset of dl_dstate_p (AddStateForce order)
    struct PlanarContact {
        double n[3] = {0, 0, 1}, off = 0;
        int p = 2;
        double kn = 0, kf = 0, mu = 0;
        int poff = 0;
    };
    std::vector<PlanarContact> pcontacts;
    // ArcContactStateForce (arc_contact_state_force.cpp:18-30): the inside of a
    // cylindrical arc of radius R about the axis through c along unit `dir`,
    // covering the angles (0, angle) from unit `start` toward aux = dir x start
    struct ArcContact {
        double c[3], dir[3], start[3], aux[3], R = 0, angle = 0;
        int p = 2;
        double kn = 0, kf = 0, mu = 0;
        int poff = 0;
    };
    std::vector<ArcContact> acontacts;
    // MeshContactStateForce (mesh_contact_state_force.cpp:101-142): penalty contact
    // between the vertices of surface A and the triangles of surface B (kn_left) and
    // the reverse (kn_right), within contact_radius; parameters (kn_left, kn_right, kf, mu)
    struct MeshContact {
        double radius = 0, kn_l = 0, kn_r = 0, kf = 0, mu = 0;
        std::vector<std::array<int, 3>> fa, fb;
        std::vector<int> va, vb;  // sorted vertex sets of the faces
        int poff = 0;
    };
    std::vector<MeshContact> mcontacts;
    // BilliardBallStateForce (billiard_ball_state_force.cpp:6-18): the body is
    // ball_num consecutive blocks of svn vertices; the balls (centres = vertex
    // means) push apart within 2 radius, with a sideways friction in the x-y plane;
    // parameters (stiffness per ball, friction coefficient per ball)
    struct BilliardBall {
        double R = 0;
        int svn = 0, nb = 0;
        std::vector<double> k, mu;
        int poff = 0;
    };
    std::vector<BilliardBall> bballs;
    // HydrodynamicsStateForce (hydrodynamics_state_force.cpp:11-42): drag and
    // thrust of water moving at v_water on surface faces (quads of a hex mesh,
    // triangles of a tet mesh), with cubic Bezier coefficient curves Cd(phi), Ct(phi)
    // of the angle of attack and a per-entry clamp at max_force; parameters (the
    // Cd points (4 x 2), the Ct points (4 x 2))
    struct Hydro {
        double rho = 0, vw[3] = {0, 0, 0}, P[16] = {0}, fmax = 0;
        int fd = 4;
        std::vector<int> faces;  // fd per face
        int poff = 0;
    };
    std::vector<Hydro> hydros;
    int gravity_off = 0, n_state_params = 0;
    Boundary boundary;
    std::map<int, int> fb_vertices;  // frictional_boundary_vertex_indices_ (sorted)

    // cached solver state (mutable members of the reference Deformable)
    bool pd_ready = false;
    double pd_dt = 0;
    SparseLDLT ldlt[3];
    std::vector<AlgPhdReplayCache> tape;
    std::vector<int> prev_candidates;
    Vec prev_lambda, last_lambda, last_lt1, last_lt2;
    int refine_steps = 0;
    std::array<std::vector<int>, 3> csr_p, csr_i;
    std::array<std::vector<double>, 3> csr_x;
    std::vector<int> frame_iters, frame_converged;
    int last_bwd_iters = 0;
    bool last_bwd_from_tape = true;
    struct BackwardStats {  // of the last backward call
        double backbone_iters = 0, backbone_converged = 0, clipped = 0, woodbury_iters = 0, basis_rows = 0,
               filtered = 0, samples = 0, db_norm = 0, trust_stage = 0, from_tape = 1;
    } last_bwd;

    const double* G(int e, int s) const { return &grad[((size_t)e * samples + s) * edim * 3]; }

    // Mesh<3,8>/<3,4>::ComputeElementVolume (mesh.cpp:111-129) and
    // InitializeAfterMesh (deformable.cpp:33-43).
    void init(int edim_, int nv_, const double* Xv, int ne_, const int* el, double density_) {
        edim = edim_; nv = nv_; ne = ne_; dofs = 3 * nv;
        X.assign(Xv, Xv + 3 * nv);
        elems.assign(el, el + (size_t)edim * ne);
        density = density_;
        samples = (edim == 8) ? 8 : 1;
        double vol_sum = 0;
        for (int e = 0; e < ne; ++e) {
            const int* v = &elems[(size_t)e * edim];
            M3 A;
            if (edim == 8) {
                for (int i = 0; i < 3; ++i) {
                    A(i, 0) = X[3 * v[1] + i] - X[3 * v[0] + i];
                    A(i, 1) = X[3 * v[2] + i] - X[3 * v[0] + i];
                    A(i, 2) = X[3 * v[4] + i] - X[3 * v[0] + i];
                }
                vol_sum += std::fabs(det3(A));
            } else {
                for (int i = 0; i < 3; ++i) {
                    A(i, 0) = X[3 * v[1] + i] - X[3 * v[0] + i];
                    A(i, 1) = X[3 * v[2] + i] - X[3 * v[0] + i];
                    A(i, 2) = X[3 * v[3] + i] - X[3 * v[0] + i];
                }
                vol_sum += std::fabs(det3(A)) / 6;
            }
        }
        element_volume = vol_sum / ne;
        const double dx = std::pow(element_volume, 1.0 / 3.0);
        grad.assign((size_t)ne * samples * edim * 3, 0.0);
        if (edim == 8) {
            // HexDeformable::InitializeFiniteElementSamples, hex_deformable.cpp:3-50.
            const double r = 1 / std::sqrt(3.0), inv = 1 / dx;
            double g8[8][8][3];
            for (int s = 0; s < 8; ++s) {
                const int bx = (s >> 2) & 1, by = (s >> 1) & 1, bz = s & 1;
                const double x = ((bx - 0.5) * r + 0.5) * dx, y = ((by - 0.5) * r + 0.5) * dx, z = ((bz - 0.5) * r + 0.5) * dx;
                const double nx = x * inv, ny = y * inv, nz = z * inv, cx = 1 - nx, cy = 1 - ny, cz = 1 - nz;
                const double rows[8][3] = {
                    {-inv * cy * cz, cx * -inv * cz, cx * cy * -inv}, {-inv * cy * nz, cx * -inv * nz, cx * cy * inv},
                    {-inv * ny * cz, cx * inv * cz, cx * ny * -inv},  {-inv * ny * nz, cx * inv * nz, cx * ny * inv},
                    {inv * cy * cz, nx * -inv * cz, nx * cy * -inv},  {inv * cy * nz, nx * -inv * nz, nx * cy * inv},
                    {inv * ny * cz, nx * inv * cz, nx * ny * -inv},   {inv * ny * nz, nx * inv * nz, nx * ny * inv}};
                for (int k = 0; k < 8; ++k) for (int j = 0; j < 3; ++j) g8[s][k][j] = rows[k][j];
            }
            for (int e = 0; e < ne; ++e)
                for (int s = 0; s < 8; ++s)
                    std::memcpy(&grad[((size_t)e * 8 + s) * 24], &g8[s][0][0], sizeof(double) * 24);
        } else {
            // TetDeformable::InitializeFiniteElementSamples, tet_deformable.cpp:3-41.
            for (int e = 0; e < ne; ++e) {
                const int* v = &elems[(size_t)e * 4];
                M3 Bi;
                for (int i = 0; i < 3; ++i) {
                    Bi(i, 0) = X[3 * v[1] + i] - X[3 * v[0] + i];
                    Bi(i, 1) = X[3 * v[2] + i] - X[3 * v[0] + i];
                    Bi(i, 2) = X[3 * v[3] + i] - X[3 * v[0] + i];
                }
                const M3 B = inv3(Bi);
                double* g = &grad[(size_t)e * 12];
                for (int j = 0; j < 3; ++j) {
                    g[0 * 3 + j] = -B(0, j) - B(1, j) - B(2, j);
                    g[1 * 3 + j] = B(0, j);
                    g[2 * 3 + j] = B(1, j);
                    g[3 * 3 + j] = B(2, j);
                }
            }
        }
        pd_ready = false;
    }

    // F = x_e G (DeformationGradient / pd_A * x_e, finite_element_sample.cpp:11-24).
    M3 deformation_gradient(const double* q, int e, int s) const {
        const int* v = &elems[(size_t)e * edim];
        const double* g = G(e, s);
        M3 F = M3::zero();
        for (int k = 0; k < edim; ++k)
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) F(i, j) += q[3 * v[k] + i] * g[k * 3 + j];
        return F;
    }
    // out_e[3k+i] += w * (A^T vec(P))_{k,i} = w * sum_j P(i,j) G(k,j).
    void add_At(const M3& P, int e, int s, double w, double* out) const {
        const double* g = G(e, s);
        for (int k = 0; k < edim; ++k)
            for (int i = 0; i < 3; ++i) {
                double acc = 0;
                for (int j = 0; j < 3; ++j) acc += P(i, j) * g[k * 3 + j];
                out[3 * k + i] += w * acc;
            }
    }

    // AddPdEnergy, deformable_pd_energy.cpp:42-153.
    void add_pd_energy(const std::string& type, const Vec& params, const std::vector<int>& idx) {
        pd_ready = false;
        Energy en;
        const int np = (int)params.size();
        if (type == "corotated" || type == "volume") {
            check(np == 1 || np == ne, "Inconsistent param size.");
            en.type = (type == "corotated") ? kCorotated : kVolume;
            if (np == 1) en.stiffness = params[0];
            else { en.elem_stiff = params; en.stiffness = std::accumulate(params.begin(), params.end(), 0.0) / np; }
        } else if (type == "pd_neohookean") {
            check(np == 2 || np == 2 * ne, "pd_neohookean expects params = [mu, lambda] or per-element.");
            en.type = kNeoHookean;
            if (np == 2) { en.mu = params[0]; en.la = params[1]; en.stiffness = 2 * params[0]; }
            else {
                double ms = 0, ls = 0;
                en.elem_stiff.resize(ne);
                for (int i = 0; i < ne; ++i) { ms += params[2 * i]; ls += params[2 * i + 1]; en.elem_stiff[i] = 2 * params[2 * i]; }
                en.mu = ms / ne; en.la = ls / ne;
                en.stiffness = std::accumulate(en.elem_stiff.begin(), en.elem_stiff.end(), 0.0) / ne;
            }
        } else {
            check(false, "Unsupported PD energy: " + type);
        }
        en.member.assign(ne, 0);
        if (idx.empty()) std::fill(en.member.begin(), en.member.end(), 1);
        for (int i : idx) { check(0 <= i && i < ne, "Element index out of bound."); en.member[i] = 1; }
        energies.push_back(en);
    }

    // AddActuation (deformable_actuation.cpp:5-26) with PdMuscleEnergy::Initialize
    // (pd_muscle_energy.cpp:5-24): the fiber direction is normalized.
    void add_actuation(double stiffness, const double* fiber, const std::vector<int>& idx) {
        pd_ready = false;
        Muscle mu;
        mu.stiffness = stiffness;
        const double nrm = std::sqrt(fiber[0] * fiber[0] + fiber[1] * fiber[1] + fiber[2] * fiber[2]);
        check(nrm > std::numeric_limits<double>::epsilon(), "Singular fiber_direction.");
        for (int i = 0; i < 3; ++i) mu.m[i] = fiber[i] / nrm;
        std::vector<char> seen(ne, 0);
        for (int i : idx) {
            check(0 <= i && i < ne, "Element index out of bound.");
            check(!seen[i], "Duplicated elements.");
            seen[i] = 1;
            mu.elems.push_back(i);
        }
        muscles.push_back(mu);
        act_dofs += (int)idx.size();
    }
    void set_actuation(const Vec& a) {
        check((int)a.size() == act_dofs, "Inconsistent actuation size.");
        act = a;
    }
    // Muscle projection Bp = a / |F m| F m (ProjectToManifold,
    // pd_muscle_energy.cpp:118-125), zero when |F m| <= eps.
    static void muscle_fm(const M3& F, const double* m, double Fm[3], double& l) {
        for (int i = 0; i < 3; ++i) Fm[i] = F(i, 0) * m[0] + F(i, 1) * m[1] + F(i, 2) * m[2];
        l = std::sqrt(Fm[0] * Fm[0] + Fm[1] * Fm[1] + Fm[2] * Fm[2]);
    }

    // SetFrictionalBoundary, deformable_friction.cpp:7-89 (planar/spherical) +
    // heightfield / press.
    void set_boundary(const std::string& type, const Vec& p, const std::vector<int>& idx) {
        pd_ready = false;
        Boundary b;
        if (type == "planar") {
            check(p.size() == 4, "Incompatible parameter number.");
            const double nn = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
            check(nn > std::numeric_limits<double>::epsilon(), "Singular normal.");
            b.type = kPlanar;
            for (int i = 0; i < 3; ++i) b.n[i] = p[i] / nn;
            b.off = p[3] / nn;
        } else if (type == "spherical") {
            check(p.size() == 4, "Incompatible parameter number.");
            b.type = kSpherical;
            for (int i = 0; i < 3; ++i) b.c[i] = p[i];
            b.r = p[3];
        } else if (type == "heightfield" || type == "press") {
            // params: [z0, nwaves, (amp, kx, ky, phase) * nwaves] (+ [plane_z] for press)
            check(p.size() >= 2, "Incompatible parameter number.");
            const int nw = (int)p[1];
            const size_t need = 2 + 4 * (size_t)nw + (type == "press" ? 1 : 0);
            check(p.size() == need, "Incompatible parameter number.");
            b.type = (type == "press") ? kPress : kHeightfield;
            b.z0 = p[0];
            b.waves.assign(p.begin() + 2, p.begin() + 2 + 4 * nw);
            if (type == "press") b.plane_z = p[2 + 4 * nw];
        } else if (type == "mesh") {
            // deformable_friction.cpp:53-73: [radius, split, n_left, n_right, faces_left, faces_right]
            check(p.size() >= 4, "Incompatible parameter number.");
            b.type = kMesh;
            b.radius = p[0];
            b.split = (int)p[1];
            b.nfl = (int)p[2];
            b.nfr = (int)p[3];
            check(b.nfl >= 0 && b.nfr >= 0 && p.size() == 4 + 3 * (size_t)(b.nfl + b.nfr),
                  "Inconsistent params for MeshFrictionalBoundary.");
            b.faces.resize(3 * (size_t)(b.nfl + b.nfr));
            for (size_t i = 0; i < b.faces.size(); ++i) b.faces[i] = (int)p[4 + i];
        } else {
            check(false, "Unsupported frictional boundary type: " + type);
        }
        std::set<int> uniq(idx.begin(), idx.end());
        check(uniq.size() == idx.size(), "Duplicated vertex elements.");
        boundary = b;
        fb_vertices.clear();
        int cnt = 0;
        for (int i : uniq) fb_vertices[i] = cnt++;
    }

    // AddStateForce("planar_contact", [n (3), offset, p, kn, kf, mu])
    // (deformable_state_force.cpp:20-31, planar_contact_state_force.cpp:3-14).
    void add_planar_contact(const Vec& prm) {
        check(prm.size() == 8, "Inconsistent params for PlanarContactStateForce.");
        PlanarContact c;
        const double nn = std::sqrt(prm[0] * prm[0] + prm[1] * prm[1] + prm[2] * prm[2]);
        check(nn > 1e-5, "Singular normal.");
        for (int i = 0; i < 3; ++i) c.n[i] = prm[i] / nn;
        c.off = prm[3] / nn;
        c.p = (int)prm[4];
        c.kn = prm[5]; c.kf = prm[6]; c.mu = prm[7];
        c.poff = n_state_params;
        n_state_params += 3;
        pcontacts.push_back(c);
    }
    // AddStateForce("arc_contact", [center (3), dir (3), start (3), radius, angle,
    // p, kn, kf, mu]) (deformable_state_force.cpp:33-49, arc_contact_state_force.cpp:18-30).
    void add_arc_contact(const Vec& prm) {
        check(prm.size() == 15, "Inconsistent params for ArcContactStateForce.");
        ArcContact a;
        const double dn = std::sqrt(prm[3] * prm[3] + prm[4] * prm[4] + prm[5] * prm[5]);
        const double sn = std::sqrt(prm[6] * prm[6] + prm[7] * prm[7] + prm[8] * prm[8]);
        for (int i = 0; i < 3; ++i) { a.c[i] = prm[i]; a.dir[i] = prm[3 + i] / dn; a.start[i] = prm[6 + i] / sn; }
        a.aux[0] = a.dir[1] * a.start[2] - a.dir[2] * a.start[1];
        a.aux[1] = a.dir[2] * a.start[0] - a.dir[0] * a.start[2];
        a.aux[2] = a.dir[0] * a.start[1] - a.dir[1] * a.start[0];
        a.R = prm[9]; a.angle = prm[10]; a.p = (int)prm[11];
        a.kn = prm[12]; a.kf = prm[13]; a.mu = prm[14];
        a.poff = n_state_params;
        n_state_params += 3;
        acontacts.push_back(a);
    }
    // One vertex against an arc (arc_contact_state_force.cpp:32-62): qp = the part of
    // x - c normal to the axis; in contact when its angle from `start` lies in
    // (0, angle); d = R - |qp| (negative: penetration), n = -qp / (|qp| + eps).
    static bool arc_geometry(const ArcContact& a, const double* x, double qp[3], double& r, double n[3]) {
        double ql[3];
        for (int k = 0; k < 3; ++k) ql[k] = x[k] - a.c[k];
        const double ax = ql[0] * a.dir[0] + ql[1] * a.dir[1] + ql[2] * a.dir[2];
        for (int k = 0; k < 3; ++k) qp[k] = ql[k] - ax * a.dir[k];
        const double th = std::atan2(qp[0] * a.aux[0] + qp[1] * a.aux[1] + qp[2] * a.aux[2],
                                     qp[0] * a.start[0] + qp[1] * a.start[1] + qp[2] * a.start[2]);
        if (!(0 < th && th < a.angle)) return false;
        r = std::sqrt(qp[0] * qp[0] + qp[1] * qp[1] + qp[2] * qp[2]);
        const double eps = std::numeric_limits<double>::epsilon();
        for (int k = 0; k < 3; ++k) n[k] = -qp[k] / (r + eps);
        return true;
    }
    // AddStateForce("mesh_contact", [radius, kn_left, kn_right, kf, mu, #faces A,
    // #faces B, faces A (3 each), faces B]) (deformable_state_force.cpp:86-112).
    void add_mesh_contact(const Vec& prm) {
        check(prm.size() >= 7, "Inconsistent params for MeshContactStateForce.");
        const int na = (int)prm[5], nb = (int)prm[6];
        check(na >= 0 && nb >= 0 && prm.size() == 7 + 3 * (size_t)(na + nb), "Inconsistent params for MeshContactStateForce.");
        MeshContact m;
        m.radius = prm[0]; m.kn_l = prm[1]; m.kn_r = prm[2]; m.kf = prm[3]; m.mu = prm[4];
        std::set<int> sa, sb;
        for (int f = 0; f < na + nb; ++f) {
            std::array<int, 3> t;
            for (int k = 0; k < 3; ++k) {
                t[k] = (int)prm[7 + 3 * f + k];
                check(t[k] >= 0 && t[k] < nv, "Mesh contact face index out of range.");
            }
            if (f < na) { m.fa.push_back(t); sa.insert(t.begin(), t.end()); }
            else { m.fb.push_back(t); sb.insert(t.begin(), t.end()); }
        }
        m.va.assign(sa.begin(), sa.end());
        m.vb.assign(sb.begin(), sb.end());
        m.poff = n_state_params;
        n_state_params += 4;
        mcontacts.push_back(m);
    }
    // MeshContactStateForce::ForwardForce (mesh_contact_state_force.cpp:144-215): for
    // each vertex of one surface and each triangle of the other within the radius,
    // f = kn (radius - dist) n + friction -min(kf, mu f_n / (|v_t| + eps)) v_t on the
    // velocity relative to the triangle's point (only when |v_t| > eps), applied to
    // the vertex and, with the barycentric weights, opposite to the triangle's corners.
    // Nothing when the surfaces' bounding boxes lie more than the radius apart.
    void mesh_contact_force(const MeshContact& m, const Vec& q, const Vec& v, Vec& f) const {
        if (m.fa.empty() || m.fb.empty()) return;
        const double eps = std::numeric_limits<double>::epsilon(), r2 = m.radius * m.radius;
        auto box = [&](const std::vector<int>& vs, double* lo, double* hi) {
            for (int k = 0; k < 3; ++k) { lo[k] = INFINITY; hi[k] = -INFINITY; }
            for (int i : vs)
                for (int k = 0; k < 3; ++k) { lo[k] = std::min(lo[k], q[3 * i + k]); hi[k] = std::max(hi[k], q[3 * i + k]); }
        };
        double la[3], ha[3], lb[3], hb[3];
        box(m.va, la, ha);
        box(m.vb, lb, hb);
        double gap2 = 0;
        for (int k = 0; k < 3; ++k) {
            const double g = ha[k] < lb[k] ? lb[k] - ha[k] : (hb[k] < la[k] ? la[k] - hb[k] : 0.0);
            gap2 += g * g;
        }
        if (gap2 > r2) return;
        auto side = [&](const std::vector<int>& verts, const std::vector<std::array<int, 3>>& faces, double kn) {
            for (int i : verts) {
                const double* p = &q[3 * i];
                for (const auto& t : faces) {
                    const double *a = &q[3 * t[0]], *b = &q[3 * t[1]], *c = &q[3 * t[2]];
                    double pt[3], w[3];
                    tri_closest(p, a, b, c, pt, w);  // the method of mesh_contact_state_force.cpp:44-97
                    const double d2 = (p[0] - pt[0]) * (p[0] - pt[0]) + (p[1] - pt[1]) * (p[1] - pt[1]) +
                                      (p[2] - pt[2]) * (p[2] - pt[2]);
                    if (d2 >= r2) continue;
                    const double dist = std::sqrt(d2);
                    double n[3];
                    if (dist > eps) {
                        for (int k = 0; k < 3; ++k) n[k] = (p[k] - pt[k]) / dist;
                    } else {
                        const double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
                        const double e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
                        n[0] = e1[1] * e2[2] - e1[2] * e2[1];
                        n[1] = e1[2] * e2[0] - e1[0] * e2[2];
                        n[2] = e1[0] * e2[1] - e1[1] * e2[0];
                        const double nn = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
                        if (nn <= eps) continue;
                        for (int k = 0; k < 3; ++k) n[k] /= nn;
                    }
                    if (n[0] * (p[0] - pt[0]) + n[1] * (p[1] - pt[1]) + n[2] * (p[2] - pt[2]) < 0)
                        for (int k = 0; k < 3; ++k) n[k] = -n[k];
                    const double fn = kn * (m.radius - dist);
                    double vr[3];
                    for (int k = 0; k < 3; ++k)
                        vr[k] = v[3 * i + k] - (w[0] * v[3 * t[0] + k] + w[1] * v[3 * t[1] + k] + w[2] * v[3 * t[2] + k]);
                    const double vn = vr[0] * n[0] + vr[1] * n[1] + vr[2] * n[2];
                    double vt[3];
                    for (int k = 0; k < 3; ++k) vt[k] = vr[k] - n[k] * vn;
                    const double vtm = std::sqrt(vt[0] * vt[0] + vt[1] * vt[1] + vt[2] * vt[2]);
                    const double cf = vtm > eps ? std::min(m.kf, m.mu * fn / (vtm + eps)) : 0.0;
                    for (int k = 0; k < 3; ++k) {
                        const double tot = fn * n[k] - cf * vt[k];
                        f[3 * i + k] += tot;
                        for (int j = 0; j < 3; ++j) f[3 * t[j] + k] -= tot * w[j];
                    }
                }
            }
        };
        side(m.va, m.fb, m.kn_l);
        side(m.vb, m.fa, m.kn_r);
    }
    // AddStateForce("billiard_ball", [radius, svn, stiffness (ball_num), friction
    // coefficients (ball_num)]) (deformable_state_force.cpp:72-85).
    void add_billiard_ball(const Vec& prm) {
        check(prm.size() >= 2 && prm.size() % 2 == 0, "Inconsistent params for BilliardBallStateForce.");
        BilliardBall b;
        b.R = prm[0];
        b.svn = (int)prm[1];
        b.nb = (int)(prm.size() - 2) / 2;
        check(b.svn > 0 && nv % b.svn == 0 && nv / b.svn == b.nb, "Inconsistent ball_num.");
        for (int i = 0; i < b.nb; ++i) { b.k.push_back(prm[2 + i]); b.mu.push_back(prm[2 + b.nb + i]); }
        b.poff = n_state_params;
        n_state_params += 2 * b.nb;
        bballs.push_back(b);
    }
    // Ball centres and centre velocities (vertex means).
    void ball_centres(const BilliardBall& b, const Vec& x, std::vector<std::array<double, 3>>& c) const {
        c.assign(b.nb, {0, 0, 0});
        for (int i = 0; i < b.nb; ++i) {
            for (int j = 0; j < b.svn; ++j)
                for (int k = 0; k < 3; ++k) c[i][k] += x[3 * ((size_t)i * b.svn + j) + k];
            for (int k = 0; k < 3; ++k) c[i][k] /= b.svn;
        }
    }
    // ForwardForce (billiard_ball_state_force.cpp:20-86): for balls i < j with
    // centres closer than 2 radius, ball j gets k_j (2R - |cj - ci|) along i->j and ball
    // i gets k_i (...) against it; friction of magnitude mu times that, along i->j
    // turned by +90 degrees in x-y for ball i when (v_i - v_j) x i2j (z part) > 0,
    // else by -90 degrees, and the opposite for ball j; each ball's force is spread
    // evenly over its vertices.
    void billiard_force(const BilliardBall& b, const Vec& q, const Vec& v, Vec& f) const {
        std::vector<std::array<double, 3>> c, vc;
        ball_centres(b, q, c);
        ball_centres(b, v, vc);
        for (int i = 0; i < b.nb; ++i)
            for (int j = i + 1; j < b.nb; ++j) {
                double dir[3], u[3], rot[3];
                for (int k = 0; k < 3; ++k) dir[k] = c[j][k] - c[i][k];
                const double dist = std::sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
                const double pen = std::max(2 * b.R - dist, 0.0);
                const double fj = pen * b.k[j], fi = pen * b.k[i];
                for (int k = 0; k < 3; ++k) u[k] = dir[k] / dist;
                const double rx = vc[i][0] - vc[j][0], ry = vc[i][1] - vc[j][1];
                if (rx * u[1] - ry * u[0] > 0) { rot[0] = -u[1]; rot[1] = u[0]; }
                else { rot[0] = u[1]; rot[1] = -u[0]; }
                rot[2] = 0;
                for (int k = 0; k < 3; ++k) {
                    const double Fi = (-u[k] * fi + rot[k] * fi * b.mu[i]) / b.svn;
                    const double Fj = (u[k] * fj - rot[k] * fj * b.mu[j]) / b.svn;
                    for (int p = 0; p < b.svn; ++p) {
                        f[3 * ((size_t)i * b.svn + p) + k] += Fi;
                        f[3 * ((size_t)j * b.svn + p) + k] += Fj;
                    }
                }
            }
    }
    // BackwardForce (billiard_ball_state_force.cpp:88-229): dl/dq through the
    // centres, dl/dv = 0 (the friction direction is piecewise constant in v), and
    // dl/d(k, mu). The reference sizes the friction magnitude's parameter gradient by
    // 2 svn instead of 2 ball_num, so its result is defined only for svn == ball_num;
    // other layouts raise an error here.
    void billiard_backward(const BilliardBall& b, const Vec& q, const Vec& v, const Vec& g, Vec& dl_dq, Vec& dl_dp) const {
        check(b.svn == b.nb, "BilliardBallStateForce backward needs as many vertices per ball as balls.");
        std::vector<std::array<double, 3>> c, vc;
        ball_centres(b, q, c);
        ball_centres(b, v, vc);
        for (int i = 0; i < b.nb; ++i)
            for (int j = i + 1; j < b.nb; ++j) {
                double dir[3], u[3];
                for (int k = 0; k < 3; ++k) dir[k] = c[j][k] - c[i][k];
                const double dist = std::sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
                check(dist > 0.1 * b.R, "Balls are too close to each other");
                const double pen = std::max(2 * b.R - dist, 0.0);
                const double fj = pen * b.k[j], fi = pen * b.k[i];
                for (int k = 0; k < 3; ++k) u[k] = dir[k] / dist;
                // d dist / d c_j = u = -d dist / d c_i; d u / d c_j = (I - u u^T) / dist = -d u / d c_i
                double Du[3][3];
                for (int r = 0; r < 3; ++r)
                    for (int s = 0; s < 3; ++s) Du[r][s] = ((r == s) - u[r] * u[s]) / dist;
                const double dfj = fj > 0 ? -b.k[j] : 0.0, dfi = fi > 0 ? -b.k[i] : 0.0;  // d f / d dist
                const double rx = vc[i][0] - vc[j][0], ry = vc[i][1] - vc[j][1];
                const double sg = rx * u[1] - ry * u[0] > 0 ? 1.0 : -1.0;
                // rot = sg * (-u_y, u_x, 0): its Jacobian rows are sg * (-Du[1], Du[0], 0)
                const double rot[3] = {-sg * u[1], sg * u[0], 0.0};
                double Drot[3][3];
                for (int s = 0; s < 3; ++s) { Drot[0][s] = -sg * Du[1][s]; Drot[1][s] = sg * Du[0][s]; Drot[2][s] = 0; }
                // summed dl/df over each ball's vertices, / svn
                double gi[3] = {0, 0, 0}, gj[3] = {0, 0, 0};
                for (int p = 0; p < b.svn; ++p)
                    for (int k = 0; k < 3; ++k) {
                        gi[k] += g[3 * ((size_t)i * b.svn + p) + k] / b.svn;
                        gj[k] += g[3 * ((size_t)j * b.svn + p) + k] / b.svn;
                    }
                // F_i = fi (-u + mu_i rot), F_j = fj (u - mu_j rot); gradient in c_j (c_i is its negative)
                double dcj[3] = {0, 0, 0};
                const double giu = gi[0] * u[0] + gi[1] * u[1] + gi[2] * u[2];
                const double gju = gj[0] * u[0] + gj[1] * u[1] + gj[2] * u[2];
                const double gir = gi[0] * rot[0] + gi[1] * rot[1];
                const double gjr = gj[0] * rot[0] + gj[1] * rot[1];
                for (int s = 0; s < 3; ++s) {
                    double t = 0;
                    for (int r = 0; r < 3; ++r)
                        t += gi[r] * fi * (-Du[r][s] + b.mu[i] * Drot[r][s]) + gj[r] * fj * (Du[r][s] - b.mu[j] * Drot[r][s]);
                    t += u[s] * (dfi * (-giu + b.mu[i] * gir) + dfj * (gju - b.mu[j] * gjr));
                    dcj[s] = t;
                }
                for (int p = 0; p < b.svn; ++p)
                    for (int k = 0; k < 3; ++k) {
                        dl_dq[3 * ((size_t)i * b.svn + p) + k] += -dcj[k] / b.svn;
                        dl_dq[3 * ((size_t)j * b.svn + p) + k] += dcj[k] / b.svn;
                    }
                dl_dp[b.poff + i] += pen * (-giu + b.mu[i] * gir);
                dl_dp[b.poff + j] += pen * (gju - b.mu[j] * gjr);
                dl_dp[b.poff + b.nb + i] += fi * gir;
                dl_dp[b.poff + b.nb + j] += -fj * gjr;
            }
    }
    // AddStateForce("hydrodynamics", [rho, v_water (3), Cd points (8), Ct points (8),
    // max_force, faces (fd each)]) (deformable_state_force.cpp:50-71); fd = 4 (hex) or 3 (tet).
    void add_hydrodynamics(const Vec& prm) {
        const int fd = edim == 8 ? 4 : 3;
        check(prm.size() >= 21 + (size_t)fd && (prm.size() - 21) % fd == 0, "Inconsistent params for HydrodynamicsStateForce.");
        Hydro h;
        h.rho = prm[0];
        for (int i = 0; i < 3; ++i) h.vw[i] = prm[1 + i];
        for (int i = 0; i < 16; ++i) h.P[i] = prm[4 + i];
        h.fmax = prm[20];
        h.fd = fd;
        for (size_t i = 21; i < prm.size(); ++i) {
            h.faces.push_back((int)prm[i]);
            check(h.faces.back() >= 0 && h.faces.back() < nv, "Hydrodynamics face index out of range.");
        }
        h.poff = n_state_params;
        n_state_params += 16;
        hydros.push_back(h);
    }
    // Cd(phi) and Ct(phi) (hydrodynamics_state_force.cpp:101-162): cubic Bezier
    // curves in t = |phi| / (pi / 2) (Cd) and t = phi / pi + 1/2 (Ct) through the y values
    // of the points; their phi-derivatives and weights on the four y values.
    static void bezier(double t, const double* y, double& val, double& dval_dt, double w[4]) {
        const double s = 1 - t;
        w[0] = s * s * s; w[1] = 3 * s * s * t; w[2] = 3 * s * t * t; w[3] = t * t * t;
        val = w[0] * y[0] + w[1] * y[1] + w[2] * y[2] + w[3] * y[3];
        dval_dt = 3 * (-s * s * y[0] + (s * s - 2 * s * t) * y[1] + (2 * s * t - t * t) * y[2] + t * t * y[3]);
    }
    void hydro_coeffs(const Hydro& h, double phi, double& cd, double& dcd, double wcd[4], double& ct, double& dct,
                      double wct[4]) const {
        const double pi = 3.14159265358979323846;
        const double ycd[4] = {h.P[1], h.P[3], h.P[5], h.P[7]}, yct[4] = {h.P[9], h.P[11], h.P[13], h.P[15]};
        const double tn = phi / (pi / 2);
        bezier(tn < 0 ? -tn : tn, ycd, cd, dcd, wcd);
        dcd *= tn < 0 ? -2 / pi : 2 / pi;
        bezier(tn * 0.5 + 0.5, yct, ct, dct, wct);
        dct *= 1 / pi;
    }
    // One face: its area A, normal n (zero when degenerate), the summed triangle
    // vectors un, and the triangles (vertex slots) that make it up: a quad is the
    // four triangles 012, 023, 013, 312 with A the mean of their |cross| (:625-640),
    // a triangle is 012 with A = |cross| (:466-470).
    struct HydroFace { double A, n[3], un[3], unn; double Aj[4][3]; int nt; };
    HydroFace hydro_face(const Hydro& h, const Vec& q, const int* f) const {
        static const int quad[4][3] = {{0, 1, 2}, {0, 2, 3}, {0, 1, 3}, {3, 1, 2}};
        HydroFace F{};
        F.nt = h.fd == 4 ? 4 : 1;
        double sumA = 0;
        for (int t = 0; t < F.nt; ++t) {
            const double* a = &q[3 * f[quad[t][0]]];
            const double* b = &q[3 * f[quad[t][1]]];
            const double* c = &q[3 * f[quad[t][2]]];
            const double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, e2[3] = {c[0] - b[0], c[1] - b[1], c[2] - b[2]};
            F.Aj[t][0] = e1[1] * e2[2] - e1[2] * e2[1];
            F.Aj[t][1] = e1[2] * e2[0] - e1[0] * e2[2];
            F.Aj[t][2] = e1[0] * e2[1] - e1[1] * e2[0];
            for (int k = 0; k < 3; ++k) F.un[k] += F.Aj[t][k];
            sumA += std::sqrt(F.Aj[t][0] * F.Aj[t][0] + F.Aj[t][1] * F.Aj[t][1] + F.Aj[t][2] * F.Aj[t][2]);
        }
        F.A = h.fd == 4 ? sumA / 4 : sumA;
        F.unn = std::sqrt(F.un[0] * F.un[0] + F.un[1] * F.un[1] + F.un[2] * F.un[2]);
        const double eps = std::numeric_limits<double>::epsilon();
        for (int k = 0; k < 3; ++k) F.n[k] = F.unn > eps ? F.un[k] / F.unn : 0.0;
        return F;
    }
    // ForwardForce (:457-499 tet, :614-673 hex): per face with A > eps, relative
    // water velocity w = v_water - mean node velocity (|w| > eps) and n . w/|w| < 0:
    // phi = acos(n . d) - pi/2, F = 1/2 rho A (Cd |w| w - Ct |w|^2 n), F / fd on each
    // node; then every entry clamped to [-max_force, max_force].
    Vec hydro_force(const Hydro& h, const Vec& q, const Vec& v) const {
        Vec f(dofs, 0.0);
        const double eps = std::numeric_limits<double>::epsilon(), pi = 3.14159265358979323846;
        const int nf = (int)h.faces.size() / h.fd;
        for (int i = 0; i < nf; ++i) {
            const int* fc = &h.faces[(size_t)i * h.fd];
            const HydroFace F = hydro_face(h, q, fc);
            if (F.A <= eps) continue;
            double w[3];
            for (int k = 0; k < 3; ++k) {
                double m = 0;
                for (int j = 0; j < h.fd; ++j) m += v[3 * fc[j] + k];
                w[k] = h.vw[k] - (h.fd == 4 ? m * 0.25 : m / 3);
            }
            const double wl = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
            if (wl <= eps) continue;
            const double nd = (F.n[0] * w[0] + F.n[1] * w[1] + F.n[2] * w[2]) / wl;
            if (nd >= 0.0) continue;
            const double phi = std::acos(nd) - pi / 2;
            double cd, dcd, wcd[4], ct, dct, wct[4];
            hydro_coeffs(h, phi, cd, dcd, wcd, ct, dct, wct);
            for (int k = 0; k < 3; ++k) {
                const double Fk = 0.5 * h.rho * F.A * cd * wl * w[k] - 0.5 * h.rho * F.A * ct * wl * wl * F.n[k];
                const double fn = h.fd == 4 ? Fk * 0.25 : Fk / 3;
                for (int j = 0; j < h.fd; ++j) f[3 * fc[j] + k] += fn;
            }
        }
        for (auto& x : f) x = std::min(std::max(x, -h.fmax), h.fmax);
        return f;
    }
    // BackwardForce (:501-612 tet, :675-829 hex), with the derivatives written out
    // as the reference forms them: dA/dp from the triangles' unit cross products,
    // dn/dp = (I - n n^T)/|un| dun/dp, dphi = -1/sqrt(1 - (n.d)^2) d(n.d), the Bezier
    // weights for the parameters; entries of dl/df where the clamp is active are dropped.
    void hydro_backward(const Hydro& h, const Vec& q, const Vec& v, const Vec& dl_df, Vec& dl_dq, Vec& dl_dv, Vec& dl_dp) const {
        static const int quad[4][3] = {{0, 1, 2}, {0, 2, 3}, {0, 1, 3}, {3, 1, 2}};
        const double eps = std::numeric_limits<double>::epsilon(), pi = 3.14159265358979323846;
        const Vec f = hydro_force(h, q, v);
        Vec g = dl_df;
        for (int i = 0; i < dofs; ++i) if (f[i] == h.fmax || f[i] == -h.fmax) g[i] = 0;
        const int nf = (int)h.faces.size() / h.fd;
        const double inv_fd = h.fd == 4 ? 0.25 : 1.0 / 3.0;
        for (int i = 0; i < nf; ++i) {
            const int* fc = &h.faces[(size_t)i * h.fd];
            const HydroFace F = hydro_face(h, q, fc);
            if (F.A <= eps) continue;
            double w[3];
            for (int k = 0; k < 3; ++k) {
                double m = 0;
                for (int j = 0; j < h.fd; ++j) m += v[3 * fc[j] + k];
                w[k] = h.vw[k] - (h.fd == 4 ? m * 0.25 : m / 3);
            }
            const double wl = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
            if (wl <= eps) continue;
            double d[3];
            for (int k = 0; k < 3; ++k) d[k] = w[k] / wl;
            const double nd = F.n[0] * d[0] + F.n[1] * d[1] + F.n[2] * d[2];
            if (nd >= 0.0) continue;
            const double phi = std::acos(nd) - pi / 2, dacos = -1 / std::sqrt(1 - nd * nd);
            double cd, dcd, wcd[4], ct, dct, wct[4];
            hydro_coeffs(h, phi, cd, dcd, wcd, ct, dct, wct);
            // dA/dp_m and dun/dp_m (3x3) per face node m
            double dA[4][3] = {}, dun[4][3][3] = {};
            const double cA = h.fd == 4 ? 0.25 : 1.0;
            for (int t = 0; t < F.nt; ++t) {
                const int s0 = quad[t][0], s1 = quad[t][1], s2 = quad[t][2];
                const double* a = &q[3 * fc[s0]];
                const double* b = &q[3 * fc[s1]];
                const double* c = &q[3 * fc[s2]];
                // cross = (b - a) x (c - b): d/da = [c - b]x, d/db = [a - c]x, d/dc = [b - a]x
                double e[3][3];
                for (int k = 0; k < 3; ++k) { e[0][k] = c[k] - b[k]; e[1][k] = a[k] - c[k]; e[2][k] = b[k] - a[k]; }
                const double an = std::sqrt(F.Aj[t][0] * F.Aj[t][0] + F.Aj[t][1] * F.Aj[t][1] + F.Aj[t][2] * F.Aj[t][2]);
                double u[3] = {0, 0, 0};
                if (an > eps) for (int k = 0; k < 3; ++k) u[k] = F.Aj[t][k] / an;
                const int slot[3] = {s0, s1, s2};
                for (int r = 0; r < 3; ++r) {
                    double S[3][3] = {{0, -e[r][2], e[r][1]}, {e[r][2], 0, -e[r][0]}, {-e[r][1], e[r][0], 0}};
                    for (int x = 0; x < 3; ++x) {
                        for (int y = 0; y < 3; ++y) dun[slot[r]][x][y] += S[x][y];
                        // dA/dp = cA * S^T u
                        dA[slot[r]][x] += cA * (S[0][x] * u[0] + S[1][x] * u[1] + S[2][x] * u[2]);
                    }
                }
            }
            double dn[4][3][3] = {};
            if (F.unn > eps)
                for (int m = 0; m < h.fd; ++m)
                    for (int x = 0; x < 3; ++x)
                        for (int y = 0; y < 3; ++y) {
                            double t = 0;
                            for (int z = 0; z < 3; ++z) t += (((x == z) - F.n[x] * F.n[z]) / F.unn) * dun[m][z][y];
                            dn[m][x][y] = t;
                        }
            // G = the masked dl/df summed over the face's nodes (each node gets F / fd)
            double G[3] = {0, 0, 0};
            for (int j = 0; j < h.fd; ++j) for (int k = 0; k < 3; ++k) G[k] += g[3 * fc[j] + k];
            const double r = 0.5 * h.rho;
            for (int m = 0; m < h.fd; ++m) {
                double dphi[3];
                for (int y = 0; y < 3; ++y) dphi[y] = dacos * (d[0] * dn[m][0][y] + d[1] * dn[m][1][y] + d[2] * dn[m][2][y]);
                for (int y = 0; y < 3; ++y) {
                    double col = 0;  // G . dF/dp_m[:, y]
                    for (int x = 0; x < 3; ++x) {
                        const double ddrag = r * wl * w[x] * (dA[m][y] * cd + F.A * dcd * dphi[y]);
                        const double dthr = -r * wl * wl * (dn[m][x][y] * F.A * ct + F.n[x] * dA[m][y] * ct + F.n[x] * F.A * dct * dphi[y]);
                        col += G[x] * (ddrag + dthr);
                    }
                    dl_dq[3 * fc[m] + y] += col * inv_fd;
                }
            }
            // velocities: w = v_water - mean, d|w|/dv_m = -d / fd, dd/dv_m = -(I - d d^T)/(fd |w|)
            double dphi_v[3];
            for (int y = 0; y < 3; ++y) {
                double t = 0;
                for (int x = 0; x < 3; ++x) t += F.n[x] * (-((x == y) - d[x] * d[y]) * inv_fd / wl);
                dphi_v[y] = dacos * t;
            }
            for (int y = 0; y < 3; ++y) {
                double col = 0;
                for (int x = 0; x < 3; ++x) {
                    const double ddrag = r * F.A * (-(x == y) * cd * wl * inv_fd + w[x] * dcd * dphi_v[y] * wl - w[x] * cd * d[y] * inv_fd);
                    const double dthr = -r * F.A * F.n[x] * (dct * dphi_v[y] * wl * wl - 2 * ct * w[y] * inv_fd);
                    col += G[x] * (ddrag + dthr);
                }
                for (int m = 0; m < h.fd; ++m) dl_dv[3 * fc[m] + y] += col * inv_fd;
            }
            const double Gw = G[0] * w[0] + G[1] * w[1] + G[2] * w[2], Gn = G[0] * F.n[0] + G[1] * F.n[1] + G[2] * F.n[2];
            for (int j = 0; j < 4; ++j) {
                dl_dp[h.poff + 2 * j + 1] += r * F.A * wl * Gw * wcd[j] * inv_fd;
                dl_dp[h.poff + 8 + 2 * j + 1] += -r * F.A * wl * wl * Gn * wct[j] * inv_fd;
            }
        }
    }
    // ForwardStateForce (deformable_state_force.cpp:114-125): gravity f_i = mass g
    // (gravitational_state_force.cpp:14-22) plus each planar contact's penalty
    // kn (-min(0, d))^(p-1) n and friction -min(kf, mu f_n / (|u_s| + eps)) u_s,
    // u_s the tangential velocity (planar_contact_state_force.cpp:16-37).
    Vec state_force(const Vec& q, const Vec& v) const {
        Vec f(dofs, 0.0);
        if (gravity.size() == 3) {
            const double m = density * element_volume;
            for (int i = 0; i < dofs; ++i) f[i] = m * gravity[i % 3];
        }
        const double eps = std::numeric_limits<double>::epsilon();
        for (const auto& c : pcontacts)
            for (int i = 0; i < nv; ++i) {
                const double* x = &q[3 * i];
                const double* vi = &v[3 * i];
                const double d = c.n[0] * x[0] + c.n[1] * x[1] + c.n[2] * x[2] + c.off;
                const double fn = c.kn * std::pow(-std::min(0.0, d), c.p - 1);
                const double vn = c.n[0] * vi[0] + c.n[1] * vi[1] + c.n[2] * vi[2];
                double us[3];
                for (int k = 0; k < 3; ++k) us[k] = vi[k] - vn * c.n[k];
                const double um = std::sqrt(us[0] * us[0] + us[1] * us[1] + us[2] * us[2]);
                const double cf = std::min(c.kf, c.mu * fn / (um + eps));
                for (int k = 0; k < 3; ++k) f[3 * i + k] += fn * c.n[k] - cf * us[k];
            }
        for (const auto& a : acontacts)
            for (int i = 0; i < nv; ++i) {
                double qp[3], r, n[3];
                if (!arc_geometry(a, &q[3 * i], qp, r, n)) continue;
                const double fn = a.kn * std::pow(-std::min(0.0, a.R - r), a.p - 1);
                const double* vi = &v[3 * i];
                const double vn = vi[0] * n[0] + vi[1] * n[1] + vi[2] * n[2];
                double us[3];
                for (int k = 0; k < 3; ++k) us[k] = vi[k] - vn * n[k];
                const double um = std::sqrt(us[0] * us[0] + us[1] * us[1] + us[2] * us[2]);
                const double cf = std::min(a.kf, a.mu * fn / (um + eps));
                for (int k = 0; k < 3; ++k) f[3 * i + k] += fn * n[k] - cf * us[k];
            }
        // each force is evaluated into its own vector and then added, as the
        // reference sums ForwardForce over its state forces
        for (const auto& m : mcontacts) {
            Vec fm(dofs, 0.0);
            mesh_contact_force(m, q, v, fm);
            for (int i = 0; i < dofs; ++i) f[i] += fm[i];
        }
        for (const auto& b : bballs) {
            Vec fb(dofs, 0.0);
            billiard_force(b, q, v, fb);
            for (int i = 0; i < dofs; ++i) f[i] += fb[i];
        }
        for (const auto& h : hydros) {
            const Vec fh = hydro_force(h, q, v);
            for (int i = 0; i < dofs; ++i) f[i] += fh[i];
        }
        return f;
    }
    // BackwardStateForce (deformable_state_force.cpp:127-145): gravity's
    // dl/dg = mass sum dl/df per axis (gravitational_state_force.cpp:24-33) and the
    // planar contacts' derivatives (planar_contact_state_force.cpp:39-84).
    void state_force_backward(const Vec& q, const Vec& v, const Vec& dl_df, Vec& dl_dq, Vec& dl_dv, Vec& dl_dp) const {
        dl_dq.assign(dofs, 0.0);
        dl_dv.assign(dofs, 0.0);
        dl_dp.assign(n_state_params, 0.0);
        if (gravity.size() == 3) {
            const double m = density * element_volume;
            for (int i = 0; i < dofs; ++i) dl_dp[gravity_off + i % 3] += m * dl_df[i];
        }
        const double eps = std::numeric_limits<double>::epsilon();
        for (const auto& c : pcontacts)
            for (int i = 0; i < nv; ++i) {
                const double* x = &q[3 * i];
                const double* vi = &v[3 * i];
                const double* g = &dl_df[3 * i];
                const double d = c.n[0] * x[0] + c.n[1] * x[1] + c.n[2] * x[2] + c.off;
                const double pen = -std::min(0.0, d);
                const double fn = c.kn * std::pow(pen, c.p - 1);
                // d fn / dq = kn (p-1) pen^(p-2) (-n) inside the plane, 0 outside
                const double dfn = d < 0 ? -c.kn * (c.p - 1) * std::pow(pen, c.p - 2) : 0.0;
                const double vn = c.n[0] * vi[0] + c.n[1] * vi[1] + c.n[2] * vi[2];
                double us[3];
                for (int k = 0; k < 3; ++k) us[k] = vi[k] - vn * c.n[k];
                const double um = std::sqrt(us[0] * us[0] + us[1] * us[1] + us[2] * us[2]);
                const double ue = um + eps;
                const double gn = g[0] * c.n[0] + g[1] * c.n[1] + g[2] * c.n[2];
                const double gu = g[0] * us[0] + g[1] * us[1] + g[2] * us[2];
                // normal part: f_n n
                for (int k = 0; k < 3; ++k) dl_dq[3 * i + k] += gn * dfn * c.n[k];
                dl_dp[c.poff] += gn * fn / c.kn;
                // (I - n n^T) g: the tangential projection of dl/df
                double gt[3];
                for (int k = 0; k < 3; ++k) gt[k] = g[k] - gn * c.n[k];
                if (c.kf <= c.mu * fn / ue) {
                    // ff = -kf u_s
                    for (int k = 0; k < 3; ++k) dl_dv[3 * i + k] += -c.kf * gt[k];
                    dl_dp[c.poff + 1] += -gu;
                } else {
                    // ff = -mu fn u_s / (|u_s| + eps)
                    for (int k = 0; k < 3; ++k) dl_dq[3 * i + k] += -c.mu * gu / ue * dfn * c.n[k];
                    // d(u_s / ue)/dv^T g = gt / ue - u_s (u_s . g) / ue^3: the reference takes
                    // d|u_s|/dv as P u_s / (|u_s| + eps)
                    for (int k = 0; k < 3; ++k) dl_dv[3 * i + k] += -c.mu * fn * (gt[k] / ue - us[k] * gu / (ue * ue * ue));
                    dl_dp[c.poff] += -c.mu * gu / ue * (fn / c.kn);
                    dl_dp[c.poff + 2] += -gu / ue * fn;
                }
            }
        // arcs (arc_contact_state_force.cpp:65-131), as 3x3 Jacobians J[k][j] = d out_k / d in_j
        for (const auto& a : acontacts)
            for (int i = 0; i < nv; ++i) {
                double qp[3], r, n[3];
                if (!arc_geometry(a, &q[3 * i], qp, r, n)) continue;
                const double* vi = &v[3 * i];
                const double* g = &dl_df[3 * i];
                const double re = r + eps, d = a.R - r, pen = -std::min(0.0, d);
                const double fn = a.kn * std::pow(pen, a.p - 1);
                double P[3][3], dn[3][3], dfn_q[3], dus_q[3][3], dus_v[3][3];
                for (int k = 0; k < 3; ++k)
                    for (int j = 0; j < 3; ++j) P[k][j] = (k == j) - a.dir[k] * a.dir[j];
                // d|qp|/dq = P^T qp / (|qp| + eps); d f_n / dq = kn (p-1) pen^(p-2) d|qp|/dq inside
                for (int j = 0; j < 3; ++j) {
                    const double dr = (P[0][j] * qp[0] + P[1][j] * qp[1] + P[2][j] * qp[2]) / re;
                    dfn_q[j] = d < 0 ? a.kn * (a.p - 1) * std::pow(pen, a.p - 2) * dr : 0.0;
                }
                // dn/dq = -(I - n n^T) P / (|qp| + eps)
                for (int k = 0; k < 3; ++k)
                    for (int j = 0; j < 3; ++j) {
                        double t = 0;
                        for (int m = 0; m < 3; ++m) t += ((k == m) - n[k] * n[m]) * P[m][j];
                        dn[k][j] = -t / re;
                    }
                const double vn = vi[0] * n[0] + vi[1] * n[1] + vi[2] * n[2];
                double us[3];
                for (int k = 0; k < 3; ++k) us[k] = vi[k] - vn * n[k];
                // du_s/dq = -(n v^T + (v . n) I) dn/dq, du_s/dv = I - n n^T
                for (int k = 0; k < 3; ++k)
                    for (int j = 0; j < 3; ++j) {
                        double t = 0;
                        for (int m = 0; m < 3; ++m) t += (n[k] * vi[m] + (k == m) * vn) * dn[m][j];
                        dus_q[k][j] = -t;
                        dus_v[k][j] = (k == j) - n[k] * n[j];
                    }
                const double um = std::sqrt(us[0] * us[0] + us[1] * us[1] + us[2] * us[2]), ue = um + eps;
                double Jq[3][3], Jv[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, Jp[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
                for (int k = 0; k < 3; ++k) {
                    for (int j = 0; j < 3; ++j) Jq[k][j] = n[k] * dfn_q[j] + dn[k][j] * fn;  // d(f_n n)/dq
                    Jp[k][0] = n[k] * fn / a.kn;
                }
                if (a.kf <= a.mu * fn / ue) {  // ff = -kf u_s
                    for (int k = 0; k < 3; ++k) {
                        for (int j = 0; j < 3; ++j) { Jq[k][j] += -a.kf * dus_q[k][j]; Jv[k][j] = -a.kf * dus_v[k][j]; }
                        Jp[k][1] = -us[k];
                    }
                } else {  // ff = -mu f_n u_s / (|u_s| + eps)
                    // the reference forms d|u_s|/dq as (du_s/dq) u_s / (|u_s| + eps), not its transpose
                    double dum_q[3], dum_v[3];
                    for (int k = 0; k < 3; ++k) {
                        dum_q[k] = (dus_q[k][0] * us[0] + dus_q[k][1] * us[1] + dus_q[k][2] * us[2]) / ue;
                        dum_v[k] = (dus_v[k][0] * us[0] + dus_v[k][1] * us[1] + dus_v[k][2] * us[2]) / ue;
                    }
                    for (int k = 0; k < 3; ++k) {
                        for (int j = 0; j < 3; ++j) {
                            Jq[k][j] += -a.mu * (us[k] * dfn_q[j] / ue + fn * (dus_q[k][j] / ue - us[k] * dum_q[j] / (ue * ue)));
                            Jv[k][j] = -a.mu * fn * (dus_v[k][j] / ue - us[k] * dum_v[j] / (ue * ue));
                        }
                        Jp[k][0] += -us[k] / ue * a.mu * fn / a.kn;
                        Jp[k][2] = -us[k] / ue * fn;
                    }
                }
                for (int j = 0; j < 3; ++j) {
                    dl_dq[3 * i + j] += g[0] * Jq[0][j] + g[1] * Jq[1][j] + g[2] * Jq[2][j];
                    dl_dv[3 * i + j] += g[0] * Jv[0][j] + g[1] * Jv[1][j] + g[2] * Jv[2][j];
                    dl_dp[a.poff + j] += g[0] * Jp[0][j] + g[1] * Jp[1][j] + g[2] * Jp[2][j];
                }
            }
        for (const auto& b : bballs) billiard_backward(b, q, v, dl_df, dl_dq, dl_dp);
        for (const auto& h : hydros) hydro_backward(h, q, v, dl_df, dl_dq, dl_dv, dl_dp);
    }

    // SetupProjectiveDynamicsSolver, deformable_projective_dynamics_forward.cpp:8-183
    // (material 'none' -> base_w = 0; the unused Acc_/AinvIc_ are omitted).
    void setup(double dt) {
        if (pd_ready) return;
        const double mass = density * element_volume;
        const double inv_h2m = mass / (dt * dt);
        std::vector<double> w_e(ne, 0.0);
        for (int e = 0; e < ne; ++e)
            for (auto& en : energies) if (en.member[e]) w_e[e] += en.elem_w(e);
        std::vector<std::map<int, double>> rows[3];
        for (int d = 0; d < 3; ++d) rows[d].assign(nv, {});
        for (int d = 0; d < 3; ++d)
            for (int i = 0; i < nv; ++i) rows[d][i][i] += dirichlet.count(3 * i + d) ? 1.0 : inv_h2m;
        for (int e = 0; e < ne; ++e) {
            const int* v = &elems[(size_t)e * edim];
            for (int s = 0; s < samples; ++s) {
                const double* g = G(e, s);
                const double w = w_e[e] * element_volume / samples;
                for (int k = 0; k < edim; ++k)
                    for (int l = 0; l < edim; ++l) {
                        const double val = w * (g[k * 3] * g[l * 3] + g[k * 3 + 1] * g[l * 3 + 1] + g[k * 3 + 2] * g[l * 3 + 2]);
                        for (int d = 0; d < 3; ++d) {
                            if (dirichlet.count(3 * v[k] + d) || dirichlet.count(3 * v[l] + d)) continue;
                            rows[d][v[k]][v[l]] += val;
                        }
                    }
            }
        }
        // PdMuscleEnergy terms w A^T M^T M A (forward.cpp:112-150): with
        // M vec(F) = F m the (k, l) block is w (m.g_k)(m.g_l) I_3
        for (const Muscle& mu : muscles) {
            const double w = mu.stiffness * element_volume / samples;
            for (int e : mu.elems) {
                const int* v = &elems[(size_t)e * edim];
                for (int s = 0; s < samples; ++s) {
                    const double* g = G(e, s);
                    for (int k = 0; k < edim; ++k)
                        for (int l = 0; l < edim; ++l) {
                            const double mk = g[k * 3] * mu.m[0] + g[k * 3 + 1] * mu.m[1] + g[k * 3 + 2] * mu.m[2];
                            const double ml = g[l * 3] * mu.m[0] + g[l * 3 + 1] * mu.m[1] + g[l * 3 + 2] * mu.m[2];
                            for (int d = 0; d < 3; ++d) {
                                if (dirichlet.count(3 * v[k] + d) || dirichlet.count(3 * v[l] + d)) continue;
                                rows[d][v[k]][v[l]] += w * mk * ml;
                            }
                        }
                }
            }
        }
        // Ordering from the pattern of axis 0 (identical Dirichlet sets give identical patterns;
        // a per-axis ordering is computed otherwise).
        for (int d = 0; d < 3; ++d) {
            std::vector<int> Ap(nv + 1, 0), Ai;
            std::vector<double> Ax;
            std::vector<std::vector<int>> adj(nv);
            for (int i = 0; i < nv; ++i) {
                for (auto& kv : rows[d][i]) {
                    Ai.push_back(kv.first);
                    Ax.push_back(kv.second);
                    if (kv.first != i) adj[i].push_back(kv.first);
                }
                Ap[i + 1] = (int)Ai.size();
            }
            std::vector<int> all(nv), perm, side(nv, 0);
            std::iota(all.begin(), all.end(), 0);
            rcb_order(all, X, adj, perm, side);
            ldlt[d].factor(nv, Ap, Ai, Ax, perm);
            csr_p[d] = Ap;
            csr_i[d] = Ai;
            csr_x[d] = Ax;
        }
        pd_ready = true;
        pd_dt = dt;
        lhs_valid = false;  // the baseline's frozen-node factors belong to the old matrix
    }

    Vec solve_axis(int d, const Vec& b) const {
        Vec x(nv);
        ldlt[d].solve(b.data(), x.data());
        // Optional iterative refinement with an extended-precision residual: a check
        // on how much of a difference comes from the conditioning of A alone.
        for (int it = 0; it < refine_steps; ++it) {
            Vec r(nv), dx(nv);
            for (int i = 0; i < nv; ++i) {
                long double acc = b[i];
                for (int p = csr_p[d][i]; p < csr_p[d][i + 1]; ++p) acc -= (long double)csr_x[d][p] * x[csr_i[d][p]];
                r[i] = (double)acc;
            }
            ldlt[d].solve(r.data(), dx.data());
            for (int i = 0; i < nv; ++i) x[i] += dx[i];
        }
        return x;
    }

    // ProjectiveDynamicsLocalStep, deformable_projective_dynamics_forward.cpp:186-292.
    Vec local_step(const Vec& q) const {
        Vec qb(dofs, 0.0);
        for (auto& kv : dirichlet) qb[kv.first] = kv.second;
        std::vector<double> acc((size_t)ne * edim * 3, 0.0);
        for (auto& en : energies) {
#pragma omp parallel for schedule(static)
            for (int e = 0; e < ne; ++e) {
                if (!en.member[e]) continue;
                const double w = en.elem_w(e) * element_volume / samples;
                for (int s = 0; s < samples; ++s) {
                    const M3 F = deformation_gradient(q.data(), e, s);
                    const M3 Fb = deformation_gradient(qb.data(), e, s);
                    FAux a;
                    a.init(F);
                    const M3 Bp = project(en, a);
                    add_At(add(Bp, scale(Fb, -1.0)), e, s, w, &acc[(size_t)e * edim * 3]);
                }
            }
        }
        // PdMuscleEnergy (forward.cpp:240-270): w A^T (M^T Bp - M^T M F_bound),
        // i.e. A^T of the matrix (Bp - F_bound m) m^T
        check((int)act.size() == act_dofs, "Inconsistent actuation size.");
        int base = 0;
        for (const Muscle& mu : muscles) {
            const double w = mu.stiffness * element_volume / samples;
            const int cnt = (int)mu.elems.size();
#pragma omp parallel for schedule(static)
            for (int ei = 0; ei < cnt; ++ei) {
                const int e = mu.elems[ei];
                const double a = act[base + ei];
                for (int s = 0; s < samples; ++s) {
                    const M3 F = deformation_gradient(q.data(), e, s);
                    const M3 Fb = deformation_gradient(qb.data(), e, s);
                    double Fm[3], l, Fbm[3], lb;
                    muscle_fm(F, mu.m, Fm, l);
                    muscle_fm(Fb, mu.m, Fbm, lb);
                    const double r = l <= std::numeric_limits<double>::epsilon() ? 0.0 : a / l;
                    M3 P;
                    for (int i = 0; i < 3; ++i)
                        for (int j = 0; j < 3; ++j) P(i, j) = (r * Fm[i] - Fbm[i]) * mu.m[j];
                    add_At(P, e, s, w, &acc[(size_t)e * edim * 3]);
                }
            }
            base += cnt;
        }
        Vec rhs(dofs, 0.0);
        for (int e = 0; e < ne; ++e)
            for (int k = 0; k < edim; ++k)
                for (int i = 0; i < 3; ++i) rhs[3 * elems[(size_t)e * edim + k] + i] += acc[((size_t)e * edim + k) * 3 + i];
        return rhs;
    }

    // SetupProjectiveDynamicsLocalStepDifferential, deformable_projective_dynamics_backward.cpp:10-80:
    // per element the 3*edim square matrix sum_s A_s^T (sum_energies w dBp) A_s with
    // w = stiffness() (mean) * vol / s, or the element's own stiffness when
    // `exact` (exact-adjoint mode).
    // filter/stage/rho: spectral filter of each sample's proximal-map block
    // (filter_params, filter_prox_jacobian); filtered (optional) receives the
    // number of samples whose block moved by more than 1e-9 in mu.
    std::vector<double> local_differential(const Vec& q, bool exact, int filter = 0, double rho = 0.9,
                                           long long* filtered = nullptr, int stage = 1) const {
        double theta = 1, lo = 0;
        bool factive = false;
        filter_params(filter, stage, rho, theta, lo, factive);
        for (auto& kv : dirichlet) check(q[kv.first] == kv.second, "Boundary conditions violated.");
        long long nfilt = 0;
        const int nd = 3 * edim;
        std::vector<double> K((size_t)ne * nd * nd, 0.0);
        // (muscle group, actuation index) of every element
        std::vector<std::vector<std::pair<int, int>>> mus_of(ne);
        {
            int base = 0;
            for (int gi = 0; gi < (int)muscles.size(); ++gi) {
                for (int ei = 0; ei < (int)muscles[gi].elems.size(); ++ei) mus_of[muscles[gi].elems[ei]].push_back({gi, base + ei});
                base += (int)muscles[gi].elems.size();
            }
            check((int)act.size() == act_dofs, "Inconsistent actuation size.");
        }
#pragma omp parallel for schedule(static) reduction(+ : nfilt)
        for (int e = 0; e < ne; ++e) {
            double* Ke = &K[(size_t)e * nd * nd];
            for (int s = 0; s < samples; ++s) {
                const M3 F = deformation_gradient(q.data(), e, s);
                FAux a;
                a.init(F);
                double wdBp[81] = {0};  // 9x9, column c = d(vec p)/d(vec F)_c
                bool any = false;
                for (auto& en : energies) {
                    if (!en.member[e]) continue;
                    any = true;
                    const double w = (exact ? en.elem_w(e) : en.stiffness) * element_volume / samples;
                    for (int c = 0; c < 9; ++c) {
                        M3 dF = M3::zero();
                        dF.a[c] = 1;
                        const M3 dp = project_diff(en, a, dF);
                        for (int r = 0; r < 9; ++r) wdBp[r + 9 * c] += w * dp.a[r];
                    }
                }
                if (any && factive) {
                    double aw = 0;
                    for (auto& en : energies) if (en.member[e]) aw += en.elem_w(e) * element_volume / samples;
                    if (filter_prox_jacobian(wdBp, aw, theta, lo) > 1e-9) ++nfilt;
                }
                // PdMuscleEnergy blocks w A^T M^T dBp/dF A (backward.cpp:54-80,
                // ProjectToManifoldDifferential, pd_muscle_energy.cpp:127-165),
                // outside the filter: d(a Fm / l) = (a / l)(I - n n^T) dF m with
                // n = Fm / l, so entry (i + 3j, p + 3q) is w (a/l)(d_ip - n_i n_p) m_q m_j
                for (const auto& gm : mus_of[e]) {
                    const Muscle& mu = muscles[gm.first];
                    any = true;
                    double Fm[3], l;
                    muscle_fm(F, mu.m, Fm, l);
                    if (l <= std::numeric_limits<double>::epsilon()) continue;
                    const double c = mu.stiffness * element_volume / samples * act[gm.second] / l;
                    const double n[3] = {Fm[0] / l, Fm[1] / l, Fm[2] / l};
                    for (int p = 0; p < 3; ++p)
                        for (int qq = 0; qq < 3; ++qq)
                            for (int i = 0; i < 3; ++i)
                                for (int j = 0; j < 3; ++j)
                                    wdBp[(i + 3 * j) + 9 * (p + 3 * qq)] +=
                                        c * ((i == p ? 1.0 : 0.0) - n[i] * n[p]) * mu.m[qq] * mu.m[j];
                }
                if (!any) continue;
                // A: 9 x nd with A[i+3j, i+3k] = G(k,j).
                const double* g = G(e, s);
                for (int col = 0; col < nd; ++col) {
                    const int kc = col / 3, ic = col % 3;
                    double t[9];  // wdBp * A[:, col]
                    for (int r = 0; r < 9; ++r) {
                        double acc = 0;
                        for (int j = 0; j < 3; ++j) acc += wdBp[r + 9 * (ic + 3 * j)] * g[kc * 3 + j];
                        t[r] = acc;
                    }
                    for (int row = 0; row < nd; ++row) {
                        const int kr = row / 3, ir = row % 3;
                        double acc = 0;
                        for (int j = 0; j < 3; ++j) acc += g[kr * 3 + j] * t[ir + 3 * j];
                        Ke[(size_t)row * nd + col] += acc;
                    }
                }
            }
        }
        if (filtered) *filtered = nfilt;
        return K;
    }

    // ApplyProjectiveDynamicsLocalStepDifferential, deformable_projective_dynamics_backward.cpp:82-139.
    Vec apply_local_differential(const std::vector<double>& K, const Vec& dq) const {
        const int nd = 3 * edim;
        Vec out(dofs, 0.0);
        std::vector<double> tmp((size_t)ne * nd, 0.0);
#pragma omp parallel for schedule(static)
        for (int e = 0; e < ne; ++e) {
            const int* v = &elems[(size_t)e * edim];
            double x[24];
            for (int k = 0; k < edim; ++k) for (int i = 0; i < 3; ++i) x[3 * k + i] = dq[3 * v[k] + i];
            const double* Ke = &K[(size_t)e * nd * nd];
            for (int r = 0; r < nd; ++r) {
                double acc = 0;
                for (int c = 0; c < nd; ++c) acc += Ke[(size_t)r * nd + c] * x[c];
                tmp[(size_t)e * nd + r] = acc;
            }
        }
        for (int e = 0; e < ne; ++e)
            for (int r = 0; r < nd; ++r) out[3 * elems[(size_t)e * edim + r / 3] + r % 3] += tmp[(size_t)e * nd + r];
        return out;
    }

    void forward(const Vec& q, const Vec& v, const Vec& f_ext, double dt, const std::map<std::string, double>& opt,
                 Vec& q_next, Vec& v_next, std::vector<int>& active);

    // ---- the DiffPD baseline forward (ForwardProjectiveDynamics with a
    // pd_eigen method other than pd_eigen_alg_phd, forward.cpp:294-664) ----
    std::vector<int> lhs_frozen;  // frozen nodes of the cached factors below
    SparseLDLT lhs_ldlt[3];
    bool lhs_valid = false;
    double baseline_energy(const Vec& q) const;
    Vec baseline_force(const Vec& q) const;
    Vec pd_lhs_solve(const Vec& rhs, const std::vector<int>& frozen);
    Vec pd_nonlinear_solve(const Vec& q_init, double inv_h2m, const Vec& rhs, const std::map<int, double>& extra,
                           const std::map<std::string, double>& opt, int* iters);
    void forward_baseline(const std::string& method, const Vec& q, const Vec& v, const Vec& f_ext, double dt,
                          const std::map<std::string, double>& opt, Vec& q_next, Vec& v_next, std::vector<int>& active);
    int last_baseline_contact_iters = 0, last_baseline_pd_iters = 0;
    Vec pd_lhs_op(const Vec& x, const std::vector<int>& frozen) const;
    Vec pd_force_differential(const Vec& q, const Vec& dq) const;
    void backward_baseline(const Vec& q, const Vec& v, const Vec& f_ext, double dt, const Vec& q_next,
                           const std::vector<int>& active, const Vec& dl_dq_next, const Vec& dl_dv_next,
                           const std::map<std::string, double>& opt, Vec& dl_dq, Vec& dl_dv, Vec& dl_df_ext,
                           Vec& dl_dmat_w, Vec& dl_dstate_p);
    int last_baseline_bwd_iters = 0, last_baseline_bwd_newton = 0;
    AlgPhdReplayCache replay_fixed_active_set(const Vec& q, const Vec& v, const Vec& f_ext, double dt,
                                              const std::vector<int>& active, const std::map<std::string, double>& opt);
    void backward(const Vec& q, const Vec& v, const Vec& f_ext, double dt, const Vec& q_next, const Vec& v_next,
                  const std::vector<int>& active, const Vec& dl_dq_next, const Vec& dl_dv_next,
                  const std::map<std::string, double>& opt, Vec& dl_dq, Vec& dl_dv, Vec& dl_df_ext, Vec& dl_dmat_w,
                  Vec& dl_dstate_p, Vec& dl_dw_elem);
};

static double opt_or(const std::map<std::string, double>& o, const char* k, double dflt) {
    auto it = o.find(k);
    return it == o.end() ? dflt : it->second;
}
static double opt_req(const std::map<std::string, double>& o, const char* k) {
    auto it = o.find(k);
    check(it != o.end(), std::string("Missing option ") + k + ".");
    return it->second;
}

// Type-II Anderson acceleration (Walker & Ni 2011) with the reference's
// regularization and restart rule, forward_gpu.cpp:1346-1379 and
// backward_gpu.cpp:822-851. Returns true if the extrapolated iterate is used.
struct Anderson {
    int window;
    std::deque<Vec> xh, gh;
    // replay: the fixed-active-set replay's variant (backward_gpu.cpp:381-405):
    // reg = 1e-10 max(tr(dG^T dG) / m, 1) and no restart
    bool replay = false;
    explicit Anderson(int w, bool replay_ = false) : window(w), replay(replay_) {}
    // x_prev: previous iterate, g_k = x_new - x_prev. On acceptance writes candidate.
    bool step(const Vec& x_prev, const Vec& g_k, Vec& candidate) {
        xh.push_back(x_prev);
        gh.push_back(g_k);
        const int m = (int)gh.size() - 1;
        bool used = false;
        if (m >= 1) {
            const size_t n = g_k.size();
            std::vector<Vec> dG(m, Vec(n)), dQ(m, Vec(n));
            double sq = 0;
            for (int i = 0; i < m; ++i)
                for (size_t j = 0; j < n; ++j) {
                    dG[i][j] = gh[i + 1][j] - gh[i][j];
                    dQ[i][j] = xh[i + 1][j] - xh[i][j];
                    sq += dG[i][j] * dG[i][j];
                }
            std::vector<double> GtG(m * m);
            Vec rhs(m);
            for (int a = 0; a < m; ++a) {
                for (int b = 0; b < m; ++b) GtG[a * m + b] = dot(dG[a], dG[b]);
                rhs[a] = dot(dG[a], g_k);
            }
            const double reg = replay ? 1e-10 * std::max(sq / std::max(m, 1), 1.0) : std::max(1e-6 * sq / std::max(m, 1), 1e-12);
            for (int a = 0; a < m; ++a) GtG[a * m + a] += reg;
            const Vec gamma = dense_ldlt_solve(GtG, m, rhs);
            if (!replay && norm(gamma) >= 10.0) {
                xh.clear();
                gh.clear();
            } else {
                candidate.assign(n, 0.0);
                for (size_t j = 0; j < n; ++j) {
                    double c = x_prev[j] + g_k[j];
                    for (int i = 0; i < m; ++i) c -= (dQ[i][j] + dG[i][j]) * gamma[i];
                    candidate[j] = c;
                }
                used = true;
            }
        }
        while ((int)xh.size() > window) { xh.pop_front(); gh.pop_front(); }
        return used;
    }
};

// ForwardProjectiveDynamicsAlgPhd, deformable_projective_dynamics_forward_gpu.cpp:387-1594.
void Body::forward(const Vec& q, const Vec& v, const Vec& f_ext, double dt, const std::map<std::string, double>& opt,
                   Vec& q_next, Vec& v_next, std::vector<int>& active) {
    const int max_pd_iter = (int)opt_req(opt, "max_pd_iter");
    const double abs_tol = opt_req(opt, "abs_tol");
    const double rel_tol = opt_req(opt, "rel_tol");
    const int verbose = (int)opt_req(opt, "verbose");
    opt_req(opt, "thread_ct");
    const int aa_window = (int)opt_or(opt, "aa_window", 5);
    check((int)q.size() == dofs && (int)v.size() == dofs && (int)f_ext.size() == dofs, "Inconsistent state size.");
    setup(dt);
    const double h = dt, mass = element_volume * density, h2m = h * h / mass, inv_h2m = mass / (h * h);
    const Vec fs = state_force(q, v);
    Vec q_tilde(dofs);
    for (int i = 0; i < dofs; ++i) q_tilde[i] = q[i] + h * v[i] + h2m * f_ext[i] + h2m * fs[i];

    // New-rollout detection for the forward tape (:437-449).
    if (!tape.empty()) {
        const auto& last = tape.back();
        bool seq = last.q_sol.size() == q.size();
        if (seq) {
            Vec d(dofs);
            for (int i = 0; i < dofs; ++i) d[i] = last.q_sol[i] - q[i];
            seq = norm(d) <= 1e-8 * std::max(1.0, std::max(norm(last.q_sol), norm(q)));
        }
        if (!seq) tape.clear();
    }

    // Contact candidates and frames (:456-528).
    // The mesh boundary keeps the candidates within 5 radii of the opposite
    // surface at the frame start, anchored at their closest points (:460-509).
    const bool mesh = boundary.type == kMesh;
    std::vector<int> cand;
    std::vector<std::array<double, 3>> nrm, t1, t2;
    for (auto& kv : fb_vertices) {
        std::array<double, 3> a, b, c;
        if (mesh) {
            const double du = boundary.mesh_frame(&q[3 * kv.first], q, kv.first, a.data(), b.data(), c.data());
            if (du > boundary.radius * 5.0) continue;
        } else {
            boundary.frame(&q[3 * kv.first], a.data(), b.data(), c.data());
        }
        cand.push_back(kv.first);
        t1.push_back(a);
        t2.push_back(b);
        nrm.push_back(c);
    }
    const int K = (int)cand.size();

    // Delassus W via per-column A^-1 (:840-1013).
    std::vector<std::array<Vec, 3>> Acol(K), At1(K), At2(K);
    for (int d = 0; d < 3; ++d)
        for (int c = 0; c < K; ++c) {
            Vec rhs(nv, 0.0);
            rhs[cand[c]] = nrm[c][d];
            Acol[c][d] = solve_axis(d, rhs);
        }
    std::vector<double> W((size_t)K * K, 0.0);
    for (int a = 0; a < K; ++a)
        for (int b = a; b < K; ++b) {
            double w = 0;
            for (int d = 0; d < 3; ++d) w += nrm[b][d] * Acol[a][d][cand[b]];
            W[(size_t)a * K + b] = W[(size_t)b * K + a] = w;
        }
    const double fmu = opt_or(opt, "friction_mu", 0.0);
    const double restitution = opt_or(opt, "restitution", 0.0);
    const bool do_fric = (K > 0) && (fmu > 0);
    std::vector<std::array<double, 4>> Wtt(K, {0, 0, 0, 0});
    if (do_fric) {
        for (int d = 0; d < 3; ++d)
            for (int c = 0; c < K; ++c) {
                Vec r1(nv, 0.0), r2(nv, 0.0);
                r1[cand[c]] = t1[c][d];
                r2[cand[c]] = t2[c][d];
                At1[c][d] = solve_axis(d, r1);
                At2[c][d] = solve_axis(d, r2);
            }
        for (int c = 0; c < K; ++c) {
            double w11 = 0, w12 = 0, w22 = 0;
            const int ni = cand[c];
            for (int d = 0; d < 3; ++d) {
                w11 += t1[c][d] * At1[c][d][ni];
                w12 += t1[c][d] * At2[c][d][ni];
                w22 += t2[c][d] * At2[c][d][ni];
            }
            Wtt[c] = {w11, w12, w12, w22};
        }
    }
    Vec rprec(K);
    for (int c = 0; c < K; ++c) rprec[c] = std::max(h * h * W[(size_t)c * K + c], 1e-10);

    Vec q_sol = q, q_prev = q;
    Vec lambda(K, 0.0), lt1(K, 0.0), lt2(K, 0.0), omega_last(K, 0.0);
    if (!prev_candidates.empty() && prev_lambda.size() == prev_candidates.size()) {
        std::unordered_map<int, int> pidx;
        for (size_t i = 0; i < prev_candidates.size(); ++i) pidx[prev_candidates[i]] = (int)i;
        for (int c = 0; c < K; ++c) {
            auto it = pidx.find(cand[c]);
            if (it != pidx.end()) lambda[c] = prev_lambda[it->second];
        }
    }
    Anderson aa(aa_window);
    Vec prev_pd_rhs(dofs, 0.0);
    bool converged = false;
    int iters = 0;
    for (int iter = 0; iter < max_pd_iter; ++iter) {
        const Vec pd_rhs = local_step(q_sol);
        Vec b_full(dofs);
        for (int i = 0; i < dofs; ++i) b_full[i] = inv_h2m * q_tilde[i] + pd_rhs[i];
        Vec b_bc = b_full;
        for (auto& kv : dirichlet) b_bc[kv.first] = 0;
        q_prev = q_sol;
        if (K > 0) {
            Vec omega(K), E(K), hvec(K);
            for (int c = 0; c < K; ++c) {
                const double dj = mesh ? boundary.mesh_signed(&q_sol[3 * cand[c]], q_sol, cand[c], nrm[c].data())
                                       : boundary.distance(&q_sol[3 * cand[c]]);
                const double lj = lambda[c], rj = rprec[c];
                double den = std::sqrt(dj * dj + rj * rj * lj * lj);
                if (den < 1e-14) den = 1e-14;
                omega[c] = 1.0 - dj / den;
                E[c] = (1.0 - rj * lj / den) * rj;
                const double phi = dj + rj * lj - den;
                double Hq = 0;
                for (int d = 0; d < 3; ++d) Hq += nrm[c][d] * q_sol[3 * cand[c] + d];
                hvec[c] = -phi + omega[c] * Hq;
            }
            omega_last = omega;
            std::vector<double> M((size_t)K * K);
            for (int a = 0; a < K; ++a)
                for (int b = 0; b < K; ++b) M[(size_t)a * K + b] = omega[a] * W[(size_t)a * K + b] * omega[b];
            for (int a = 0; a < K; ++a) M[(size_t)a * K + a] += E[a];
            // A^-1 b_full per axis (:1199-1210; note b_full, not b_bc).
            std::array<Vec, 3> Ab;
            for (int d = 0; d < 3; ++d) {
                Vec bd(nv);
                for (int i = 0; i < nv; ++i) bd[i] = b_full[3 * i + d];
                Ab[d] = solve_axis(d, bd);
            }
            Vec Ag(dofs);
            for (int d = 0; d < 3; ++d) {
                Vec gd = Ab[d];
                for (int c = 0; c < K; ++c) {
                    const double s = omega[c] * lambda[c];
                    for (int i = 0; i < nv; ++i) gd[i] += s * Acol[c][d][i];
                    if (do_fric)
                        for (int i = 0; i < nv; ++i) gd[i] += lt1[c] * At1[c][d][i] + lt2[c] * At2[c][d][i];
                }
                for (int i = 0; i < nv; ++i) Ag[3 * i + d] = gd[i];
            }
            Vec rhs(K);
            for (int c = 0; c < K; ++c) {
                double val = 0;
                for (int d = 0; d < 3; ++d) val += nrm[c][d] * Ag[3 * cand[c] + d];
                rhs[c] = hvec[c] - omega[c] * val;
            }
            const Vec dl = dense_ldlt_solve(M, K, rhs);
            if (verbose > 1)
                for (int c = 0; c < K; ++c) {
                    if (omega[c] == 0.0) continue;
                    double nu = 0, nhv = 0;
                    for (int d = 0; d < 3; ++d) { nu += nrm[c][d] * Ab[d][cand[c]]; nhv += nrm[c][d] * (Ag[3 * cand[c] + d] - Ab[d][cand[c]]); }
                    fprintf(stderr, "contact %d c=%d d=%.17g om=%.17g E=%.17g h=%.17g nu=%.17g nhv=%.17g Hcc=%.17g dl=%.17g\n", iter, c,
                            boundary.distance(&q_sol[3 * cand[c]]), omega[c], E[c], hvec[c], nu, nhv, W[(size_t)c * K + c], dl[c]);
                }
            for (int c = 0; c < K; ++c) lambda[c] = std::min(std::max(lambda[c] + dl[c], 0.0), 1e8);
            Vec wl(K);
            for (int c = 0; c < K; ++c) wl[c] = omega[c] * lambda[c];
            if (do_fric) {
                for (int c = 0; c < K; ++c) {
                    if (lambda[c] <= 0.0) { lt1[c] = lt2[c] = 0.0; continue; }
                    const int node = cand[c];
                    double qn[3], step[3];
                    for (int d = 0; d < 3; ++d) {
                        qn[d] = Ab[d][node];
                        for (int b = 0; b < K; ++b) qn[d] += wl[b] * Acol[b][d][node];
                        step[d] = qn[d] - q[3 * node + d];
                    }
                    const double sn = step[0] * nrm[c][0] + step[1] * nrm[c][1] + step[2] * nrm[c][2];
                    double ut[3];
                    for (int d = 0; d < 3; ++d) ut[d] = step[d] - sn * nrm[c][d];
                    const double u1 = ut[0] * t1[c][0] + ut[1] * t1[c][1] + ut[2] * t1[c][2];
                    const double u2 = ut[0] * t2[c][0] + ut[1] * t2[c][1] + ut[2] * t2[c][2];
                    coulomb_step(Wtt[c].data(), u1, u2, fmu, lambda[c], lt1[c], lt2[c]);
                }
            }
            for (int d = 0; d < 3; ++d) {
                Vec qd = Ab[d];
                for (int c = 0; c < K; ++c) {
                    for (int i = 0; i < nv; ++i) qd[i] += wl[c] * Acol[c][d][i];
                    if (do_fric)
                        for (int i = 0; i < nv; ++i) qd[i] += lt1[c] * At1[c][d][i] + lt2[c] * At2[c][d][i];
                }
                for (int i = 0; i < nv; ++i) q_sol[3 * i + d] = qd[i];
            }
        } else {
            for (int d = 0; d < 3; ++d) {
                Vec bd(nv);
                for (int i = 0; i < nv; ++i) bd[i] = b_bc[3 * i + d];
                const Vec qd = solve_axis(d, bd);
                for (int i = 0; i < nv; ++i) q_sol[3 * i + d] = qd[i];
            }
        }
        ++iters;
        for (auto& kv : dirichlet) q_sol[kv.first] = kv.second;
        if (aa_window > 0) {
            Vec gk(dofs), cand_q;
            for (int i = 0; i < dofs; ++i) gk[i] = q_sol[i] - q_prev[i];
            if (aa.step(q_prev, gk, cand_q)) {
                q_sol = cand_q;
                for (auto& kv : dirichlet) q_sol[kv.first] = kv.second;
            }
        }
        Vec dq(dofs), dr(dofs);
        for (int i = 0; i < dofs; ++i) { dq[i] = q_sol[i] - q_prev[i]; dr[i] = pd_rhs[i] - prev_pd_rhs[i]; }
        const double abs_err = norm(dq), rhs_norm = norm(q_prev), rc = norm(dr), rr = norm(prev_pd_rhs);
        prev_pd_rhs = pd_rhs;
        if (verbose > 0) fprintf(stderr, "alg_phd PD iter %d  dq=%g  dpd_rhs=%g\n", iter, abs_err, rc);
        if (iter > 0 && abs_err <= rel_tol * rhs_norm + abs_tol && rc <= rel_tol * rr + abs_tol) { converged = true; break; }
    }
    frame_iters.push_back(iters);
    frame_converged.push_back(converged ? 1 : 0);

    // Restitution post-pass (:1483-1531).
    if (restitution > 0) {
        // mesh: the opposite surface where the last UpdateVertices left it (the
        // last iterate's start), slop 0.1 radius (:1485-1505)
        const double slop = opt_or(opt, "friction_contact_slop", mesh ? 0.1 * boundary.radius : 1e-8);
        for (int c = 0; c < K; ++c) {
            if (lambda[c] <= 0) continue;
            const int node = cand[c];
            double sd, nh[3];
            if (mesh) {
                double cl[3], fn[3];
                const double* x = &q_sol[3 * node];
                if (!boundary.mesh_closest(x, q_prev, node, cl, fn)) continue;
                const double dp[3] = {x[0] - cl[0], x[1] - cl[1], x[2] - cl[2]};
                sd = dp[0] * nrm[c][0] + dp[1] * nrm[c][1] + dp[2] * nrm[c][2] - boundary.radius;
                const double dl = std::sqrt(dp[0] * dp[0] + dp[1] * dp[1] + dp[2] * dp[2]);
                for (int d = 0; d < 3; ++d) nh[d] = dl > 1e-12 ? dp[d] / dl : fn[d];
                if (nh[0] * nrm[c][0] + nh[1] * nrm[c][1] + nh[2] * nrm[c][2] < 0) for (int d = 0; d < 3; ++d) nh[d] = -nh[d];
            } else {
                sd = boundary.distance(&q_sol[3 * node]);
                boundary.normal(&q_sol[3 * node], nh);
            }
            const double nn = std::sqrt(nh[0] * nh[0] + nh[1] * nh[1] + nh[2] * nh[2]);
            if (nn <= 1e-12) continue;
            for (int d = 0; d < 3; ++d) nh[d] /= nn;
            if (sd > slop) continue;
            const double vin = v[3 * node] * nh[0] + v[3 * node + 1] * nh[1] + v[3 * node + 2] * nh[2];
            if (vin < 0) {
                double vout = 0;
                for (int d = 0; d < 3; ++d) vout += (q_sol[3 * node + d] - q[3 * node + d]) / h * nh[d];
                const double dv = -restitution * vin - vout;
                if (dv > 0) for (int d = 0; d < 3; ++d) q_sol[3 * node + d] += h * dv * nh[d];
            }
        }
    }
    q_next = q_sol;
    v_next.assign(dofs, 0.0);
    for (int i = 0; i < dofs; ++i) v_next[i] = (q_next[i] - q[i]) / h;
    active.clear();
    for (int c = 0; c < K; ++c) if (lambda[c] > 0) active.push_back(cand[c]);
    prev_candidates = cand;
    prev_lambda = lambda;
    last_lambda = lambda;
    last_lt1 = lt1;
    last_lt2 = lt2;
    AlgPhdReplayCache fc;
    fc.contact_candidates = active;
    fc.q_tilde = q_tilde;
    fc.q_sol = q_next;
    const int na = (int)active.size();
    fc.lambda.assign(na, 0.0);
    fc.omega.assign(na, 0.0);
    fc.normals.assign(na, {0, 0, 0});
    std::unordered_map<int, int> cidx;
    for (int c = 0; c < K; ++c) cidx[cand[c]] = c;
    for (int a = 0; a < na; ++a) {
        const int c = cidx[active[a]];
        fc.normals[a] = nrm[c];
        fc.omega[a] = omega_last[c];
        fc.lambda[a] = lambda[c];
    }
    tape.push_back(std::move(fc));
    if (tape.size() > 8192) tape.erase(tape.begin(), tape.begin() + (tape.size() - 8192));
}

// BackwardProjectiveDynamicsAlgPhd, deformable_projective_dynamics_backward_gpu.cpp:648-1250.
// The tape must match (the fixed-active-set replay at :113-418 is not restated).
// ReplayAlgPhdFixedActiveSet (deformable_projective_dynamics_backward_gpu.cpp:113-418):
// the backward's own PD loop over a given active set, used when no forward tape
// entry matches the frame. It differs from the forward: only the active
// contacts, no friction, lambda from 0, normals re-evaluated at every iterate,
// W(ci, cj) = sum_d n_ci[d]^2 (A_d^-1)(node_cj, node_ci) for ci <= cj (mirrored;
// the reference multiplies by n_ci twice, :203-220), r_j = max(W_jj, 1e-10)
// without the h^2 of the forward, lambda clamped below only, q = A^-1 (b_full +
// J^T Omega lambda) directly, and Anderson acceleration with reg =
// 1e-10 max(tr / m, 1) and no restart (:381-405).
AlgPhdReplayCache Body::replay_fixed_active_set(const Vec& q, const Vec& v, const Vec& f_ext, double dt,
                                                const std::vector<int>& active,
                                                const std::map<std::string, double>& opt) {
    check(boundary.type != kMesh, "The fixed-active-set replay does not support the mesh boundary in this build.");
    const int max_pd_iter = (int)opt_req(opt, "max_pd_iter");
    const double abs_tol = opt_req(opt, "abs_tol");
    const double rel_tol = opt_req(opt, "rel_tol");
    const int aa_window = (int)opt_or(opt, "aa_window", 5);
    const double h = dt, mass = element_volume * density, h2m = h * h / mass, inv_h2m = mass / (h * h);
    AlgPhdReplayCache cache;
    cache.contact_candidates = active;
    const int na = (int)active.size();
    const Vec fs = state_force(q, v);
    cache.q_tilde.assign(dofs, 0.0);
    for (int i = 0; i < dofs; ++i) cache.q_tilde[i] = q[i] + h * v[i] + h2m * f_ext[i] + h2m * fs[i];
    cache.q_sol = q;
    if (na == 0) return cache;
    cache.lambda.assign(na, 0.0);
    cache.omega.assign(na, 0.0);
    cache.normals.assign(na, {0, 0, 0});
    for (int c = 0; c < na; ++c) boundary.normal(&q[3 * active[c]], cache.normals[c].data());
    std::vector<double> W((size_t)na * na, 0.0);
    for (int d = 0; d < 3; ++d)
        for (int ci = 0; ci < na; ++ci) {
            Vec e(nv, 0.0);
            e[active[ci]] = cache.normals[ci][d];
            const Vec col = solve_axis(d, e);
            for (int cj = ci; cj < na; ++cj) W[(size_t)ci * na + cj] += cache.normals[ci][d] * col[active[cj]];
        }
    for (int ci = 0; ci < na; ++ci)
        for (int cj = ci + 1; cj < na; ++cj) W[(size_t)cj * na + ci] = W[(size_t)ci * na + cj];
    Anderson aa(aa_window, true);
    Vec prev_pd_rhs(dofs, 0.0);
    auto ainv = [&](const Vec& b) {
        Vec out(dofs);
        for (int d = 0; d < 3; ++d) {
            Vec bd(nv);
            for (int i = 0; i < nv; ++i) bd[i] = b[3 * i + d];
            const Vec x = solve_axis(d, bd);
            for (int i = 0; i < nv; ++i) out[3 * i + d] = x[i];
        }
        return out;
    };
    for (int iter = 0; iter < max_pd_iter; ++iter) {
        const Vec pd_rhs = local_step(cache.q_sol);
        Vec b_full(dofs);
        for (int i = 0; i < dofs; ++i) b_full[i] = inv_h2m * cache.q_tilde[i] + pd_rhs[i];
        const Vec q_prev = cache.q_sol;
        Vec delta(na), E(na), hvec(na);
        for (int c = 0; c < na; ++c) {
            const double* x = &cache.q_sol[3 * active[c]];
            delta[c] = boundary.distance(x);
            boundary.normal(x, cache.normals[c].data());
        }
        for (int c = 0; c < na; ++c) {
            const double dj = delta[c], lj = cache.lambda[c], rj = std::max(W[(size_t)c * na + c], 1e-10);
            double den = std::sqrt(dj * dj + rj * rj * lj * lj);
            if (den < 1e-14) den = 1e-14;
            cache.omega[c] = 1.0 - dj / den;
            E[c] = (1.0 - rj * lj / den) * rj;
            const double phi = dj + rj * lj - den;
            double Hq = 0;
            for (int d = 0; d < 3; ++d) Hq += cache.normals[c][d] * cache.q_sol[3 * active[c] + d];
            hvec[c] = -phi + cache.omega[c] * Hq;
        }
        std::vector<double> M((size_t)na * na);
        for (int a = 0; a < na; ++a)
            for (int b = 0; b < na; ++b) M[(size_t)a * na + b] = cache.omega[a] * W[(size_t)a * na + b] * cache.omega[b];
        for (int a = 0; a < na; ++a) M[(size_t)a * na + a] += E[a];
        Vec g = b_full;
        for (int c = 0; c < na; ++c)
            for (int d = 0; d < 3; ++d) g[3 * active[c] + d] += cache.normals[c][d] * cache.omega[c] * cache.lambda[c];
        const Vec Ag = ainv(g);
        Vec rhs(na);
        for (int c = 0; c < na; ++c) {
            double val = 0;
            for (int d = 0; d < 3; ++d) val += cache.normals[c][d] * Ag[3 * active[c] + d];
            rhs[c] = hvec[c] - cache.omega[c] * val;
        }
        const Vec dl = dense_ldlt_solve(M, na, rhs);
        for (int c = 0; c < na; ++c) cache.lambda[c] = std::max(cache.lambda[c] + dl[c], 0.0);
        Vec bc = b_full;
        for (int c = 0; c < na; ++c)
            for (int d = 0; d < 3; ++d) bc[3 * active[c] + d] += cache.normals[c][d] * cache.omega[c] * cache.lambda[c];
        cache.q_sol = ainv(bc);
        for (auto& kv : dirichlet) cache.q_sol[kv.first] = kv.second;
        if (aa_window > 0) {
            Vec gk(dofs);
            for (int i = 0; i < dofs; ++i) gk[i] = cache.q_sol[i] - q_prev[i];
            Vec cand;
            if (aa.step(q_prev, gk, cand)) {
                cache.q_sol = cand;
                for (auto& kv : dirichlet) cache.q_sol[kv.first] = kv.second;
            }
        }
        Vec dq(dofs), dr(dofs);
        for (int i = 0; i < dofs; ++i) { dq[i] = cache.q_sol[i] - q_prev[i]; dr[i] = pd_rhs[i] - prev_pd_rhs[i]; }
        const double abs_err = norm(dq), ref_err = norm(q_prev), rc = norm(dr), rr = norm(prev_pd_rhs);
        prev_pd_rhs = pd_rhs;
        if (iter > 0 && abs_err <= rel_tol * ref_err + abs_tol && rc <= rel_tol * rr + abs_tol) break;
    }
    return cache;
}

void Body::backward(const Vec& q, const Vec& v, const Vec& f_ext, double dt, const Vec& q_next, const Vec& v_next,
                    const std::vector<int>& active, const Vec& dl_dq_next, const Vec& dl_dv_next,
                    const std::map<std::string, double>& opt, Vec& dl_dq, Vec& dl_dv, Vec& dl_df_ext, Vec& dl_dmat_w,
                    Vec& dl_dstate_p, Vec& dl_dw_elem) {
    (void)v; (void)f_ext; (void)v_next;
    const int max_pd_iter = (int)opt_req(opt, "max_pd_iter");
    const double abs_tol = opt_req(opt, "abs_tol");
    const double rel_tol = opt_req(opt, "rel_tol");
    opt_req(opt, "verbose");
    opt_req(opt, "thread_ct");
    const double h = dt, inv_h = 1 / h, mass = element_volume * density, h2m = h * h / mass, inv_h2m = mass / (h * h);
    setup(dt);
    const int nE = (int)energies.size();
    dl_dq.assign(dofs, 0.0);
    dl_dv.assign(dofs, 0.0);
    dl_df_ext.assign(dofs, 0.0);
    dl_dmat_w.assign(nE, 0.0);
    dl_dstate_p.assign(n_state_params, 0.0);
    dl_dw_elem.assign((size_t)nE * ne, 0.0);
    Vec rhs_qbar(dofs);
    for (int i = 0; i < dofs; ++i) {
        dl_dq[i] += -dl_dv_next[i] * inv_h;
        rhs_qbar[i] = dl_dq_next[i] + dl_dv_next[i] * inv_h;
    }
    const bool exact = opt_or(opt, "alg_phd_bwd_exact_kpd", 0.0) > 0.5;
    const int filter = (int)opt_or(opt, "alg_phd_bwd_kpd_filter", 0.0);
    const double rho = opt_or(opt, "alg_phd_bwd_trust_radius", 0.9);
    check(filter >= 0 && filter <= 3, "alg_phd_bwd_kpd_filter must be 0 (none), 1 (clamp), 2 (abs) or 3 (trust region).");
    check(!(filter == 3) || (rho > 0 && rho < 1), "alg_phd_bwd_trust_radius must lie in (0, 1).");
    long long nfilt = 0;
    // trust region (mode 3): stage 0 is the unfiltered operator; the backbone
    // moves to stage 1 (clamp) and then 2 (abs), both with the lower bound -rho,
    // when its fixed point diverges
    int stage = filter == 3 ? 0 : 1;
    std::vector<double> Kpd = local_differential(q_next, exact, filter, rho, &nfilt, stage);
    last_bwd = BackwardStats();
    last_bwd.filtered = (double)nfilt;
    last_bwd.samples = (double)ne * samples;

    AlgPhdReplayCache replay;
    bool from_tape = false;
    if (!tape.empty()) {
        const auto& tail = tape.back();
        bool match = tail.contact_candidates == active && tail.q_sol.size() == q_next.size();
        if (match) {
            Vec d(dofs);
            for (int i = 0; i < dofs; ++i) d[i] = tail.q_sol[i] - q_next[i];
            match = norm(d) <= 1e-8 * std::max(1.0, std::max(norm(tail.q_sol), norm(q_next)));
        }
        if (match) { replay = tail; tape.pop_back(); from_tape = true; }
        else if (tail.q_sol.size() == q_next.size()) {  // stale tape (:730-737)
            Vec d(dofs);
            for (int i = 0; i < dofs; ++i) d[i] = tail.q_sol[i] - q_next[i];
            if (norm(d) / std::max(1.0, std::max(norm(tail.q_sol), norm(q_next))) > 1e-6) tape.clear();
        }
    }
    if (!from_tape) replay = replay_fixed_active_set(q, v, f_ext, dt, active, opt);
    last_bwd_from_tape = from_tape;
    last_bwd.from_tape = from_tape ? 1 : 0;

    auto apply_ainv = [&](const Vec& rhs) {
        Vec out(dofs);
        for (int d = 0; d < 3; ++d) {
            Vec bd(nv);
            for (int i = 0; i < nv; ++i) bd[i] = rhs[3 * i + d];
            const Vec x = solve_axis(d, bd);
            for (int i = 0; i < nv; ++i) out[3 * i + d] = x[i];
        }
        for (auto& kv : dirichlet) out[kv.first] = 0;
        return out;
    };
    const int aa_window = (int)opt_or(opt, "aa_window", 5);
    const double clip = opt_or(opt, "alg_phd_bwd_clip_norm", 1e6);
    auto safe_clip = [&](Vec x) {
        for (double& t : x) if (!std::isfinite(t)) t = 0;
        const double n = norm(x);
        if (clip > 0 && std::isfinite(n) && n > clip) for (double& t : x) t *= clip / n;
        return x;
    };
    // Fixed point s <- safe_clip(A^-1 (rhs + K_pd s)) with AA (backward_gpu.cpp:818-861).
    // div (optional): stop early and set it when the iteration diverges: a
    // non-finite residual, an iterate at the clip norm, or a residual 1e3 x its
    // smallest so far after 10 iterations
    auto fixed_point = [&](const Vec& rhs, int max_it, int* used_iters, bool* conv = nullptr, bool* div = nullptr) {
        Vec s(dofs, 0.0);
        Anderson aa(aa_window);
        int it_used = 0;
        if (conv) *conv = false;
        if (div) *div = false;
        double err_min = std::numeric_limits<double>::infinity();
        for (int i = 0; i < max_it; ++i) {
            const Vec kpd = apply_local_differential(Kpd, s);
            Vec t(dofs);
            for (int j = 0; j < dofs; ++j) t[j] = rhs[j] + kpd[j];
            Vec s_new = safe_clip(apply_ainv(t));
            Vec diff(dofs);
            for (int j = 0; j < dofs; ++j) diff[j] = s_new[j] - s[j];
            const double err = norm(diff);
            if (aa_window > 0) {
                Vec cand_x;
                if (aa.step(s, diff, cand_x)) s_new = safe_clip(cand_x);
            }
            s = s_new;
            it_used = i + 1;
            if (i > 0 && err < rel_tol * norm(s) + abs_tol) { if (conv) *conv = true; break; }
            if (div) {
                err_min = std::min(err_min, err);
                if (!std::isfinite(err) || (clip > 0 && norm(s) >= 0.999 * clip) || (i >= 10 && err > 1e3 * err_min)) {
                    *div = true;
                    break;
                }
            }
        }
        if (used_iters) *used_iters = it_used;
        return s;
    };
    int bb_iters = 0, bb_total = 0;
    bool bb_conv = false;
    Vec db;
    for (;;) {
        bool div = false;
        db = fixed_point(rhs_qbar, max_pd_iter, &bb_iters, &bb_conv, (filter == 3 && stage < 2) ? &div : nullptr);
        bb_total += bb_iters;
        if (!div) break;
        ++stage;
        Kpd = local_differential(q_next, exact, filter, rho, &nfilt, stage);
        last_bwd.filtered = (double)nfilt;
    }
    bb_iters = bb_total;
    last_bwd.trust_stage = filter == 3 ? stage : 0;
    last_bwd_iters = bb_iters;
    const bool diverged = clip > 0 && norm(db) >= clip * 0.999;
    last_bwd.backbone_iters = bb_iters;
    last_bwd.backbone_converged = bb_conv ? 1 : 0;
    last_bwd.clipped = diverged ? 1 : 0;
    last_bwd.db_norm = norm(db);
    // Cascade injection (backward_gpu.cpp:878-916) from the backbone adjoint db,
    // or, with alg_phd_bwd_state_woodbury = 1 (no reference counterpart), from the
    // Woodbury-corrected adjoint z, which is the exact adjoint of the step with
    // its active contacts held (the reference routes only dl_dmat_w through z).
    auto inject = [&](const Vec& src) {
        Vec inj(dofs);
        for (int i = 0; i < dofs; ++i) inj[i] = inv_h2m * src[i];
        const double cap = norm(rhs_qbar) + 1.0, in = norm(inj);
        if (in > cap) for (double& t : inj) t *= cap / in;
        for (int i = 0; i < dofs; ++i) {
            dl_dq[i] += inj[i];
            dl_df_ext[i] += h2m * inj[i];
            dl_dv[i] += h * inj[i];
        }
        // through h2m f_state(q, v) (backward_gpu.cpp:904-915)
        Vec dl_df(dofs), sq, sv, sp;
        for (int i = 0; i < dofs; ++i) dl_df[i] = h2m * inj[i];
        state_force_backward(q, v, dl_df, sq, sv, sp);
        for (int i = 0; i < dofs; ++i) { dl_dq[i] += sq[i]; dl_dv[i] += sv[i]; }
        for (int k = 0; k < n_state_params; ++k) dl_dstate_p[k] += sp[k];
    };
    const bool state_wb = opt_or(opt, "alg_phd_bwd_state_woodbury", 0.0) > 0.5;
    if (!diverged && !state_wb) inject(db);
    Vec z = db;
    const bool skip_wb = opt_or(opt, "skip_contact_woodbury", 0.0) > 0.5;
    const int nl = (int)replay.contact_candidates.size();
    if (!diverged && !skip_wb && nl > 0) {
        const double fmu = opt_or(opt, "friction_mu", 0.0);
        const bool tb = fmu > 0;
        std::vector<int> act_ci;
        for (int c = 0; c < nl; ++c) if (replay.omega[c] >= 0.5) act_ci.push_back(c);
        std::vector<std::array<double, 3>> T1(nl), T2(nl);
        if (tb)
            for (int c = 0; c < nl; ++c) {
                const auto& n = replay.normals[c];
                int axis = 0;
                for (int d = 1; d < 3; ++d) if (std::fabs(n[d]) < std::fabs(n[axis])) axis = d;
                double e[3] = {0, 0, 0};
                e[axis] = 1;
                const double ne_ = n[0] * e[0] + n[1] * e[1] + n[2] * e[2];
                double a1[3] = {e[0] - ne_ * n[0], e[1] - ne_ * n[1], e[2] - ne_ * n[2]};
                double an = std::sqrt(a1[0] * a1[0] + a1[1] * a1[1] + a1[2] * a1[2]);
                if (an > 1e-12) for (double& t : a1) t /= an;
                else { a1[0] = 1; a1[1] = 0; a1[2] = 0; }
                double a2[3] = {n[1] * a1[2] - n[2] * a1[1], n[2] * a1[0] - n[0] * a1[2], n[0] * a1[1] - n[1] * a1[0]};
                double bn = std::sqrt(a2[0] * a2[0] + a2[1] * a2[1] + a2[2] * a2[2]);
                if (bn > 1e-12) for (double& t : a2) t /= bn;
                else { a2[0] = 0; a2[1] = 1; a2[2] = 0; }
                for (int d = 0; d < 3; ++d) { T1[c][d] = a1[d]; T2[c][d] = a2[d]; }
            }
        struct BC { int ci, basis; };
        std::vector<BC> full;
        const double tlmin = opt_or(opt, "alg_phd_bwd_tangent_lambda_min", 1e-8);
        for (int c : act_ci) {
            full.push_back({c, 0});
            if (tb && replay.lambda[c] > tlmin) { full.push_back({c, 1}); full.push_back({c, 2}); }
        }
        auto dir = [&](const BC& b, double out[3]) {
            for (int d = 0; d < 3; ++d)
                out[d] = b.basis == 0 ? replay.omega[b.ci] * replay.normals[b.ci][d]
                                      : (b.basis == 1 ? T1[b.ci][d] : T2[b.ci][d]);
        };
        auto project_val = [&](const BC& b, const Vec& x) {
            const int node = replay.contact_candidates[b.ci];
            double dd[3];
            dir(b, dd);
            return dd[0] * x[3 * node] + dd[1] * x[3 * node + 1] + dd[2] * x[3 * node + 2];
        };
        const int nbf = (int)full.size();
        if (nbf > 0) {
            Vec rf(nbf);
            double rmax = 0;
            for (int i = 0; i < nbf; ++i) { rf[i] = project_val(full[i], z); rmax = std::max(rmax, std::fabs(rf[i])); }
            const double gate = std::max(opt_or(opt, "alg_phd_bwd_rhs_prune_abs", 1e-12),
                                         opt_or(opt, "alg_phd_bwd_rhs_prune_rel", 1e-3) * rmax);
            std::vector<int> sel;
            for (int i = 0; i < nbf; ++i) if (std::fabs(rf[i]) >= gate) sel.push_back(i);
            const int topk = (int)opt_or(opt, "alg_phd_bwd_basis_topk", 0);
            if (topk > 0 && (int)sel.size() > topk) {
                std::nth_element(sel.begin(), sel.begin() + topk, sel.end(),
                                 [&](int a, int b) { return std::fabs(rf[a]) > std::fabs(rf[b]); });
                sel.resize(topk);
            }
            const int nb = (int)sel.size();
            if (nb > 0) {
                const int wt_cap = (int)opt_or(opt, "alg_phd_bwd_wt_max_iter", 120);
                const int wt_max = std::min(max_pd_iter, std::max(1, wt_cap));
                std::vector<Vec> WT(nb);
                last_bwd.basis_rows = nb;
                for (int k = 0; k < nb; ++k) {
                    const BC b = full[sel[k]];
                    Vec j(dofs, 0.0);
                    double dd[3];
                    dir(b, dd);
                    const int node = replay.contact_candidates[b.ci];
                    for (int d = 0; d < 3; ++d) j[3 * node + d] = dd[d];
                    int used = 0;
                    WT[k] = fixed_point(j, wt_max, &used);
                    last_bwd.woodbury_iters = std::max(last_bwd.woodbury_iters, (double)used);
                }
                std::vector<double> GT((size_t)nb * nb);
                Vec rb(nb);
                for (int k = 0; k < nb; ++k) {
                    rb[k] = rf[sel[k]];
                    for (int l = 0; l < nb; ++l) GT[(size_t)l * nb + k] = project_val(full[sel[l]], WT[k]);
                }
                // G_T.ldlt() (:1186): G_T is symmetric up to the fixed-point tolerance.
                const Vec alpha = dense_ldlt_solve(GT, nb, rb);
                for (int k = 0; k < nb; ++k)
                    for (int i = 0; i < dofs; ++i) z[i] -= alpha[k] * WT[k][i];
                z = safe_clip(z);
            }
        }
    }
    if (!diverged && state_wb) inject(z);
    // dl/da and dl/d(act_w) (:1209-1228) with ActuationForceDifferential
    // (deformable_actuation.cpp:120-177): df_k/da = -(vol/s) A^T vec(dP/da),
    // df_k/dw = -(vol/s) A^T vec(P / stiffness), P = k (1 - a/l) F m m^T
    // (StressTensor / StressTensorDifferential, pd_muscle_energy.cpp:33-116)
    last_dl_da.assign(act_dofs, 0.0);
    last_dl_dact_w.assign(muscles.size(), 0.0);
    if (!diverged && act_dofs > 0) {
        check((int)act.size() == act_dofs, "Inconsistent actuation size.");
        Vec da(act_dofs, 0.0), dw(muscles.size(), 0.0);
        int base = 0;
        for (int gi = 0; gi < (int)muscles.size(); ++gi) {
            const Muscle& mu = muscles[gi];
            for (int ei = 0; ei < (int)mu.elems.size(); ++ei) {
                const int e = mu.elems[ei];
                const double a = act[base + ei];
                const int* vv = &elems[(size_t)e * edim];
                double fa[24] = {0}, fw[24] = {0};
                for (int s = 0; s < samples; ++s) {
                    const M3 F = deformation_gradient(q_next.data(), e, s);
                    double Fm[3], l;
                    muscle_fm(F, mu.m, Fm, l);
                    if (l <= std::numeric_limits<double>::epsilon()) continue;
                    M3 dPda, Pw;
                    for (int i = 0; i < 3; ++i)
                        for (int j = 0; j < 3; ++j) {
                            dPda(i, j) = -mu.stiffness / l * Fm[i] * mu.m[j];
                            Pw(i, j) = (1 - a / l) * Fm[i] * mu.m[j];
                        }
                    add_At(dPda, e, s, -element_volume / samples, fa);
                    add_At(Pw, e, s, -element_volume / samples, fw);
                }
                double sa = 0, sw = 0;
                for (int k = 0; k < edim; ++k)
                    for (int i = 0; i < 3; ++i) {
                        sa += z[3 * vv[k] + i] * fa[3 * k + i];
                        sw += z[3 * vv[k] + i] * fw[3 * k + i];
                    }
                da[base + ei] += sa;
                dw[gi] += sw;
            }
            base += (int)mu.elems.size();
        }
        bool ok = true;
        for (double x : da) if (!std::isfinite(x)) ok = false;
        if (ok) { last_dl_da = da; last_dl_dact_w = dw; }
    }
    // dl/d(mat_w) (:1231-1249) with PdEnergyForceDifferential(require_dw),
    // deformable_pd_energy.cpp:368-403: df/dw = -(vol/s) A^T vec(P / stiffness_).
    const bool skip_mat = opt_or(opt, "skip_material_grad", 0.0) > 0.5;
    if (!skip_mat && !diverged) {
        Vec contrib(nE, 0.0);
        std::vector<double> per((size_t)nE * ne, 0.0);
        bool ok = true;
        for (int ei = 0; ei < nE; ++ei) {
            const Energy& en = energies[ei];
            const double inv_w = 1 / en.stiffness;
            for (int e = 0; e < ne; ++e) {
                if (!en.member[e]) continue;
                const int* vv = &elems[(size_t)e * edim];
                double acc[24] = {0};
                for (int s = 0; s < samples; ++s) {
                    const M3 F = deformation_gradient(q_next.data(), e, s);
                    FAux a;
                    a.init(F);
                    const M3 p = project(en, a);
                    const M3 P = stress(en, a, p);
                    add_At(P, e, s, -inv_w * element_volume / samples, acc);
                }
                double sum = 0;
                for (int k = 0; k < edim; ++k)
                    for (int i = 0; i < 3; ++i) sum += z[3 * vv[k] + i] * acc[3 * k + i];
                per[(size_t)ei * ne + e] = sum;
                contrib[ei] += sum;
            }
        }
        for (double c : contrib) if (!std::isfinite(c)) ok = false;
        if (ok) {
            for (int ei = 0; ei < nE; ++ei) dl_dmat_w[ei] += contrib[ei];
            dl_dw_elem = per;
        }
    }
}


// ---------------------------------------------------------------------------
// The DiffPD baseline forward.
// ---------------------------------------------------------------------------
// ComputePdEnergy (deformable_pd_energy.cpp:156-189) and ActuationEnergy
// (deformable_actuation.cpp:28-46): element energies w/2 |F - p(F)|^2 with the
// energy's stiffness() (the mean for per-element parameters, as
// PdElementEnergy::EnergyDensity reads stiffness_), summed per element over the
// energies, then over the elements; muscles k/2 (|F m| - a)^2.
double Body::baseline_energy(const Vec& q) const {
    std::vector<double> ee(ne, 0.0);
    for (const auto& en : energies) {
        check(en.type != kNeoHookean, "The baseline forward supports the corotated and volume energies.");
#pragma omp parallel for schedule(static)
        for (int e = 0; e < ne; ++e) {
            if (!en.member[e]) continue;
            for (int s = 0; s < samples; ++s) {
                const M3 F = deformation_gradient(q.data(), e, s);
                FAux a;
                a.init(F);
                const M3 d = add(F, scale(project(en, a), -1.0));
                double n2 = 0;
                for (int i = 0; i < 9; ++i) n2 += d.a[i] * d.a[i];
                ee[e] += en.stiffness * 0.5 * n2 * element_volume / samples;
            }
        }
    }
    double tot = 0;
    for (double x : ee) tot += x;
    int base = 0;
    for (const Muscle& mu : muscles) {
        for (int ei = 0; ei < (int)mu.elems.size(); ++ei)
            for (int s = 0; s < samples; ++s) {
                double Fm[3], l;
                muscle_fm(deformation_gradient(q.data(), mu.elems[ei], s), mu.m, Fm, l);
                const double a = act[base + ei];
                tot += mu.stiffness * 0.5 * (l - a) * (l - a) * element_volume / samples;
            }
        base += (int)mu.elems.size();
    }
    return tot;
}

// PdEnergyForce (deformable_pd_energy.cpp:192-240) + ActuationForce
// (deformable_actuation.cpp:48-77): f = -(vol/s) sum A^T vec(P), P = k (F - p(F))
// for the element energies and k (1 - a/l) F m m^T for a muscle.
Vec Body::baseline_force(const Vec& q) const {
    std::vector<double> acc((size_t)ne * edim * 3, 0.0);
    for (const auto& en : energies) {
        check(en.type != kNeoHookean, "The baseline forward supports the corotated and volume energies.");
#pragma omp parallel for schedule(static)
        for (int e = 0; e < ne; ++e) {
            if (!en.member[e]) continue;
            for (int s = 0; s < samples; ++s) {
                const M3 F = deformation_gradient(q.data(), e, s);
                FAux a;
                a.init(F);
                const M3 P = scale(add(F, scale(project(en, a), -1.0)), en.stiffness);
                add_At(P, e, s, -element_volume / samples, &acc[(size_t)e * edim * 3]);
            }
        }
    }
    int base = 0;
    for (const Muscle& mu : muscles) {
        for (int ei = 0; ei < (int)mu.elems.size(); ++ei) {
            const int e = mu.elems[ei];
            for (int s = 0; s < samples; ++s) {
                double Fm[3], l;
                muscle_fm(deformation_gradient(q.data(), e, s), mu.m, Fm, l);
                if (l <= std::numeric_limits<double>::epsilon()) continue;
                M3 P;
                const double r = mu.stiffness * (1 - act[base + ei] / l);
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) P(i, j) = r * Fm[i] * mu.m[j];
                add_At(P, e, s, -element_volume / samples, &acc[(size_t)e * edim * 3]);
            }
        }
        base += (int)mu.elems.size();
    }
    Vec f(dofs, 0.0);
    for (int e = 0; e < ne; ++e)
        for (int k = 0; k < edim; ++k)
            for (int i = 0; i < 3; ++i) f[3 * elems[(size_t)e * edim + k] + i] += acc[((size_t)e * edim + k) * 3 + i];
    return f;
}

// PdLhsSolve (forward.cpp:690-936): the operator PdLhsMatrixOp (:666-688) keeps
// the rows and columns of the frozen nodes out of A (x_c = rhs_c on them), so the
// solve is x_f = A_ff^-1 rhs_f per axis. The reference reaches it by a Woodbury
// update of A's factor; here A_ff is factored directly (cached per frozen set).
Vec Body::pd_lhs_solve(const Vec& rhs, const std::vector<int>& frozen) {
    Vec x(dofs, 0.0);
    if (frozen.empty()) {
        for (int d = 0; d < 3; ++d) {
            Vec b(nv);
            for (int i = 0; i < nv; ++i) b[i] = rhs[3 * i + d];
            const Vec y = solve_axis(d, b);
            for (int i = 0; i < nv; ++i) x[3 * i + d] = y[i];
        }
        return x;
    }
    if (!lhs_valid || lhs_frozen != frozen) {
        std::vector<char> fz(nv, 0);
        for (int i : frozen) fz[i] = 1;
        for (int d = 0; d < 3; ++d) {
            std::vector<int> Ap(nv + 1, 0), Ai;
            std::vector<double> Ax;
            std::vector<std::vector<int>> adj(nv);
            for (int i = 0; i < nv; ++i) {
                if (fz[i]) { Ai.push_back(i); Ax.push_back(1.0); }
                else
                    for (int p = csr_p[d][i]; p < csr_p[d][i + 1]; ++p) {
                        const int j = csr_i[d][p];
                        if (fz[j]) continue;
                        Ai.push_back(j);
                        Ax.push_back(csr_x[d][p]);
                        if (j != i) adj[i].push_back(j);
                    }
                Ap[i + 1] = (int)Ai.size();
            }
            std::vector<int> all(nv), perm, side(nv, 0);
            std::iota(all.begin(), all.end(), 0);
            rcb_order(all, X, adj, perm, side);
            lhs_ldlt[d].factor(nv, Ap, Ai, Ax, perm);
        }
        lhs_frozen = frozen;
        lhs_valid = true;
    }
    std::vector<char> fz(nv, 0);
    for (int i : frozen) fz[i] = 1;
    for (int d = 0; d < 3; ++d) {
        Vec b(nv), y(nv);
        for (int i = 0; i < nv; ++i) b[i] = fz[i] ? 0.0 : rhs[3 * i + d];
        lhs_ldlt[d].solve(b.data(), y.data());
        for (int i = 0; i < nv; ++i) x[3 * i + d] = fz[i] ? rhs[3 * i + d] : y[i];
    }
    return x;
}

// PdNonlinearSolve (forward.cpp:294-526): minimizes
// 1/2 (q - rhs)^T (M/h^2)(q - rhs) + E_pd(q) + E_act(q) over the free DOFs with
// L-BFGS (initial inverse Hessian PdLhsSolve) and a backtracking Armijo line
// search, or with plain local/global PD steps when use_bfgs is 0.
Vec Body::pd_nonlinear_solve(const Vec& q_init, double inv_h2m, const Vec& rhs, const std::map<int, double>& extra,
                             const std::map<std::string, double>& opt, int* iters) {
    const int max_pd_iter = (int)opt_req(opt, "max_pd_iter");
    const double abs_tol = opt_req(opt, "abs_tol");
    const double rel_tol = opt_req(opt, "rel_tol");
    opt_req(opt, "verbose");
    const bool use_bfgs = opt_req(opt, "use_bfgs") != 0.0;
    check(max_pd_iter > 0, "Invalid max_pd_iter: " + std::to_string(max_pd_iter));
    int hist = 0, max_ls = 0;
    if (use_bfgs) {
        hist = (int)opt_req(opt, "bfgs_history_size");
        check(hist >= 1, "Invalid bfgs_history_size.");
        max_ls = (int)opt_req(opt, "max_ls_iter");
        check(max_ls > 0, "Invalid max_ls_iter: " + std::to_string(max_ls));
    }
    std::map<int, double> aug = dirichlet;
    for (auto& kv : extra) aug[kv.first] = kv.second;
    // frozen nodes of PdLhsSolve: the extra conditions come in whole vertices
    std::vector<int> frozen;
    {
        std::set<int> fn;
        for (auto& kv : extra) fn.insert(kv.first / 3);
        frozen.assign(fn.begin(), fn.end());
    }
    Vec q_sol = q_init, sel(dofs, 1.0);
    for (auto& kv : aug) { q_sol[kv.first] = kv.second; sel[kv.first] = 0.0; }
    auto eval_obj = [&](const Vec& qc, double ec) {
        double t = 0;
        for (int i = 0; i < dofs; ++i) t += (qc[i] - rhs[i]) * inv_h2m * (qc[i] - rhs[i]);
        return 0.5 * t + ec;
    };
    Vec force = baseline_force(q_sol);
    double energy = baseline_energy(q_sol);
    double obj = eval_obj(q_sol, energy);
    Vec grad(dofs);
    for (int i = 0; i < dofs; ++i) grad[i] = (inv_h2m * (q_sol[i] - rhs[i]) - force[i]) * sel[i];
    double rhs_norm = 0;
    for (int i = 0; i < dofs; ++i) rhs_norm += (sel[i] * inv_h2m * rhs[i]) * (sel[i] * inv_h2m * rhs[i]);
    rhs_norm = std::sqrt(rhs_norm);
    std::deque<Vec> xh, gh, sh, yh;
    for (int it = 0; it < max_pd_iter; ++it) {
        if (use_bfgs) {
            Vec dir;
            const int bsize = (int)xh.size();
            if (bsize == 0) {
                xh.push_back(q_sol);
                gh.push_back(grad);
                dir = pd_lhs_solve(grad, frozen);
            } else {
                const Vec ql = xh.back(), gl = gh.back();
                xh.push_back(q_sol);
                gh.push_back(grad);
                Vec si(dofs), yi(dofs);
                for (int i = 0; i < dofs; ++i) { si[i] = q_sol[i] - ql[i]; yi[i] = grad[i] - gl[i]; }
                sh.push_back(si);
                yh.push_back(yi);
                if (bsize == hist + 1) { xh.pop_front(); gh.pop_front(); sh.pop_front(); yh.pop_front(); }
                Vec qv = grad;
                std::deque<double> rho, alpha;
                for (int k = (int)sh.size() - 1; k >= 0; --k) {
                    const double r = 1 / dot(yh[k], sh[k]);
                    const double a = r * dot(sh[k], qv);
                    rho.push_front(r);
                    alpha.push_front(a);
                    for (int i = 0; i < dofs; ++i) qv[i] -= a * yh[k][i];
                }
                Vec z = pd_lhs_solve(qv, frozen);
                for (int k = 0; k < (int)sh.size(); ++k) {
                    const double beta = rho[k] * dot(yh[k], z);
                    for (int i = 0; i < dofs; ++i) z[i] += sh[k][i] * (alpha[k] - beta);
                }
                dir = z;
            }
            if (dot(dir, grad) <= 0)
                for (int i = 0; i < dofs; ++i) dir[i] = grad[i] * sel[i];
            double step = 1;
            Vec qn(dofs);
            for (int i = 0; i < dofs; ++i) qn[i] = q_sol[i] - step * dir[i];
            double en = baseline_energy(qn), on = eval_obj(qn, en);
            const double gamma = 1e-4, gd = dot(grad, dir);
            for (int j = 0; j < max_ls; ++j) {
                const double cond = obj - gamma * step * gd;
                if (!std::isnan(on) && on < cond + std::numeric_limits<double>::epsilon()) break;
                step /= 2;
                for (int i = 0; i < dofs; ++i) qn[i] = q_sol[i] - step * dir[i];
                en = baseline_energy(qn);
                on = eval_obj(qn, en);
            }
            q_sol = qn;
            energy = en;
            obj = on;
        } else {
            // local step with the augmented conditions, then the global step
            std::map<int, double> saved = dirichlet;
            dirichlet = aug;
            Vec pd = local_step(q_sol);
            dirichlet = saved;
            Vec b(dofs);
            for (int i = 0; i < dofs; ++i) b[i] = (inv_h2m * rhs[i] + pd[i]) * sel[i];
            q_sol = pd_lhs_solve(b, frozen);
            for (auto& kv : aug) q_sol[kv.first] = kv.second;
        }
        force = baseline_force(q_sol);
        for (int i = 0; i < dofs; ++i) grad[i] = (inv_h2m * (q_sol[i] - rhs[i]) - force[i]) * sel[i];
        if (norm(grad) <= rel_tol * rhs_norm + abs_tol) {
            if (iters) *iters = it + 1;
            return q_sol;
        }
    }
    check(false, "PD method fails to converge.");
    return Vec(dofs, 0.0);
}

// PdEnergyForceDifferential(q, dq, 0) (deformable_pd_energy.cpp:243-300) +
// ActuationForceDifferential(q, a, dq, 0, 0) (deformable_actuation.cpp:79-114):
// df = -(vol/s) sum A^T vec(dP), dP = k (dF - dp/dF[dF]) with each energy's
// stiffness(), and for a muscle k (dr F + r dF) m m^T with r = 1 - a/l,
// dr = a (n . dF m) / l^2 (StressTensorDifferential, pd_muscle_energy.cpp:47-64).
Vec Body::pd_force_differential(const Vec& q, const Vec& dq) const {
    std::vector<double> acc((size_t)ne * edim * 3, 0.0);
    for (const auto& en : energies) {
#pragma omp parallel for schedule(static)
        for (int e = 0; e < ne; ++e) {
            if (!en.member[e]) continue;
            for (int s = 0; s < samples; ++s) {
                const M3 F = deformation_gradient(q.data(), e, s);
                const M3 dF = deformation_gradient(dq.data(), e, s);
                FAux a;
                a.init(F);
                const M3 dP = scale(add(dF, scale(project_diff(en, a, dF), -1.0)), en.stiffness);
                add_At(dP, e, s, -element_volume / samples, &acc[(size_t)e * edim * 3]);
            }
        }
    }
    int base = 0;
    for (const Muscle& mu : muscles) {
        for (int ei = 0; ei < (int)mu.elems.size(); ++ei) {
            const int e = mu.elems[ei];
            const double a = act[base + ei];
            for (int s = 0; s < samples; ++s) {
                double Fm[3], l, dFm[3], dl_unused;
                muscle_fm(deformation_gradient(q.data(), e, s), mu.m, Fm, l);
                muscle_fm(deformation_gradient(dq.data(), e, s), mu.m, dFm, dl_unused);
                if (l <= std::numeric_limits<double>::epsilon()) continue;
                const double dl = (Fm[0] * dFm[0] + Fm[1] * dFm[1] + Fm[2] * dFm[2]) / l;
                const double r = 1 - a / l, dr = a * dl / (l * l);
                M3 dP;
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) dP(i, j) = mu.stiffness * (dr * Fm[i] + r * dFm[i]) * mu.m[j];
                add_At(dP, e, s, -element_volume / samples, &acc[(size_t)e * edim * 3]);
            }
        }
        base += (int)mu.elems.size();
    }
    Vec f(dofs, 0.0);
    for (int e = 0; e < ne; ++e)
        for (int k = 0; k < edim; ++k)
            for (int i = 0; i < 3; ++i) f[3 * elems[(size_t)e * edim + k] + i] += acc[((size_t)e * edim + k) * 3 + i];
    return f;
}

// PdLhsMatrixOp (forward.cpp:666-688): A x per axis with the frozen nodes'
// entries of x zeroed, then x itself on the frozen rows.
Vec Body::pd_lhs_op(const Vec& x, const std::vector<int>& frozen) const {
    std::vector<char> fz(nv, 0);
    for (int i : frozen) fz[i] = 1;
    Vec out(dofs, 0.0);
    for (int d = 0; d < 3; ++d)
        for (int i = 0; i < nv; ++i) {
            double acc = 0;
            for (int p = csr_p[d][i]; p < csr_p[d][i + 1]; ++p) {
                const int j = csr_i[d][p];
                if (!fz[j]) acc += csr_x[d][p] * x[3 * j + d];
            }
            out[3 * i + d] = fz[i] ? x[3 * i + d] : acc;
        }
    return out;
}

// BackwardProjectiveDynamics for a pd_eigen baseline method
// (deformable_projective_dynamics_backward.cpp:141-679), its dense-solver path
// (material 'none', no MAS): L-BFGS on 1/2 x^T S x - b^T x with
// S = PdLhsMatrixOp - K_pd (mean-stiffness K_pd, backward.cpp:10-80) over the
// free DOFs, b = dl/dq_next + dl/dv_next / h; when it meets an indefinite
// direction or the iteration cap, the Newton system is solved instead (dense
// LDL^T of S here, which is the Newton matrix for energies of uniform
// stiffness). Then the fixed DOFs' share through NewtonMatrixOp, the actuation
// and material gradients with the adjoint, and the chain back through
// rhs = q + h v + h^2/m (f_ext + f_state).
void Body::backward_baseline(const Vec& q, const Vec& v, const Vec& f_ext, double dt, const Vec& q_next,
                             const std::vector<int>& active, const Vec& dl_dq_next, const Vec& dl_dv_next,
                             const std::map<std::string, double>& opt, Vec& dl_dq, Vec& dl_dv, Vec& dl_df_ext,
                             Vec& dl_dmat_w, Vec& dl_dstate_p) {
    (void)v; (void)f_ext;
    const int max_pd_iter = (int)opt_req(opt, "max_pd_iter");
    const double abs_tol = opt_req(opt, "abs_tol"), rel_tol = opt_req(opt, "rel_tol");
    opt_req(opt, "verbose");
    opt_req(opt, "thread_ct");
    const bool use_bfgs = opt_req(opt, "use_bfgs") != 0.0;
    check(max_pd_iter > 0, "Invalid max_pd_iter: " + std::to_string(max_pd_iter));
    int hist = 0, max_ls = 0;
    if (use_bfgs) {
        hist = (int)opt_req(opt, "bfgs_history_size");
        check(hist >= 1, "Invalid bfgs_history_size.");
        max_ls = (int)opt_req(opt, "max_ls_iter");
        check(max_ls > 0, "Invalid max_ls_iter: " + std::to_string(max_ls));
    }
    setup(dt);
    const int nE = (int)energies.size();
    dl_dq.assign(dofs, 0.0);
    dl_dv.assign(dofs, 0.0);
    dl_df_ext.assign(dofs, 0.0);
    dl_dmat_w.assign(nE, 0.0);
    dl_dstate_p.assign(n_state_params, 0.0);
    last_dl_da.assign(act_dofs, 0.0);
    last_dl_dact_w.assign(muscles.size(), 0.0);
    const double h = dt, inv_h = 1 / h, mass = element_volume * density, h2m = h * h / mass, inv_h2m = mass / (h * h);
    const std::vector<double> Kpd = local_differential(q_next, false);
    std::map<int, double> aug = dirichlet;
    std::set<int> fset(active.begin(), active.end());
    const std::vector<int> frozen(fset.begin(), fset.end());
    for (int idx : frozen) for (int i = 0; i < 3; ++i) aug[3 * idx + i] = q[3 * idx + i];
    Vec agg(dofs), sel(dofs, 1.0);
    for (int i = 0; i < dofs; ++i) {
        dl_dq[i] += -dl_dv_next[i] * inv_h;
        agg[i] = dl_dq_next[i] + dl_dv_next[i] * inv_h;
    }
    for (auto& kv : aug) sel[kv.first] = 0.0;
    auto Sop = [&](const Vec& x) {
        Vec a = pd_lhs_op(x, frozen);
        const Vec k = apply_local_differential(Kpd, x);
        for (int i = 0; i < dofs; ++i) a[i] -= k[i];
        return a;
    };
    Vec x(dofs, 0.0);
    Vec Sx = Sop(x), grad(dofs);
    for (int i = 0; i < dofs; ++i) grad[i] = (Sx[i] - agg[i]) * sel[i];
    double obj = 0.5 * dot(x, Sx) - dot(agg, x);
    const double bn = norm(agg);
    bool success = false;
    std::deque<Vec> xh, gh, sh, yh;
    int it_used = 0;
    for (int it = 0; it < max_pd_iter && !success; ++it) {
        it_used = it + 1;
        if (use_bfgs) {
            Vec dir;
            const int bsize = (int)xh.size();
            if (bsize == 0) {
                xh.push_back(x);
                gh.push_back(grad);
                dir = pd_lhs_solve(grad, frozen);
            } else {
                const Vec xl = xh.back(), gl = gh.back();
                xh.push_back(x);
                gh.push_back(grad);
                Vec si(dofs), yi(dofs);
                for (int i = 0; i < dofs; ++i) { si[i] = x[i] - xl[i]; yi[i] = grad[i] - gl[i]; }
                sh.push_back(si);
                yh.push_back(yi);
                if (bsize == hist + 1) { xh.pop_front(); gh.pop_front(); sh.pop_front(); yh.pop_front(); }
                Vec qv = grad;
                std::deque<double> rho, alpha;
                for (int k = (int)sh.size() - 1; k >= 0; --k) {
                    const double r = 1 / dot(yh[k], sh[k]);
                    const double a = r * dot(sh[k], qv);
                    rho.push_front(r);
                    alpha.push_front(a);
                    for (int i = 0; i < dofs; ++i) qv[i] -= a * yh[k][i];
                }
                Vec z = pd_lhs_solve(qv, frozen);
                for (int k = 0; k < (int)sh.size(); ++k) {
                    const double beta = rho[k] * dot(yh[k], z);
                    for (int i = 0; i < dofs; ++i) z[i] += sh[k][i] * (alpha[k] - beta);
                }
                dir = z;
            }
            for (int i = 0; i < dofs; ++i) dir[i] *= sel[i];
            const double gd = dot(dir, grad);
            if (gd < -1e-4) break;  // indefinite: the Newton solve below
            double step = 1;
            Vec xn(dofs);
            for (int i = 0; i < dofs; ++i) xn[i] = x[i] - step * dir[i];
            Vec Sxn = Sop(xn);
            double on = 0.5 * dot(xn, Sxn) - dot(agg, xn);
            for (int j = 0; j < max_ls; ++j) {
                const double cond = obj - 1e-4 * step * gd;
                if (!std::isnan(on) && on < cond + std::numeric_limits<double>::epsilon()) break;
                step /= 2;
                for (int i = 0; i < dofs; ++i) xn[i] = x[i] - step * dir[i];
                Sxn = Sop(xn);
                on = 0.5 * dot(xn, Sxn) - dot(agg, xn);
            }
            x = xn;
            Sx = Sxn;
            obj = on;
        } else {
            const Vec k = apply_local_differential(Kpd, x);
            Vec b(dofs);
            for (int i = 0; i < dofs; ++i) b[i] = (agg[i] + k[i]) * sel[i];
            x = pd_lhs_solve(b, frozen);
            for (int i = 0; i < dofs; ++i) x[i] *= sel[i];
            Sx = Sop(x);
            obj = 0.5 * dot(x, Sx) - dot(agg, x);
        }
        for (int i = 0; i < dofs; ++i) grad[i] = (Sx[i] - agg[i]) * sel[i];
        if (norm(grad) <= rel_tol * bn + abs_tol) {
            success = true;
            for (auto& kv : aug) x[kv.first] = agg[kv.first];
        }
    }
    last_baseline_bwd_iters = it_used;
    last_baseline_bwd_newton = success ? 0 : 1;
    if (!success) {
        // Newton system with the augmented rows and columns as identities
        std::vector<double> M((size_t)dofs * dofs, 0.0);
        for (int c = 0; c < dofs; ++c) {
            Vec e(dofs, 0.0);
            if (sel[c] == 0.0) { M[(size_t)c * dofs + c] = 1.0; continue; }
            e[c] = 1.0;
            const Vec col = Sop(e);
            for (int r = 0; r < dofs; ++r) M[(size_t)r * dofs + c] = sel[r] == 0.0 ? 0.0 : col[r];
        }
        x = dense_ldlt_solve(M, dofs, agg);
    }
    // equation (13) and the fixed DOFs' share (:614-630)
    Vec dl_drhs(dofs);
    for (int i = 0; i < dofs; ++i) dl_drhs[i] = x[i] * inv_h2m;
    for (auto& kv : aug) dl_drhs[kv.first] = agg[kv.first];
    Vec adjoint = x;
    for (auto& kv : aug) adjoint[kv.first] = 0.0;
    {
        // NewtonMatrixOp(q_next, a, inv_h2m, {}, -adjoint) = inv_h2m (-adjoint) - df(-adjoint), with
        // df the PD + muscle force differential (each energy's stiffness(), deformable_pd_energy.cpp:243-300)
        Vec mv(dofs);
        for (int i = 0; i < dofs; ++i) mv[i] = -adjoint[i];
        Vec df = pd_force_differential(q_next, mv);
        for (auto& kv : aug) dl_drhs[kv.first] += inv_h2m * mv[kv.first] - df[kv.first];
    }
    // dl/da, dl/d(act_w) (:632-638) and dl/d(mat_w) (:643-647) with the adjoint
    {
        Vec da(act_dofs, 0.0), dw(muscles.size(), 0.0);
        int base = 0;
        for (int gi = 0; gi < (int)muscles.size(); ++gi) {
            const Muscle& mu = muscles[gi];
            for (int ei = 0; ei < (int)mu.elems.size(); ++ei) {
                const int e = mu.elems[ei];
                const double a = act[base + ei];
                const int* vv = &elems[(size_t)e * edim];
                double fa[24] = {0}, fw[24] = {0};
                for (int s = 0; s < samples; ++s) {
                    double Fm[3], l;
                    muscle_fm(deformation_gradient(q_next.data(), e, s), mu.m, Fm, l);
                    if (l <= std::numeric_limits<double>::epsilon()) continue;
                    M3 dPda, Pw;
                    for (int i = 0; i < 3; ++i)
                        for (int j = 0; j < 3; ++j) {
                            dPda(i, j) = -mu.stiffness / l * Fm[i] * mu.m[j];
                            Pw(i, j) = (1 - a / l) * Fm[i] * mu.m[j];
                        }
                    add_At(dPda, e, s, -element_volume / samples, fa);
                    add_At(Pw, e, s, -element_volume / samples, fw);
                }
                for (int k = 0; k < edim; ++k)
                    for (int i = 0; i < 3; ++i) {
                        da[base + ei] += adjoint[3 * vv[k] + i] * fa[3 * k + i];
                        dw[gi] += adjoint[3 * vv[k] + i] * fw[3 * k + i];
                    }
            }
            base += (int)mu.elems.size();
        }
        last_dl_da = da;
        last_dl_dact_w = dw;
        for (int ei = 0; ei < nE; ++ei) {
            const Energy& en = energies[ei];
            const double inv_w = 1 / en.stiffness;
            for (int e = 0; e < ne; ++e) {
                if (!en.member[e]) continue;
                const int* vv = &elems[(size_t)e * edim];
                double acc[24] = {0};
                for (int s = 0; s < samples; ++s) {
                    const M3 F = deformation_gradient(q_next.data(), e, s);
                    FAux a;
                    a.init(F);
                    const M3 p = project(en, a);
                    add_At(stress(en, a, p), e, s, -inv_w * element_volume / samples, acc);
                }
                for (int k = 0; k < edim; ++k)
                    for (int i = 0; i < 3; ++i) dl_dmat_w[ei] += adjoint[3 * vv[k] + i] * acc[3 * k + i];
            }
        }
    }
    // rhs = q (frozen nodes) / dirichlet values / q + h v + h2m (f_ext + f_state)
    Vec basic = dl_drhs;
    for (int idx : frozen)
        for (int i = 0; i < 3; ++i) {
            basic[3 * idx + i] = 0;
            dl_dq[3 * idx + i] += dl_drhs[3 * idx + i];
        }
    for (auto& kv : dirichlet) basic[kv.first] = 0;
    for (int i = 0; i < dofs; ++i) {
        dl_dq[i] += basic[i];
        dl_dv[i] += basic[i] * h;
        dl_df_ext[i] += basic[i] * h2m;
    }
    {
        Vec dl_df(dofs), sq, sv, sp;
        for (int i = 0; i < dofs; ++i) dl_df[i] = basic[i] * h2m;
        state_force_backward(q, v, dl_df, sq, sv, sp);
        for (int i = 0; i < dofs; ++i) { dl_dq[i] += sq[i]; dl_dv[i] += sv[i]; }
        for (int k = 0; k < n_state_params; ++k) dl_dstate_p[k] += sp[k];
    }
}

// SameSet / ProposeNewSet (common.cpp:173-216): keep a /\ b plus the first half
// (in order) of each one's own part; false when both own parts have at most one
// element.
static bool propose_new_set(const std::set<int>& a, const std::set<int>& b, std::set<int>& out) {
    std::set<int> inter, ac, bc;
    for (int e : a) (b.count(e) ? inter : ac).insert(e);
    for (int e : b) if (!inter.count(e)) bc.insert(e);
    out = inter;
    if ((int)ac.size() <= 1 && (int)bc.size() <= 1) return false;
    int k = 0;
    for (int e : ac) { if (k == (int)ac.size() / 2) break; out.insert(e); ++k; }
    k = 0;
    for (int e : bc) { if (k == (int)bc.size() / 2) break; out.insert(e); ++k; }
    return true;
}

// ForwardProjectiveDynamics (forward.cpp:529-664): the active set of contact
// nodes, each frozen where it is at the frame start, is corrected until no
// frozen node pulls on its obstacle and no free node penetrates (10 rounds at
// most, ProposeNewSet against cycles). `active` carries the previous frame's
// set in and this frame's out.
void Body::forward_baseline(const std::string& method, const Vec& q, const Vec& v, const Vec& f_ext, double dt,
                            const std::map<std::string, double>& opt, Vec& q_next, Vec& v_next,
                            std::vector<int>& active) {
    opt_req(opt, "max_pd_iter");
    opt_req(opt, "abs_tol");
    opt_req(opt, "rel_tol");
    opt_req(opt, "verbose");
    opt_req(opt, "thread_ct");
    const bool use_bfgs = opt_req(opt, "use_bfgs") != 0.0;
    if (use_bfgs) {
        check(opt_req(opt, "bfgs_history_size") > 1, "Invalid bfgs_history_size.");
        check(opt_req(opt, "max_ls_iter") > 0, "Invalid max_ls_iter.");
    }
    check(boundary.type == kPlanar || boundary.type == kSpherical || boundary.type == 0,
          "The baseline forward supports the planar and spherical boundaries.");
    check((int)q.size() == dofs && (int)v.size() == dofs && (int)f_ext.size() == dofs, "Inconsistent state size.");
    setup(dt);
    const double h = dt, mass = element_volume * density, h2m = h * h / mass, inv_h2m = mass / (h * h);
    const Vec fs = state_force(q, v);
    Vec rhs(dofs);
    for (int i = 0; i < dofs; ++i) rhs[i] = q[i] + h * v[i] + h2m * f_ext[i] + h2m * fs[i];
    auto finish = [&](const Vec& qs) {
        q_next = qs;
        v_next.assign(dofs, 0.0);
        for (int i = 0; i < dofs; ++i) v_next[i] = (q_next[i] - q[i]) / h;
    };
    last_baseline_pd_iters = 0;
    if (method.size() >= 13 && method.compare(method.size() - 13, 13, "fixed_contact") == 0) {
        std::map<int, double> extra;
        for (int idx : active) for (int i = 0; i < 3; ++i) extra[3 * idx + i] = q[3 * idx + i];
        int it = 0;
        finish(pd_nonlinear_solve(q, inv_h2m, rhs, extra, opt, &it));
        last_baseline_pd_iters = it;
        last_baseline_contact_iters = 1;
        return;
    }
    const int max_contact_iter = 10;
    std::vector<std::set<int>> history;
    for (int ci = 0; ci < max_contact_iter; ++ci) {
        std::map<int, double> extra;
        for (int idx : active) for (int i = 0; i < 3; ++i) extra[3 * idx + i] = q[3 * idx + i];
        int it = 0;
        const Vec q_sol = pd_nonlinear_solve(q, inv_h2m, rhs, extra, opt, &it);
        last_baseline_pd_iters += it;
        const Vec f_sol = baseline_force(q_sol);
        const std::set<int> past(active.begin(), active.end());
        history.push_back(past);
        active.clear();
        bool good = true;
        for (auto& kv : fb_vertices) {
            const int node = kv.first;
            const double* x = &q_sol[3 * node];
            const double dist = boundary.distance(x);
            double nn[3];
            boundary.normal(x, nn);
            double cf = 0;
            for (int d = 0; d < 3; ++d) cf += nn[d] * (inv_h2m * (q_sol[3 * node + d] - rhs[3 * node + d]) - f_sol[3 * node + d]);
            if (past.count(node)) {
                if (cf >= 0) active.push_back(node);
                else good = false;
            } else if (dist < 0) {
                active.push_back(node);
                good = false;
            }
        }
        bool early = false;
        if (ci >= 1) {
            const std::set<int> cur(active.begin(), active.end());
            if (history[ci - 1] == cur) {
                std::set<int> ns;
                if (propose_new_set(history[ci], cur, ns)) active.assign(ns.begin(), ns.end());
                else early = true;
            }
        }
        if (good || ci == max_contact_iter - 1 || early) {
            finish(q_sol);
            active.assign(past.begin(), past.end());
            last_baseline_contact_iters = ci + 1;
            return;
        }
    }
}

}  // namespace orc

// =============================================================================
// C ABI for the test harness (ctypes).
// =============================================================================
namespace {
struct Handle {
    orc::Body body;
    std::string err;
};
std::map<std::string, double> make_opts(const char* const* keys, const double* vals, int n) {
    std::map<std::string, double> o;
    for (int i = 0; i < n; ++i) o[keys[i]] = vals[i];
    return o;
}
template <class F>
int guard(Handle* h, F&& f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        h->err = e.what();
        return 1;
    }
}
}  // namespace

extern "C" {
void* orc_create() { return new Handle(); }
void orc_destroy(void* h) { delete static_cast<Handle*>(h); }
const char* orc_last_error(void* h) { return static_cast<Handle*>(h)->err.c_str(); }

int orc_init_mesh(void* hp, int edim, int nv, const double* X, int ne, const int* elems, double density) {
    Handle* h = static_cast<Handle*>(hp);
    return guard(h, [&] { h->body.init(edim, nv, X, ne, elems, density); });
}
int orc_add_pd_energy(void* hp, const char* type, const double* p, int np, const int* idx, int ni) {
    Handle* h = static_cast<Handle*>(hp);
    return guard(h, [&] { h->body.add_pd_energy(type, orc::Vec(p, p + np), std::vector<int>(idx, idx + ni)); });
}
int orc_set_dirichlet(void* hp, int dof, double val) {
    Handle* h = static_cast<Handle*>(hp);
    return guard(h, [&] { h->body.dirichlet[dof] = val; h->body.pd_ready = false; });
}
int orc_add_gravity(void* hp, const double* g) {
    Handle* h = static_cast<Handle*>(hp);
    return guard(h, [&] {
        h->body.gravity.assign(g, g + 3);
        h->body.gravity_off = h->body.n_state_params;
        h->body.n_state_params += 3;
    });
}
int orc_set_frictional_boundary(void* hp, const char* type, const double* p, int np, const int* idx, int ni) {
    Handle* h = static_cast<Handle*>(hp);
    return guard(h, [&] { h->body.set_boundary(type, orc::Vec(p, p + np), std::vector<int>(idx, idx + ni)); });
}
int orc_set_press_plane(void* hp, double z) {
    Handle* h = static_cast<Handle*>(hp);
    return guard(h, [&] { h->body.boundary.plane_z = z; });
}
int orc_add_actuation(void* hp, double stiffness, const double* fiber, const int* idx, int ni) {
    Handle* h = static_cast<Handle*>(hp);
    return guard(h, [&] { h->body.add_actuation(stiffness, fiber, std::vector<int>(idx, idx + ni)); });
}
int orc_set_actuation(void* hp, const double* a, int na) {
    Handle* h = static_cast<Handle*>(hp);
    return guard(h, [&] { h->body.set_actuation(orc::Vec(a, a + na)); });
}
int orc_act_dofs(void* hp) { return static_cast<Handle*>(hp)->body.act_dofs; }
int orc_num_muscles(void* hp) { return (int)static_cast<Handle*>(hp)->body.muscles.size(); }
// dl/da and dl/d(act_w) of the last backward call
int orc_actuation_grads(void* hp, double* dl_da, int na, double* dl_dact_w, int nm) {
    Handle* h = static_cast<Handle*>(hp);
    return guard(h, [&] {
        const auto& b = h->body;
        orc::check(na == (int)b.last_dl_da.size() && nm == (int)b.last_dl_dact_w.size(), "actuation gradient sizes");
        std::copy(b.last_dl_da.begin(), b.last_dl_da.end(), dl_da);
        std::copy(b.last_dl_dact_w.begin(), b.last_dl_dact_w.end(), dl_dact_w);
    });
}
// The DiffPD baseline forward (a pd_eigen method other than pd_eigen_alg_phd):
// active_in is the previous frame's active set, active_out this frame's.
int orc_forward_baseline(void* hp, const char* method, const double* q, const double* v, const double* f_ext, double dt,
                         const char* const* keys, const double* vals, int nopt, double* q_next, double* v_next,
                         const int* active_in, int n_in, int* active_out, int* n_out, int cap, int* stats) {
    Handle* h = static_cast<Handle*>(hp);
    return guard(h, [&] {
        const int n = h->body.dofs;
        orc::Vec qn, vn;
        std::vector<int> act(active_in, active_in + n_in);
        h->body.forward_baseline(method, orc::Vec(q, q + n), orc::Vec(v, v + n), orc::Vec(f_ext, f_ext + n), dt,
                                 make_opts(keys, vals, nopt), qn, vn, act);
        std::copy(qn.begin(), qn.end(), q_next);
        std::copy(vn.begin(), vn.end(), v_next);
        orc::check((int)act.size() <= cap, "active buffer too small");
        std::copy(act.begin(), act.end(), active_out);
        *n_out = (int)act.size();
        stats[0] = h->body.last_baseline_contact_iters;
        stats[1] = h->body.last_baseline_pd_iters;
    });
}
int orc_backward_baseline(void* hp, const double* q, const double* v, const double* f_ext, double dt, const double* q_next,
                          const int* active, int n_active, const double* dl_dq_next, const double* dl_dv_next,
                          const char* const* keys, const double* vals, int nopt, double* dl_dq, double* dl_dv,
                          double* dl_df_ext, double* dl_dmat_w, double* dl_dstate_p, int* stats) {
    Handle* h = static_cast<Handle*>(hp);
    return guard(h, [&] {
        const int n = h->body.dofs;
        orc::Vec a, b, c, d, e;
        h->body.backward_baseline(orc::Vec(q, q + n), orc::Vec(v, v + n), orc::Vec(f_ext, f_ext + n), dt,
                                  orc::Vec(q_next, q_next + n), std::vector<int>(active, active + n_active),
                                  orc::Vec(dl_dq_next, dl_dq_next + n), orc::Vec(dl_dv_next, dl_dv_next + n),
                                  make_opts(keys, vals, nopt), a, b, c, d, e);
        std::copy(a.begin(), a.end(), dl_dq);
        std::copy(b.begin(), b.end(), dl_dv);
        std::copy(c.begin(), c.end(), dl_df_ext);
        std::copy(d.begin(), d.end(), dl_dmat_w);
        std::copy(e.begin(), e.end(), dl_dstate_p);
        stats[0] = h->body.last_baseline_bwd_iters;
        stats[1] = h->body.last_baseline_bwd_newton;
    });
}
int orc_add_planar_contact(void* hp, const double* p, int np) {
    Handle* h = static_cast<Handle*>(hp);
    return guard(h, [&] { h->body.add_planar_contact(orc::Vec(p, p + np)); });
}
int orc_add_hydrodynamics(void* hp, const double* p, int np) {
    Handle* h = static_cast<Handle*>(hp);
    return guard(h, [&] { h->body.add_hydrodynamics(orc::Vec(p, p + np)); });
}
int orc_add_billiard_ball(void* hp, const double* p, int np) {
    Handle* h = static_cast<Handle*>(hp);
    return guard(h, [&] { h->body.add_billiard_ball(orc::Vec(p, p + np)); });
}
int orc_add_mesh_contact(void* hp, const double* p, int np) {
    Handle* h = static_cast<Handle*>(hp);
    return guard(h, [&] { h->body.add_mesh_contact(orc::Vec(p, p + np)); });
}
int orc_add_arc_contact(void* hp, const double* p, int np) {
    Handle* h = static_cast<Handle*>(hp);
    return guard(h, [&] { h->body.add_arc_contact(orc::Vec(p, p + np)); });
}
int orc_num_state_params(void* hp) { return static_cast<Handle*>(hp)->body.n_state_params; }
// f_state(q, v) and its backward for a given dl/df (test access)
int orc_state_force(void* hp, const double* q, const double* v, double* f) {
    Handle* h = static_cast<Handle*>(hp);
    return guard(h, [&] {
        const int n = h->body.dofs;
        const orc::Vec r = h->body.state_force(orc::Vec(q, q + n), orc::Vec(v, v + n));
        std::copy(r.begin(), r.end(), f);
    });
}
int orc_state_force_backward(void* hp, const double* q, const double* v, const double* dl_df, double* dq, double* dv,
                             double* dp) {
    Handle* h = static_cast<Handle*>(hp);
    return guard(h, [&] {
        const int n = h->body.dofs;
        orc::Vec a, b, c;
        h->body.state_force_backward(orc::Vec(q, q + n), orc::Vec(v, v + n), orc::Vec(dl_df, dl_df + n), a, b, c);
        std::copy(a.begin(), a.end(), dq);
        std::copy(b.begin(), b.end(), dv);
        std::copy(c.begin(), c.end(), dp);
    });
}
int orc_dofs(void* hp) { return static_cast<Handle*>(hp)->body.dofs; }
int orc_num_energies(void* hp) { return (int)static_cast<Handle*>(hp)->body.energies.size(); }
int orc_num_elements(void* hp) { return static_cast<Handle*>(hp)->body.ne; }

int orc_forward(void* hp, const double* q, const double* v, const double* f_ext, double dt, const char* const* keys,
                const double* vals, int nopt, double* q_next, double* v_next, int* active, int* n_active, int cap) {
    Handle* h = static_cast<Handle*>(hp);
    return guard(h, [&] {
        const int n = h->body.dofs;
        orc::Vec qn, vn;
        std::vector<int> act;
        h->body.forward(orc::Vec(q, q + n), orc::Vec(v, v + n), orc::Vec(f_ext, f_ext + n), dt, make_opts(keys, vals, nopt),
                        qn, vn, act);
        std::copy(qn.begin(), qn.end(), q_next);
        std::copy(vn.begin(), vn.end(), v_next);
        orc::check((int)act.size() <= cap, "active buffer too small");
        std::copy(act.begin(), act.end(), active);
        *n_active = (int)act.size();
    });
}
int orc_backward(void* hp, const double* q, const double* v, const double* f_ext, double dt, const double* q_next,
                 const double* v_next, const int* active, int n_active, const double* dl_dq_next, const double* dl_dv_next,
                 const char* const* keys, const double* vals, int nopt, double* dl_dq, double* dl_dv, double* dl_df_ext,
                 double* dl_dmat_w, double* dl_dstate_p, double* dl_dw_elem) {
    Handle* h = static_cast<Handle*>(hp);
    return guard(h, [&] {
        const int n = h->body.dofs;
        orc::Vec a, b, c, d, e, f;
        h->body.backward(orc::Vec(q, q + n), orc::Vec(v, v + n), orc::Vec(f_ext, f_ext + n), dt, orc::Vec(q_next, q_next + n),
                         orc::Vec(v_next, v_next + n), std::vector<int>(active, active + n_active),
                         orc::Vec(dl_dq_next, dl_dq_next + n), orc::Vec(dl_dv_next, dl_dv_next + n),
                         make_opts(keys, vals, nopt), a, b, c, d, e, f);
        std::copy(a.begin(), a.end(), dl_dq);
        std::copy(b.begin(), b.end(), dl_dv);
        std::copy(c.begin(), c.end(), dl_df_ext);
        std::copy(d.begin(), d.end(), dl_dmat_w);
        std::copy(e.begin(), e.end(), dl_dstate_p);
        std::copy(f.begin(), f.end(), dl_dw_elem);
    });
}
// x = A_d^-1 b per axis on an AoS 3N vector (solver set up for dt).
int orc_apply_ainv(void* hp, double dt, const double* b, double* x) {
    Handle* h = static_cast<Handle*>(hp);
    return guard(h, [&] {
        h->body.setup(dt);
        const int nv = h->body.nv;
        for (int d = 0; d < 3; ++d) {
            orc::Vec bd(nv);
            for (int i = 0; i < nv; ++i) bd[i] = b[3 * i + d];
            const orc::Vec xd = h->body.solve_axis(d, bd);
            for (int i = 0; i < nv; ++i) x[3 * i + d] = xd[i];
        }
    });
}
int orc_local_step(void* hp, const double* q, double* out) {
    Handle* h = static_cast<Handle*>(hp);
    return guard(h, [&] {
        const orc::Vec r = h->body.local_step(orc::Vec(q, q + h->body.dofs));
        std::copy(r.begin(), r.end(), out);
    });
}
// Per-element K_pd blocks (ne x 3edim x 3edim, row-major).
int orc_local_differential(void* hp, const double* q, int exact, int filter, double rho, int stage, double* out) {
    Handle* h = static_cast<Handle*>(hp);
    return guard(h, [&] {
        const std::vector<double> K =
            h->body.local_differential(orc::Vec(q, q + h->body.dofs), exact != 0, filter, rho, nullptr, stage);
        std::copy(K.begin(), K.end(), out);
    });
}
int orc_frame_stats(void* hp, int* iters, int* conv, int cap) {
    Handle* h = static_cast<Handle*>(hp);
    const int n = (int)h->body.frame_iters.size();
    for (int i = 0; i < n && i < cap; ++i) { iters[i] = h->body.frame_iters[i]; conv[i] = h->body.frame_converged[i]; }
    return n;
}
void orc_reset_frame_stats(void* hp) {
    Handle* h = static_cast<Handle*>(hp);
    h->body.frame_iters.clear();
    h->body.frame_converged.clear();
}
void orc_set_refinement(void* hp, int steps) { static_cast<Handle*>(hp)->body.refine_steps = steps; }
int orc_last_contact_multipliers(void* hp, double* l, double* a, double* b, int cap) {
    Handle* h = static_cast<Handle*>(hp);
    const int n = (int)h->body.last_lambda.size();
    for (int i = 0; i < n && i < cap; ++i) { l[i] = h->body.last_lambda[i]; a[i] = h->body.last_lt1[i]; b[i] = h->body.last_lt2[i]; }
    return n;
}
int orc_last_backward_iters(void* hp) { return static_cast<Handle*>(hp)->body.last_bwd_iters; }
// [backbone iterations (all trust stages), backbone converged, clipped, max Woodbury
//  basis iterations, basis rows, filtered samples, samples, |db|, trust stage, from tape]
int orc_last_backward_stats(void* hp, double* out, int cap) {
    const auto& b = static_cast<Handle*>(hp)->body.last_bwd;
    const double v[10] = {b.backbone_iters, b.backbone_converged, b.clipped, b.woodbury_iters,
                          b.basis_rows, b.filtered, b.samples, b.db_norm, b.trust_stage, b.from_tape};
    for (int i = 0; i < 10 && i < cap; ++i) out[i] = v[i];
    return 10;
}
void orc_coulomb_step(const double* Wt, double u1, double u2, double fmu, double lam, double* out2) {
    orc::coulomb_step(Wt, u1, u2, fmu, lam, out2[0], out2[1]);
}
// Closest point of triangle (a, b, c) to p (the mesh boundary's query).
void orc_tri_closest(const double* p, const double* a, const double* b, const double* c, double* out3, double* w3) {
    orc::tri_closest(p, a, b, c, out3, w3);
}
// Frictional boundary of the body at q (3): out = [distance, n (3), t1 (3), t2 (3),
// governing obstacle of the press: 0 ground, 1 plane (-1 for other types)].
int orc_boundary_query(void* hp, const double* q, double* out) {
    Handle* h = static_cast<Handle*>(hp);
    return guard(h, [&] {
        const auto& b = h->body.boundary;
        out[0] = b.distance(q);
        b.frame(q, out + 4, out + 7, out + 1);
        double dg = 0, dp = 0;
        out[10] = b.type == orc::kPress ? b.press_side(q, dg, dp) : -1;
    });
}
double orc_filter_prox_jacobian(double* M, double aw, double theta, double lo) {
    return orc::filter_prox_jacobian(M, aw, theta, lo);
}
void orc_jacobi_eigen(int n, double* A, double* w, double* Z) { orc::jacobi_eigen(n, A, w, Z); }

// --- element-level primitives for unit tests -------------------------------
void orc_svd3(const double* F, double* U, double* sig, double* V) {
    orc::M3 f, u, v;
    std::memcpy(f.a, F, sizeof(double) * 9);
    orc::svd3(f, u, sig, v);
    std::memcpy(U, u.a, sizeof(double) * 9);
    std::memcpy(V, v.a, sizeof(double) * 9);
}
// type: 0 corotated, 1 volume, 2 neohookean (stiffness k, mu, la).
void orc_project(int type, double k, double mu, double la, const double* F, double* out) {
    orc::Energy en;
    en.type = type; en.stiffness = k; en.mu = mu; en.la = la;
    orc::M3 f;
    std::memcpy(f.a, F, sizeof(double) * 9);
    orc::FAux a;
    a.init(f);
    const orc::M3 p = orc::project(en, a);
    std::memcpy(out, p.a, sizeof(double) * 9);
}
void orc_project_diff(int type, double k, double mu, double la, const double* F, const double* dF, double* out) {
    orc::Energy en;
    en.type = type; en.stiffness = k; en.mu = mu; en.la = la;
    orc::M3 f, df;
    std::memcpy(f.a, F, sizeof(double) * 9);
    std::memcpy(df.a, dF, sizeof(double) * 9);
    orc::FAux a;
    a.init(f);
    const orc::M3 p = orc::project_diff(en, a, df);
    std::memcpy(out, p.a, sizeof(double) * 9);
}
void orc_nh_prox(double k, double mu, double la, const double* sF, double* s) {
    orc::Energy en;
    en.type = orc::kNeoHookean; en.stiffness = k; en.mu = mu; en.la = la;
    orc::nh_solve_prox(en, sF, s);
}
}
