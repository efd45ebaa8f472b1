// Dense SPD solve on the device for the per-iteration contact system
// M dl = rhs (the reference factors M_sys = Omega W Omega + E with a host LDL^T
// every PD iteration, deformable_projective_dynamics_forward_gpu.cpp:1199-1240).
//
// Blocked right-looking Cholesky in 64 x 64 blocks. The right-hand side rides
// along as an extra block row, so L^-1 rhs comes out of the factorization; a
// blocked backward substitution with the inverted diagonal blocks (formed by
// the panel kernel next to the block rows) then gives x = L^-T L^-1 rhs.
//
// Layout (column-major, lda = np + 64, np = 64 * ceil(n / 64)):
//   rows [0, np)        lower triangle of M, identity on the padding;
//   rows [np, np + 64)  right-hand-side block: row np = rhs^T, the rest zero.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace phd {

constexpr int kCholBlock = 64;

inline int chol_padded(int n) { return kCholBlock * ((n + kCholBlock - 1) / kCholBlock); }
inline int64_t chol_lda(int n) { return (int64_t)chol_padded(n) + kCholBlock; }
inline int64_t chol_elems(int n) { return chol_lda(n) * chol_padded(n); }
// y (np) + the inverted diagonal blocks L_kk^-T (np / 64 blocks of 64 x 64)
inline int64_t chol_work(int n) { return (int64_t)chol_padded(n) * (1 + kCholBlock); }

// Writes the contact system of the index set J (nJ rows) into the layout above:
// M_ab = om_i (n_i . n_j) H_ij om_j + [a == b] E_i with i = J[a], j = J[b]
// (n_i . n_j = 1 when nrm is null); H is the symmetric K x K block of A^-1 on
// the contact vertices.
template <class TV, class TH>
void launch_chol_assemble(hipStream_t s, int nJ, const int* J, const TH* H, int K, const double* nrm,
                          const double* om, const double* E, const double* rhs, TV* M);

// Factorization + both substitutions. work: device scratch of chol_work(n)
// entries; x: np entries, on return x[0, n) holds the solution. fail[0] is set
// to 1 when a pivot is not positive (the caller clears it beforehand).
template <class TV>
void launch_chol_solve(hipStream_t s, int n, TV* M, TV* work, TV* x, int* fail);

// Clears the failure flag with a kernel (safe inside captured graphs).
void launch_clear_flag(hipStream_t s, int* flag);

}  // namespace phd
