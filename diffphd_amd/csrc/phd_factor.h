// Host-side construction of the sparse-inverse factor S with S^T S = A^-1.
//
// The reference (deformable_projective_dynamics_forward_gpu.cpp:600-647 with
// RealSim SparseLDLT.cpp:226-431) builds S = D^-1/2 L^-1 P from a METIS-ordered
// LDL^T and stores it as an FP64 CSR of S^T. Here the ordering is a nested
// dissection by recursive coordinate bisection, the numeric factorization is a
// multifrontal FP64 Cholesky A = P^T L L^T P (so S = L^-1 P, identical to
// D^-1/2 L_unit^-1 P), and S is stored index-free: with the elimination tree
// postordered, row k of L^-1 is exactly the subtree of k, i.e. the contiguous
// column range [fd(k), k]. Value (k, j) lives at V[base[k] + j].
#pragma once
#include <cstdint>
#include <vector>

namespace phd {

struct Factor {
    int n = 0;
    std::vector<int> perm;    // perm[new] = old vertex
    std::vector<int> iperm;   // iperm[old] = new
    std::vector<int> parent;  // elimination tree (new indices), -1 for roots
    std::vector<int> fd;      // first descendant (postorder) of each node
    std::vector<int> depth;   // root depth 0
    // Cholesky factor L (new indices), CSC, diagonal first in each column.
    std::vector<int64_t> Lp;
    std::vector<int> Li;
    std::vector<double> Lx;
    // Index-free layout of S = L^-1 (rows = new indices).
    std::vector<int64_t> base;  // V[base[k] + j] = S(k, j) for fd[k] <= j <= k
    int64_t v_len = 0;          // padded length of V
    int64_t nnz = 0;            // sum of subtree sizes
    int max_depth = 0;
    double factor_seconds = 0, order_seconds = 0;

    // Builds the factor of the symmetric matrix given in full CSR (scalar,
    // n x n) with vertex coordinates X (3n) used by the bisection.
    void build(int n_, const std::vector<int>& Ap, const std::vector<int>& Ai, const std::vector<double>& Ax,
               const std::vector<double>& X, int threads);
};

// Work units of the two SpMVs over the S layout.
//  * rows (y = S w): units of (group of kRowGroup consecutive rows, chunk of
//    chunk_w columns), see below.
//  * cols (x = S^T y): tile t = columns [64t, 64t+63]; its items are the rows
//    intersecting the tile (the tile's own rows then the ancestors of its last
//    column), split into units of at most `unit_items` rows.
struct SpmvPlan {
    // cols
    std::vector<int> items;         // concatenated per tile
    std::vector<int> cu_tile, cu_i0, cu_i1;  // unit -> tile, item range
    std::vector<int> tile_unit_ptr;          // units of tile t
    // S w by units of (group of kRowGroup consecutive rows, chunk of kChunkW
    // columns): each unit stages w[chunk] once for all rows of the group that
    // cross it. Row k's partial for chunk c sits in slot rbase[k] + c - fd(k) / chunk_w.
    static constexpr int kRowGroup = 64;
    int chunk_w = 1024;
    std::vector<int> gu_group, gu_chunk, rbase;
    void build(const Factor& f, int unit_items, int group_chunk_w = 1024);
};

void host_factor_check(int n, const std::vector<int>& Ap, const std::vector<int>& Ai, const std::vector<double>& Ax,
                       const std::vector<double>& X, double* out);

}  // namespace phd
