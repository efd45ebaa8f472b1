// Launchers for the HIP kernels of the PD time step (phd_kernels.hip).
// All pointers are device pointers; every launcher enqueues on `stream`.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace phd {

// Per-element material description of the merged PD energies.
struct DevMaterials {
    int ne = 0, edim = 8, samples = 8;
    const int* elems = nullptr;   // SoA: elems[k * ne + e]
    const double* G = nullptr;    // grad_undeformed_sample_weight rows (edim x 3, row-major) per sample
    int g_stride = 0;             // doubles per element (0: shared by all elements)
    const double* wc = nullptr;   // corotated weight * vol / samples (nullptr: none)
    const double* wv = nullptr;   // volume weight * vol / samples
    int n_nh = 0;
    const double* wn[4] = {nullptr, nullptr, nullptr, nullptr};
    double nh_mu[4] = {0, 0, 0, 0}, nh_la[4] = {0, 0, 0, 0}, nh_k[4] = {0, 0, 0, 0};
    // muscle fibers (PdMuscleEnergy), at most one per element: weight
    // stiffness * vol / samples (0: none), unit fiber direction (SoA, 3 ne) and
    // actuation level; mw == nullptr when the body has no muscles
    const double* mw = nullptr;
    const double* mfib = nullptr;
    const double* mact = nullptr;
};

struct DotArgs {
    int m = 0;
    const double* a[32];
    const double* b[32];
};

// Boundary (frictional obstacle) description for the contact kernels.
struct DevBoundary {
    int type = 0;                 // 1 planar, 2 spherical, 3 heightfield, 4 press, 5 mesh
    double n[3] = {0, 0, 1}, off = 0;
    double c[3] = {0, 0, 0}, r = 0;
    double z0 = 0, plane_z = 0;
    int nwaves = 0;
    const double* waves = nullptr;  // device (amp, kx, ky, phase) quadruples
    // mesh: two triangle sets of the body itself (left, then right: 3 vertex
    // indices each); a vertex below `split` queries the right set, any other
    // the left set, at the body's current positions
    int split = 0, nfl = 0, nfr = 0;
    const int* faces = nullptr;
    double radius = 0;
};

// Axis stride of the S-product input w (3 axes): 64-aligned with room for the
// 16-byte vector loads past the last column.
__host__ __device__ inline int spmv_ldw(int n) { return ((n + 64 + 63) / 64) * 64; }
// Tail padding (elements) of the factor arrays V64 / V32 for the same loads.
constexpr int kVPad = 64;

// ---- factor expansion and the S / S^T products --------------------------------
// chain[i] = 1 when parent(i) == i + 1 and pattern(i) = {i} + pattern(i + 1).
void launch_expand_s(hipStream_t s, int n, int t_begin, int t_end, const int64_t* Lp, const int* Li, const double* Lx,
                     const int* parent, const int* depth, const int* fd, const int64_t* base, const unsigned char* chain,
                     double* ws, int64_t ws_stride, double* V64, float* V32);

// S w by (64-row group, chunk_w-column chunk) units with w staged in LDS; the
// per-(row, chunk) partials are then summed per row (plan: SpmvPlan::gu_*, rbase,
// built for the same chunk_w, one of kSpmvChunks).
constexpr int kSpmvChunks[3] = {512, 1024, 2048};
template <class TV>
void launch_spmv_rows_grp(hipStream_t s, int nunits, const int* gu_group, const int* gu_chunk, const int* fd,
                          const int64_t* base, const int* rbase, const TV* V, const TV* w, int n, TV* pslot, int chunk_w);
template <class TV>
void launch_spmv_rows_slots(hipStream_t s, int n, const int* rbase, const TV* pslot, TV* y);
template <class TV>
void launch_spmv_cols(hipStream_t s, int nunits, const int* cu_tile, const int* cu_i0, const int* cu_i1, const int* items,
                      const int* fd, const int64_t* base, const TV* V, const TV* y, int n, TV* xpart);
// out[3 perm[j] + d] = (add ? add[...] : 0) + sum of tile partials; vertices with dmask set keep dval.
template <class TV>
void launch_spmv_cols_fin(hipStream_t s, int n, const int* tile_unit_ptr, const TV* xpart, const int* perm,
                          const double* add, const unsigned char* dmask, const double* dval, double* out);
// w[d * n + j] = r[3 perm[j] + d] (SoA, postorder), zero at dmask vertices.
template <class TV>
void launch_perm_in(hipStream_t s, int n, const int* perm, const double* r, const unsigned char* dmask, TV* w);

// ---- local step, gather, residual --------------------------------------------
// fp32: the SVD and the projections in FP32 (F and the sums in FP64)
void launch_local_step(hipStream_t s, const DevMaterials& m, const double* q, double* elem_out, int block = 128,
                       bool fp32 = false);
// pd_rhs = sum of element contributions - c_dir; b = inv_h2m * q_tilde + pd_rhs
void launch_gather_rhs(hipStream_t s, int nv, const int* inc_ptr, const int* inc_idx, const double* elem_out,
                       const double* c_dir, const double* q_tilde, double inv_h2m, double* pd_rhs, double* b);
// Per-DOF Dirichlet (the bordered correction of a global solve on the shared
// factor; PdirAxes: axis d's fixed vertices are fv[off[d] .. off[d + 1]) and
// its G_d^-1 starts at ginv + goff[d]).
struct PdirAxes {
    int off[4];
    long long goff[3];
};
void launch_pdir_coef(hipStream_t s, int np, const PdirAxes& ax, const int* map, const int* fv, const double* ginv,
                      const double* y, const double* t, double* coef);
void launch_pdir_apply(hipStream_t s, int nv, int np, const double* Z, const double* coef, const unsigned char* pmask,
                       const double* o, double* y);
void launch_pdir_columns(hipStream_t s, int nv, int p0, int na, const double* x, double* Z);
void launch_pdir_gather(hipStream_t s, int np, int nv, const int* pv, const double* Z, double* out);
// r = b - A q for the three axes (AoS); A in ELL form (entry j of row v at
// j * nv + v, width entries per row); dmask: whole-vertex Dirichlet rows
// (r = 0); pmask (per DOF, may be null): DOFs fixed on their own (r = 0, and
// their couplings are left out on that axis)
void launch_residual(hipStream_t s, int nv, int width, const int* Ai, const double* Ax, const double* b, const double* q,
                     const unsigned char* dmask, double* r,
                     const unsigned char* pmask = nullptr);
void launch_gather_vertex(hipStream_t s, int nv, const int* inc_ptr, const int* inc_idx, const double* elem_out, double* out);

// ---- reductions and vector updates -------------------------------------------
void launch_dots(hipStream_t s, int n, const DotArgs& args, double* partial, double* out);
int dots_partial_size();
// out = x + alpha * (y - z) ... generic axpby: out = a*x + b*y (+ c*z)
// Anderson history push: dQ = x - x_last, dG = g - g_last (skipped when dQ is null),
// then x_last = x, g_last = g.
void launch_aa_push(hipStream_t s, int n, const double* x, const double* g, double* x_last, double* g_last, double* dQ,
                    double* dG);
void launch_axpbypcz(hipStream_t s, int n, double a, const double* x, double b, const double* y, double c,
                     const double* z, double* out);
// out = q_prev + g - sum_i gamma_i (dQ_i + dG_i)
void launch_aa_combine(hipStream_t s, int n, const double* q_prev, const double* g, int m, const double* const* dQ,
                       const double* const* dG, const double* gamma, double* out);
void launch_set_dirichlet(hipStream_t s, int nd, const int* ddof, const double* dval, double* q);
void launch_zero_dirichlet(hipStream_t s, int nd, const int* ddof, double* q);
// NaN/Inf -> 0 (returned count ignored), scaling done separately
void launch_nan_to_zero(hipStream_t s, int n, double* x);
void launch_scale(hipStream_t s, int n, double a, double* x);
// x *= clip / |x| if |x| = sqrt(norm2[0]) is finite and > clip (norm2 on the device)
void launch_clip_scale(hipStream_t s, int n, double clip, const double* norm2, double* x);

// ---- contact -----------------------------------------------------------------
// normals / tangents at the frame-start positions (GetLocalFrame)
// Mesh boundary (type 5): the frames are anchored at the closest point of the
// opposite triangle set at q (normal (p - closest) / |p - closest|, tangents by
// Gram-Schmidt from the axis least parallel to it), and enabled_out[c] marks
// the candidates within 5 contact radii of that set; the distance is the
// projection of p - closest on the frame-start normal refn minus the radius.
void launch_contact_frames(hipStream_t s, int K, const DevBoundary& b, const int* nodes, const double* q, double* nrm,
                           double* t1, double* t2, unsigned char* enabled_out = nullptr);
void launch_contact_distance(hipStream_t s, int K, const DevBoundary& b, const int* nodes, const double* q, double* delta,
                             const double* refn = nullptr);
// Lists (candidate paths, or rows crossed by paths) cut into pieces of at most
// kContactPiece entries: piece q covers entries [b[q], e[q]) and the pieces of
// list l are [ptr[l], ptr[l + 1]) (device arrays).
constexpr int kContactPiece = 1024;
struct ContactPieces {
    int nlists = 0, npieces = 0;
    const int* b = nullptr;
    const int* e = nullptr;
    const int* ptr = nullptr;
};
// u[c] = sum_p val[p] * y[d * n + row[p]] for the three axes: out[3c + d]
// (part: 3 doubles per piece of scratch)
template <class TV>
void launch_contact_gather(hipStream_t s, const ContactPieces& pc, const int* cp_row, const TV* cp_val, const TV* y, int n,
                           double* part, double* out);
// y[d * n + row] += sum over the row's contacts of val * f[3c + d]
template <class TV>
void launch_contact_scatter(hipStream_t s, const ContactPieces& pc, const int* rc_row, const int* rc_c, const TV* rc_val,
                            const double* f, double* part, int n, TV* y);
// add f[3c + d] at node c: r[3 node + d] += f
void launch_add_node_force(hipStream_t s, int K, const int* nodes, const double* f, double* r);
// M[i][j] = om_i (n_i . n_j) H[i][j] om_j + (i == j) E_i  (row-major, TM = float or double)
template <class TM>
void launch_assemble_msys(hipStream_t s, int K, const double* H, const double* nrm, const double* om, const double* E, TM* M);
// out[3c + d] = sum_j H[c][j] f[3j + d]
void launch_h_apply(hipStream_t s, int K, const double* H, const double* f, double* out);

void launch_gather_vals(hipStream_t s, int64_t m, const int* rows, const int* cols, const int64_t* base,
                        const double* V64, double* o64, float* o32);
void launch_scatter_dense(hipStream_t s, int64_t m, const int* ridx, const int* cidx, const double* val, int K, double* D);
// mirrors the lower triangle (column-major, as written by rocblas_dsyrk) into the upper one
void launch_symmetrize_lower(hipStream_t s, int K, double* H);

void launch_gather_nodes(hipStream_t s, int K, const int* nodes, const double* q, double* out);
template <class TG>
void launch_assemble_g(hipStream_t s, int nb, const int* cidx, const double* dirs, const double* H, int K, TG* G);
// ub[a] = dirs_a . u[cidx[a]] (3-vectors per candidate)
// ub[a] = d_a . (u + u2)_{c_a} (u2 may be null)
void launch_basis_project(hipStream_t s, int nb, const int* cidx, const double* dirs, const double* u, const double* u2,
                          double* ub);
// F[3c + d] = -sum over the basis rows a of candidate c (clist[cptr[c] .. cptr[c+1]), increasing a) of g[a] dirs_a[d]
void launch_basis_expand(hipStream_t s, int K, const int* cptr, const int* clist, const double* dirs, const double* g,
                         double* F);
// Gs (ld x ld, column-major, ld a multiple of 64, zero past n) = full symmetric
// matrix from the lower triangle of G (n x n, column-major).
template <class TG, class TV>
void launch_sym_full(hipStream_t s, int n, int64_t ld, const TG* G, TV* Gs);
// y = Gs x for the full symmetric Gs above.
template <class TV>
void launch_gemv_sym(hipStream_t s, int n, int64_t ld, const TV* Gs, const double* x, double* y);

// out[3a + d] = sum_b H[rows[a]][cols[b]] f[3 cols[b] + d]
// ---- NCP iteration on the device (forward_gpu.cpp:1149-1302) ----
// flag bits of a candidate: in J (omega != 0), carries an impulse (F != 0),
// carries an impulse of g (omega lambda != 0 or friction), in P (J or F);
// friction sets: lambda > 0, omega lambda != 0
constexpr int kNcpJ = 1, kNcpF = 2, kNcpG = 4, kNcpP = 8, kNcpL = 1, kNcpN = 2;
// enabled (may be null): candidates outside the set take no part (omega 0, no impulse)
void launch_ncp_sets(hipStream_t s, int K, const unsigned char* enabled, const double* delta, const double* lambda,
                     const double* rprec, const double* nrm, const double* t1, const double* t2, const double* qC, double* uC,
                     const double* F, const double* lt1, const double* lt2, int do_fric, double* om, double* E, double* hvec,
                     double* Fg, unsigned char* flags);
// lists (nl <= 4, K entries each) of the candidates whose flags contain masks[i];
// counts[i] their sizes; pos[c] = position of c in list 0 or -1
void launch_ncp_compact(hipStream_t s, int K, const unsigned char* flags, int nl, const int* masks, int* lists, int* counts,
                        int* pos);
void launch_ncp_rhs(hipStream_t s, int nP, const int* P, const double* hprev, const int* posJ, const double* hg,
                    const double* hvec, const double* om, const double* nrm, double* uC, double* rJ);
template <class TV>
void launch_ncp_lambda(hipStream_t s, int K, const int* posJ, const TV* x, const double* hvec, const double* E,
                       const double* om, double lam_max, double* lambda);
void launch_fric_sets(hipStream_t s, int K, const double* lambda, const double* om, const double* nrm, double* Fn,
                      unsigned char* flags);
void launch_fric_coulomb(hipStream_t s, int K, const double* lambda, const int* posL, const double* hn, const double* uC,
                         const double* qframe, const double* nrm, const double* t1, const double* t2, const double* Hdiag,
                         double fmu, double* lt1, double* lt2);
void launch_ncp_impulses(hipStream_t s, int K, const double* om, const double* lambda, const double* nrm, const double* lt1,
                         const double* lt2, const double* t1, const double* t2, int do_fric, double* F, double* dF);
// H_rows,cols f with both counts read on the device; grid for max_rows rows.
// TH: the stored precision of H (FP64, or its FP32 copy), sums in FP64.
template <class TH>
void launch_h_apply_idx_dev(hipStream_t s, int max_rows, const int* nr_dev, const int* rows, const int* nc_dev,
                            const int* cols, const TH* H, int K, const double* f, double* out);
template <class TH>
void launch_h_apply_idx(hipStream_t s, int nr, const int* rows, int nc, const int* cols, const TH* H, int K,
                        const double* f, double* out);
void launch_to_float(hipStream_t s, int64_t n, const double* a, float* b);
// contact system restricted to the index set J (row-major nJ x nJ)
template <class TM>
void launch_assemble_msys_idx(hipStream_t s, int nJ, const int* J, const double* H, int K, const double* nrm,
                              const double* om, const double* E, TM* M);

// streaming copy of n16 16-byte words (byte-counter calibration)
void launch_copy(hipStream_t s, int64_t n16, const void* a, void* b);

// ---- backward ----------------------------------------------------------------
// per-sample 9x9 proximal-map Jacobians weighted by the K_pd weights
template <class TW>
// filter != 0: each sample block's eigenvalues mu (relative to the sample's
// weight in A) above 1 move to 1 + theta (mu - 1), then up to lo; filter = 0
// keeps the reference's K_pd. nfilt (may be null) counts the samples whose
// block moved by more than 1e-9.
void launch_kpd_build(hipStream_t s, const DevMaterials& m, const double* kc, const double* kv, const double* const* kn,
                      const double* q, int filter, double theta, double lo, int* nfilt, TW* W);
template <class TW>
void launch_kpd_apply(hipStream_t s, const DevMaterials& m, const TW* W, const double* dq, double* elem_out);
// g[type][e]: z_e^T (-(vol/s) sum_s A_s^T vec(P_s / stiffness)) per merged energy type
void launch_dmatw(hipStream_t s, const DevMaterials& m, const double* q, const double* z, double vol_over_s, double* g_cor,
                  double* g_vol, double* g_nh);
// the baseline forward's per-element PD + muscle energy and element forces
void launch_pd_energy_force(hipStream_t s, const DevMaterials& m, const double* kc, const double* kv, const double* q,
                            double* energy, double* elem_out);
// x += Z[zcol[c]] mu[c] per axis over the nc frozen nodes; dst = src on the nodes
void launch_frozen_apply(hipStream_t s, int nv, int nc, const double* Z, const int* zcol, const double* mu, double* x);
void launch_copy_nodes(hipStream_t s, int nc, const int* nodes, const double* src, double* dst);
// G[i nc + j] = Z[zcol[j]][nodes[i]]
void launch_gather_z(hipStream_t s, int nc, int nv, const int* nodes, const int* zcol, const double* Z, double* G);
// per muscle element (m.mw != nullptr): g_da[e] = z^T df/da, g_dw[e] = z^T df/d(act_w) (0 elsewhere)
void launch_dact(hipStream_t s, const DevMaterials& m, const double* q, const double* z, double vol_over_s, double* g_da,
                 double* g_dw);

}  // namespace phd
