// Host orchestration of the MI355X PD time step: the product counterpart of
// the reference's Deformable<3, 8/4> state (cpp/core/include/fem/deformable.h)
// restricted to the `pd_eigen_alg_phd` path.
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <cfloat>
#include <cmath>
#include <vector>

#include "phd_factor.h"
#include "phd_kernels.h"
#include "phd_math.h"

namespace phd {

void hip_check(hipError_t e, const char* what);

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void alloc(size_t count) {
        if (count == n && p) return;
        release();
        if (count == 0) return;
        hip_check(hipMalloc(&p, count * sizeof(T)), "hipMalloc");
        n = count;
    }
    void upload(const std::vector<T>& h) {
        alloc(h.size());
        if (!h.empty()) hip_check(hipMemcpy(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice), "upload");
    }
    void upload(const T* h, size_t count) {
        alloc(count);
        if (count) hip_check(hipMemcpy(p, h, count * sizeof(T), hipMemcpyHostToDevice), "upload");
    }
    void download(T* h, size_t count) const {
        if (count) hip_check(hipMemcpy(h, p, count * sizeof(T), hipMemcpyDeviceToHost), "download");
    }
    std::vector<T> to_host() const {
        std::vector<T> h(n);
        download(h.data(), n);
        return h;
    }
};

enum EnergyKind { kCorotated = 0, kVolume = 1, kNeoHookean = 2 };

struct Energy {
    int type = kCorotated;
    double stiffness = 0;             // mean stiffness (stiffness_())
    std::vector<double> elem_stiff;   // per-element stiffness (empty: homogeneous)
    double mu = 0, la = 0;            // Neo-Hookean mean Lame parameters
    std::vector<char> member;
    double elem_w(int e) const { return elem_stiff.empty() ? stiffness : elem_stiff[e]; }
};

struct HostBoundary {
    int type = 0;  // 0 none, 1 planar, 2 spherical, 3 heightfield, 4 press, 5 mesh
    double n[3] = {0, 0, 1}, off = 0, c[3] = {0, 0, 0}, r = 0, z0 = 0, plane_z = 0;
    std::vector<double> waves;
    // mesh: contact radius, split index, left then right faces (3 indices each)
    double radius = 0;
    int split = 0, nfl = 0, nfr = 0;
    std::vector<int> faces;

    // Mesh: closest point to p of the triangle set opposite to vertex `node`,
    // the faces at the positions X (3N); false when that set is empty. fn: the
    // face's unit normal ((0, 0, 1) for a degenerate face).
    bool mesh_closest(const double* p, const double* X, int node, double closest[3], double fn[3]) const {
        const bool right = node < split;
        const int f0 = right ? nfl : 0, nf = right ? nfr : nfl;
        double best = INFINITY;
        int bf = -1;
        for (int f = 0; f < nf; ++f) {
            const int* fv = &faces[3 * (size_t)(f0 + f)];
            double cl[3];
            closest_on_triangle(p, &X[3 * (size_t)fv[0]], &X[3 * (size_t)fv[1]], &X[3 * (size_t)fv[2]], cl);
            double d2 = 0;
            for (int d = 0; d < 3; ++d) d2 += (p[d] - cl[d]) * (p[d] - cl[d]);
            if (d2 < best) { best = d2; bf = f; for (int d = 0; d < 3; ++d) closest[d] = cl[d]; }
        }
        if (bf < 0) return false;
        const int* fv = &faces[3 * (size_t)(f0 + bf)];
        const double *A = &X[3 * (size_t)fv[0]], *B = &X[3 * (size_t)fv[1]], *C = &X[3 * (size_t)fv[2]];
        const double e1[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]}, e2[3] = {C[0] - A[0], C[1] - A[1], C[2] - A[2]};
        fn[0] = e1[1] * e2[2] - e1[2] * e2[1]; fn[1] = e1[2] * e2[0] - e1[0] * e2[2]; fn[2] = e1[0] * e2[1] - e1[1] * e2[0];
        const double fl = std::sqrt(fn[0] * fn[0] + fn[1] * fn[1] + fn[2] * fn[2]);
        if (fl > DBL_EPSILON) for (int d = 0; d < 3; ++d) fn[d] /= fl;
        else { fn[0] = 0; fn[1] = 0; fn[2] = 1; }
        return true;
    }

    // Signed distance and outward unit normal of the governing obstacle at q
    // (the host twin of the contact kernels' boundary evaluation).
    void eval(const double* q, double& dist, double nn[3]) const {
        if (type == 1) {
            dist = q[0] * n[0] + q[1] * n[1] + q[2] * n[2] + off;
            for (int d = 0; d < 3; ++d) nn[d] = n[d];
            return;
        }
        if (type == 2) {
            const double d0 = q[0] - c[0], d1 = q[1] - c[1], d2 = q[2] - c[2];
            const double rho = std::sqrt(d0 * d0 + d1 * d1 + d2 * d2);
            dist = rho - r;
            nn[0] = d0 / rho; nn[1] = d1 / rho; nn[2] = d2 / rho;
            return;
        }
        double h = z0, hx = 0, hy = 0;
        for (size_t i = 0; i + 3 < waves.size(); i += 4) {
            const double arg = waves[i + 1] * q[0] + waves[i + 2] * q[1] + waves[i + 3];
            h += waves[i] * std::sin(arg);
            hx += waves[i] * waves[i + 1] * std::cos(arg);
            hy += waves[i] * waves[i + 2] * std::cos(arg);
        }
        const double s = std::sqrt(1 + hx * hx + hy * hy);
        const double dg = (q[2] - h) / s;
        if (type == 4 && plane_z - q[2] < dg) {
            dist = plane_z - q[2];
            nn[0] = 0; nn[1] = 0; nn[2] = -1;
            return;
        }
        dist = dg;
        nn[0] = -hx / s; nn[1] = -hy / s; nn[2] = 1 / s;
    }
};

struct ReplayFrame {  // AlgPhdReplayCache fields used by the backward (deformable.h:336-355)
    std::vector<int> active;
    std::vector<double> q_sol;
    std::vector<double> lambda, omega;
    std::vector<std::array<double, 3>> normals;
};

struct StepStats {
    int iters = 0;
    int converged = 0;
    double seconds = 0;
};

class WarmStartNet;

struct DevAnderson;

class Body {
public:
    explicit Body(int device);
    ~Body();

    // ---- scene (Deformable mutators) ----
    void init_mesh(int edim, int nv, const double* X, int ne, const int* elems, double density);
    void add_pd_energy(const std::string& type, const std::vector<double>& params, const std::vector<int>& idx);
    void set_dirichlet(int dof, double val);
    void add_gravity(const double* g);
    // AddStateForce("planar_contact", [n (3), offset, p, kn, kf, mu]) (PlanarContactStateForce)
    void add_planar_contact(const std::vector<double>& p);
    void add_arc_contact(const std::vector<double>& p);
    void add_mesh_contact(const std::vector<double>& p);
    void add_billiard_ball(const std::vector<double>& p);
    void add_hydrodynamics(const std::vector<double>& p);
    int num_state_params() const { return n_state_params_; }
    // ForwardStateForce / BackwardStateForce at (q, v) on the host (3N vectors)
    void state_force(const double* q, const double* v, double* f) const;
    void state_force_backward(const double* q, const double* v, const double* dl_df, double* dl_dq, double* dl_dv,
                              double* dl_dp) const;
    void set_boundary(const std::string& type, const std::vector<double>& p, const std::vector<int>& idx);
    void set_press_plane(double z) { boundary_.plane_z = z; }
    // PdMuscleEnergy groups (AddActuation) and the actuation levels a of the
    // following forward/backward calls (one per group element, group order)
    void add_actuation(double stiffness, const double* fiber, const std::vector<int>& idx);
    void set_actuation(const double* a, int na);
    // dl/da and dl/d(act_w) of the last backward call
    const std::vector<double>& last_dl_da() const { return last_dl_da_; }
    const std::vector<double>& last_dl_dact_w() const { return last_dl_dact_w_; }

    int dofs() const { return dofs_; }
    int device() const { return device_; }
    int num_elements() const { return ne_; }
    int num_energies() const { return (int)energies_.size(); }
    int act_dofs() const { return act_dofs_; }
    int num_muscles() const { return (int)muscles_.size(); }
    int num_boundary_vertices() const { return (int)fb_vertices_.size(); }

    // ---- hot path ----
    void forward(const double* q, const double* v, const double* f_ext, double dt,
                 const std::map<std::string, double>& opt, double* q_next, double* v_next, std::vector<int>& active);
    void backward(const double* q, const double* v, const double* f_ext, double dt, const double* q_next,
                  const double* v_next, const std::vector<int>& active, const double* dl_dq_next, const double* dl_dv_next,
                  const std::map<std::string, double>& opt, double* dl_dq, double* dl_dv, double* dl_df_ext,
                  double* dl_dmat_w, double* dl_dstate_p, double* dl_dw_elem);
    // The DiffPD baseline forward (ForwardProjectiveDynamics with a pd_eigen
    // method other than pd_eigen_alg_phd, forward.cpp:294-664): `active` carries
    // the previous frame's active contact set in and this frame's out.
    void forward_baseline(const std::string& method, const double* q, const double* v, const double* f_ext, double dt,
                          const std::map<std::string, double>& opt, double* q_next, double* v_next,
                          std::vector<int>& active);
    struct BaselineStats {
        int contact_iters = 0, pd_iters = 0, ls_steps = 0, bwd_iters = 0;
    } last_baseline;
    // BackwardProjectiveDynamics for the baseline (backward.cpp:141-679, its
    // dense-solver path); dl_da / dl_dact_w go to last_dl_da() / last_dl_dact_w().
    void backward_baseline(const double* q, const double* v, const double* f_ext, double dt, const double* q_next,
                           const std::vector<int>& active, const double* dl_dq_next, const double* dl_dv_next,
                           const std::map<std::string, double>& opt, double* dl_dq, double* dl_dv, double* dl_df_ext,
                           double* dl_dmat_w, double* dl_dstate_p);
    // x = A^-1 b (AoS 3N) with the FP32 (precision=1) or FP64 (0) factor and `refine` FP64 refinement steps.
    void apply_ainv(double dt, const double* b, double* x, int precision, int refine);
    // The same passes on caller-owned device vectors, every intermediate in the
    // library's fixed device scratch array: allocates nothing (factor must be built).
    void apply_ainv_noalloc(const double* b_dev, double* x_dev, int precision, int refine);
    size_t noalloc_scratch_bytes(int precision) const;
    void setup(double dt, int threads);

    // ---- autotuner of the launch shapes (factor built for dt first) ----
    // Times `reps` launches of each candidate on this body's factor and mesh
    // and keeps the fastest: the column-chunk width of S w (FP32, the mixed
    // path's product), the unit height of S^T y, the local step's block size.
    struct TuneResult {
        int chunk_w = 0, unit_items = 0, local_block = 0;
        double rows_ms[3] = {0}, cols_ms[4] = {0}, local_ms[3] = {0};
    };
    static constexpr int kTuneChunks[3] = {512, 1024, 2048};
    static constexpr int kTuneUnits[4] = {256, 512, 1024, 2048};
    static constexpr int kTuneBlocks[3] = {64, 128, 256};
    TuneResult autotune(double dt, int threads, int reps);
    // sets the three shapes directly (each from its candidate list); the
    // factor must be built
    void set_launch_shapes(int chunk_w, int unit_items, int local_block);
    int chunk_w() const { return chunk_w_; }
    int unit_items() const { return unit_items_; }
    int local_block() const { return local_block_; }

    // ---- statistics ----
    std::vector<StepStats> frame_stats;
    std::vector<double> last_lambda_, last_lt1_, last_lt2_;
    int last_backward_iters = 0;
    struct BackwardStats {  // of the last backward call
        double backbone_iters = 0, backbone_converged = 0, clipped = 0, woodbury_iters = 0, woodbury_converged = 0,
               basis_rows = 0, filtered = 0, samples = 0, db_norm = 0, trust_stage = 0, from_tape = 1,
               gram_fp64_fallbacks = 0;
    } last_bwd;
    long long factor_bytes() const;
    long long factor_nnz() const { return factor_.nnz; }
    int cached_n() const { return ready_ ? nv_ : -1; }
    double setup_seconds = 0;
    // device-event timing of the S and S^T kernels (milliseconds, accumulated)
    double t_spmv_rows_ms = 0, t_spmv_cols_ms = 0;
    long long n_spmv_rows = 0, n_spmv_cols = 0;
    bool timing = false;
    int timing_level = 0;  // 2: synchronize at phase boundaries for the wall-clock split
    // host wall-clock per phase (seconds, accumulated while `timing` is on); see PhaseId
    enum PhaseId { kFwdTotal = 0, kFwdLocal, kFwdSolve, kFwdContact, kFwdDense, kFwdAA, kFwdGate,
                   kBwdTotal, kBwdKpd, kBwdBackbone, kBwdBasisFactor, kBwdWoodbury, kBwdBasisSolve, kBwdMatGrad,
                   kDenseCalls, kDenseRows, kDenseFp32Fail, kBasisRows, kFwdCStage, kFwdCSets, kFwdCApply, kFwdFric, kNumPhases };
    double phase[kNumPhases] = {0};
    struct PhaseTimer {
        Body* b; int id; double t0;
        PhaseTimer(Body* b_, int id_);
        ~PhaseTimer();
    };
    hipStream_t stream() const { return stream_; }
    void collect_timing();
    const Factor& factor() const { return factor_; }
    const SpmvPlan& plan() const { return plan_; }
    int contact_candidates() const { return K_; }
    int64_t contact_entries() const { return contact_entries_; }  // entries of the candidate paths of S
    int contact_rows() const { return rc_nrows_; }

    // ---- learned warm start (option phd_warm_start = 1) ----
    WarmStartNet& net();
    // raw (standardize = false) or standardized features of state (q, v) into
    // d_feat_; needs the solver set up for dt and this frame's inertial target
    // in d_qtilde_ (set_inertial_target)
    void mlp_features(const double* q_dev, const double* v_dev, double dt, bool standardize);
    void set_inertial_target(const double* q, const double* v, const double* f_ext, double dt);
    const float* feat_dev() const { return d_feat_.p; }
    // forget the previous frame's contact multipliers (the NCP warm start and
    // the network's contact flags), as in a newly constructed body
    void reset_contact_history() { prev_candidates_.clear(); prev_lambda_.clear(); }
    int warm_used = 0, warm_fallback = 0;

private:
    // launch shapes (defaults; autotune() replaces them)
    int chunk_w_ = 1024, unit_items_ = 512, local_block_ = 128;
    bool local_fp32_ = false;  // forward option phd_local_fp32
    bool h32_ = false;         // this frame reads the FP32 copy of H (mixed precision, phd_contact_h32)
    void upload_plan();  // plan_ buffers and the partial-sum arrays sized by it
    int device_ = 0;
    hipStream_t stream_ = nullptr;
    void* scratch_ = nullptr;  // device address of the library's scratch array on device_
    template <class TV>
    void ainv_passes(const double* b_dev, double* x_dev, int refine, char* scr);
    // Pinned host memory for the few scalars read back per iteration (the
    // contact set sizes, the dense solve's pivot flag, the dot products).
    static constexpr size_t kPinSmallBytes = 4096;
    void* pin_small_ = nullptr;
    // scene
    int edim_ = 8, nv_ = 0, ne_ = 0, dofs_ = 0, samples_ = 8;
    std::vector<double> X_;
    std::vector<int> elems_;
    double density_ = 0, element_volume_ = 0;
    std::vector<double> G_;
    int g_stride_ = 0;
    std::vector<Energy> energies_;
    struct Muscle {
        double stiffness = 0;
        double m[3] = {0, 0, 0};  // unit fiber direction
        std::vector<int> elems;
    };
    std::vector<Muscle> muscles_;
    int act_dofs_ = 0;
    std::vector<int> mus_act_;     // per element: index into a, or -1 (no muscle)
    std::vector<double> act_;      // actuation levels of the current call
    std::vector<double> last_dl_da_, last_dl_dact_w_;
    std::map<int, double> dirichlet_;
    std::vector<double> gravity_;
    struct PlanarContactForce {
        double n[3] = {0, 0, 1}, off = 0;
        int p = 2;
        double kn = 0, kf = 0, mu = 0;
        int poff = 0;  // its parameters' offset in dl_dstate_p
    };
    std::vector<PlanarContactForce> pcontacts_;
    struct ArcContactForce {  // inside of a cylindrical arc about the axis (c, dir)
        double c[3], dir[3], start[3], aux[3], R = 0, angle = 0;
        int p = 2;
        double kn = 0, kf = 0, mu = 0;
        int poff = 0;
    };
    std::vector<ArcContactForce> acontacts_;
    struct MeshContactForce {  // vertices of surface A against triangles of B and the reverse
        double radius = 0, kn_l = 0, kn_r = 0, kf = 0, mu = 0;
        std::vector<std::array<int, 3>> fa, fb;
        std::vector<int> va, vb;  // sorted vertex sets of fa and fb
        int poff = 0;
    };
    std::vector<MeshContactForce> mcontacts_;
    struct BilliardBallForce {  // consecutive blocks of svn vertices as balls
        double R = 0;
        int svn = 0, nb = 0;
        std::vector<double> k, mu;
        int poff = 0;
    };
    std::vector<BilliardBallForce> bballs_;
    struct HydroForce {  // water drag and thrust on surface faces (fd nodes each)
        double rho = 0, vw[3] = {0, 0, 0}, P[16] = {0}, fmax = 0;
        int fd = 4;
        std::vector<int> faces;
        int poff = 0;
    };
    std::vector<HydroForce> hydros_;
    int gravity_off_ = 0, n_state_params_ = 0;
    HostBoundary boundary_;
    std::map<int, int> fb_vertices_;

    // solver cache
    bool ready_ = false;
    double ready_dt_ = 0;
    Factor factor_;
    SpmvPlan plan_;
    std::vector<int> Ap_, Ai_;
    std::vector<double> Ax_;
    std::vector<unsigned char> dmask_;
    // per-DOF Dirichlet: DOFs fixed on their own (their vertex keeps a free
    // DOF), the vertices carrying them, and the bordered correction's data
    static constexpr int kMaxPartialVertices = 4096;
    std::vector<unsigned char> pmask_;
    std::vector<int> pverts_;
    DevBuf<unsigned char> d_pmask_;
    DevBuf<double> d_pZ_, d_pginv_, d_pcoef_;
    DevBuf<int> d_pmap_, d_pfv_;
    PdirAxes pax_{};
    void build_partial_dirichlet();
    // y (AoS 3N, the result of a global solve y = t + A^-1 r): constrains the
    // solve to leave y = t on the partially fixed DOFs, then writes o there
    // (t, o may be null: zero)
    void partial_dirichlet(double* y, const double* t, const double* o);
    std::vector<int> ddof_;
    std::vector<double> dval_, dval_full_;
    std::vector<double> c_dir_;
    // device
    DevBuf<double> V64_;
    DevBuf<float> V32_;
    DevBuf<int64_t> d_base_;
    DevBuf<int> d_fd_, d_perm_, d_items_, d_cu_tile_, d_cu_i0_, d_cu_i1_, d_tile_unit_ptr_, d_gu_group_, d_gu_chunk_,
        d_rbase_;
    int ell_w_ = 0;            // A in ELL form: column / value of entry j of row v at j * nv + v
    DevBuf<int> d_Aie_;
    DevBuf<double> d_Axe_;
    DevBuf<unsigned char> d_dmask_;
    DevBuf<int> d_ddof_;
    DevBuf<double> d_dval_, d_dval_full_;
    DevBuf<double> d_c_dir_;
    DevBuf<int> d_elems_, d_inc_ptr_, d_inc_idx_;
    DevBuf<double> d_G_, d_wc_, d_wv_, d_wn_[4];
    DevBuf<double> d_kc_, d_kv_, d_kn_[4];
    DevBuf<double> d_mw_, d_mfib_, d_mact_, d_gda_, d_gdw_;
    DevMaterials mat_;
    int n_nh_ = 0;
    std::vector<int> nh_energy_;  // energy index of each NH slot
    bool has_cor_ = false, has_vol_ = false;
    bool kpd_exact_ = false;
    // work vectors (3N FP64)
    DevBuf<double> d_q_, d_qprev_, d_qtilde_, d_b_, d_r_, d_pdrhs_, d_pdrhs_prev_, d_elem_, d_g_, d_tmp_, d_kz_;
    // warm start: network, per-vertex feature inputs, guesses
    std::unique_ptr<WarmStartNet> net_;
    DevBuf<float> d_logk_, d_cflag_, d_cdist_, d_feat_, d_mlpout_, d_fmean_, d_fstd_, d_adiag_;
    DevBuf<double> d_Xrest_, d_v_, d_guess_;
    void warm_start(const double* v, double dt, double inv_h2m, bool learned);
    DevBuf<double> d_hist_q_[8], d_hist_g_[8];
    DevBuf<double> d_dot_partial_, d_dot_out_;
    DevBuf<float> w32_, y32_, part32_, xpart32_;
    DevBuf<double> w64_, y64_, part64_, xpart64_;
    // contact
    int K_ = 0;
    std::vector<int> cand_;
    DevBuf<int> d_cand_;
    DevBuf<double> d_H_;
    DevBuf<float> d_H32_;  // FP32 copy of H for the mixed path's applies and assembly        // A^-1 restricted to the candidates (K x K, FP64)
    std::vector<double> Hdiag_;
    DevBuf<int> d_cp_ptr_, d_cp_row_, d_rc_ptr_, d_rc_row_, d_rc_c_;
    DevBuf<double> d_cp_val64_, d_rc_val64_;
    DevBuf<float> d_cp_val32_, d_rc_val32_;
    int rc_nrows_ = 0;
    int64_t contact_entries_ = 0;
    ContactPieces gather_pieces_, scatter_pieces_;
    DevBuf<int> d_gp_b_, d_gp_e_, d_gp_ptr_, d_sp_b_, d_sp_e_, d_sp_ptr_;
    DevBuf<double> d_piece_part_;
    DevBuf<double> d_waves_;
    DevBuf<int> d_mesh_faces_;
    DevBuf<double> d_nrm_, d_t1_, d_t2_, d_delta_, d_F_, d_dF_, d_uC_, d_tmpK3_, d_om_, d_E_;
    // per-candidate NCP state of the frame (device-resident)
    DevBuf<double> d_rprec_, d_lambda_, d_lt1_, d_lt2_, d_qframeC_, d_hvec_, d_Fg_, d_Fn_, d_hprev_, d_Hdiag_;
    DevBuf<unsigned char> d_ncp_flags_;
    DevBuf<int> d_ncp_lists_, d_ncp_counts_, d_posJ_, d_posL_;
    DevBuf<double> d_C64_, d_cy64_, d_cx64_;  // dense contact system (blocked Cholesky layout)
    DevBuf<float> d_C32_, d_cy32_, d_cx32_;
    DevBuf<double> d_rhsK_;
    DevBuf<float> d_rhsK32_;
    DevBuf<int> d_info_;
    DevBuf<int> d_idxJ_, d_idxP_, d_idxF_;
    void* blas_ = nullptr;
    std::vector<hipEvent_t> ev_pool_;
    size_t ev_used_ = 0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_rows_, ev_cols_;
    hipEvent_t next_event();
    // backward
    DevBuf<double> d_W64_;
    DevBuf<float> d_W32_;
    DevBuf<int> d_nfilt_;  // samples whose proximal-map block the filter changed
    DevBuf<double> d_gcor_, d_gvol_, d_gnh_;
    // state kept between calls (mutable members of the reference Deformable)
    std::vector<ReplayFrame> tape_;
    std::vector<int> prev_candidates_;
    std::vector<double> prev_lambda_;

    void build_materials();
    void setup_contact(int threads);
    DevBoundary dev_boundary();
    // one global solve pass: r (AoS 3N FP64) -> y (S r, factor precision); returns nothing
    template <class TV>
    void spmv_S(const double* r, DevBuf<TV>& w, DevBuf<TV>& y, DevBuf<TV>& part, const TV* V);
    template <class TV>
    void spmv_ST(DevBuf<TV>& y, DevBuf<TV>& xpart, const TV* V, const double* add, const double* dval, double* out);
    template <class TV>
    void contact_gather(DevBuf<TV>& y, double* out);
    template <class TV>
    void contact_scatter(const double* f, DevBuf<TV>& y);
    void local_rhs(const double* q, double inv_h2m);
    void kpd_apply(const double* x, double* out, int precision);
    std::vector<double> dots(const std::vector<std::pair<const double*, const double*>>& pairs);
    bool anderson_step(DevAnderson& aa, const double* xprev, const double* g, double* out,
                       const std::vector<std::pair<const double*, const double*>>& extra, std::vector<double>* extra_out);
    bool solve_contact_system(int n, int precision, const int* J, bool plain_w = false);
    void forward_impl(const double* q, const double* v, const double* f_ext, double dt,
                      const std::map<std::string, double>& opt, double* q_next, double* v_next, std::vector<int>& active,
                      const std::vector<int>* replay_active, ReplayFrame* replay_out);
    DevBuf<unsigned char> d_enabled_;  // the replay's contacts among the candidates
    // baseline forward (phd_baseline.cpp)
    double baseline_energy(const double* q);
    void baseline_force(const double* q, double* f);
    void ainv64(const double* r, double* x);
    void pd_lhs_solve(const double* r, double* x, const std::vector<int>& frozen);
    int pd_nonlinear_solve(const double* q_init, double inv_h2m, const double* rhs, const std::vector<int>& frozen,
                           const std::vector<double>& frozen_vals, const std::map<std::string, double>& opt,
                           double* q_sol);
    std::map<int, int> zcol_;                  // frozen node -> column of d_Z_ (A^-1 e_node)
    DevBuf<double> d_Z_, d_bl_energy_, d_bl_mu_;
    DevBuf<int> d_bl_zcol_, d_bl_nodes_, d_bl_fix_;
    DevBuf<double> d_bl_fixv_;
    std::vector<int> g_frozen_;                // frozen set of the factored G below
    std::vector<double> g_llt_;                // Cholesky factor of G = Z restricted to the frozen nodes
    std::vector<std::unique_ptr<DevBuf<double>>> d_bl_;  // work vectors (3N)
};

}  // namespace phd
