/*
 * phd.h -- C ABI of the MI355X-native projective-dynamics (PD) time step.
 *
 * Drop-in boundary for DiffPhD's `pd_eigen_alg_phd` path. Plain pointers and
 * sizes only; every function returns 0 on success and a non-zero code on
 * failure, with the message available from phd_last_error(ctx) (the binding
 * raises it as a Python RuntimeError, as the reference's SWIG %exception does
 * for std::runtime_error, cpp/core/src/py_diff_pd_core.i:17-27). No exception
 * crosses the ABI. Host buffers belong to the caller; device memory belongs
 * to the context. One context per body (and per thread): unlike the
 * reference's file-scope GPU cache (forward_gpu.cpp:280-298) contexts are
 * independent. All positions/velocities/forces are FP64 AoS vectors
 * [x0, y0, z0, x1, ...] of length 3N in the mesh's vertex order.
 */
#ifndef PHD_H
#define PHD_H

#ifdef __cplusplus
extern "C" {
#endif

typedef void* phd_ctx;

/* Context on HIP device `device`. Replaces the file-scope cuSPARSE handle and
 * factor cache of forward_gpu.cpp:55-70, 280-298. NULL on failure. */
phd_ctx phd_ctx_create(int device);
void phd_ctx_destroy(phd_ctx ctx);
/* Last error message of the context (valid until the next call). */
const char* phd_last_error(phd_ctx ctx);

/* Mesh + density: Deformable::Initialize(vertices, elements, density, 'none', E, nu),
 * deformable.cpp:24-43 with mesh.cpp:36-57. X: 3*nv (xyz per vertex), elems:
 * edim*ne (edim = 8 hex in the 000..111 node order, or 4 tet). */
int phd_init_mesh(phd_ctx ctx, int edim, int nv, const double* X, int ne, const int* elems, double density);
/* Deformable::AddPdEnergy (deformable_pd_energy.cpp:42-153): type "corotated",
 * "volume" (params 1 or ne) or "pd_neohookean" (params 2 or 2*ne); idx = element
 * indices (ni = 0: all elements). */
int phd_add_pd_energy(phd_ctx ctx, const char* type, const double* params, int np, const int* idx, int ni);
/* Deformable::SetDirichletBoundaryCondition (deformable.h:70-73). */
int phd_set_dirichlet(phd_ctx ctx, int dof, double val);
/* Deformable::AddStateForce (deformable_state_force.cpp:10-31): "gravity" [gx, gy, gz]
 * or "planar_contact" [nx, ny, nz, offset, p, kn, kf, mu] (a penalty floor with
 * friction, planar_contact_state_force.cpp), or "arc_contact" [center (3), dir (3),
 * start (3), radius, angle, p, kn, kf, mu] (the inside of a cylindrical arc,
 * deformable_state_force.cpp:33-49, arc_contact_state_force.cpp), or
 * "mesh_contact" [radius, kn_left, kn_right, kf, mu, #faces A, #faces B, faces A,
 * faces B] (two surfaces of the body, deformable_state_force.cpp:86-112,
 * mesh_contact_state_force.cpp; forward only, its backward is zero), or
 * "billiard_ball" [radius, vertices per ball, stiffness per ball, friction per
 * ball] (deformable_state_force.cpp:72-85, billiard_ball_state_force.cpp), or
 * "hydrodynamics" [rho, v_water (3), Cd points (8), Ct points (8), max_force,
 * faces (4 per quad on hex bodies, 3 per triangle on tet bodies)]
 * (deformable_state_force.cpp:50-71, hydrodynamics_state_force.cpp). */
int phd_add_state_force(phd_ctx ctx, const char* type, const double* params, int np);
/* Deformable::NumOfStateForceParameters(): the length of dl_dstate_p (3 per
 * state force, in AddStateForce order). */
int phd_num_state_force_parameters(phd_ctx ctx);
/* Deformable::SetFrictionalBoundary (deformable_friction.cpp:7-89): "planar"
 * [nx, ny, nz, offset], "spherical" [cx, cy, cz, r], and the two boundaries with
 * no reference counterpart used by the press scene: "heightfield"
 * [z0, nw, (amp, kx, ky, phase) * nw] and "press" (heightfield params + plane
 * height; the plane's normal is -z). */
int phd_set_frictional_boundary(phd_ctx ctx, const char* type, const double* params, int np, const int* idx, int ni);
/* Moves the press plane (prescribed motion between frames). No reference counterpart. */
int phd_set_press_plane(phd_ctx ctx, double z);
/* Deformable::AddActuation (deformable_actuation.cpp:5-26): one PdMuscleEnergy
 * group (pd_muscle_energy.cpp:5-24) of stiffness, fiber direction fiber[3]
 * (normalized here) and element indices idx[ni]. Each element may belong to one
 * group only (the reference lets groups overlap). */
int phd_add_actuation(phd_ctx ctx, double stiffness, const double* fiber, const int* idx, int ni);
/* The actuation levels `a` of Deformable::Forward / Backward
 * (deformable.h:116-125) for the following phd_step_forward / phd_step_backward
 * calls: na == phd_act_dofs(ctx), one level per group element in group order. */
int phd_set_actuation(phd_ctx ctx, const double* a, int na);
/* Deformable::act_dofs() and NumOfPdMuscleEnergies() (deformable.h:47, 63-65). */
int phd_act_dofs(phd_ctx ctx);
int phd_num_pd_muscle_energies(phd_ctx ctx);
/* dl_da (na = act_dofs) and dl_dact_w (nm = muscle groups) of the last
 * phd_step_backward call: the outputs of Deformable::Backward that
 * deformable_projective_dynamics_backward_gpu.cpp:1209-1228 fills. */
int phd_actuation_gradients(phd_ctx ctx, double* dl_da, int na, double* dl_dact_w, int nm);

/* SetupProjectiveDynamicsSolver (deformable_projective_dynamics_forward.cpp:8-183)
 * plus the sparse-inverse factor build of forward_gpu.cpp:584-764: assembles A,
 * orders it by nested dissection, factors it in FP64 on the host and expands
 * S = L^-1 on the device (FP64 and FP32 copies). Cached until the scene changes. */
int phd_factor_build(phd_ctx ctx, double dt, int threads);

/* x = A^-1 b = S^T (S b) on the three axes; supersedes AlgPhd_GpuApplyAinv
 * (forward_gpu.cpp:339-361). precision: 0 FP64 factor, 1 FP32 factor;
 * refine: number of FP64 iterative-refinement steps. */
int phd_apply_ainv(phd_ctx ctx, double dt, const double* b, double* x, int precision, int refine);

/* phd_apply_ainv on DEVICE vectors (b_dev, x_dev: 3N doubles each, AoS, on the
 * context's device, not aliased) that allocates nothing: every intermediate
 * (permuted input, S b, the row-chunk and column-tile partials, the FP64
 * residual) lives in one fixed device array of the library, at most 64 MiB,
 * which each GPU holds until the process ends. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * stream capture, tight serving loops). The factor must already be built
 * (phd_factor_build); a factor whose intermediates exceed 64 MiB is rejected
 * with an error. Work is enqueued on the context's stream and the call returns
 * after that stream has finished. Calls that overlap on one GPU (two contexts
 * or two threads on the same device) share the array and race. */
int phd_apply_ainv_noalloc(phd_ctx ctx, const double* b_dev, double* x_dev, int precision, int refine);

/* One PD step, ForwardProjectiveDynamicsAlgPhd (forward_gpu.cpp:387-1594),
 * options as the reference's std::map<string, real> (keys/vals, n entries;
 * required: max_pd_iter, abs_tol, rel_tol, verbose, thread_ct). active: caller
 * buffer of capacity *n_active on entry, at least the number of frictional-
 * boundary vertices (checked before the step); on exit *n_active = number of
 * contact nodes with lambda > 0 written to active. */
int phd_step_forward(phd_ctx ctx, const double* q, const double* v, const double* f_ext, double dt,
                     const char* const* keys, const double* vals, int n, double* q_next, double* v_next, int* active,
                     int* n_active);

/* Adjoint of one PD step, BackwardProjectiveDynamicsAlgPhd
 * (deformable_projective_dynamics_backward_gpu.cpp:648-1250). Options beyond the
 * reference's: alg_phd_bwd_kpd_filter (0 none = the reference's K_pd, 1 clamp,
 * 2 abs: the eigenvalue filters of deformable_newton_forward.cpp:393-433 applied
 * to each sample's proximal-map block; 3 trust region: the unfiltered
 * operator first, then, if the backbone fixed point diverges, the clamp and
 * then the abs filter with the block's spectrum relative to the sample's weight
 * in A bounded below by -rho, which makes the backbone a contraction),
 * alg_phd_bwd_trust_radius (rho, default 0.9), alg_phd_bwd_state_woodbury
 * (1: dl_dq, dl_dv, dl_df_ext from the
 * Woodbury-corrected adjoint, the exact adjoint with the active contacts held;
 * default 0 keeps the reference's routing). Outputs (3N each):
 * dl_dq, dl_dv, dl_df_ext; dl_dmat_w (one per PD element energy); dl_dstate_p
 * (3 when gravity is set); dl_dw_elem (energies x elements, may be NULL): the
 * per-element stiffness gradient whose sum over an energy's elements is dl_dmat_w. */
int phd_step_backward(phd_ctx ctx, const double* q, const double* v, const double* f_ext, double dt,
                      const double* q_next, const double* v_next, const int* active, int n_active,
                      const double* dl_dq_next, const double* dl_dv_next, const char* const* keys, const double* vals,
                      int n, double* dl_dq, double* dl_dv, double* dl_df_ext, double* dl_dmat_w, double* dl_dstate_p,
                      double* dl_dw_elem);

/* Sizes. */
/* ForwardProjectiveDynamics with a pd_eigen method other than pd_eigen_alg_phd
 * (deformable_projective_dynamics_forward.cpp:294-664): the DiffPD baseline
 * (L-BFGS PD, PdNonlinearSolve, inside the active-set contact loop; method
 * "..._fixed_contact" keeps active_in frozen). active_in holds the previous
 * frame's active contact set (the reference's in/out argument), active_out this
 * frame's; stats (3 ints, may be NULL): contact rounds, PD iterations, line-search
 * halvings. Host arrays; FP64; corotated, volume and muscle energies. */
int phd_step_forward_baseline(phd_ctx ctx, const char* method, const double* q, const double* v, const double* f_ext,
                              double dt, const char* const* keys, const double* vals, int n, double* q_next,
                              double* v_next, const int* active_in, int n_in, int* active_out, int* n_active,
                              int* stats);
/* BackwardProjectiveDynamics for a pd_eigen baseline method
 * (deformable_projective_dynamics_backward.cpp:141-679, the path of bodies
 * without an elastic material): L-BFGS on the adjoint system over the free DOFs,
 * then the frozen nodes' and Dirichlet DOFs' shares; dl_da / dl_dact_w through
 * phd_actuation_gradients. active: the frame's active set (the forward's
 * output). Energies of uniform stiffness; the reference's Newton-Cholesky
 * fallback is not built (an indefinite operator or the iteration cap is an
 * error). stats (1 int, may be NULL): L-BFGS iterations. */
int phd_step_backward_baseline(phd_ctx ctx, const double* q, const double* v, const double* f_ext, double dt,
                               const double* q_next, const int* active, int n_active, const double* dl_dq_next,
                               const double* dl_dv_next, const char* const* keys, const double* vals, int n,
                               double* dl_dq, double* dl_dv, double* dl_df_ext, double* dl_dmat_w, double* dl_dstate_p,
                               int* stats);
/* phd_step_forward / phd_step_backward with the state vectors (q, v, f_ext,
 * q_next, v_next, dl_dq_next, dl_dv_next, dl_dq, dl_dv, dl_df_ext) in device
 * memory of the body's GPU, e.g. the data pointers of torch tensors handed over
 * by DLPack; active, dl_dmat_w, dl_dstate_p and dl_dw_elem stay host arrays.
 * They serve the zero-copy torch wrapper that the reference's sim.py:9-85
 * lacked. Each call copies the inputs to the host once (the step's host logic
 * reads them) and writes the outputs straight into the caller's buffers; a
 * pointer that is not device memory of this GPU is an error. */
int phd_step_forward_device(phd_ctx ctx, const double* q, const double* v, const double* f_ext, double dt,
                            const char* const* keys, const double* vals, int n, double* q_next, double* v_next,
                            int* active, int* n_active);
int phd_step_backward_device(phd_ctx ctx, const double* q, const double* v, const double* f_ext, double dt,
                             const double* q_next, const double* v_next, const int* active, int n_active,
                             const double* dl_dq_next, const double* dl_dv_next, const char* const* keys,
                             const double* vals, int n, double* dl_dq, double* dl_dv, double* dl_df_ext,
                             double* dl_dmat_w, double* dl_dstate_p, double* dl_dw_elem);
int phd_dofs(phd_ctx ctx);
int phd_num_elements(phd_ctx ctx);
int phd_num_pd_element_energies(phd_ctx ctx);

/* Statistics (alg_phd_gpu_memory.h:9-27, per context): factor bytes on the
 * device, nnz of S, cached N, per-frame PD iterations and convergence flags. */
long long phd_s_memory_bytes(phd_ctx ctx);
long long phd_s_nnz(phd_ctx ctx);
int phd_cached_n(phd_ctx ctx);
int phd_frame_stats(phd_ctx ctx, int* iters, int* converged, double* seconds, int cap);
void phd_reset_frame_stats(phd_ctx ctx);
/* Drops the previous frame's contact multipliers, which warm-start the next
 * frame's NCP iteration (the reference keeps them in its GPU state,
 * alg_phd_gpu_memory.h, so a new Deformable starts without them): rollouts
 * from different initial states on one body then start alike. */
int phd_reset_contact_history(phd_ctx ctx);
int phd_last_backward_iters(phd_ctx ctx);
/* Last backward step (no reference counterpart; the reference prints these):
 * [0] backbone fixed-point iterations, [1] backbone converged (0/1), [2] backbone
 * clipped, i.e. |db| reached 0.999 x alg_phd_bwd_clip_norm so the cascade, the
 * Woodbury correction and dl_dmat_w were skipped (backward_gpu.cpp:809-866),
 * [3] Woodbury fixed-point iterations, [4] Woodbury converged, [5] basis rows,
 * [6] samples whose proximal-map block the K_pd filter changed, [7] samples,
 * [8] |db|, [9] trust-region stage that produced db (0 unfiltered, 1 clamp,
 * 2 abs), [10] 1 when the contact data came from the forward tape, 0 when the
 * frame was replayed over the given active set (backward_gpu.cpp:113-418),
 * [11] 1 when the mixed path's FP32 factor of the basis Gram matrix broke down
 * and the FP64 one was used. Returns the number of entries (12). [0] counts
 * every stage. */
int phd_last_backward_stats(phd_ctx ctx, double* out, int cap);
/* Factor description: out[0] nnz(S), [1] padded value count, [2] max etree depth,
 * [3] ordering seconds, [4] factor seconds, [5] setup seconds, [6] row units,
 * [7] column units, [8] nnz(L), [9] contact candidates K, [10] entries of the
 * candidate paths of S, [11] rows crossed by those paths, [12] column-chunk
 * width of S w, [13] unit height of S^T y, [14] local-step block size (up to
 * `cap` values). */
int phd_factor_info(phd_ctx ctx, double* out, int cap);
/* Autotuner of the launch shapes (no counterpart in the reference, whose
 * cuSPARSE products have none to tune): builds the factor for dt if needed,
 * times `reps` launches of each candidate on this body and keeps the fastest
 * of the column-chunk width of S w (512, 1024, 2048; FP32 product), the unit
 * height of S^T y (256, 512, 1024, 2048 rows) and the local step's block size
 * (64, 128, 256). Results are independent of the choice up to the summation
 * order of the partial sums. out: [0] chunk width, [1] unit height, [2] block,
 * [3..5] ms per S w, [6..9] ms per S^T y, [10..12] ms per local step, one per
 * candidate (up to `cap` values; out may be NULL). */
int phd_autotune(phd_ctx ctx, double dt, int threads, int reps, double* out, int cap);
/* Sets the three launch shapes directly (values from the candidate lists
 * above), e.g. to reuse a tuned choice; the factor must be built. */
int phd_set_launch_shapes(phd_ctx ctx, int chunk_w, int unit_items, int local_block);
/* Device-event timing of the S (rows) and S^T (columns) kernels: enable = 1 starts
 * recording; phd_timing_read writes [rows_ms, rows_launches, cols_ms, cols_launches]
 * accumulated since the last reset (reset = 1 clears after reading). */
int phd_timing_enable(phd_ctx ctx, int enable);
int phd_timing_read(phd_ctx ctx, double* out4, int reset);

/* Host wall-clock split of forward/backward (seconds, accumulated while timing
 * is enabled at level 2, which synchronizes the stream at phase boundaries).
 * Order: fwd total, fwd local step, fwd global solve, fwd contact block, fwd
 * dense contact solve, fwd Anderson, fwd gate, bwd total, bwd K_pd build, bwd
 * backbone, bwd basis factor, bwd Woodbury fixed point, bwd basis solves, bwd
 * material gradient; counters: dense solves, dense rows summed, FP32 dense
 * failures, basis rows summed. Returns the number of entries. */
int phd_phase_times(phd_ctx ctx, double* out, int cap, int reset);

/* Contact multipliers of the last forward frame, per candidate (sorted by vertex):
 * normal lambda and tangential lt1, lt2. Returns the candidate count. */
int phd_last_contact_multipliers(phd_ctx ctx, double* lambda, double* lt1, double* lt2, int cap);

/* Dense SPD solve A x = b (A: n x n column-major, host) through the blocked
 * Cholesky used for the contact system; precision 0 = FP64, 1 = FP32.
 * Returns 0, 1 when a pivot was not positive, -1 on error (phd_last_error(NULL)).
 * Test entry point; no counterpart in the reference. */
int phd_dense_spd_solve(int n, const double* A, const double* b, double* x, int precision);

/* ---- Learned warm start of the PD iteration (no counterpart in the reference:
 * its initial iterate is q_t, deformable_projective_dynamics_forward_gpu.cpp:1069).
 * A per-vertex MLP 13 -> 128 -> 128 -> 128 -> 3 (GELU, BF16 weights on MFMA,
 * FP32 accumulate) predicts (q_next - q - h v) / tscale from standardized
 * features; forward option phd_warm_start = 1 starts the PD loop from the
 * prediction when its PD residual is below that of q + h v. */
int phd_mlp_num_params(void);
int phd_mlp_init(phd_ctx ctx, int seed);
/* feature means / scales (13 each) and the output scale */
int phd_mlp_set_normalization(phd_ctx ctx, const float* fmean, const float* fstd, float tscale);
int phd_mlp_get_params(phd_ctx ctx, float* out);       /* phd_mlp_num_params() floats */
int phd_mlp_set_params(phd_ctx ctx, const float* params);
/* feat: n x 16 standardized features (columns 13..15 zero); out: n x 3 */
int phd_mlp_forward(phd_ctx ctx, int n, const float* feat, float* out);
/* full-batch MSE mean((out - target)^2) over n x 3 and its parameter gradient */
int phd_mlp_loss_grad(phd_ctx ctx, int n, const float* feat, const float* target, float* grad, double* loss);
/* `steps` Adam steps (lr, beta 0.9 / 0.999) on minibatches of a seeded permutation; losses: steps floats or NULL */
int phd_mlp_train(phd_ctx ctx, int n, const float* feat, const float* target, int steps, int batch, float lr, int seed,
                  float* losses);
/* raw (unstandardized) features of state (q, v, f_ext may be NULL) for this body: n x 16 */
int phd_mlp_features(phd_ctx ctx, const double* q, const double* v, const double* f_ext, double dt, float* feat);
/* forward frames that started from the network / fell back to q + h v */
int phd_warm_start_stats(phd_ctx ctx, int* used, int* fallback, int reset);
/* D = A B for 16 x 16 row-major inputs rounded to BF16, through one MFMA (layout probe) */
int phd_mlp_mfma_probe(const float* A, const float* B, float* D);

/* Streams `nbytes` through a plain device copy kernel `reps` times and writes the
 * average duration in ms (calibrates the HBM byte counters and the bandwidth peak). */
int phd_copy_probe(phd_ctx ctx, long long nbytes, int reps, double* ms_out);

/* Host-only self-check of the factor construction (no device needed; small n):
 * builds the factor of the symmetric CSR matrix (full pattern) with vertex
 * coordinates X and writes [max|S^T S A - I|, max S-product plan error,
 * max S^T-product plan error, nnz(S), max depth, nnz(L), structure violations]. */
int phd_host_factor_check(int n, const int* Ap, const int* Ai, const double* Ax, const double* X, double* out7);

#ifdef __cplusplus
}
#endif
#endif /* PHD_H */
