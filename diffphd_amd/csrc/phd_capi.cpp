// C ABI (include/phd.h) over phd::Body. Exceptions stop here and become
// return codes plus phd_last_error().
#include "../../include/phd.h"

#include <map>
#include <stdexcept>
#include <string>

#include "phd_solver.h"
#include "phd_dense.h"
#include "phd_mlp.h"

namespace {
struct Ctx {
    std::unique_ptr<phd::Body> body;
    std::string err;
};
thread_local std::string g_create_error;

template <class F>
int guard(phd_ctx c, F&& f) {
    Ctx* x = static_cast<Ctx*>(c);
    if (!x) return 2;
    try {
        f(*x->body);
        x->err.clear();
        return 0;
    } catch (const std::exception& e) {
        x->err = e.what();
        return 1;
    } catch (...) {
        x->err = "Unknown error.";
        return 1;
    }
}
// Copies of caller vectors in device memory of the body's GPU (the
// device-pointer entry points): the step's host-side logic (tape, candidate
// sets, restitution) reads its inputs on the host.
std::vector<double> from_device(const double* p, size_t n) {
    std::vector<double> h(n);
    if (n) phd::hip_check(hipMemcpy(h.data(), p, n * sizeof(double), hipMemcpyDeviceToHost), "device input");
    return h;
}
void to_device(double* p, const std::vector<double>& h) {
    if (!h.empty()) phd::hip_check(hipMemcpy(p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice), "device output");
}
void check_device_ptr(const phd::Body& b, const void* p, const char* what) {
    hipPointerAttribute_t attr;
    const hipError_t e = hipPointerGetAttributes(&attr, p);
    const bool dev = e == hipSuccess && attr.type == hipMemoryTypeDevice && attr.device == b.device();
    if (e != hipSuccess) (void)hipGetLastError();
    if (!dev)
        throw std::runtime_error(std::string(what) + " is not device memory of the body's GPU (" + std::to_string(b.device()) +
                                 ")");
}
std::map<std::string, double> opts(const char* const* keys, const double* vals, int n) {
    std::map<std::string, double> o;
    for (int i = 0; i < n; ++i) o[keys[i]] = vals[i];
    return o;
}
}  // namespace

extern "C" {

phd_ctx phd_ctx_create(int device) {
    try {
        Ctx* c = new Ctx();
        c->body.reset(new phd::Body(device));
        return c;
    } catch (const std::exception& e) {
        g_create_error = e.what();
        return nullptr;
    }
}
void phd_ctx_destroy(phd_ctx ctx) { delete static_cast<Ctx*>(ctx); }
const char* phd_last_error(phd_ctx ctx) {
    if (!ctx) return g_create_error.c_str();
    return static_cast<Ctx*>(ctx)->err.c_str();
}

int phd_init_mesh(phd_ctx ctx, int edim, int nv, const double* X, int ne, const int* elems, double density) {
    return guard(ctx, [&](phd::Body& b) { b.init_mesh(edim, nv, X, ne, elems, density); });
}
int phd_add_pd_energy(phd_ctx ctx, const char* type, const double* params, int np, const int* idx, int ni) {
    return guard(ctx, [&](phd::Body& b) {
        b.add_pd_energy(type, std::vector<double>(params, params + np), std::vector<int>(idx, idx + ni));
    });
}
int phd_set_dirichlet(phd_ctx ctx, int dof, double val) {
    return guard(ctx, [&](phd::Body& b) { b.set_dirichlet(dof, val); });
}
int phd_add_state_force(phd_ctx ctx, const char* type, const double* params, int np) {
    return guard(ctx, [&](phd::Body& b) {
        const std::string t(type);
        if (t == "gravity") {
            if (np != 3) throw std::runtime_error("Inconsistent params for GravitionalStateForce.");
            b.add_gravity(params);
        } else if (t == "planar_contact") {
            b.add_planar_contact(std::vector<double>(params, params + np));
        } else if (t == "arc_contact") {
            b.add_arc_contact(std::vector<double>(params, params + np));
        } else if (t == "mesh_contact") {
            b.add_mesh_contact(std::vector<double>(params, params + np));
        } else if (t == "billiard_ball") {
            b.add_billiard_ball(std::vector<double>(params, params + np));
        } else if (t == "hydrodynamics") {
            b.add_hydrodynamics(std::vector<double>(params, params + np));
        } else {
            throw std::runtime_error(std::string("Unsupported state force type: ") + type);
        }
    });
}
int phd_num_state_force_parameters(phd_ctx ctx) { return ctx ? static_cast<Ctx*>(ctx)->body->num_state_params() : -1; }
int phd_set_frictional_boundary(phd_ctx ctx, const char* type, const double* params, int np, const int* idx, int ni) {
    return guard(ctx, [&](phd::Body& b) {
        b.set_boundary(type, std::vector<double>(params, params + np), std::vector<int>(idx, idx + ni));
    });
}
int phd_set_press_plane(phd_ctx ctx, double z) {
    return guard(ctx, [&](phd::Body& b) { b.set_press_plane(z); });
}
int phd_factor_build(phd_ctx ctx, double dt, int threads) {
    return guard(ctx, [&](phd::Body& b) { b.setup(dt, threads); });
}
int phd_apply_ainv(phd_ctx ctx, double dt, const double* bvec, double* x, int precision, int refine) {
    return guard(ctx, [&](phd::Body& b) { b.apply_ainv(dt, bvec, x, precision, refine); });
}
int phd_apply_ainv_noalloc(phd_ctx ctx, const double* b_dev, double* x_dev, int precision, int refine) {
    return guard(ctx, [&](phd::Body& b) { b.apply_ainv_noalloc(b_dev, x_dev, precision, refine); });
}
int phd_step_forward(phd_ctx ctx, const double* q, const double* v, const double* f_ext, double dt,
                     const char* const* keys, const double* vals, int n, double* q_next, double* v_next, int* active,
                     int* n_active) {
    return guard(ctx, [&](phd::Body& b) {
        // checked before the step: a rejected call leaves the tape and the
        // lambda warm start as they were
        if (!n_active || *n_active < b.num_boundary_vertices())
            throw std::runtime_error("active_contact_idx buffer too small: it needs room for every frictional-boundary vertex (" +
                                     std::to_string(b.num_boundary_vertices()) + ")");
        std::vector<int> act;
        b.forward(q, v, f_ext, dt, opts(keys, vals, n), q_next, v_next, act);
        if ((int)act.size() > *n_active) throw std::runtime_error("active_contact_idx buffer too small");
        for (size_t i = 0; i < act.size(); ++i) active[i] = act[i];
        *n_active = (int)act.size();
    });
}
int phd_step_backward(phd_ctx ctx, const double* q, const double* v, const double* f_ext, double dt,
                      const double* q_next, const double* v_next, const int* active, int n_active,
                      const double* dl_dq_next, const double* dl_dv_next, const char* const* keys, const double* vals,
                      int n, double* dl_dq, double* dl_dv, double* dl_df_ext, double* dl_dmat_w, double* dl_dstate_p,
                      double* dl_dw_elem) {
    return guard(ctx, [&](phd::Body& b) {
        b.backward(q, v, f_ext, dt, q_next, v_next, std::vector<int>(active, active + n_active), dl_dq_next, dl_dv_next,
                   opts(keys, vals, n), dl_dq, dl_dv, dl_df_ext, dl_dmat_w, dl_dstate_p, dl_dw_elem);
    });
}
int phd_step_forward_baseline(phd_ctx ctx, const char* method, const double* q, const double* v, const double* f_ext,
                              double dt, const char* const* keys, const double* vals, int n, double* q_next,
                              double* v_next, const int* active_in, int n_in, int* active_out, int* n_active,
                              int* stats) {
    return guard(ctx, [&](phd::Body& b) {
        if (!n_active || *n_active < b.num_boundary_vertices())
            throw std::runtime_error("active_contact_idx buffer too small: it needs room for every frictional-boundary vertex (" +
                                     std::to_string(b.num_boundary_vertices()) + ")");
        std::vector<int> act(active_in, active_in + n_in);
        b.forward_baseline(method, q, v, f_ext, dt, opts(keys, vals, n), q_next, v_next, act);
        for (size_t i = 0; i < act.size(); ++i) active_out[i] = act[i];
        *n_active = (int)act.size();
        if (stats) {
            stats[0] = b.last_baseline.contact_iters;
            stats[1] = b.last_baseline.pd_iters;
            stats[2] = b.last_baseline.ls_steps;
        }
    });
}
int phd_step_backward_baseline(phd_ctx ctx, const double* q, const double* v, const double* f_ext, double dt,
                               const double* q_next, const int* active, int n_active, const double* dl_dq_next,
                               const double* dl_dv_next, const char* const* keys, const double* vals, int n,
                               double* dl_dq, double* dl_dv, double* dl_df_ext, double* dl_dmat_w, double* dl_dstate_p,
                               int* stats) {
    return guard(ctx, [&](phd::Body& b) {
        b.backward_baseline(q, v, f_ext, dt, q_next, std::vector<int>(active, active + n_active), dl_dq_next, dl_dv_next,
                            opts(keys, vals, n), dl_dq, dl_dv, dl_df_ext, dl_dmat_w, dl_dstate_p);
        if (stats) stats[0] = b.last_baseline.bwd_iters;
    });
}
int phd_step_forward_device(phd_ctx ctx, const double* q, const double* v, const double* f_ext, double dt,
                            const char* const* keys, const double* vals, int n, double* q_next, double* v_next,
                            int* active, int* n_active) {
    return guard(ctx, [&](phd::Body& b) {
        const size_t n3 = (size_t)b.dofs();
        for (auto pv : {std::make_pair((const void*)q, "q"), std::make_pair((const void*)v, "v"),
                        std::make_pair((const void*)f_ext, "f_ext"), std::make_pair((const void*)q_next, "q_next"),
                        std::make_pair((const void*)v_next, "v_next")})
            check_device_ptr(b, pv.first, pv.second);
        if (!n_active || *n_active < b.num_boundary_vertices())
            throw std::runtime_error("active_contact_idx buffer too small: it needs room for every frictional-boundary vertex (" +
                                     std::to_string(b.num_boundary_vertices()) + ")");
        const auto hq = from_device(q, n3), hv = from_device(v, n3), hf = from_device(f_ext, n3);
        std::vector<double> qn(n3), vn(n3);
        std::vector<int> act;
        b.forward(hq.data(), hv.data(), hf.data(), dt, opts(keys, vals, n), qn.data(), vn.data(), act);
        to_device(q_next, qn);
        to_device(v_next, vn);
        for (size_t i = 0; i < act.size(); ++i) active[i] = act[i];
        *n_active = (int)act.size();
    });
}
int phd_step_backward_device(phd_ctx ctx, const double* q, const double* v, const double* f_ext, double dt,
                             const double* q_next, const double* v_next, const int* active, int n_active,
                             const double* dl_dq_next, const double* dl_dv_next, const char* const* keys,
                             const double* vals, int n, double* dl_dq, double* dl_dv, double* dl_df_ext,
                             double* dl_dmat_w, double* dl_dstate_p, double* dl_dw_elem) {
    return guard(ctx, [&](phd::Body& b) {
        const size_t n3 = (size_t)b.dofs();
        const void* dev[] = {q, v, f_ext, q_next, v_next, dl_dq_next, dl_dv_next, dl_dq, dl_dv, dl_df_ext};
        const char* names[] = {"q", "v", "f_ext", "q_next", "v_next", "dl_dq_next", "dl_dv_next", "dl_dq", "dl_dv", "dl_df_ext"};
        for (int i = 0; i < 10; ++i) check_device_ptr(b, dev[i], names[i]);
        const auto hq = from_device(q, n3), hv = from_device(v, n3), hf = from_device(f_ext, n3);
        const auto hqn = from_device(q_next, n3), hvn = from_device(v_next, n3);
        const auto hdq = from_device(dl_dq_next, n3), hdv = from_device(dl_dv_next, n3);
        std::vector<double> oq(n3), ov(n3), of(n3);
        b.backward(hq.data(), hv.data(), hf.data(), dt, hqn.data(), hvn.data(), std::vector<int>(active, active + n_active),
                   hdq.data(), hdv.data(), opts(keys, vals, n), oq.data(), ov.data(), of.data(), dl_dmat_w, dl_dstate_p,
                   dl_dw_elem);
        to_device(dl_dq, oq);
        to_device(dl_dv, ov);
        to_device(dl_df_ext, of);
    });
}
int phd_dofs(phd_ctx ctx) { return ctx ? static_cast<Ctx*>(ctx)->body->dofs() : -1; }
int phd_num_elements(phd_ctx ctx) { return ctx ? static_cast<Ctx*>(ctx)->body->num_elements() : -1; }
int phd_num_pd_element_energies(phd_ctx ctx) { return ctx ? static_cast<Ctx*>(ctx)->body->num_energies() : -1; }
int phd_add_actuation(phd_ctx ctx, double stiffness, const double* fiber, const int* idx, int ni) {
    return guard(ctx, [&](phd::Body& b) { b.add_actuation(stiffness, fiber, std::vector<int>(idx, idx + ni)); });
}
int phd_set_actuation(phd_ctx ctx, const double* a, int na) {
    return guard(ctx, [&](phd::Body& b) { b.set_actuation(a, na); });
}
int phd_act_dofs(phd_ctx ctx) { return ctx ? static_cast<Ctx*>(ctx)->body->act_dofs() : -1; }
int phd_num_pd_muscle_energies(phd_ctx ctx) { return ctx ? static_cast<Ctx*>(ctx)->body->num_muscles() : -1; }
int phd_actuation_gradients(phd_ctx ctx, double* dl_da, int na, double* dl_dact_w, int nm) {
    return guard(ctx, [&](phd::Body& b) {
        const auto& ga = b.last_dl_da();
        const auto& gw = b.last_dl_dact_w();
        if (na != (int)ga.size() || nm != (int)gw.size())
            throw std::runtime_error("phd_actuation_gradients: sizes differ from the last backward call (" +
                                     std::to_string(ga.size()) + " levels, " + std::to_string(gw.size()) + " groups)");
        std::copy(ga.begin(), ga.end(), dl_da);
        std::copy(gw.begin(), gw.end(), dl_dact_w);
    });
}
long long phd_s_memory_bytes(phd_ctx ctx) { return ctx ? static_cast<Ctx*>(ctx)->body->factor_bytes() : 0; }
long long phd_s_nnz(phd_ctx ctx) { return ctx ? static_cast<Ctx*>(ctx)->body->factor_nnz() : 0; }
int phd_cached_n(phd_ctx ctx) { return ctx ? static_cast<Ctx*>(ctx)->body->cached_n() : -1; }
int phd_frame_stats(phd_ctx ctx, int* iters, int* converged, double* seconds, int cap) {
    if (!ctx) return -1;
    const auto& st = static_cast<Ctx*>(ctx)->body->frame_stats;
    const int n = (int)st.size();
    for (int i = 0; i < n && i < cap; ++i) {
        if (iters) iters[i] = st[i].iters;
        if (converged) converged[i] = st[i].converged;
        if (seconds) seconds[i] = st[i].seconds;
    }
    return n;
}
void phd_reset_frame_stats(phd_ctx ctx) {
    if (ctx) static_cast<Ctx*>(ctx)->body->frame_stats.clear();
}
int phd_reset_contact_history(phd_ctx ctx) {
    return guard(ctx, [&](phd::Body& b) { b.reset_contact_history(); });
}
int phd_last_backward_iters(phd_ctx ctx) { return ctx ? static_cast<Ctx*>(ctx)->body->last_backward_iters : -1; }
int phd_last_backward_stats(phd_ctx ctx, double* out, int cap) {
    if (!ctx) return -1;
    const auto& s = static_cast<Ctx*>(ctx)->body->last_bwd;
    const double v[12] = {s.backbone_iters, s.backbone_converged, s.clipped, s.woodbury_iters, s.woodbury_converged,
                          s.basis_rows, s.filtered, s.samples, s.db_norm, s.trust_stage, s.from_tape, s.gram_fp64_fallbacks};
    for (int i = 0; i < 12 && i < cap; ++i) out[i] = v[i];
    return 12;
}
int phd_factor_info(phd_ctx ctx, double* out, int cap) {
    return guard(ctx, [&](phd::Body& b) {
        const phd::Factor& f = b.factor();
        const double v[15] = {(double)f.nnz, (double)f.v_len, (double)f.max_depth, f.order_seconds, f.factor_seconds,
                              b.setup_seconds, (double)b.plan().gu_group.size(), (double)b.plan().cu_tile.size(),
                              (double)f.Lx.size(), (double)b.contact_candidates(), (double)b.contact_entries(),
                              (double)b.contact_rows(), (double)b.chunk_w(), (double)b.unit_items(),
                              (double)b.local_block()};
        for (int i = 0; i < cap && i < 15; ++i) out[i] = v[i];
    });
}
int phd_set_launch_shapes(phd_ctx ctx, int chunk_w, int unit_items, int local_block) {
    return guard(ctx, [&](phd::Body& b) { b.set_launch_shapes(chunk_w, unit_items, local_block); });
}
int phd_autotune(phd_ctx ctx, double dt, int threads, int reps, double* out, int cap) {
    return guard(ctx, [&](phd::Body& b) {
        const phd::Body::TuneResult r = b.autotune(dt, threads, reps);
        const double v[13] = {(double)r.chunk_w, (double)r.unit_items, (double)r.local_block,
                              r.rows_ms[0], r.rows_ms[1], r.rows_ms[2], r.cols_ms[0], r.cols_ms[1], r.cols_ms[2],
                              r.cols_ms[3], r.local_ms[0], r.local_ms[1], r.local_ms[2]};
        for (int i = 0; out && i < cap && i < 13; ++i) out[i] = v[i];
    });
}
int phd_timing_enable(phd_ctx ctx, int enable) {
    return guard(ctx, [&](phd::Body& b) {
        b.timing = enable != 0;
        b.timing_level = enable;
    });
}
int phd_timing_read(phd_ctx ctx, double* out4, int reset) {
    return guard(ctx, [&](phd::Body& b) {
        b.collect_timing();
        out4[0] = b.t_spmv_rows_ms;
        out4[1] = (double)b.n_spmv_rows;
        out4[2] = b.t_spmv_cols_ms;
        out4[3] = (double)b.n_spmv_cols;
        if (reset) { b.t_spmv_rows_ms = b.t_spmv_cols_ms = 0; b.n_spmv_rows = b.n_spmv_cols = 0; }
    });
}

int phd_phase_times(phd_ctx ctx, double* out, int cap, int reset) {
    if (!ctx) return -1;
    phd::Body& b = *static_cast<Ctx*>(ctx)->body;
    const int n = (int)phd::Body::kNumPhases;
    for (int i = 0; i < n && i < cap; ++i) out[i] = b.phase[i];
    if (reset) for (int i = 0; i < n; ++i) b.phase[i] = 0;
    return n;
}

int phd_last_contact_multipliers(phd_ctx ctx, double* lambda, double* lt1, double* lt2, int cap) {
    if (!ctx) return -1;
    phd::Body& b = *static_cast<Ctx*>(ctx)->body;
    const int n = (int)b.last_lambda_.size();
    for (int i = 0; i < n && i < cap; ++i) { lambda[i] = b.last_lambda_[i]; lt1[i] = b.last_lt1_[i]; lt2[i] = b.last_lt2_[i]; }
    return n;
}

int phd_dense_spd_solve(int n, const double* A, const double* b, double* x, int precision) {
    try {
        if (n <= 0) throw std::runtime_error("n must be positive");
        const int np = phd::chol_padded(n);
        const int64_t lda = phd::chol_lda(n);
        std::vector<double> L((size_t)lda * np, 0.0);
        for (int c = 0; c < np; ++c)
            for (int r = c; r < np; ++r) L[r + (size_t)c * lda] = (r < n && c < n) ? A[r + (size_t)c * n] : (r == c ? 1.0 : 0.0);
        for (int c = 0; c < n; ++c) L[np + (size_t)c * lda] = b[c];
        int fail = 0;
        auto run = [&](auto tag) {
            using TV = decltype(tag);
            std::vector<TV> Lt(L.begin(), L.end()), xt(np);
            phd::DevBuf<TV> dM, dy, dx;
            phd::DevBuf<int> df;
            dM.upload(Lt);
            dy.alloc((size_t)phd::chol_work(n));
            dx.alloc(np);
            df.alloc(1);
            phd::launch_clear_flag(nullptr, df.p);
            phd::launch_chol_solve<TV>(nullptr, n, dM.p, dy.p, dx.p, df.p);
            phd::hip_check(hipDeviceSynchronize(), "dense solve");
            dx.download(xt.data(), np);
            df.download(&fail, 1);
            for (int i = 0; i < n; ++i) x[i] = (double)xt[i];
        };
        if (precision == 1) run(float{}); else run(double{});
        return fail ? 1 : 0;
    } catch (const std::exception& e) {
        g_create_error = e.what();
        return -1;
    }
}

int phd_mlp_num_params(void) { return phd::MlpLayout::total; }

int phd_mlp_init(phd_ctx ctx, int seed) {
    return guard(ctx, [&](phd::Body& b) { b.net().init((uint64_t)seed, b.stream()); });
}
int phd_mlp_set_normalization(phd_ctx ctx, const float* fmean, const float* fstd, float tscale) {
    return guard(ctx, [&](phd::Body& b) {
        phd::WarmStartNet& n = b.net();
        n.fmean.assign(fmean, fmean + phd::kMlpIn);
        n.fstd.assign(fstd, fstd + phd::kMlpIn);
        for (float s : n.fstd) if (!(s > 0)) throw std::runtime_error("feature scales must be positive");
        n.tscale = tscale;
    });
}
int phd_mlp_get_params(phd_ctx ctx, float* out) {
    return guard(ctx, [&](phd::Body& b) { b.net().get_params(out, b.stream()); });
}
int phd_mlp_set_params(phd_ctx ctx, const float* p) {
    return guard(ctx, [&](phd::Body& b) { b.net().set_params(p, b.stream()); });
}
int phd_mlp_forward(phd_ctx ctx, int n, const float* feat, float* out) {
    return guard(ctx, [&](phd::Body& b) {
        phd::DevBuf<float> f, o;
        f.upload(feat, (size_t)n * phd::kMlpInPad);
        o.alloc((size_t)n * phd::kMlpOut);
        b.net().forward(b.stream(), n, f.p, o.p);
        phd::hip_check(hipStreamSynchronize(b.stream()), "mlp forward");
        o.download(out, (size_t)n * phd::kMlpOut);
    });
}
int phd_mlp_loss_grad(phd_ctx ctx, int n, const float* feat, const float* target, float* grad, double* loss) {
    return guard(ctx, [&](phd::Body& b) {
        phd::DevBuf<float> f, t, g;
        f.upload(feat, (size_t)n * phd::kMlpInPad);
        t.upload(target, (size_t)n * phd::kMlpOut);
        g.alloc(phd::MlpLayout::total);
        *loss = b.net().loss_grad(b.stream(), n, f.p, t.p, g.p);
        g.download(grad, phd::MlpLayout::total);
    });
}
int phd_mlp_train(phd_ctx ctx, int n, const float* feat, const float* target, int steps, int batch, float lr, int seed,
                  float* losses) {
    return guard(ctx, [&](phd::Body& b) {
        phd::DevBuf<float> f, t;
        f.upload(feat, (size_t)n * phd::kMlpInPad);
        t.upload(target, (size_t)n * phd::kMlpOut);
        b.net().train(b.stream(), n, f.p, t.p, steps, batch, lr, (uint64_t)seed, losses);
    });
}
int phd_mlp_features(phd_ctx ctx, const double* q, const double* v, const double* f_ext, double dt, float* feat) {
    return guard(ctx, [&](phd::Body& b) {
        b.setup(dt, 8);
        b.set_inertial_target(q, v, f_ext, dt);
        phd::DevBuf<double> dq, dv;
        dq.upload(q, (size_t)b.dofs());
        dv.upload(v, (size_t)b.dofs());
        b.mlp_features(dq.p, dv.p, dt, false);
        phd::hip_check(hipMemcpyAsync(feat, b.feat_dev(), sizeof(float) * (size_t)(b.dofs() / 3) * phd::kMlpInPad,
                                      hipMemcpyDeviceToHost, b.stream()), "features D2H");
        phd::hip_check(hipStreamSynchronize(b.stream()), "features");
    });
}
int phd_warm_start_stats(phd_ctx ctx, int* used, int* fallback, int reset) {
    return guard(ctx, [&](phd::Body& b) {
        *used = b.warm_used;
        *fallback = b.warm_fallback;
        if (reset) b.warm_used = b.warm_fallback = 0;
    });
}
int phd_mlp_mfma_probe(const float* A, const float* B, float* D) {
    try {
        phd::mlp_mfma_probe(A, B, D);
        return 0;
    } catch (const std::exception& e) {
        g_create_error = e.what();
        return -1;
    }
}

int phd_copy_probe(phd_ctx ctx, long long nbytes, int reps, double* ms_out) {
    return guard(ctx, [&](phd::Body& b) {
        phd::DevBuf<char> x, y;
        const long long n16 = nbytes / 16;
        x.alloc((size_t)n16 * 16);
        y.alloc((size_t)n16 * 16);
        phd::hip_check(hipMemset(x.p, 0, (size_t)n16 * 16), "memset");
        hipEvent_t e0, e1;
        phd::hip_check(hipEventCreate(&e0), "event");
        phd::hip_check(hipEventCreate(&e1), "event");
        phd::hip_check(hipEventRecord(e0, b.stream()), "record");
        for (int r = 0; r < reps; ++r) phd::launch_copy(b.stream(), n16, x.p, y.p);
        phd::hip_check(hipEventRecord(e1, b.stream()), "record");
        phd::hip_check(hipEventSynchronize(e1), "sync");
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        *ms_out = ms / reps;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    });
}

int phd_host_factor_check(int n, const int* Ap, const int* Ai, const double* Ax, const double* X, double* out7) {
    try {
        std::vector<int> ap(Ap, Ap + n + 1), ai(Ai, Ai + Ap[n]);
        std::vector<double> ax(Ax, Ax + Ap[n]), x(X, X + 3 * (size_t)n);
        phd::host_factor_check(n, ap, ai, ax, x, out7);
        return 0;
    } catch (const std::exception& e) {
        g_create_error = e.what();
        return 1;
    }
}

}  // extern "C"
