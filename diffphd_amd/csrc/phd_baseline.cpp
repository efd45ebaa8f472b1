// The DiffPD baseline forward on the device: ForwardProjectiveDynamics with a
// pd_eigen method other than pd_eigen_alg_phd
// (deformable_projective_dynamics_forward.cpp:294-664). PdNonlinearSolve
// minimizes 1/2 (q - rhs)^T (M/h^2)(q - rhs) + E_pd(q) + E_act(q) over the free
// DOFs by L-BFGS whose initial inverse Hessian is PdLhsSolve (A^-1 with the
// active contact nodes frozen), with a backtracking Armijo line search; the
// active-set loop around it freezes contact nodes where they stand at the frame
// start and corrects the set until no frozen node pulls on its obstacle and no
// free node penetrates. Vectors stay on the device; the line-search and BFGS
// scalars, the small frozen block of A^-1 and the contact test run on the host.
#include "phd_solver.h"

#include <algorithm>
#include <cmath>
#include <deque>
#include <memory>
#include <set>

namespace phd {

namespace {
double req(const std::map<std::string, double>& o, const char* k) {
    auto it = o.find(k);
    if (it == o.end()) throw std::runtime_error(std::string("Missing option ") + k + ".");
    return it->second;
}
void require(bool ok, const std::string& msg) {
    if (!ok) throw std::runtime_error(msg);
}
// dense Cholesky (row-major, lower) and the solve with it
bool chol(std::vector<double>& A, int n) {
    for (int j = 0; j < n; ++j) {
        double d = A[j * n + j];
        for (int k = 0; k < j; ++k) d -= A[j * n + k] * A[j * n + k];
        if (!(d > 0)) return false;
        d = std::sqrt(d);
        A[j * n + j] = d;
        for (int i = j + 1; i < n; ++i) {
            double s = A[i * n + j];
            for (int k = 0; k < j; ++k) s -= A[i * n + k] * A[j * n + k];
            A[i * n + j] = s / d;
        }
    }
    return true;
}
void chol_solve(const std::vector<double>& L, int n, double* b) {
    for (int i = 0; i < n; ++i) {
        double s = b[i];
        for (int k = 0; k < i; ++k) s -= L[i * n + k] * b[k];
        b[i] = s / L[i * n + i];
    }
    for (int i = n - 1; i >= 0; --i) {
        double s = b[i];
        for (int k = i + 1; k < n; ++k) s -= L[k * n + i] * b[k];
        b[i] = s / L[i * n + i];
    }
}
// SameSet / ProposeNewSet (common.cpp:173-216)
bool propose_new_set(const std::set<int>& a, const std::set<int>& b, std::set<int>& out) {
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
}  // namespace

// ComputePdEnergy + ActuationEnergy: per-element energies summed on the host in
// element order.
double Body::baseline_energy(const double* q) {
    d_bl_energy_.alloc(ne_);
    if (d_elem_.n < (size_t)ne_ * edim_ * 3) d_elem_.alloc((size_t)ne_ * edim_ * 3);
    launch_pd_energy_force(stream_, mat_, has_cor_ ? d_kc_.p : nullptr, has_vol_ ? d_kv_.p : nullptr, q, d_bl_energy_.p,
                           d_elem_.p);
    hip_check(hipStreamSynchronize(stream_), "baseline energy");
    std::vector<double> e(ne_);
    d_bl_energy_.download(e.data(), ne_);
    double t = 0;
    for (double x : e) t += x;
    return t;
}

// PdEnergyForce + ActuationForce at q into f (3N, device).
void Body::baseline_force(const double* q, double* f) {
    d_bl_energy_.alloc(ne_);
    if (d_elem_.n < (size_t)ne_ * edim_ * 3) d_elem_.alloc((size_t)ne_ * edim_ * 3);
    launch_pd_energy_force(stream_, mat_, has_cor_ ? d_kc_.p : nullptr, has_vol_ ? d_kv_.p : nullptr, q, d_bl_energy_.p,
                           d_elem_.p);
    launch_gather_vertex(stream_, nv_, d_inc_ptr_.p, d_inc_idx_.p, d_elem_.p, f);
}

// PdLhsSolve (forward.cpp:690-936): x_f = A_ff^-1 r_f with the frozen nodes'
// rows and columns out of A, x = r on them. On the shared factor:
// x = A^-1 r + Z mu with Z = A^-1 E (E: the frozen nodes' unit vectors, the same
// on the three axes) and mu = -G^-1 (A^-1 r)_C, G = E^T A^-1 E, so that x_C = 0
// while A x - r stays on the frozen rows; then x_C = r_C.
void Body::pd_lhs_solve(const double* r, double* x, const std::vector<int>& frozen) {
    ainv64(r, x);
    const int nc = (int)frozen.size();
    if (nc == 0) return;
    // A^-1 columns of new frozen nodes, three per global solve (one per axis)
    std::vector<int> fresh;
    for (int v : frozen) if (!zcol_.count(v)) fresh.push_back(v);
    if (!fresh.empty()) {
        const size_t cap = fb_vertices_.size();
        require(zcol_.size() + fresh.size() <= cap, "frozen node outside the frictional boundary");
        require(cap * (size_t)nv_ <= (size_t(1) << 29), "the baseline forward's A^-1 columns at the " +
                                                            std::to_string(cap) + " boundary vertices exceed 4 GiB");
        if (d_Z_.n < cap * (size_t)nv_) {
            DevBuf<double> nz;
            nz.alloc(cap * (size_t)nv_);
            if (!zcol_.empty())
                hip_check(hipMemcpyAsync(nz.p, d_Z_.p, zcol_.size() * (size_t)nv_ * sizeof(double),
                                         hipMemcpyDeviceToDevice, stream_), "Z grow");
            hip_check(hipStreamSynchronize(stream_), "Z grow");
            std::swap(d_Z_.p, nz.p);
            std::swap(d_Z_.n, nz.n);
        }
        std::vector<double> e((size_t)dofs_);
        DevBuf<double> de, dx;
        dx.alloc(dofs_);
        for (size_t k = 0; k < fresh.size(); k += 3) {
            const int na = (int)std::min<size_t>(3, fresh.size() - k);
            std::fill(e.begin(), e.end(), 0.0);
            for (int a = 0; a < na; ++a) e[3 * (size_t)fresh[k + a] + a] = 1.0;
            de.upload(e);
            ainv64(de.p, dx.p);
            const int p0 = (int)zcol_.size();
            launch_pdir_columns(stream_, nv_, p0, na, dx.p, d_Z_.p);
            for (int a = 0; a < na; ++a) zcol_[fresh[k + a]] = p0 + a;
            hip_check(hipStreamSynchronize(stream_), "Z columns");
        }
    }
    if (g_frozen_ != frozen) {
        std::vector<int> zc(nc);
        for (int i = 0; i < nc; ++i) zc[i] = zcol_.at(frozen[i]);
        d_bl_nodes_.upload(frozen);
        d_bl_zcol_.upload(zc);
        DevBuf<double> dG;
        dG.alloc((size_t)nc * nc);
        launch_gather_z(stream_, nc, nv_, d_bl_nodes_.p, d_bl_zcol_.p, d_Z_.p, dG.p);
        hip_check(hipStreamSynchronize(stream_), "G");
        g_llt_.assign((size_t)nc * nc, 0.0);
        dG.download(g_llt_.data(), (size_t)nc * nc);
        for (int i = 0; i < nc; ++i)  // symmetrize (A^-1 from the factor is symmetric to round-off)
            for (int j = 0; j < i; ++j) g_llt_[i * nc + j] = g_llt_[j * nc + i] = 0.5 * (g_llt_[i * nc + j] + g_llt_[j * nc + i]);
        require(chol(g_llt_, nc), "PdLhsSolve: the frozen block of A^-1 is not positive definite.");
        g_frozen_ = frozen;
    }
    d_bl_mu_.alloc(3 * (size_t)nc);
    launch_gather_nodes(stream_, nc, d_bl_nodes_.p, x, d_bl_mu_.p);
    std::vector<double> y(3 * (size_t)nc), mu(3 * (size_t)nc);
    hip_check(hipStreamSynchronize(stream_), "frozen rows");
    d_bl_mu_.download(y.data(), 3 * (size_t)nc);
    for (int d = 0; d < 3; ++d) {
        std::vector<double> b(nc);
        for (int i = 0; i < nc; ++i) b[i] = -y[3 * i + d];
        chol_solve(g_llt_, nc, b.data());
        for (int i = 0; i < nc; ++i) mu[3 * i + d] = b[i];
    }
    d_bl_mu_.upload(mu);
    launch_frozen_apply(stream_, nv_, nc, d_Z_.p, d_bl_zcol_.p, d_bl_mu_.p, x);
    launch_copy_nodes(stream_, nc, d_bl_nodes_.p, r, x);
}

// PdNonlinearSolve (forward.cpp:294-526). Returns the iterations used; throws
// when max_pd_iter is reached, as the reference does.
int Body::pd_nonlinear_solve(const double* q_init, double inv_h2m, const double* rhs, const std::vector<int>& frozen,
                             const std::vector<double>& frozen_vals, const std::map<std::string, double>& opt,
                             double* q_sol) {
    const int max_pd_iter = (int)req(opt, "max_pd_iter");
    const double abs_tol = req(opt, "abs_tol"), rel_tol = req(opt, "rel_tol");
    req(opt, "verbose");
    const bool use_bfgs = req(opt, "use_bfgs") != 0.0;
    require(max_pd_iter > 0, "Invalid max_pd_iter: " + std::to_string(max_pd_iter));
    int hist = 0, max_ls = 0;
    if (use_bfgs) {
        hist = (int)req(opt, "bfgs_history_size");
        require(hist >= 1, "Invalid bfgs_history_size.");
        max_ls = (int)req(opt, "max_ls_iter");
        require(max_ls > 0, "Invalid max_ls_iter: " + std::to_string(max_ls));
    }
    const int n3 = dofs_;
    auto vec = [&](size_t i) -> double* {
        while (d_bl_.size() <= i) d_bl_.emplace_back(new DevBuf<double>());
        d_bl_[i]->alloc(n3);
        return d_bl_[i]->p;
    };
    double *grad = vec(0), *force = vec(1), *dir = vec(2), *qn = vec(3), *tmp = vec(4);
    // augmented Dirichlet DOFs: the body's own and the frozen nodes' (at their values)
    std::map<int, double> aug = dirichlet_;
    for (size_t i = 0; i < frozen.size(); ++i)
        for (int d = 0; d < 3; ++d) aug[3 * frozen[i] + d] = frozen_vals[3 * i + d];
    std::vector<int> fix;
    std::vector<double> fixv;
    for (auto& kv : aug) { fix.push_back(kv.first); fixv.push_back(kv.second); }
    const int nfix = (int)fix.size();
    d_bl_fix_.upload(fix);
    d_bl_fixv_.upload(fixv);
    auto objective = [&](const double* qc, double e) {
        launch_axpbypcz(stream_, n3, 1.0, qc, -1.0, rhs, 0.0, nullptr, tmp);
        return 0.5 * inv_h2m * dots({{tmp, tmp}})[0] + e;
    };
    auto gradient = [&](const double* qc) {  // grad = (inv_h2m (q - rhs) - f) on the free DOFs
        baseline_force(qc, force);
        launch_axpbypcz(stream_, n3, inv_h2m, qc, -inv_h2m, rhs, -1.0, force, grad);
        launch_zero_dirichlet(stream_, nfix, d_bl_fix_.p, grad);
    };
    hip_check(hipMemcpyAsync(q_sol, q_init, sizeof(double) * n3, hipMemcpyDeviceToDevice, stream_), "q_init");
    launch_set_dirichlet(stream_, nfix, d_bl_fix_.p, d_bl_fixv_.p, q_sol);
    double energy = baseline_energy(q_sol);
    double obj = objective(q_sol, energy);
    gradient(q_sol);
    launch_axpbypcz(stream_, n3, inv_h2m, rhs, 0.0, nullptr, 0.0, nullptr, tmp);
    launch_zero_dirichlet(stream_, nfix, d_bl_fix_.p, tmp);
    const double rhs_norm = std::sqrt(dots({{tmp, tmp}})[0]);
    // L-BFGS pairs (s_k, y_k) in a ring of work vectors; x_k, g_k of the last iterate
    std::deque<int> sh, yh;
    int next_slot = 7;
    std::vector<int> free_slots;
    auto slot = [&]() {
        if (!free_slots.empty()) { const int k = free_slots.back(); free_slots.pop_back(); return k; }
        return next_slot++;
    };
    double* xlast = vec(5);
    double* glast = vec(6);
    bool have_last = false;
    for (int it = 0; it < max_pd_iter; ++it) {
        if (use_bfgs) {
            if (!have_last) {
                pd_lhs_solve(grad, dir, frozen);
            } else {
                const int ks = slot(), ky = slot();
                double* s = vec(ks);
                double* y = vec(ky);
                launch_axpbypcz(stream_, n3, 1.0, q_sol, -1.0, xlast, 0.0, nullptr, s);
                launch_axpbypcz(stream_, n3, 1.0, grad, -1.0, glast, 0.0, nullptr, y);
                sh.push_back(ks);
                yh.push_back(ky);
                if ((int)sh.size() > hist) {
                    free_slots.push_back(sh.front());
                    free_slots.push_back(yh.front());
                    sh.pop_front();
                    yh.pop_front();
                }
                // two-loop recursion, newest pair first
                launch_axpbypcz(stream_, n3, 1.0, grad, 0.0, nullptr, 0.0, nullptr, tmp);
                const int m = (int)sh.size();
                std::vector<double> rho(m), alpha(m);
                for (int k = m - 1; k >= 0; --k) {
                    const double* sk = d_bl_[sh[k]]->p;
                    const double* yk = d_bl_[yh[k]]->p;
                    const auto d = dots({{yk, sk}, {sk, tmp}});
                    rho[k] = 1 / d[0];
                    alpha[k] = rho[k] * d[1];
                    launch_axpbypcz(stream_, n3, 1.0, tmp, -alpha[k], yk, 0.0, nullptr, tmp);
                }
                pd_lhs_solve(tmp, dir, frozen);
                for (int k = 0; k < m; ++k) {
                    const double* sk = d_bl_[sh[k]]->p;
                    const double* yk = d_bl_[yh[k]]->p;
                    const double beta = rho[k] * dots({{yk, dir}})[0];
                    launch_axpbypcz(stream_, n3, 1.0, dir, alpha[k] - beta, sk, 0.0, nullptr, dir);
                }
            }
            hip_check(hipMemcpyAsync(xlast, q_sol, sizeof(double) * n3, hipMemcpyDeviceToDevice, stream_), "x_last");
            hip_check(hipMemcpyAsync(glast, grad, sizeof(double) * n3, hipMemcpyDeviceToDevice, stream_), "g_last");
            have_last = true;
            double gd = dots({{dir, grad}})[0];
            if (gd <= 0) {  // not a descent direction: the gradient itself (already zero on fixed DOFs)
                hip_check(hipMemcpyAsync(dir, grad, sizeof(double) * n3, hipMemcpyDeviceToDevice, stream_), "dir");
                gd = dots({{dir, grad}})[0];
            }
            // backtracking line search: obj(q - t d) < obj - 1e-4 t g.d (+ eps)
            double step = 1;
            launch_axpbypcz(stream_, n3, 1.0, q_sol, -step, dir, 0.0, nullptr, qn);
            double en = baseline_energy(qn), on = objective(qn, en);
            for (int j = 0; j < max_ls; ++j) {
                const double cond = obj - 1e-4 * step * gd;
                if (!std::isnan(on) && on < cond + DBL_EPSILON) break;
                step /= 2;
                launch_axpbypcz(stream_, n3, 1.0, q_sol, -step, dir, 0.0, nullptr, qn);
                en = baseline_energy(qn);
                on = objective(qn, en);
                ++last_baseline.ls_steps;
            }
            hip_check(hipMemcpyAsync(q_sol, qn, sizeof(double) * n3, hipMemcpyDeviceToDevice, stream_), "accept");
            energy = en;
            obj = on;
        } else {
            // local step with the augmented conditions' constant part, then the global step
            require(frozen.empty() && pverts_.empty(),
                    "The baseline forward without L-BFGS supports no frictional contact in this build.");
            local_rhs(q_sol, inv_h2m);  // b = inv_h2m q~ + pd_rhs, with q~ = rhs on this call
            launch_axpbypcz(stream_, n3, 1.0, d_b_.p, 0.0, nullptr, 0.0, nullptr, tmp);
            launch_zero_dirichlet(stream_, nfix, d_bl_fix_.p, tmp);
            pd_lhs_solve(tmp, q_sol, frozen);
            launch_set_dirichlet(stream_, nfix, d_bl_fix_.p, d_bl_fixv_.p, q_sol);
        }
        gradient(q_sol);
        if (std::sqrt(dots({{grad, grad}})[0]) <= rel_tol * rhs_norm + abs_tol) return it + 1;
    }
    throw std::runtime_error("PD method fails to converge.");
}

void Body::forward_baseline(const std::string& method, const double* q, const double* v, const double* f_ext, double dt,
                            const std::map<std::string, double>& opt, double* q_next, double* v_next,
                            std::vector<int>& active) {
    req(opt, "max_pd_iter");
    req(opt, "abs_tol");
    req(opt, "rel_tol");
    req(opt, "verbose");
    const int threads = (int)req(opt, "thread_ct");
    if (req(opt, "use_bfgs") != 0.0) {
        require(req(opt, "bfgs_history_size") > 1, "Invalid bfgs_history_size.");
        require(req(opt, "max_ls_iter") > 0, "Invalid max_ls_iter.");
    }
    require(boundary_.type != 5, "The baseline forward does not support the mesh boundary in this build.");
    require(std::none_of(energies_.begin(), energies_.end(), [](const Energy& e) { return e.type == kNeoHookean; }),
            "The baseline forward supports the corotated and volume energies.");
    require((int)act_.size() == act_dofs_, "Inconsistent actuation size.");
    setup(dt, threads);
    require(pverts_.empty(), "The baseline forward supports whole-vertex Dirichlet conditions only.");
    if (kpd_exact_) { kpd_exact_ = false; build_materials(); }  // energies read each energy's stiffness()
    last_baseline = BaselineStats();
    const int n3 = dofs_;
    const double h = dt, mass = element_volume_ * density_, h2m = h * h / mass, inv_h2m = mass / (h * h);
    set_inertial_target(q, v, f_ext, dt);  // rhs = q + h v + h2m (f_ext + f_state) in d_qtilde_
    std::vector<double> rhs(n3);
    d_qtilde_.download(rhs.data(), n3);
    DevBuf<double> dq0, dqs, dfs;
    dq0.upload(q, n3);
    dqs.alloc(n3);
    dfs.alloc(n3);
    auto finish = [&](const std::vector<double>& qs) {
        for (int i = 0; i < n3; ++i) { q_next[i] = qs[i]; v_next[i] = (qs[i] - q[i]) / h; }
    };
    auto frozen_of = [&](const std::vector<int>& nodes, std::vector<int>& fr, std::vector<double>& vals) {
        std::set<int> s(nodes.begin(), nodes.end());
        fr.assign(s.begin(), s.end());
        vals.clear();
        for (int v0 : fr) for (int d = 0; d < 3; ++d) vals.push_back(q[3 * v0 + d]);
    };
    std::vector<double> qs(n3), fs(n3);
    if (method.size() >= 13 && method.compare(method.size() - 13, 13, "fixed_contact") == 0) {
        std::vector<int> fr;
        std::vector<double> vals;
        frozen_of(active, fr, vals);
        last_baseline.pd_iters = pd_nonlinear_solve(dq0.p, inv_h2m, d_qtilde_.p, fr, vals, opt, dqs.p);
        hip_check(hipStreamSynchronize(stream_), "baseline");
        dqs.download(qs.data(), n3);
        finish(qs);
        last_baseline.contact_iters = 1;
        return;
    }
    const int max_contact_iter = 10;
    std::vector<std::set<int>> history;
    for (int ci = 0; ci < max_contact_iter; ++ci) {
        std::vector<int> fr;
        std::vector<double> vals;
        frozen_of(active, fr, vals);
        last_baseline.pd_iters += pd_nonlinear_solve(dq0.p, inv_h2m, d_qtilde_.p, fr, vals, opt, dqs.p);
        baseline_force(dqs.p, dfs.p);
        hip_check(hipStreamSynchronize(stream_), "baseline");
        dqs.download(qs.data(), n3);
        dfs.download(fs.data(), n3);
        const std::set<int> past(active.begin(), active.end());
        history.push_back(past);
        active.clear();
        bool good = true;
        for (auto& kv : fb_vertices_) {
            const int node = kv.first;
            double dist, nn[3];
            boundary_.eval(&qs[3 * (size_t)node], dist, nn);
            double cf = 0;
            for (int d = 0; d < 3; ++d) {
                const size_t i = 3 * (size_t)node + d;
                cf += nn[d] * (inv_h2m * (qs[i] - rhs[i]) - fs[i]);
            }
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
            finish(qs);
            active.assign(past.begin(), past.end());
            last_baseline.contact_iters = ci + 1;
            return;
        }
    }
}

}  // namespace phd

namespace phd {

// BackwardProjectiveDynamics for a pd_eigen baseline method
// (deformable_projective_dynamics_backward.cpp:141-679), the path the reference
// takes for bodies without an elastic material: L-BFGS on 1/2 x^T S x - b^T x,
// S = PdLhsMatrixOp - K_pd (each energy's stiffness() in K_pd, backward.cpp:10-80),
// over the free DOFs with b = dl/dq_next + dl/dv_next / h; then
// dl/drhs = x M/h^2 on the free DOFs, b + NewtonMatrixOp(-x) on the fixed ones
// (for energies of uniform stiffness NewtonMatrixOp = A - K_pd), the actuation and
// material gradients with the adjoint x, and the chain back through the frozen
// nodes (rhs = q), the Dirichlet values and rhs = q + h v + h^2/m (f_ext + f_state).
// The reference falls back to a sparse Cholesky of the Newton matrix when L-BFGS
// meets an indefinite direction or its cap; that fallback is not built here.
void Body::backward_baseline(const double* q, const double* v, const double* f_ext, double dt, const double* q_next,
                             const std::vector<int>& active, const double* dl_dq_next, const double* dl_dv_next,
                             const std::map<std::string, double>& opt, double* dl_dq, double* dl_dv, double* dl_df_ext,
                             double* dl_dmat_w, double* dl_dstate_p) {
    (void)f_ext;
    const int max_pd_iter = (int)req(opt, "max_pd_iter");
    const double abs_tol = req(opt, "abs_tol"), rel_tol = req(opt, "rel_tol");
    req(opt, "verbose");
    const int threads = (int)req(opt, "thread_ct");
    const bool use_bfgs = req(opt, "use_bfgs") != 0.0;
    require(max_pd_iter > 0, "Invalid max_pd_iter: " + std::to_string(max_pd_iter));
    int hist = 0, max_ls = 0;
    if (use_bfgs) {
        hist = (int)req(opt, "bfgs_history_size");
        require(hist >= 1, "Invalid bfgs_history_size.");
        max_ls = (int)req(opt, "max_ls_iter");
        require(max_ls > 0, "Invalid max_ls_iter: " + std::to_string(max_ls));
    }
    require(std::none_of(energies_.begin(), energies_.end(), [](const Energy& e) { return e.type == kNeoHookean; }),
            "The baseline backward supports the corotated and volume energies.");
    require(std::all_of(energies_.begin(), energies_.end(), [](const Energy& e) { return e.elem_stiff.empty(); }),
            "The baseline backward needs energies of uniform stiffness (per-element parameters: use one energy per "
            "element group).");
    require((int)act_.size() == act_dofs_, "Inconsistent actuation size.");
    setup(dt, threads);
    require(pverts_.empty(), "The baseline backward supports whole-vertex Dirichlet conditions only.");
    if (kpd_exact_) { kpd_exact_ = false; build_materials(); }
    const int n3 = dofs_, nE = (int)energies_.size();
    const double h = dt, inv_h = 1 / h, mass = element_volume_ * density_, h2m = h * h / mass, inv_h2m = mass / (h * h);
    // K_pd at q_next (unfiltered, FP64)
    d_q_.upload(q_next, n3);
    d_W64_.alloc((size_t)ne_ * samples_ * 81);
    d_nfilt_.alloc(1);
    const double* kn[4] = {d_kn_[0].p, d_kn_[1].p, d_kn_[2].p, d_kn_[3].p};
    launch_kpd_build<double>(stream_, mat_, has_cor_ ? d_kc_.p : nullptr, has_vol_ ? d_kv_.p : nullptr, kn, d_q_.p, 0, 1.0,
                             -INFINITY, d_nfilt_.p, d_W64_.p);
    // augmented conditions: the body's Dirichlet DOFs and the frozen contact nodes (at q)
    std::set<int> fset(active.begin(), active.end());
    const std::vector<int> frozen(fset.begin(), fset.end());
    std::vector<char> fixed(n3, 0);
    for (auto& kv : dirichlet_) fixed[kv.first] = 1;
    for (int v0 : frozen) for (int d = 0; d < 3; ++d) fixed[3 * (size_t)v0 + d] = 1;
    std::vector<int> fix;
    for (int i = 0; i < n3; ++i) if (fixed[i]) fix.push_back(i);
    const int nfix = (int)fix.size();
    d_bl_fix_.upload(fix);
    d_bl_nodes_.upload(frozen);
    std::vector<double> agg(n3);
    for (int i = 0; i < n3; ++i) {
        dl_dq[i] = -dl_dv_next[i] * inv_h;
        dl_dv[i] = 0;
        dl_df_ext[i] = 0;
        agg[i] = dl_dq_next[i] + dl_dv_next[i] * inv_h;
    }
    auto vec = [&](size_t i) -> double* {
        while (d_bl_.size() <= i) d_bl_.emplace_back(new DevBuf<double>());
        d_bl_[i]->alloc(n3);
        return d_bl_[i]->p;
    };
    double *b = vec(0), *x = vec(1), *Sx = vec(2), *grad = vec(3), *dir = vec(4), *xn = vec(5), *Sxn = vec(6),
           *tmp = vec(7), *xl = vec(8), *gl = vec(9), *kx = vec(10), *zero = vec(11);
    hip_check(hipMemcpy(b, agg.data(), sizeof(double) * n3, hipMemcpyHostToDevice), "b");
    hip_check(hipMemsetAsync(zero, 0, sizeof(double) * n3, stream_), "zero");
    hip_check(hipMemsetAsync(x, 0, sizeof(double) * n3, stream_), "x = 0");
    // S x = PdLhsMatrixOp(x) - K_pd x: A x with the frozen entries of x zeroed,
    // x itself on the frozen rows
    auto apply_S = [&](const double* xin, double* out) {
        hip_check(hipMemcpyAsync(tmp, xin, sizeof(double) * n3, hipMemcpyDeviceToDevice, stream_), "S x");
        launch_zero_dirichlet(stream_, nfix, d_bl_fix_.p, tmp);  // the body's Dirichlet entries are 0 already
        launch_residual(stream_, nv_, ell_w_, d_Aie_.p, d_Axe_.p, zero, tmp, d_dmask_.p, out, nullptr);  // -A x
        launch_axpbypcz(stream_, n3, -1.0, out, 0.0, nullptr, 0.0, nullptr, out);
        launch_copy_nodes(stream_, (int)frozen.size(), d_bl_nodes_.p, xin, out);
        kpd_apply(xin, kx, 0);
        launch_axpbypcz(stream_, n3, 1.0, out, -1.0, kx, 0.0, nullptr, out);
    };
    auto objective = [&](const double* xc, const double* Sxc) {
        const auto d = dots({{xc, Sxc}, {b, xc}});
        return 0.5 * d[0] - d[1];
    };
    auto gradient = [&](const double* Sxc, double* g) {
        launch_axpbypcz(stream_, n3, 1.0, Sxc, -1.0, b, 0.0, nullptr, g);
        launch_zero_dirichlet(stream_, nfix, d_bl_fix_.p, g);
    };
    apply_S(x, Sx);
    gradient(Sx, grad);
    double obj = objective(x, Sx);
    const double bn = std::sqrt(dots({{b, b}})[0]);
    bool success = false;
    std::deque<int> sh, yh;
    std::vector<int> free_slots;
    int next_slot = 12;
    auto slot = [&]() {
        if (!free_slots.empty()) { const int k = free_slots.back(); free_slots.pop_back(); return k; }
        return next_slot++;
    };
    bool have_last = false;
    int it_used = 0;
    for (int it = 0; it < max_pd_iter && !success; ++it) {
        it_used = it + 1;
        if (use_bfgs) {
            if (!have_last) {
                pd_lhs_solve(grad, dir, frozen);
            } else {
                const int ks = slot(), ky = slot();
                double* s = vec(ks);
                double* y = vec(ky);
                launch_axpbypcz(stream_, n3, 1.0, x, -1.0, xl, 0.0, nullptr, s);
                launch_axpbypcz(stream_, n3, 1.0, grad, -1.0, gl, 0.0, nullptr, y);
                sh.push_back(ks);
                yh.push_back(ky);
                if ((int)sh.size() > hist) {
                    free_slots.push_back(sh.front());
                    free_slots.push_back(yh.front());
                    sh.pop_front();
                    yh.pop_front();
                }
                double* qv = xn;  // scratch for the two-loop vector
                launch_axpbypcz(stream_, n3, 1.0, grad, 0.0, nullptr, 0.0, nullptr, qv);
                const int m = (int)sh.size();
                std::vector<double> rho(m), alpha(m);
                for (int k = m - 1; k >= 0; --k) {
                    const double* sk = d_bl_[sh[k]]->p;
                    const double* yk = d_bl_[yh[k]]->p;
                    const auto d = dots({{yk, sk}, {sk, qv}});
                    rho[k] = 1 / d[0];
                    alpha[k] = rho[k] * d[1];
                    launch_axpbypcz(stream_, n3, 1.0, qv, -alpha[k], yk, 0.0, nullptr, qv);
                }
                pd_lhs_solve(qv, dir, frozen);
                for (int k = 0; k < m; ++k) {
                    const double* sk = d_bl_[sh[k]]->p;
                    const double* yk = d_bl_[yh[k]]->p;
                    const double beta = rho[k] * dots({{yk, dir}})[0];
                    launch_axpbypcz(stream_, n3, 1.0, dir, alpha[k] - beta, sk, 0.0, nullptr, dir);
                }
            }
            hip_check(hipMemcpyAsync(xl, x, sizeof(double) * n3, hipMemcpyDeviceToDevice, stream_), "x_last");
            hip_check(hipMemcpyAsync(gl, grad, sizeof(double) * n3, hipMemcpyDeviceToDevice, stream_), "g_last");
            have_last = true;
            launch_zero_dirichlet(stream_, nfix, d_bl_fix_.p, dir);
            const double gd = dots({{dir, grad}})[0];
            require(gd >= -1e-4, "The baseline backward's L-BFGS met an indefinite operator; the reference's Newton "
                                 "fallback is not built.");
            double step = 1;
            launch_axpbypcz(stream_, n3, 1.0, x, -step, dir, 0.0, nullptr, xn);
            apply_S(xn, Sxn);
            double on = objective(xn, Sxn);
            for (int j = 0; j < max_ls; ++j) {
                const double cond = obj - 1e-4 * step * gd;
                if (!std::isnan(on) && on < cond + DBL_EPSILON) break;
                step /= 2;
                launch_axpbypcz(stream_, n3, 1.0, x, -step, dir, 0.0, nullptr, xn);
                apply_S(xn, Sxn);
                on = objective(xn, Sxn);
            }
            std::swap(x, xn);
            std::swap(Sx, Sxn);
            obj = on;
        } else {
            kpd_apply(x, kx, 0);
            launch_axpbypcz(stream_, n3, 1.0, b, 1.0, kx, 0.0, nullptr, tmp);
            launch_zero_dirichlet(stream_, nfix, d_bl_fix_.p, tmp);
            pd_lhs_solve(tmp, x, frozen);
            launch_zero_dirichlet(stream_, nfix, d_bl_fix_.p, x);
            apply_S(x, Sx);
            obj = objective(x, Sx);
        }
        gradient(Sx, grad);
        if (std::sqrt(dots({{grad, grad}})[0]) <= rel_tol * bn + abs_tol) success = true;
    }
    require(success, "The baseline backward's L-BFGS did not converge in max_pd_iter; the reference's Newton "
                     "fallback is not built.");
    last_baseline.bwd_iters = it_used;
    // adjoint = x with the fixed entries 0 (x is 0 there); NewtonMatrixOp(-adjoint) = -(A - K_pd) x
    hip_check(hipMemcpyAsync(tmp, x, sizeof(double) * n3, hipMemcpyDeviceToDevice, stream_), "adj");
    launch_residual(stream_, nv_, ell_w_, d_Aie_.p, d_Axe_.p, zero, tmp, d_dmask_.p, Sxn, nullptr);  // -A x
    kpd_apply(x, kx, 0);
    launch_axpbypcz(stream_, n3, 1.0, Sxn, 1.0, kx, 0.0, nullptr, Sxn);  // -(A - K_pd) x = NewtonMatrixOp(-x)
    std::vector<double> hx(n3), hn(n3);
    hip_check(hipStreamSynchronize(stream_), "baseline backward");
    hip_check(hipMemcpy(hx.data(), x, sizeof(double) * n3, hipMemcpyDeviceToHost), "x");
    hip_check(hipMemcpy(hn.data(), Sxn, sizeof(double) * n3, hipMemcpyDeviceToHost), "newton");
    std::vector<double> dl_drhs(n3);
    for (int i = 0; i < n3; ++i) dl_drhs[i] = fixed[i] ? agg[i] + hn[i] : hx[i] * inv_h2m;
    // actuation and material gradients with the adjoint x
    last_dl_da_.assign(act_dofs_, 0.0);
    last_dl_dact_w_.assign(muscles_.size(), 0.0);
    if (act_dofs_ > 0) {
        d_gda_.alloc(ne_); d_gdw_.alloc(ne_);
        launch_dact(stream_, mat_, d_q_.p, x, element_volume_ / samples_, d_gda_.p, d_gdw_.p);
        hip_check(hipStreamSynchronize(stream_), "dact");
        std::vector<double> ga(ne_), gw(ne_);
        d_gda_.download(ga.data(), ne_);
        d_gdw_.download(gw.data(), ne_);
        for (size_t gi = 0; gi < muscles_.size(); ++gi)
            for (int e : muscles_[gi].elems) {
                last_dl_da_[mus_act_[e]] = ga[e];
                last_dl_dact_w_[gi] += gw[e];
            }
    }
    for (int i = 0; i < nE; ++i) dl_dmat_w[i] = 0;
    if (nE > 0) {
        d_gcor_.alloc(ne_); d_gvol_.alloc(ne_); d_gnh_.alloc(1);
        launch_dmatw(stream_, mat_, d_q_.p, x, element_volume_ / samples_, has_cor_ ? d_gcor_.p : nullptr,
                     has_vol_ ? d_gvol_.p : nullptr, nullptr);
        hip_check(hipStreamSynchronize(stream_), "dmatw");
        std::vector<double> gc(ne_, 0.0), gv(ne_, 0.0);
        if (has_cor_) d_gcor_.download(gc.data(), ne_);
        if (has_vol_) d_gvol_.download(gv.data(), ne_);
        for (int ei = 0; ei < nE; ++ei) {
            const Energy& en = energies_[ei];
            const std::vector<double>& src = en.type == kCorotated ? gc : gv;
            double s = 0;
            for (int e = 0; e < ne_; ++e) if (en.member[e]) s += src[e];
            dl_dmat_w[ei] = s;
        }
    }
    // rhs = q on the frozen nodes, the Dirichlet values, q + h v + h2m (f_ext + f_state)
    std::vector<double> basic = dl_drhs;
    for (int v0 : frozen)
        for (int d = 0; d < 3; ++d) {
            const size_t i = 3 * (size_t)v0 + d;
            basic[i] = 0;
            dl_dq[i] += dl_drhs[i];
        }
    for (auto& kv : dirichlet_) basic[kv.first] = 0;
    std::vector<double> dl_df(n3), sq(n3), sv(n3);
    for (int i = 0; i < n3; ++i) {
        dl_dq[i] += basic[i];
        dl_dv[i] += basic[i] * h;
        dl_df_ext[i] += basic[i] * h2m;
        dl_df[i] = basic[i] * h2m;
    }
    state_force_backward(q, v, dl_df.data(), sq.data(), sv.data(), dl_dstate_p);
    for (int i = 0; i < n3; ++i) { dl_dq[i] += sq[i]; dl_dv[i] += sv[i]; }
}

}  // namespace phd
