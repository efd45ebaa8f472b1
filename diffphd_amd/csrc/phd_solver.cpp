// Host orchestration of the PD forward/backward time step on one MI355X.
//
// The forward restates ForwardProjectiveDynamicsAlgPhd
// (cpp/core/src/fem/deformable_projective_dynamics_forward_gpu.cpp:387-1594) and
// the backward BackwardProjectiveDynamicsAlgPhd
// (cpp/core/src/fem/deformable_projective_dynamics_backward_gpu.cpp:648-1250);
// every vector of size 3N lives in HBM and only K-sized contact vectors and a
// handful of reduction results cross PCIe per iteration.
//
// Each global solve is written in residual form, x_{k+1} = x_k + S^T S r_k with
// r_k evaluated in FP64 by the 27-point stencil of A, so the FP32 factor acts
// as a preconditioner and the fixed point keeps FP64 accuracy; with the FP64
// factor the form is algebraically identical to x = A^-1 b.
#include "phd_solver.h"
#include "phd_dense.h"
#include "phd_mlp.h"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <numeric>
#include <set>
#include <mutex>
#include <stdexcept>
#include <unordered_map>

// Scratch of the allocation-free A^-1 entry point (Body::apply_ainv_noalloc):
// one file-scope device array, which every GPU the library runs on holds at
// its own address until the process ends. Host code uses only the address
// that hipGetSymbolAddress returns for the current device (scratch_address).
static constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace phd {

namespace {
std::mutex g_scratch_mu;
constexpr int kMaxDevices = 64;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

// Device address of `scratch` on `device`, which must be the current device;
// resolved once per device.
static void* scratch_address(int device) {
    if (device < 0 || device >= kMaxDevices) throw std::runtime_error("device ordinal out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (!p) throw std::runtime_error("hipGetSymbolAddress(scratch) returned a null address");
        g_scratch_ptr[device] = p;
    }
    return g_scratch_ptr[device];
}

Body::PhaseTimer::PhaseTimer(Body* b_, int id_) : b(b_), id(id_), t0(0) {
    if (b->timing_level >= 2) {
        hip_check(hipStreamSynchronize(b->stream_), "phase sync");
        t0 = std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
    }
}
Body::PhaseTimer::~PhaseTimer() {
    if (b->timing_level >= 2) {
        (void)hipStreamSynchronize(b->stream_);
        b->phase[id] += std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count() - t0;
    }
}


void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string("HIP error in ") + what + ": " + hipGetErrorString(e));
}
static void check(bool ok, const std::string& msg) {
    if (!ok) throw std::runtime_error(msg);
}
static double opt_or(const std::map<std::string, double>& o, const char* k, double d) {
    auto it = o.find(k);
    return it == o.end() ? d : it->second;
}
static double opt_req(const std::map<std::string, double>& o, const char* k) {
    auto it = o.find(k);
    check(it != o.end(), std::string("Missing option ") + k + ".");
    return it->second;
}
static double nrm2(const std::vector<double>& v) {
    double s = 0;
    for (double x : v) s += x * x;
    return std::sqrt(s);
}
// Dense symmetric solve by LDL^T (small systems: AA Gram). Row-major.
static std::vector<double> small_ldlt_solve(std::vector<double> M, int n, std::vector<double> x) {
    std::vector<double> d(n);
    for (int j = 0; j < n; ++j) {
        double djj = M[j * n + j];
        for (int k = 0; k < j; ++k) djj -= M[j * n + k] * M[j * n + k] * d[k];
        d[j] = djj;
        for (int i = j + 1; i < n; ++i) {
            double s = M[i * n + j];
            for (int k = 0; k < j; ++k) s -= M[i * n + k] * M[j * n + k] * d[k];
            M[i * n + j] = djj != 0 ? s / djj : 0;
        }
    }
    for (int i = 0; i < n; ++i) for (int k = 0; k < i; ++k) x[i] -= M[i * n + k] * x[k];
    for (int i = 0; i < n; ++i) x[i] = d[i] != 0 ? x[i] / d[i] : 0;
    for (int i = n - 1; i >= 0; --i) for (int k = i + 1; k < n; ++k) x[i] -= M[k * n + i] * x[k];
    return x;
}
static double det3(const double* A) {  // column-major
    return A[0] * (A[4] * A[8] - A[7] * A[5]) - A[3] * (A[1] * A[8] - A[7] * A[2]) + A[6] * (A[1] * A[5] - A[4] * A[2]);
}

Body::Body(int device) : device_(device) {
    hip_check(hipSetDevice(device_), "hipSetDevice");
    scratch_ = scratch_address(device_);
    hip_check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
    hip_check(hipHostMalloc((void**)&pin_small_, kPinSmallBytes, hipHostMallocDefault), "hipHostMalloc");
    rocblas_handle h;
    check(rocblas_create_handle(&h) == rocblas_status_success, "rocblas_create_handle failed");
    rocblas_set_stream(h, stream_);
    blas_ = h;
}
Body::~Body() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (pin_small_) (void)hipHostFree(pin_small_);
    for (auto e : ev_pool_) (void)hipEventDestroy(e);
    if (blas_) rocblas_destroy_handle((rocblas_handle)blas_);
    if (stream_) (void)hipStreamDestroy(stream_);
}

// ---------------------------------------------------------------------------
// Scene: Mesh::Initialize (mesh.cpp:4-32), HexDeformable/TetDeformable
// ::InitializeFiniteElementSamples (hex_deformable.cpp:3-50, tet_deformable.cpp:3-41).
// ---------------------------------------------------------------------------
void Body::init_mesh(int edim, int nv, const double* X, int ne, const int* elems, double density) {
    check(edim == 8 || edim == 4, "Only hex (8) and tet (4) elements are supported.");
    edim_ = edim; nv_ = nv; ne_ = ne; dofs_ = 3 * nv; density_ = density;
    samples_ = edim == 8 ? 8 : 1;
    X_.assign(X, X + 3 * (size_t)nv);
    elems_.assign(elems, elems + (size_t)edim * ne);
    double vsum = 0;
    for (int e = 0; e < ne; ++e) {
        const int* v = &elems_[(size_t)e * edim];
        const int c[3] = {1, 2, edim == 8 ? 4 : 3};
        double A[9];
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 3; ++i) A[i + 3 * j] = X_[3 * v[c[j]] + i] - X_[3 * v[0] + i];
        vsum += std::fabs(det3(A)) / (edim == 8 ? 1.0 : 6.0);
    }
    element_volume_ = vsum / ne;
    const double dx = std::pow(element_volume_, 1.0 / 3.0);
    if (edim == 8) {
        g_stride_ = 0;
        G_.assign(8 * 24, 0.0);
        const double r = 1 / std::sqrt(3.0), inv = 1 / dx;
        for (int s = 0; s < 8; ++s) {
            const int bx = (s >> 2) & 1, by = (s >> 1) & 1, bz = s & 1;
            const double x = ((bx - 0.5) * r + 0.5) * dx, y = ((by - 0.5) * r + 0.5) * dx, z = ((bz - 0.5) * r + 0.5) * dx;
            const double nx = x * inv, ny = y * inv, nz = z * inv, cx = 1 - nx, cy = 1 - ny, cz = 1 - nz;
            const double rows[8][3] = {
                {-inv * cy * cz, cx * -inv * cz, cx * cy * -inv}, {-inv * cy * nz, cx * -inv * nz, cx * cy * inv},
                {-inv * ny * cz, cx * inv * cz, cx * ny * -inv},  {-inv * ny * nz, cx * inv * nz, cx * ny * inv},
                {inv * cy * cz, nx * -inv * cz, nx * cy * -inv},  {inv * cy * nz, nx * -inv * nz, nx * cy * inv},
                {inv * ny * cz, nx * inv * cz, nx * ny * -inv},   {inv * ny * nz, nx * inv * nz, nx * ny * inv}};
            for (int k = 0; k < 8; ++k)
                for (int j = 0; j < 3; ++j) G_[s * 24 + k * 3 + j] = rows[k][j];
        }
    } else {
        g_stride_ = 12;
        G_.assign((size_t)ne * 12, 0.0);
        for (int e = 0; e < ne; ++e) {
            const int* v = &elems_[(size_t)e * 4];
            double Bi[9];
            for (int j = 0; j < 3; ++j)
                for (int i = 0; i < 3; ++i) Bi[i + 3 * j] = X_[3 * v[j + 1] + i] - X_[3 * v[0] + i];
            const double d = det3(Bi);
            double B[9];  // inverse, column-major
            B[0] = (Bi[4] * Bi[8] - Bi[7] * Bi[5]) / d; B[3] = (Bi[6] * Bi[5] - Bi[3] * Bi[8]) / d; B[6] = (Bi[3] * Bi[7] - Bi[6] * Bi[4]) / d;
            B[1] = (Bi[7] * Bi[2] - Bi[1] * Bi[8]) / d; B[4] = (Bi[0] * Bi[8] - Bi[6] * Bi[2]) / d; B[7] = (Bi[6] * Bi[1] - Bi[0] * Bi[7]) / d;
            B[2] = (Bi[1] * Bi[5] - Bi[4] * Bi[2]) / d; B[5] = (Bi[3] * Bi[2] - Bi[0] * Bi[5]) / d; B[8] = (Bi[0] * Bi[4] - Bi[3] * Bi[1]) / d;
            double* g = &G_[(size_t)e * 12];
            for (int j = 0; j < 3; ++j) {
                g[0 * 3 + j] = -B[0 + 3 * j] - B[1 + 3 * j] - B[2 + 3 * j];
                g[1 * 3 + j] = B[0 + 3 * j];
                g[2 * 3 + j] = B[1 + 3 * j];
                g[3 * 3 + j] = B[2 + 3 * j];
            }
        }
    }
    energies_.clear();
    pcontacts_.clear();
    acontacts_.clear();
    mcontacts_.clear();
    bballs_.clear();
    hydros_.clear();
    gravity_off_ = n_state_params_ = 0;
    muscles_.clear();
    act_dofs_ = 0;
    mus_act_.assign(ne, -1);
    act_.clear();
    last_dl_da_.clear();
    last_dl_dact_w_.clear();
    dirichlet_.clear();
    gravity_.clear();
    boundary_ = HostBoundary();
    fb_vertices_.clear();
    tape_.clear();
    ready_ = false;
}

// AddPdEnergy, deformable_pd_energy.cpp:42-153 (element energies).
void Body::add_pd_energy(const std::string& type, const std::vector<double>& params, const std::vector<int>& idx) {
    ready_ = false;
    Energy en;
    const int np = (int)params.size();
    if (type == "corotated" || type == "volume") {
        check(np == 1 || np == ne_, "Inconsistent param size: expected 1 (homogeneous) or element_num=" +
                                       std::to_string(ne_) + " (per-element), got " + std::to_string(np));
        en.type = type == "corotated" ? kCorotated : kVolume;
        if (np == 1) en.stiffness = params[0];
        else { en.elem_stiff = params; en.stiffness = std::accumulate(params.begin(), params.end(), 0.0) / np; }
    } else if (type == "pd_neohookean") {
        check(np == 2 || np == 2 * ne_, "pd_neohookean expects params = [mu, lambda] or per-element [mu_1, la_1, ..., mu_n, la_n] of length 2*element_num=" +
                                            std::to_string(2 * ne_) + ", got " + std::to_string(np));
        en.type = kNeoHookean;
        if (np == 2) { en.mu = params[0]; en.la = params[1]; en.stiffness = 2 * params[0]; }
        else {
            double ms = 0, ls = 0;
            en.elem_stiff.resize(ne_);
            for (int i = 0; i < ne_; ++i) { ms += params[2 * i]; ls += params[2 * i + 1]; en.elem_stiff[i] = 2 * params[2 * i]; }
            en.mu = ms / ne_; en.la = ls / ne_;
            en.stiffness = std::accumulate(en.elem_stiff.begin(), en.elem_stiff.end(), 0.0) / ne_;
        }
        check((int)std::count_if(energies_.begin(), energies_.end(), [](const Energy& x) { return x.type == kNeoHookean; }) < 4,
              "At most 4 pd_neohookean energies are supported.");
    } else {
        throw std::runtime_error("Unsupported PD energy: " + type);
    }
    en.member.assign(ne_, 0);
    if (idx.empty()) std::fill(en.member.begin(), en.member.end(), 1);
    for (int i : idx) { check(0 <= i && i < ne_, "Element index out of bound."); en.member[i] = 1; }
    energies_.push_back(en);
}

// AddActuation (deformable_actuation.cpp:5-26) with PdMuscleEnergy::Initialize
// (pd_muscle_energy.cpp:5-24). Each element carries at most one fiber here
// (the reference lets groups overlap).
void Body::add_actuation(double stiffness, const double* fiber, const std::vector<int>& idx) {
    const double nrm = std::sqrt(fiber[0] * fiber[0] + fiber[1] * fiber[1] + fiber[2] * fiber[2]);
    check(nrm > DBL_EPSILON, "Singular fiber_direction.");
    std::vector<char> seen(ne_, 0);
    for (int i : idx) {
        check(0 <= i && i < ne_, "Element index out of bound.");
        check(!seen[i], "Duplicated elements.");
        seen[i] = 1;
        check(mus_act_[i] < 0, "Element " + std::to_string(i) + " already belongs to a muscle: each element carries at "
                               "most one fiber in this build.");
    }
    Muscle mu;
    mu.stiffness = stiffness;
    for (int d = 0; d < 3; ++d) mu.m[d] = fiber[d] / nrm;
    mu.elems = idx;
    for (size_t k = 0; k < idx.size(); ++k) mus_act_[idx[k]] = act_dofs_ + (int)k;
    act_dofs_ += (int)idx.size();
    muscles_.push_back(mu);
    act_.clear();
    ready_ = false;
}

// The actuation levels a of the following forward and backward calls
// (Deformable::Forward's `a`, one entry per muscle element in group order).
void Body::set_actuation(const double* a, int na) {
    check(na == act_dofs_, "Inconsistent actuation size.");
    act_.assign(a, a + na);
    if (act_dofs_ == 0) return;
    std::vector<double> pe(ne_, 0.0);
    for (int e = 0; e < ne_; ++e) if (mus_act_[e] >= 0) pe[e] = act_[mus_act_[e]];
    d_mact_.upload(pe);
    mat_.mact = d_mact_.p;
}

void Body::set_dirichlet(int dof, double val) {
    check(0 <= dof && dof < dofs_, "Dirichlet dof out of bound.");
    dirichlet_[dof] = val;
    ready_ = false;
}

void Body::add_gravity(const double* g) {
    gravity_.assign(g, g + 3);
    gravity_off_ = n_state_params_;
    n_state_params_ += 3;
}

// PlanarContactStateForce::Initialize (planar_contact_state_force.cpp:3-14).
void Body::add_planar_contact(const std::vector<double>& p) {
    check(p.size() == 8, "Inconsistent params for PlanarContactStateForce.");
    PlanarContactForce c;
    const double nn = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    check(nn > 1e-5, "Singular normal.");
    for (int i = 0; i < 3; ++i) c.n[i] = p[i] / nn;
    c.off = p[3] / nn;
    c.p = (int)p[4];
    c.kn = p[5]; c.kf = p[6]; c.mu = p[7];
    c.poff = n_state_params_;
    n_state_params_ += 3;
    pcontacts_.push_back(c);
}

// ArcContactStateForce::Initialize (arc_contact_state_force.cpp:18-30): params
// [center (3), dir (3), start (3), radius, angle, p, kn, kf, mu]
// (deformable_state_force.cpp:33-49); aux = dir x start.
void Body::add_arc_contact(const std::vector<double>& p) {
    check(p.size() == 15, "Inconsistent params for ArcContactStateForce.");
    ArcContactForce a;
    const double dn = std::sqrt(p[3] * p[3] + p[4] * p[4] + p[5] * p[5]);
    const double sn = std::sqrt(p[6] * p[6] + p[7] * p[7] + p[8] * p[8]);
    for (int i = 0; i < 3; ++i) { a.c[i] = p[i]; a.dir[i] = p[3 + i] / dn; a.start[i] = p[6 + i] / sn; }
    for (int i = 0; i < 3; ++i) a.aux[i] = a.dir[(i + 1) % 3] * a.start[(i + 2) % 3] - a.dir[(i + 2) % 3] * a.start[(i + 1) % 3];
    a.R = p[9]; a.angle = p[10]; a.p = (int)p[11];
    a.kn = p[12]; a.kf = p[13]; a.mu = p[14];
    a.poff = n_state_params_;
    n_state_params_ += 3;
    acontacts_.push_back(a);
}

// MeshContactStateForce::Initialize (mesh_contact_state_force.cpp:101-142): params
// [radius, kn_left, kn_right, kf, mu, #faces A, #faces B, faces A, faces B]
// (deformable_state_force.cpp:86-112); 4 parameters (kn_left, kn_right, kf, mu).
void Body::add_mesh_contact(const std::vector<double>& p) {
    check(p.size() >= 7, "Inconsistent params for MeshContactStateForce.");
    const int na = (int)p[5], nb = (int)p[6];
    check(na >= 0 && nb >= 0 && p.size() == 7 + 3 * (size_t)(na + nb), "Inconsistent params for MeshContactStateForce.");
    MeshContactForce m;
    m.radius = p[0]; m.kn_l = p[1]; m.kn_r = p[2]; m.kf = p[3]; m.mu = p[4];
    std::set<int> sa, sb;
    for (int f = 0; f < na + nb; ++f) {
        std::array<int, 3> t = {(int)p[7 + 3 * f], (int)p[8 + 3 * f], (int)p[9 + 3 * f]};
        for (int k : t) check(k >= 0 && k < nv_, "Mesh contact face index out of range.");
        (f < na ? m.fa : m.fb).push_back(t);
        (f < na ? sa : sb).insert(t.begin(), t.end());
    }
    m.va.assign(sa.begin(), sa.end());
    m.vb.assign(sb.begin(), sb.end());
    m.poff = n_state_params_;
    n_state_params_ += 4;
    mcontacts_.push_back(m);
}

// BilliardBallStateForce::Initialize (billiard_ball_state_force.cpp:6-18): params
// [radius, vertices per ball, stiffness per ball, friction coefficient per ball]
// (deformable_state_force.cpp:72-85); the body's vertices are the balls in order.
void Body::add_billiard_ball(const std::vector<double>& p) {
    check(p.size() >= 2 && p.size() % 2 == 0, "Inconsistent params for BilliardBallStateForce.");
    BilliardBallForce b;
    b.R = p[0];
    b.svn = (int)p[1];
    b.nb = (int)(p.size() - 2) / 2;
    check(b.svn > 0 && nv_ % b.svn == 0 && nv_ / b.svn == b.nb, "Inconsistent ball_num.");
    b.k.assign(p.begin() + 2, p.begin() + 2 + b.nb);
    b.mu.assign(p.begin() + 2 + b.nb, p.end());
    b.poff = n_state_params_;
    n_state_params_ += 2 * b.nb;
    bballs_.push_back(b);
}

// HydrodynamicsStateForce::Initialize (hydrodynamics_state_force.cpp:11-42): params
// [rho, v_water (3), Cd points (8), Ct points (8), max_force, faces]
// (deformable_state_force.cpp:50-71); faces are quads (hex) or triangles (tet).
void Body::add_hydrodynamics(const std::vector<double>& p) {
    const int fd = edim_ == 8 ? 4 : 3;
    check(p.size() >= 21 + (size_t)fd && (p.size() - 21) % fd == 0, "Inconsistent params for HydrodynamicsStateForce.");
    HydroForce h;
    h.rho = p[0];
    std::copy(p.begin() + 1, p.begin() + 4, h.vw);
    std::copy(p.begin() + 4, p.begin() + 20, h.P);
    h.fmax = p[20];
    h.fd = fd;
    for (size_t i = 21; i < p.size(); ++i) {
        h.faces.push_back((int)p[i]);
        check(h.faces.back() >= 0 && h.faces.back() < nv_, "Hydrodynamics face index out of range.");
    }
    h.poff = n_state_params_;
    n_state_params_ += 16;
    hydros_.push_back(h);
}

namespace {
inline double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline void cross3(const double* a, const double* b, double* c) {
    c[0] = a[1] * b[2] - a[2] * b[1]; c[1] = a[2] * b[0] - a[0] * b[2]; c[2] = a[0] * b[1] - a[1] * b[0];
}

// Triangles of a hydrodynamic face, by node slot: a quad is 012, 023, 013, 312
// (hydrodynamics_state_force.cpp:618-623), a triangle 012.
constexpr int kHydroTris[4][3] = {{0, 1, 2}, {0, 2, 3}, {0, 1, 3}, {3, 1, 2}};

// The state of one face (hydrodynamics_state_force.cpp:457-499, 614-673): triangle
// vectors c_t = (b - a) x (c - b), area A = mean |c_t| (quad) or |c_0| (triangle),
// un = sum c_t, n = un / |un|, relative water velocity w = v_water - mean node velocity,
// and the angle of attack phi = acos(n . w/|w|) - pi/2. False when the face takes no force.
template <class H>
struct HydroFaceState {
    const int* fc;
    double ct[4][3], A = 0, un[3] = {0, 0, 0}, unn = 0, n[3] = {0, 0, 0}, w[3], wl = 0, d[3], nd = 0, phi = 0;
    int nt = 1;
    bool eval(const H& h, const double* q, const double* v, const int* f) {
        fc = f;
        nt = h.fd == 4 ? 4 : 1;
        double sum = 0;
        for (int t = 0; t < nt; ++t) {
            const double *a = q + 3 * (size_t)f[kHydroTris[t][0]], *b = q + 3 * (size_t)f[kHydroTris[t][1]],
                         *c = q + 3 * (size_t)f[kHydroTris[t][2]];
            const double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, e2[3] = {c[0] - b[0], c[1] - b[1], c[2] - b[2]};
            cross3(e1, e2, ct[t]);
            for (int k = 0; k < 3; ++k) un[k] += ct[t][k];
            sum += std::sqrt(dot3(ct[t], ct[t]));
        }
        A = h.fd == 4 ? sum / 4 : sum;
        if (A <= DBL_EPSILON) return false;
        unn = std::sqrt(dot3(un, un));
        if (unn > DBL_EPSILON) for (int k = 0; k < 3; ++k) n[k] = un[k] / unn;
        for (int k = 0; k < 3; ++k) {
            double m = 0;
            for (int j = 0; j < h.fd; ++j) m += v[3 * (size_t)f[j] + k];
            w[k] = h.vw[k] - (h.fd == 4 ? m * 0.25 : m / 3);
        }
        wl = std::sqrt(dot3(w, w));
        if (wl <= DBL_EPSILON) return false;
        for (int k = 0; k < 3; ++k) d[k] = w[k] / wl;
        nd = dot3(n, d);
        if (nd >= 0.0) return false;
        phi = std::acos(nd) - M_PI / 2;
        return true;
    }
};

// Cd (t = |phi| / (pi/2)) and Ct (t = phi / pi + 1/2): cubic Bezier curves through the
// y values of the four points (hydrodynamics_state_force.cpp:101-162); value,
// derivative in phi and the weights of the y values.
inline void hydro_curve(const double* P, bool drag, double phi, double& val, double& dphi, double* wts) {
    const double tn = phi / (M_PI / 2);
    const double t = drag ? std::fabs(tn) : 0.5 * tn + 0.5;
    const double dt = drag ? (tn < 0 ? -2 / M_PI : 2 / M_PI) : 1 / M_PI;
    const double y0 = P[1], y1 = P[3], y2 = P[5], y3 = P[7], s = 1 - t;
    wts[0] = s * s * s; wts[1] = 3 * s * s * t; wts[2] = 3 * s * t * t; wts[3] = t * t * t;
    val = wts[0] * y0 + wts[1] * y1 + wts[2] * y2 + wts[3] * y3;
    dphi = 3 * dt * (-s * s * y0 + (s * s - 2 * s * t) * y1 + (2 * s * t - t * t) * y2 + t * t * y3);
}

// The face force F = 1/2 rho A (Cd |w| w - Ct |w|^2 n), F / fd on each node, then every
// entry clamped to [-max_force, max_force] (this force's own vector).
template <class H>
std::vector<double> hydro_forward(const H& h, int dofs, const double* q, const double* v) {
    std::vector<double> f(dofs, 0.0);
    const int nf = (int)h.faces.size() / h.fd;
    for (int i = 0; i < nf; ++i) {
        HydroFaceState<H> S;
        if (!S.eval(h, q, v, &h.faces[(size_t)i * h.fd])) continue;
        double cd, dcd, wcd[4], ct, dct, wct[4];
        hydro_curve(h.P, true, S.phi, cd, dcd, wcd);
        hydro_curve(h.P + 8, false, S.phi, ct, dct, wct);
        for (int k = 0; k < 3; ++k) {
            const double F = 0.5 * h.rho * S.A * cd * S.wl * S.w[k] - 0.5 * h.rho * S.A * ct * S.wl * S.wl * S.n[k];
            const double fn = h.fd == 4 ? F * 0.25 : F / 3;
            for (int j = 0; j < h.fd; ++j) f[3 * (size_t)S.fc[j] + k] += fn;
        }
    }
    for (auto& x : f) x = std::min(std::max(x, -h.fmax), h.fmax);
    return f;
}

// Per-ball means of x (centres of q, centre velocities of v).
template <class B>
std::vector<std::array<double, 3>> ball_means(const B& b, const double* x) {
    std::vector<std::array<double, 3>> c(b.nb, std::array<double, 3>{0, 0, 0});
    for (int i = 0; i < b.nb; ++i) {
        for (int p = 0; p < b.svn; ++p)
            for (int k = 0; k < 3; ++k) c[i][k] += x[3 * ((size_t)i * b.svn + p) + k];
        for (int k = 0; k < 3; ++k) c[i][k] /= b.svn;
    }
    return c;
}

// One pair i < j of balls (billiard_ball_state_force.cpp:51-84): unit u from
// centre i to centre j, overlap 2R - dist (clamped at 0), and the friction
// direction for ball i, u turned by +-90 degrees in x-y by the sign of the z part of
// (v_i - v_j) x u. Returns dist.
template <class B>
double ball_pair(const B& b, const std::array<double, 3>& ci, const std::array<double, 3>& cj,
                 const std::array<double, 3>& vi, const std::array<double, 3>& vj, double* u, double& pen, double* rot,
                 double& sg) {
    for (int k = 0; k < 3; ++k) u[k] = cj[k] - ci[k];
    const double dist = std::sqrt(dot3(u, u));
    for (int k = 0; k < 3; ++k) u[k] /= dist;
    pen = std::max(2 * b.R - dist, 0.0);
    sg = (vi[0] - vj[0]) * u[1] - (vi[1] - vj[1]) * u[0] > 0 ? 1.0 : -1.0;
    rot[0] = -sg * u[1]; rot[1] = sg * u[0]; rot[2] = 0.0;
    return dist;
}

// MeshContactStateForce::ForwardForce (mesh_contact_state_force.cpp:144-215), added
// into f: nothing when the two surfaces' bounding boxes are more than the radius
// apart; else, for each vertex of one surface and triangle of the other closer than
// the radius, kn (radius - dist) along the separating normal plus the friction
// -min(kf, mu f_n / (|v_t| + eps)) v_t of the velocity relative to the triangle's
// point (when |v_t| > eps), on the vertex and, barycentrically, against the corners.
template <class M>
void mesh_contact_add(const M& m, int dofs, const double* q, const double* v, double* f) {
    if (m.fa.empty() || m.fb.empty()) return;
    const double r2 = m.radius * m.radius;
    double lo[2][3], hi[2][3];
    for (int s = 0; s < 2; ++s) {
        for (int k = 0; k < 3; ++k) { lo[s][k] = INFINITY; hi[s][k] = -INFINITY; }
        for (int i : (s == 0 ? m.va : m.vb))
            for (int k = 0; k < 3; ++k) {
                lo[s][k] = std::min(lo[s][k], q[3 * (size_t)i + k]);
                hi[s][k] = std::max(hi[s][k], q[3 * (size_t)i + k]);
            }
    }
    double gap2 = 0;
    for (int k = 0; k < 3; ++k) {
        const double g = hi[0][k] < lo[1][k] ? lo[1][k] - hi[0][k] : (hi[1][k] < lo[0][k] ? lo[0][k] - hi[1][k] : 0.0);
        gap2 += g * g;
    }
    if (gap2 > r2) return;
    std::vector<double> fm(dofs, 0.0);  // this force's own vector, added to f at the end
    auto sweep = [&](const std::vector<int>& verts, const std::vector<std::array<int, 3>>& tris, double kn) {
        for (int i : verts) {
            const double* p = q + 3 * (size_t)i;
            for (const auto& t : tris) {
                const double *a = q + 3 * (size_t)t[0], *b = q + 3 * (size_t)t[1], *c = q + 3 * (size_t)t[2];
                double x[3], w[3];
                closest_on_triangle(p, a, b, c, x, w);  // the method of mesh_contact_state_force.cpp:44-97
                const double d2 = (p[0] - x[0]) * (p[0] - x[0]) + (p[1] - x[1]) * (p[1] - x[1]) + (p[2] - x[2]) * (p[2] - x[2]);
                if (d2 >= r2) continue;
                const double dist = std::sqrt(d2);
                double n[3], px[3];
                for (int k = 0; k < 3; ++k) px[k] = p[k] - x[k];
                if (dist > DBL_EPSILON) {
                    for (int k = 0; k < 3; ++k) n[k] = px[k] / dist;
                } else {  // on the triangle: its normal
                    const double u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, z[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
                    for (int k = 0; k < 3; ++k) n[k] = u[(k + 1) % 3] * z[(k + 2) % 3] - u[(k + 2) % 3] * z[(k + 1) % 3];
                    const double nn = std::sqrt(dot3(n, n));
                    if (nn <= DBL_EPSILON) continue;
                    for (int k = 0; k < 3; ++k) n[k] /= nn;
                }
                if (dot3(n, px) < 0) for (int k = 0; k < 3; ++k) n[k] = -n[k];
                const double fn = kn * (m.radius - dist);
                double vr[3];
                for (int k = 0; k < 3; ++k)
                    vr[k] = v[3 * (size_t)i + k] - (w[0] * v[3 * (size_t)t[0] + k] + w[1] * v[3 * (size_t)t[1] + k] +
                                                    w[2] * v[3 * (size_t)t[2] + k]);
                const double vn = dot3(vr, n);
                double vt[3];
                for (int k = 0; k < 3; ++k) vt[k] = vr[k] - n[k] * vn;
                const double vtm = std::sqrt(dot3(vt, vt));
                const double cf = vtm > DBL_EPSILON ? std::min(m.kf, m.mu * fn / (vtm + DBL_EPSILON)) : 0.0;
                for (int k = 0; k < 3; ++k) {
                    const double tot = fn * n[k] - cf * vt[k];
                    fm[3 * (size_t)i + k] += tot;
                    for (int j = 0; j < 3; ++j) fm[3 * (size_t)t[j] + k] -= tot * w[j];
                }
            }
        }
    };
    sweep(m.va, m.fb, m.kn_l);
    sweep(m.vb, m.fa, m.kn_r);
    for (int i = 0; i < dofs; ++i) f[i] += fm[i];
}
// A vertex x against an arc (arc_contact_state_force.cpp:32-62): qp is x - c
// without its axial part; contact when the angle of qp from `start` toward `aux`
// lies in (0, angle). Returns |qp| and n = -qp / (|qp| + eps), or false.
template <class Arc>
bool arc_hit(const Arc& a, const double* x, double* qp, double& r, double* n) {
    double ql[3];
    for (int k = 0; k < 3; ++k) ql[k] = x[k] - a.c[k];
    const double ax = dot3(ql, a.dir);
    for (int k = 0; k < 3; ++k) qp[k] = ql[k] - ax * a.dir[k];
    const double th = std::atan2(dot3(qp, a.aux), dot3(qp, a.start));
    if (!(th > 0 && th < a.angle)) return false;
    r = std::sqrt(dot3(qp, qp));
    for (int k = 0; k < 3; ++k) n[k] = -qp[k] / (r + DBL_EPSILON);
    return true;
}
}  // namespace

// ForwardStateForce (deformable_state_force.cpp:114-125): gravity m g
// (gravitational_state_force.cpp:14-22) plus each planar contact's penalty
// kn (-min(0, d))^(p-1) n and friction -min(kf, mu f_n / (|u_s| + eps)) u_s
// (planar_contact_state_force.cpp:16-37), evaluated at the frame start.
void Body::state_force(const double* q, const double* v, double* f) const {
    const double mass = element_volume_ * density_;
    for (int i = 0; i < dofs_; ++i) f[i] = gravity_.size() == 3 ? mass * gravity_[i % 3] : 0.0;
    for (const auto& c : pcontacts_)
        for (int i = 0; i < nv_; ++i) {
            const double* x = q + 3 * (size_t)i;
            const double* vi = v + 3 * (size_t)i;
            const double d = c.n[0] * x[0] + c.n[1] * x[1] + c.n[2] * x[2] + c.off;
            const double fn = c.kn * std::pow(-std::min(0.0, d), c.p - 1);
            const double vn = c.n[0] * vi[0] + c.n[1] * vi[1] + c.n[2] * vi[2];
            double us[3];
            for (int k = 0; k < 3; ++k) us[k] = vi[k] - vn * c.n[k];
            const double um = std::sqrt(us[0] * us[0] + us[1] * us[1] + us[2] * us[2]);
            const double cf = std::min(c.kf, c.mu * fn / (um + DBL_EPSILON));
            for (int k = 0; k < 3; ++k) f[3 * (size_t)i + k] += fn * c.n[k] - cf * us[k];
        }
    for (const auto& a : acontacts_)
        for (int i = 0; i < nv_; ++i) {
            double qp[3], r, n[3];
            if (!arc_hit(a, q + 3 * (size_t)i, qp, r, n)) continue;
            const double fn = a.kn * std::pow(-std::min(0.0, a.R - r), a.p - 1);
            const double* vi = v + 3 * (size_t)i;
            const double vn = dot3(vi, n);
            double us[3];
            for (int k = 0; k < 3; ++k) us[k] = vi[k] - vn * n[k];
            const double cf = std::min(a.kf, a.mu * fn / (std::sqrt(dot3(us, us)) + DBL_EPSILON));
            for (int k = 0; k < 3; ++k) f[3 * (size_t)i + k] += fn * n[k] - cf * us[k];
        }
    // mesh contacts: forward only; their BackwardForce returns zeros
    // (mesh_contact_state_force.cpp:217-223), so they add nothing to the adjoint
    for (const auto& m : mcontacts_) mesh_contact_add(m, dofs_, q, v, f);
    for (const auto& h : hydros_) {
        const std::vector<double> fh = hydro_forward(h, dofs_, q, v);
        for (int i = 0; i < dofs_; ++i) f[i] += fh[i];
    }
    // billiard balls (billiard_ball_state_force.cpp:20-86): ball i gets
    // k_i pen (-u + mu_i rot), ball j gets k_j pen (u - mu_j rot), spread over their vertices
    for (const auto& b : bballs_) {
        const auto c = ball_means(b, q), vc = ball_means(b, v);
        std::vector<double> fb(dofs_, 0.0);
        for (int i = 0; i < b.nb; ++i)
            for (int j = i + 1; j < b.nb; ++j) {
                double u[3], rot[3], pen, sg;
                ball_pair(b, c[i], c[j], vc[i], vc[j], u, pen, rot, sg);
                const double fi = pen * b.k[i], fj = pen * b.k[j];
                for (int k = 0; k < 3; ++k) {
                    const double Fi = (-u[k] * fi + rot[k] * fi * b.mu[i]) / b.svn;
                    const double Fj = (u[k] * fj - rot[k] * fj * b.mu[j]) / b.svn;
                    for (int p = 0; p < b.svn; ++p) {
                        fb[3 * ((size_t)i * b.svn + p) + k] += Fi;
                        fb[3 * ((size_t)j * b.svn + p) + k] += Fj;
                    }
                }
            }
        for (int i = 0; i < dofs_; ++i) f[i] += fb[i];
    }
}

// BackwardStateForce (deformable_state_force.cpp:127-145): dl/dq, dl/dv and
// dl/dp of a given dl/df (planar_contact_state_force.cpp:39-84, with the
// reference's d|u_s|/dv = P u_s / (|u_s| + eps); gravity: dl/dg = m sum dl/df).
void Body::state_force_backward(const double* q, const double* v, const double* dl_df, double* dl_dq, double* dl_dv,
                                double* dl_dp) const {
    for (int i = 0; i < dofs_; ++i) { dl_dq[i] = 0; dl_dv[i] = 0; }
    for (int k = 0; k < n_state_params_; ++k) dl_dp[k] = 0;
    const double mass = element_volume_ * density_;
    if (gravity_.size() == 3) for (int i = 0; i < dofs_; ++i) dl_dp[gravity_off_ + i % 3] += mass * dl_df[i];
    for (const auto& c : pcontacts_)
        for (int i = 0; i < nv_; ++i) {
            const double* x = q + 3 * (size_t)i;
            const double* vi = v + 3 * (size_t)i;
            const double* g = dl_df + 3 * (size_t)i;
            const double d = c.n[0] * x[0] + c.n[1] * x[1] + c.n[2] * x[2] + c.off;
            const double pen = -std::min(0.0, d);
            const double fn = c.kn * std::pow(pen, c.p - 1);
            const double dfn = d < 0 ? -c.kn * (c.p - 1) * std::pow(pen, c.p - 2) : 0.0;  // d f_n / d(n . q)
            const double vn = c.n[0] * vi[0] + c.n[1] * vi[1] + c.n[2] * vi[2];
            double us[3];
            for (int k = 0; k < 3; ++k) us[k] = vi[k] - vn * c.n[k];
            const double ue = std::sqrt(us[0] * us[0] + us[1] * us[1] + us[2] * us[2]) + DBL_EPSILON;
            const double gn = g[0] * c.n[0] + g[1] * c.n[1] + g[2] * c.n[2];
            const double gu = g[0] * us[0] + g[1] * us[1] + g[2] * us[2];
            double gt[3];
            for (int k = 0; k < 3; ++k) gt[k] = g[k] - gn * c.n[k];
            for (int k = 0; k < 3; ++k) dl_dq[3 * (size_t)i + k] += gn * dfn * c.n[k];
            dl_dp[c.poff] += gn * fn / c.kn;
            if (c.kf <= c.mu * fn / ue) {
                for (int k = 0; k < 3; ++k) dl_dv[3 * (size_t)i + k] += -c.kf * gt[k];
                dl_dp[c.poff + 1] += -gu;
            } else {
                for (int k = 0; k < 3; ++k) {
                    dl_dq[3 * (size_t)i + k] += -c.mu * gu / ue * dfn * c.n[k];
                    dl_dv[3 * (size_t)i + k] += -c.mu * fn * (gt[k] / ue - us[k] * gu / (ue * ue * ue));
                }
                dl_dp[c.poff] += -c.mu * gu / ue * (fn / c.kn);
                dl_dp[c.poff + 2] += -gu / ue * fn;
            }
        }
    // arcs (arc_contact_state_force.cpp:65-131) as vector-Jacobian products, with
    // P = I - dir dir^T, s = 1 / (|qp| + eps) and the reference's dn/dq = -s (I - n n^T) P,
    // so (dn/dq)^T w = -s P (w - (n . w) n)
    for (const auto& a : acontacts_)
        for (int i = 0; i < nv_; ++i) {
            double qp[3], r, n[3];
            if (!arc_hit(a, q + 3 * (size_t)i, qp, r, n)) continue;
            const double* vi = v + 3 * (size_t)i;
            const double* g = dl_df + 3 * (size_t)i;
            const double sc = 1.0 / (r + DBL_EPSILON), d = a.R - r, pen = -std::min(0.0, d);
            const double fn = a.kn * std::pow(pen, a.p - 1);
            auto proj_axis = [&](const double* w, double* out) {  // P w
                const double t = dot3(w, a.dir);
                for (int k = 0; k < 3; ++k) out[k] = w[k] - t * a.dir[k];
            };
            auto dn_T = [&](const double* w, double* out) {  // (dn/dq)^T w
                const double t = dot3(n, w);
                double u[3];
                for (int k = 0; k < 3; ++k) u[k] = w[k] - t * n[k];
                proj_axis(u, out);
                for (int k = 0; k < 3; ++k) out[k] *= -sc;
            };
            // d f_n / dq = kn (p-1) pen^(p-2) P qp / (|qp| + eps) inside the arc's radius
            double dfn[3];
            proj_axis(qp, dfn);
            const double c1 = d < 0 ? a.kn * (a.p - 1) * std::pow(pen, a.p - 2) * sc : 0.0;
            for (int k = 0; k < 3; ++k) dfn[k] *= c1;
            const double vn = dot3(vi, n), gn = dot3(g, n);
            double us[3];
            for (int k = 0; k < 3; ++k) us[k] = vi[k] - vn * n[k];
            const double ue = std::sqrt(dot3(us, us)) + DBL_EPSILON, gu = dot3(g, us);
            // normal part f_n n
            double tq[3];
            dn_T(g, tq);
            for (int k = 0; k < 3; ++k) dl_dq[3 * (size_t)i + k] += gn * dfn[k] + fn * tq[k];
            dl_dp[a.poff] += gn * fn / a.kn;
            // (du_s/dq)^T g = -(dn/dq)^T (v (n . g) + (v . n) g); du_s/dv = I - n n^T
            double w[3], dusT_g[3], gt[3];
            for (int k = 0; k < 3; ++k) w[k] = vi[k] * gn + vn * g[k];
            dn_T(w, dusT_g);
            for (int k = 0; k < 3; ++k) { dusT_g[k] = -dusT_g[k]; gt[k] = g[k] - gn * n[k]; }
            if (a.kf <= a.mu * fn / ue) {  // ff = -kf u_s
                for (int k = 0; k < 3; ++k) {
                    dl_dq[3 * (size_t)i + k] += -a.kf * dusT_g[k];
                    dl_dv[3 * (size_t)i + k] += -a.kf * gt[k];
                }
                dl_dp[a.poff + 1] += -gu;
            } else {  // ff = -mu f_n u_s / (|u_s| + eps)
                // the reference's d|u_s|/dq is (du_s/dq) u_s / (|u_s| + eps) (not transposed):
                // du_s/dq x = -(n (v . y) + (v . n) y) with y = (dn/dq) x = -s (I - n n^T) P x
                double pu[3], y[3], dum_q[3], dum_v[3];
                proj_axis(us, pu);
                const double npu = dot3(n, pu);
                for (int k = 0; k < 3; ++k) y[k] = -sc * (pu[k] - npu * n[k]);
                const double vy = dot3(vi, y), nus = dot3(n, us);
                for (int k = 0; k < 3; ++k) {
                    dum_q[k] = -(n[k] * vy + vn * y[k]) / ue;
                    dum_v[k] = (us[k] - nus * n[k]) / ue;
                }
                for (int k = 0; k < 3; ++k) {
                    dl_dq[3 * (size_t)i + k] += -a.mu * (dfn[k] * gu / ue + fn * (dusT_g[k] / ue - dum_q[k] * gu / (ue * ue)));
                    dl_dv[3 * (size_t)i + k] += -a.mu * fn * (gt[k] / ue - dum_v[k] * gu / (ue * ue));
                }
                dl_dp[a.poff] += -gu / ue * a.mu * fn / a.kn;
                dl_dp[a.poff + 2] += -gu / ue * fn;
            }
        }
    // billiard balls (billiard_ball_state_force.cpp:88-229): through the centres only
    // (dl/dv = 0: the friction direction is piecewise constant in v). The reference
    // sizes the friction magnitude's parameter gradient by 2 svn instead of 2 ball_num
    // (:165-166), so its backward is defined only when they agree; other layouts raise.
    for (const auto& b : bballs_) {
        check(b.svn == b.nb, "BilliardBallStateForce backward needs as many vertices per ball as balls.");
        const auto c = ball_means(b, q), vc = ball_means(b, v);
        for (int i = 0; i < b.nb; ++i)
            for (int j = i + 1; j < b.nb; ++j) {
                double u[3], rot[3], pen, sg;
                const double dist = ball_pair(b, c[i], c[j], vc[i], vc[j], u, pen, rot, sg);
                check(dist > 0.1 * b.R, "Balls are too close to each other");
                const double fi = pen * b.k[i], fj = pen * b.k[j];
                double gi[3] = {0, 0, 0}, gj[3] = {0, 0, 0};  // each ball's mean dl/df
                for (int p = 0; p < b.svn; ++p)
                    for (int k = 0; k < 3; ++k) {
                        gi[k] += dl_df[3 * ((size_t)i * b.svn + p) + k] / b.svn;
                        gj[k] += dl_df[3 * ((size_t)j * b.svn + p) + k] / b.svn;
                    }
                // w = weights of u and rot in g_i . F_i + g_j . F_j:
                // F_i = fi (-u + mu_i rot), F_j = fj (u - mu_j rot)
                const double ai = fi, aj = fj;
                double wu[3], wr[3];
                for (int k = 0; k < 3; ++k) {
                    wu[k] = -ai * gi[k] + aj * gj[k];
                    wr[k] = ai * b.mu[i] * gi[k] - aj * b.mu[j] * gj[k];
                }
                // d u / d c_j = (I - u u^T) / dist; d rot / d c_j = sg (-row 1, row 0, 0) of it;
                // d f / d dist = -k when the balls overlap
                const double wrr[3] = {sg * wr[1], -sg * wr[0], 0.0};  // (d rot/du)^T wr
                double t[3];
                for (int k = 0; k < 3; ++k) t[k] = wu[k] + wrr[k];
                const double tu = dot3(t, u);
                const double dfi = fi > 0 ? -b.k[i] : 0.0, dfj = fj > 0 ? -b.k[j] : 0.0;
                const double mag = dfi * (-dot3(gi, u) + b.mu[i] * dot3(gi, rot)) + dfj * (dot3(gj, u) - b.mu[j] * dot3(gj, rot));
                double dcj[3];
                for (int k = 0; k < 3; ++k) dcj[k] = (t[k] - tu * u[k]) / dist + mag * u[k];
                for (int p = 0; p < b.svn; ++p)
                    for (int k = 0; k < 3; ++k) {
                        dl_dq[3 * ((size_t)i * b.svn + p) + k] -= dcj[k] / b.svn;
                        dl_dq[3 * ((size_t)j * b.svn + p) + k] += dcj[k] / b.svn;
                    }
                dl_dp[b.poff + i] += pen * (-dot3(gi, u) + b.mu[i] * dot3(gi, rot));
                dl_dp[b.poff + j] += pen * (dot3(gj, u) - b.mu[j] * dot3(gj, rot));
                dl_dp[b.poff + b.nb + i] += fi * dot3(gi, rot);
                dl_dp[b.poff + b.nb + j] += -fj * dot3(gj, rot);
            }
    }
    // hydrodynamics (hydrodynamics_state_force.cpp:501-612, 675-829) as a scalar
    // chain: with G = the clamp-masked dl/df summed over a face's nodes / fd,
    // L = G . F = r A (Cd |w| (G.w) - Ct |w|^2 (G.n)), r = rho / 2; A, n and phi carry it
    // to the nodes' positions through the triangle vectors, |w| and w to their velocities.
    for (const auto& h : hydros_) {
        const std::vector<double> fh = hydro_forward(h, dofs_, q, v);
        const int nf = (int)h.faces.size() / h.fd;
        const double inv_fd = h.fd == 4 ? 0.25 : 1.0 / 3.0, r = 0.5 * h.rho;
        for (int i = 0; i < nf; ++i) {
            HydroFaceState<HydroForce> S;
            if (!S.eval(h, q, v, &h.faces[(size_t)i * h.fd])) continue;
            double G[3] = {0, 0, 0};
            for (int j = 0; j < h.fd; ++j)
                for (int k = 0; k < 3; ++k) {
                    const size_t e = 3 * (size_t)S.fc[j] + k;
                    if (fh[e] != h.fmax && fh[e] != -h.fmax) G[k] += dl_df[e];
                }
            for (int k = 0; k < 3; ++k) G[k] *= inv_fd;
            double cd, dcd, wcd[4], ct, dct, wct[4];
            hydro_curve(h.P, true, S.phi, cd, dcd, wcd);
            hydro_curve(h.P + 8, false, S.phi, ct, dct, wct);
            const double Gw = dot3(G, S.w), Gn = dot3(G, S.n), wl = S.wl, w2 = wl * wl;
            const double L_A = r * (cd * wl * Gw - ct * w2 * Gn);
            const double L_phi = r * S.A * (dcd * wl * Gw - dct * w2 * Gn);
            const double dacos = -1 / std::sqrt(1 - S.nd * S.nd);
            // dL/dn (direct, thrust) + through phi (dphi/dn = dacos d)
            double Ln[3], Lun[3] = {0, 0, 0};
            for (int k = 0; k < 3; ++k) Ln[k] = -r * S.A * ct * w2 * G[k] + L_phi * dacos * S.d[k];
            if (S.unn > DBL_EPSILON) {
                const double nl = dot3(S.n, Ln);
                for (int k = 0; k < 3; ++k) Lun[k] = (Ln[k] - nl * S.n[k]) / S.unn;
            }
            // each triangle vector c_t = (b - a) x (c - b): (dc_t/da)^T x = x x (c - b),
            // (dc_t/db)^T x = x x (a - c), (dc_t/dc)^T x = x x (b - a); x = dL/dun + dL/dA dA/dc_t
            const double cA = h.fd == 4 ? 0.25 : 1.0;
            for (int t = 0; t < S.nt; ++t) {
                const int sl[3] = {kHydroTris[t][0], kHydroTris[t][1], kHydroTris[t][2]};
                const double *a = q + 3 * (size_t)S.fc[sl[0]], *b = q + 3 * (size_t)S.fc[sl[1]], *c = q + 3 * (size_t)S.fc[sl[2]];
                const double cn = std::sqrt(dot3(S.ct[t], S.ct[t]));
                double x[3];
                for (int k = 0; k < 3; ++k) x[k] = Lun[k] + (cn > DBL_EPSILON ? L_A * cA * S.ct[t][k] / cn : 0.0);
                const double e[3][3] = {{c[0] - b[0], c[1] - b[1], c[2] - b[2]},
                                        {a[0] - c[0], a[1] - c[1], a[2] - c[2]},
                                        {b[0] - a[0], b[1] - a[1], b[2] - a[2]}};
                for (int m = 0; m < 3; ++m) {
                    double y[3];
                    cross3(x, e[m], y);
                    for (int k = 0; k < 3; ++k) dl_dq[3 * (size_t)S.fc[sl[m]] + k] += y[k];
                }
            }
            // velocities: dw/dv_m = -I / fd; d|w| = d . dw; dphi/dw = dacos (n - (n.d) d) / |w|
            double Lw[3];
            for (int k = 0; k < 3; ++k)
                Lw[k] = r * S.A * (cd * (wl * G[k] + Gw * S.d[k]) - ct * 2 * wl * Gn * S.d[k]) +
                        L_phi * dacos * (S.n[k] - S.nd * S.d[k]) / wl;
            for (int m = 0; m < h.fd; ++m)
                for (int k = 0; k < 3; ++k) dl_dv[3 * (size_t)S.fc[m] + k] += -inv_fd * Lw[k];
            for (int j = 0; j < 4; ++j) {
                dl_dp[h.poff + 2 * j + 1] += r * S.A * wl * Gw * wcd[j];
                dl_dp[h.poff + 8 + 2 * j + 1] += -r * S.A * w2 * Gn * wct[j];
            }
        }
    }
}

// SetFrictionalBoundary, deformable_friction.cpp:7-89, plus heightfield/press.
void Body::set_boundary(const std::string& type, const std::vector<double>& p, const std::vector<int>& idx) {
    ready_ = false;
    HostBoundary b;
    if (type == "planar") {
        check(p.size() == 4, "Incompatible parameter number.");
        const double nn = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
        check(nn > 2.220446049250313e-16, "Singular normal.");
        b.type = 1;
        for (int i = 0; i < 3; ++i) b.n[i] = p[i] / nn;
        b.off = p[3] / nn;
    } else if (type == "spherical") {
        check(p.size() == 4, "Incompatible parameter number.");
        b.type = 2;
        for (int i = 0; i < 3; ++i) b.c[i] = p[i];
        b.r = p[3];
    } else if (type == "heightfield" || type == "press") {
        check(p.size() >= 2, "Incompatible parameter number.");
        const int nw = (int)p[1];
        const size_t need = 2 + 4 * (size_t)nw + (type == "press" ? 1 : 0);
        check(p.size() == need, "Incompatible parameter number.");
        b.type = type == "press" ? 4 : 3;
        b.z0 = p[0];
        b.waves.assign(p.begin() + 2, p.begin() + 2 + 4 * nw);
        if (type == "press") b.plane_z = p[2 + 4 * nw];
    } else if (type == "mesh") {
        // deformable_friction.cpp:53-73: [radius, split, n_left, n_right, faces_left, faces_right]
        check(p.size() >= 4, "Incompatible parameter number.");
        b.type = 5;
        b.radius = p[0];
        b.split = (int)p[1];
        b.nfl = (int)p[2];
        b.nfr = (int)p[3];
        check(b.nfl >= 0 && b.nfr >= 0 && p.size() == 4 + 3 * (size_t)(b.nfl + b.nfr),
              "Inconsistent params for MeshFrictionalBoundary.");
        b.faces.resize(3 * (size_t)(b.nfl + b.nfr));
        for (size_t i = 0; i < b.faces.size(); ++i) {
            b.faces[i] = (int)p[4 + i];
            check(0 <= b.faces[i] && b.faces[i] < nv_, "Mesh boundary face vertex out of bound.");
        }
    } else {
        throw std::runtime_error("Unsupported frictional boundary type: " + type);
    }
    std::set<int> uniq(idx.begin(), idx.end());
    check(uniq.size() == idx.size(), "Duplicated vertex elements.");
    for (int i : uniq) check(0 <= i && i < nv_, "Frictional boundary vertex out of bound.");
    boundary_ = b;
    fb_vertices_.clear();
    int cnt = 0;
    for (int i : uniq) fb_vertices_[i] = cnt++;
}

DevBoundary Body::dev_boundary() {
    DevBoundary b;
    b.type = boundary_.type;
    for (int i = 0; i < 3; ++i) { b.n[i] = boundary_.n[i]; b.c[i] = boundary_.c[i]; }
    b.off = boundary_.off;
    b.r = boundary_.r;
    b.z0 = boundary_.z0;
    b.plane_z = boundary_.plane_z;
    b.nwaves = (int)boundary_.waves.size() / 4;
    b.waves = d_waves_.p;
    b.split = boundary_.split;
    b.nfl = boundary_.nfl;
    b.nfr = boundary_.nfr;
    b.faces = d_mesh_faces_.p;
    b.radius = boundary_.radius;
    return b;
}

long long Body::factor_bytes() const {
    return (long long)(V64_.n * sizeof(double) + V32_.n * sizeof(float));
}

// ---------------------------------------------------------------------------
// Per-element weights of the merged energies (local step and K_pd).
// ---------------------------------------------------------------------------
void Body::build_materials() {
    const double vs = element_volume_ / samples_;
    std::vector<double> wc(ne_, 0.0), wv(ne_, 0.0), kc(ne_, 0.0), kv(ne_, 0.0);
    has_cor_ = has_vol_ = false;
    n_nh_ = 0;
    nh_energy_.clear();
    std::vector<std::vector<double>> wn, kn;
    for (int ei = 0; ei < (int)energies_.size(); ++ei) {
        const Energy& en = energies_[ei];
        if (en.type == kNeoHookean) {
            std::vector<double> w(ne_, 0.0), k(ne_, 0.0);
            for (int e = 0; e < ne_; ++e)
                if (en.member[e]) { w[e] = en.elem_w(e) * vs; k[e] = (kpd_exact_ ? en.elem_w(e) : en.stiffness) * vs; }
            wn.push_back(w);
            kn.push_back(k);
            nh_energy_.push_back(ei);
            continue;
        }
        auto& w = en.type == kCorotated ? wc : wv;
        auto& k = en.type == kCorotated ? kc : kv;
        (en.type == kCorotated ? has_cor_ : has_vol_) = true;
        for (int e = 0; e < ne_; ++e)
            if (en.member[e]) { w[e] += en.elem_w(e) * vs; k[e] += (kpd_exact_ ? en.elem_w(e) : en.stiffness) * vs; }
    }
    n_nh_ = (int)wn.size();
    if (has_cor_) { d_wc_.upload(wc); d_kc_.upload(kc); } else { d_wc_.release(); d_kc_.release(); }
    if (has_vol_) { d_wv_.upload(wv); d_kv_.upload(kv); } else { d_wv_.release(); d_kv_.release(); }
    for (int h = 0; h < 4; ++h) {
        if (h < n_nh_) { d_wn_[h].upload(wn[h]); d_kn_[h].upload(kn[h]); }
        else { d_wn_[h].release(); d_kn_[h].release(); }
    }
    mat_ = DevMaterials();
    mat_.ne = ne_;
    mat_.edim = edim_;
    mat_.samples = samples_;
    mat_.elems = d_elems_.p;
    mat_.G = d_G_.p;
    mat_.g_stride = g_stride_;
    mat_.wc = has_cor_ ? d_wc_.p : nullptr;
    mat_.wv = has_vol_ ? d_wv_.p : nullptr;
    mat_.n_nh = n_nh_;
    if (act_dofs_ > 0) {
        std::vector<double> mw(ne_, 0.0), mf((size_t)3 * ne_, 0.0), pe(ne_, 0.0);
        for (const Muscle& mu : muscles_)
            for (int e : mu.elems) {
                mw[e] = mu.stiffness * vs;
                for (int d = 0; d < 3; ++d) mf[(size_t)d * ne_ + e] = mu.m[d];
            }
        if ((int)act_.size() == act_dofs_)
            for (int e = 0; e < ne_; ++e) if (mus_act_[e] >= 0) pe[e] = act_[mus_act_[e]];
        d_mw_.upload(mw);
        d_mfib_.upload(mf);
        d_mact_.upload(pe);
        mat_.mw = d_mw_.p;
        mat_.mfib = d_mfib_.p;
        mat_.mact = d_mact_.p;
    } else {
        d_mw_.release(); d_mfib_.release(); d_mact_.release();
    }
    for (int h = 0; h < n_nh_; ++h) {
        const Energy& en = energies_[nh_energy_[h]];
        mat_.wn[h] = d_wn_[h].p;
        mat_.nh_mu[h] = en.mu;
        mat_.nh_la[h] = en.la;
        mat_.nh_k[h] = en.stiffness;
    }
}

// ---------------------------------------------------------------------------
// SetupProjectiveDynamicsSolver (deformable_projective_dynamics_forward.cpp:8-183)
// and the sparse-inverse factor build (forward_gpu.cpp:584-764).
// ---------------------------------------------------------------------------
void Body::setup(double dt, int threads) {
    if (ready_) return;  // like pd_solver_ready_, a dt change does not rebuild (reference quirk)
    auto t0 = std::chrono::steady_clock::now();
    check(nv_ > 0, "Mesh is not initialized.");
    check(!energies_.empty(), "No PD energy was added.");
    threads = std::max(1, threads);
    // Dirichlet conditions. A vertex with its three DOFs fixed leaves the
    // shared matrix (identity row, couplings dropped) as in the reference's
    // per-axis matrices; a DOF fixed on its own stays in the shared matrix and
    // every global solve applies the bordered correction (partial_dirichlet).
    dmask_.assign(nv_, 0);
    pmask_.assign(dofs_, 0);
    pverts_.clear();
    dval_full_.assign(dofs_, 0.0);
    ddof_.clear();
    dval_.clear();
    for (auto& kv : dirichlet_) {
        dval_full_[kv.first] = kv.second;
        ddof_.push_back(kv.first);
        dval_.push_back(kv.second);
    }
    for (int v = 0; v < nv_; ++v) {
        const int nfix = (int)dirichlet_.count(3 * v) + (int)dirichlet_.count(3 * v + 1) + (int)dirichlet_.count(3 * v + 2);
        if (nfix == 3) dmask_[v] = 1;
        else if (nfix > 0) {
            for (int d = 0; d < 3; ++d) pmask_[3 * v + d] = dirichlet_.count(3 * v + d) ? 1 : 0;
            pverts_.push_back(v);
        }
    }
    check((int)pverts_.size() <= kMaxPartialVertices,
          "At most " + std::to_string(kMaxPartialVertices) + " vertices may have only some of their DOFs fixed.");
    check(pverts_.empty() || fb_vertices_.empty(),
          "Per-DOF Dirichlet conditions (a vertex with only some DOFs fixed) cannot be combined with a frictional "
          "boundary in this build.");
    const double mass = density_ * element_volume_;
    const double inv_h2m = mass / (dt * dt);
    const double vs = element_volume_ / samples_;
    // scalar A (identical for the three axes)
    std::vector<double> w_e(ne_, 0.0);
    for (int e = 0; e < ne_; ++e)
        for (auto& en : energies_) if (en.member[e]) w_e[e] += en.elem_w(e);
    std::vector<std::vector<std::pair<int, double>>> rows(nv_);
    for (int v = 0; v < nv_; ++v) rows[v].push_back({v, dmask_[v] ? 1.0 : inv_h2m});
    std::vector<double> Ke((size_t)edim_ * edim_);
    for (int e = 0; e < ne_; ++e) {
        if (g_stride_ == 0 && e > 0) { /* shared G: Ke reused */ }
        else {
            std::fill(Ke.begin(), Ke.end(), 0.0);
            for (int s = 0; s < samples_; ++s) {
                const double* g = &G_[(size_t)e * g_stride_ + (size_t)s * edim_ * 3];
                for (int k = 0; k < edim_; ++k)
                    for (int l = 0; l < edim_; ++l)
                        Ke[k * edim_ + l] += vs * (g[k * 3] * g[l * 3] + g[k * 3 + 1] * g[l * 3 + 1] + g[k * 3 + 2] * g[l * 3 + 2]);
            }
        }
        const int* v = &elems_[(size_t)e * edim_];
        for (int k = 0; k < edim_; ++k) {
            if (dmask_[v[k]]) continue;
            for (int l = 0; l < edim_; ++l) {
                if (dmask_[v[l]]) continue;
                rows[v[k]].push_back({v[l], w_e[e] * Ke[k * edim_ + l]});
            }
        }
    }
    // muscle terms w A^T M^T M A (forward.cpp:112-150): block (k, l) is
    // w (m.g_k)(m.g_l) I_3, the same on the three axes
    for (const Muscle& mu : muscles_) {
        const double w = mu.stiffness * vs;
        for (int e : mu.elems) {
            std::fill(Ke.begin(), Ke.end(), 0.0);
            for (int s = 0; s < samples_; ++s) {
                const double* g = &G_[(size_t)e * g_stride_ + (size_t)s * edim_ * 3];
                for (int k = 0; k < edim_; ++k) {
                    const double mk = g[k * 3] * mu.m[0] + g[k * 3 + 1] * mu.m[1] + g[k * 3 + 2] * mu.m[2];
                    for (int l = 0; l < edim_; ++l)
                        Ke[k * edim_ + l] += w * mk * (g[l * 3] * mu.m[0] + g[l * 3 + 1] * mu.m[1] + g[l * 3 + 2] * mu.m[2]);
                }
            }
            const int* v = &elems_[(size_t)e * edim_];
            for (int k = 0; k < edim_; ++k) {
                if (dmask_[v[k]]) continue;
                for (int l = 0; l < edim_; ++l) {
                    if (dmask_[v[l]]) continue;
                    rows[v[k]].push_back({v[l], Ke[k * edim_ + l]});
                }
            }
        }
    }
    Ap_.assign(nv_ + 1, 0);
    Ai_.clear();
    Ax_.clear();
    for (int v = 0; v < nv_; ++v) {
        auto& r = rows[v];
        std::sort(r.begin(), r.end(), [](auto& a, auto& b) { return a.first < b.first; });
        for (size_t i = 0; i < r.size(); ++i) {
            if (!Ai_.empty() && (int)Ai_.size() > Ap_[v] && Ai_.back() == r[i].first) Ax_.back() += r[i].second;
            else { Ai_.push_back(r[i].first); Ax_.push_back(r[i].second); }
        }
        Ap_[v + 1] = (int)Ai_.size();
        std::vector<std::pair<int, double>>().swap(r);
    }
    // factor (host FP64) and expansion of S = L^-1 (device FP64, FP32 copy)
    factor_.build(nv_, Ap_, Ai_, Ax_, X_, threads);
    plan_.build(factor_, unit_items_, chunk_w_);
    zcol_.clear();  // the baseline forward's A^-1 columns belong to the old factor
    g_frozen_.clear();
    {
        DevBuf<int64_t> dLp;
        DevBuf<int> dLi, dpar, ddep;
        DevBuf<double> dLx;
        dLp.upload(factor_.Lp);
        dLi.upload(factor_.Li);
        dLx.upload(factor_.Lx);
        dpar.upload(factor_.parent);
        ddep.upload(factor_.depth);
        d_fd_.upload(factor_.fd);
        d_base_.upload(factor_.base);
        V64_.alloc(factor_.v_len + kVPad);
        V32_.alloc(factor_.v_len + kVPad);
        hip_check(hipMemsetAsync(V64_.p, 0, V64_.n * sizeof(double), stream_), "memset V64");
        hip_check(hipMemsetAsync(V32_.p, 0, V32_.n * sizeof(float), stream_), "memset V32");
        // chain[i]: column i + 1 is the parent of i and pattern(i) = {i} + pattern(i + 1)
        std::vector<unsigned char> chain(nv_, 0);
        for (int i = 0; i + 1 < nv_; ++i)
            chain[i] = factor_.parent[i] == i + 1 &&
                       factor_.Lp[i + 1] - factor_.Lp[i] == factor_.Lp[i + 2] - factor_.Lp[i + 1] + 1;
        DevBuf<unsigned char> dchain;
        dchain.upload(chain);
        const int ntiles = (nv_ + 63) / 64;
        const int64_t ws_stride = (int64_t)(64 + factor_.max_depth + 1) * 64;
        // workspace of up to 8 GiB: every tile of the 131k-vertex grid in one launch
        const int64_t batch = std::max<int64_t>(1, std::min<int64_t>(ntiles, (int64_t)(8LL << 30) / (ws_stride * 8)));
        DevBuf<double> ws;
        ws.alloc((size_t)(batch * ws_stride));
        for (int t0b = 0; t0b < ntiles; t0b += (int)batch) {
            const int t1b = std::min<int>(ntiles, t0b + (int)batch);
            launch_expand_s(stream_, nv_, t0b, t1b, dLp.p, dLi.p, dLx.p, dpar.p, ddep.p, d_fd_.p, d_base_.p, dchain.p, ws.p,
                            ws_stride, V64_.p, V32_.p);
        }
        hip_check(hipStreamSynchronize(stream_), "expand S");
        hip_check(hipGetLastError(), "expand S launch");
    }
    d_perm_.upload(factor_.perm);
    {  // A in ELL form for the residual kernel (entries in CSR row order, zero padding)
        ell_w_ = 0;
        for (int v = 0; v < nv_; ++v) ell_w_ = std::max(ell_w_, Ap_[v + 1] - Ap_[v]);
        std::vector<int> ai((size_t)ell_w_ * nv_);
        std::vector<double> ax((size_t)ell_w_ * nv_, 0.0);
        for (int v = 0; v < nv_; ++v)
            for (int j = 0; j < ell_w_; ++j) {
                const int p = Ap_[v] + j;
                const bool in = p < Ap_[v + 1];
                ai[(size_t)j * nv_ + v] = in ? Ai_[p] : v;
                if (in) ax[(size_t)j * nv_ + v] = Ax_[p];
            }
        d_Aie_.upload(ai);
        d_Axe_.upload(ax);
    }
    d_dmask_.upload(dmask_);
    d_ddof_.upload(ddof_);
    d_dval_.upload(dval_);
    d_dval_full_.upload(dval_full_);
    // constant Dirichlet part of the local step: sum_e w_e (vol/s) sum_s A_s^T F_bound,s
    c_dir_.assign(dofs_, 0.0);
    if (!dirichlet_.empty()) {
        for (int e = 0; e < ne_; ++e) {
            const int* v = &elems_[(size_t)e * edim_];
            bool any = false;
            for (int k = 0; k < edim_; ++k)
                any |= dmask_[v[k]] != 0 || pmask_[3 * v[k]] || pmask_[3 * v[k] + 1] || pmask_[3 * v[k] + 2];
            if (!any || w_e[e] == 0) continue;
            for (int s = 0; s < samples_; ++s) {
                const double* g = &G_[(size_t)e * g_stride_ + (size_t)s * edim_ * 3];
                double Fb[9] = {0};
                for (int k = 0; k < edim_; ++k)
                    for (int i = 0; i < 3; ++i)
                        for (int j = 0; j < 3; ++j) Fb[i + 3 * j] += dval_full_[3 * v[k] + i] * g[k * 3 + j];
                for (int k = 0; k < edim_; ++k)
                    for (int i = 0; i < 3; ++i) {
                        double a = 0;
                        for (int j = 0; j < 3; ++j) a += Fb[i + 3 * j] * g[k * 3 + j];
                        c_dir_[3 * v[k] + i] += w_e[e] * vs * a;
                    }
            }
        }
        // muscles: w A^T M^T M F_bound = w (F_bound m)(m.g_k) at vertex k
        for (const Muscle& mu : muscles_)
            for (int e : mu.elems) {
                const int* v = &elems_[(size_t)e * edim_];
                bool any = false;
                for (int k = 0; k < edim_; ++k)
                    any |= dmask_[v[k]] != 0 || pmask_[3 * v[k]] || pmask_[3 * v[k] + 1] || pmask_[3 * v[k] + 2];
                if (!any) continue;
                for (int s = 0; s < samples_; ++s) {
                    const double* g = &G_[(size_t)e * g_stride_ + (size_t)s * edim_ * 3];
                    double Fbm[3] = {0, 0, 0};
                    for (int k = 0; k < edim_; ++k) {
                        const double gm = g[k * 3] * mu.m[0] + g[k * 3 + 1] * mu.m[1] + g[k * 3 + 2] * mu.m[2];
                        for (int i = 0; i < 3; ++i) Fbm[i] += dval_full_[3 * v[k] + i] * gm;
                    }
                    for (int k = 0; k < edim_; ++k) {
                        const double gm = g[k * 3] * mu.m[0] + g[k * 3 + 1] * mu.m[1] + g[k * 3 + 2] * mu.m[2];
                        for (int i = 0; i < 3; ++i) c_dir_[3 * v[k] + i] += mu.stiffness * vs * Fbm[i] * gm;
                    }
                }
            }
    }
    d_c_dir_.upload(c_dir_);
    // mesh data on device
    std::vector<int> soa((size_t)edim_ * ne_);
    for (int e = 0; e < ne_; ++e)
        for (int k = 0; k < edim_; ++k) soa[(size_t)k * ne_ + e] = elems_[(size_t)e * edim_ + k];
    d_elems_.upload(soa);
    d_G_.upload(G_);
    std::vector<int> cnt(nv_ + 1, 0), inc((size_t)ne_ * edim_);
    for (size_t i = 0; i < elems_.size(); ++i) cnt[elems_[i] + 1]++;
    for (int v = 0; v < nv_; ++v) cnt[v + 1] += cnt[v];
    std::vector<int> fill(cnt.begin(), cnt.end() - 1);
    for (int e = 0; e < ne_; ++e)
        for (int k = 0; k < edim_; ++k) inc[fill[elems_[(size_t)e * edim_ + k]]++] = e * edim_ + k;
    d_inc_ptr_.upload(cnt);
    d_inc_idx_.upload(inc);
    build_materials();
    // work vectors
    for (auto* b : {&d_q_, &d_qprev_, &d_qtilde_, &d_b_, &d_r_, &d_pdrhs_, &d_pdrhs_prev_, &d_g_, &d_tmp_, &d_kz_})
        b->alloc(dofs_);
    d_elem_.alloc((size_t)ne_ * edim_ * 3);
    d_dot_partial_.alloc(dots_partial_size());
    d_dot_out_.alloc(32);
    w32_.alloc((size_t)3 * spmv_ldw(nv_)); y32_.alloc((size_t)3 * nv_);
    w64_.alloc((size_t)3 * spmv_ldw(nv_)); y64_.alloc((size_t)3 * nv_);
    // k_perm_in writes columns < n only; the padding stays zero from here on
    hip_check(hipMemsetAsync(w32_.p, 0, w32_.n * sizeof(float), stream_), "w32 padding");
    hip_check(hipMemsetAsync(w64_.p, 0, w64_.n * sizeof(double), stream_), "w64 padding");
    upload_plan();
    build_partial_dirichlet();
    setup_contact(threads);
    ready_ = true;
    ready_dt_ = dt;
    setup_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// Contact paths of S restricted to the candidate columns and the candidate
// block H = A^-1_CC = S_C^T S_C (the Delassus operator's building block,
// forward_gpu.cpp:840-1066).
void Body::setup_contact(int threads) {
    (void)threads;
    cand_.clear();
    for (auto& kv : fb_vertices_) cand_.push_back(kv.first);
    K_ = (int)cand_.size();
    if (!boundary_.waves.empty()) d_waves_.upload(boundary_.waves);
    if (!boundary_.faces.empty()) d_mesh_faces_.upload(boundary_.faces);
    if (K_ == 0) return;
    d_cand_.upload(cand_);
    // paths: contact c in column iperm[v] walks its ancestors
    std::vector<int> cp_ptr(K_ + 1, 0), cp_row, cp_col;
    for (int c = 0; c < K_; ++c) {
        const int j = factor_.iperm[cand_[c]];
        for (int k = j; k != -1; k = factor_.parent[k]) { cp_row.push_back(k); cp_col.push_back(j); }
        cp_ptr[c + 1] = (int)cp_row.size();
    }
    const int64_t m = (int64_t)cp_row.size();
    contact_entries_ = m;
    DevBuf<int> drow, dcol;
    drow.upload(cp_row);
    dcol.upload(cp_col);
    d_cp_val64_.alloc(m);
    d_cp_val32_.alloc(m);
    launch_gather_vals(stream_, m, drow.p, dcol.p, d_base_.p, V64_.p, d_cp_val64_.p, d_cp_val32_.p);
    d_cp_ptr_.upload(cp_ptr);
    d_cp_row_.upload(cp_row);
    // row-compact transpose
    std::vector<int> rowid(nv_, -1), urows;
    for (int64_t p = 0; p < m; ++p)
        if (rowid[cp_row[p]] < 0) { rowid[cp_row[p]] = (int)urows.size(); urows.push_back(cp_row[p]); }
    const int R = (int)urows.size();
    std::vector<int> rc_ptr(R + 1, 0);
    for (int64_t p = 0; p < m; ++p) rc_ptr[rowid[cp_row[p]] + 1]++;
    for (int r = 0; r < R; ++r) rc_ptr[r + 1] += rc_ptr[r];
    std::vector<int> rc_c(m), rc_src(m), fillp(rc_ptr.begin(), rc_ptr.end() - 1);
    for (int c = 0; c < K_; ++c)
        for (int p = cp_ptr[c]; p < cp_ptr[c + 1]; ++p) {
            const int slot = fillp[rowid[cp_row[p]]]++;
            rc_c[slot] = c;
            rc_src[slot] = p;
        }
    std::vector<int> rc_rowk(R), rc_rows_full(m), rc_cols_full(m);
    for (int r = 0; r < R; ++r) rc_rowk[r] = urows[r];
    for (int64_t s = 0; s < m; ++s) { rc_rows_full[s] = cp_row[rc_src[s]]; rc_cols_full[s] = cp_col[rc_src[s]]; }
    DevBuf<int> dr2, dc2;
    dr2.upload(rc_rows_full);
    dc2.upload(rc_cols_full);
    d_rc_val64_.alloc(m);
    d_rc_val32_.alloc(m);
    launch_gather_vals(stream_, m, dr2.p, dc2.p, d_base_.p, V64_.p, d_rc_val64_.p, d_rc_val32_.p);
    d_rc_ptr_.upload(rc_ptr);
    d_rc_row_.upload(rc_rowk);
    d_rc_c_.upload(rc_c);
    rc_nrows_ = R;
    // both list structures cut into pieces of at most kContactPiece entries
    auto cut = [](const std::vector<int>& ptr, DevBuf<int>& db, DevBuf<int>& de, DevBuf<int>& dp, ContactPieces& pc) {
        const int L = (int)ptr.size() - 1;
        std::vector<int> b, e, pp(L + 1, 0);
        for (int l = 0; l < L; ++l) {
            for (int s0 = ptr[l]; s0 < ptr[l + 1]; s0 += kContactPiece) {
                b.push_back(s0);
                e.push_back(std::min(ptr[l + 1], s0 + kContactPiece));
            }
            pp[l + 1] = (int)b.size();
        }
        db.upload(b);
        de.upload(e);
        dp.upload(pp);
        pc.nlists = L;
        pc.npieces = (int)b.size();
        pc.b = db.p;
        pc.e = de.p;
        pc.ptr = dp.p;
    };
    cut(cp_ptr, d_gp_b_, d_gp_e_, d_gp_ptr_, gather_pieces_);
    cut(rc_ptr, d_sp_b_, d_sp_e_, d_sp_ptr_, scatter_pieces_);
    d_piece_part_.alloc((size_t)3 * std::max(gather_pieces_.npieces, scatter_pieces_.npieces) + 3);
    // H = S_C^T S_C via a dense S_C (R x K) and rocBLAS dsyrk
    {
        DevBuf<double> D;
        D.alloc((size_t)R * K_);
        hip_check(hipMemsetAsync(D.p, 0, D.n * sizeof(double), stream_), "memset S_C");
        std::vector<int> ridx(m), cidx(m);
        for (int c = 0; c < K_; ++c)
            for (int p = cp_ptr[c]; p < cp_ptr[c + 1]; ++p) { ridx[p] = rowid[cp_row[p]]; cidx[p] = c; }
        DevBuf<int> dri, dci;
        dri.upload(ridx);
        dci.upload(cidx);
        launch_scatter_dense(stream_, m, dri.p, dci.p, d_cp_val64_.p, K_, D.p);
        d_H_.alloc((size_t)K_ * K_);
        const double one = 1.0, zero = 0.0;
        // D is row-major R x K == column-major K x R (= S_C^T); H = D D^T.
        check(rocblas_dsyrk((rocblas_handle)blas_, rocblas_fill_lower, rocblas_operation_none, K_, R, &one, D.p, K_, &zero,
                            d_H_.p, K_) == rocblas_status_success, "rocblas_dsyrk failed");
        launch_symmetrize_lower(stream_, K_, d_H_.p);
        d_H32_.alloc((size_t)K_ * K_);
        launch_to_float(stream_, (int64_t)K_ * K_, d_H_.p, d_H32_.p);
        hip_check(hipStreamSynchronize(stream_), "contact H");
    }
    std::vector<double> Hdiag(K_);
    for (int c = 0; c < K_; ++c)
        hip_check(hipMemcpy(&Hdiag[c], d_H_.p + (size_t)c * K_ + c, sizeof(double), hipMemcpyDeviceToHost), "H diag");
    Hdiag_ = Hdiag;
    d_Hdiag_.upload(Hdiag);
    for (auto* b : {&d_nrm_, &d_t1_, &d_t2_, &d_F_, &d_dF_, &d_uC_, &d_tmpK3_, &d_qframeC_, &d_Fg_, &d_Fn_, &d_hprev_})
        b->alloc((size_t)3 * K_);
    for (auto* b : {&d_delta_, &d_om_, &d_E_, &d_rhsK_, &d_rprec_, &d_lambda_, &d_lt1_, &d_lt2_, &d_hvec_}) b->alloc(K_);
    for (auto* b : {&d_idxJ_, &d_idxP_, &d_idxF_, &d_posJ_, &d_posL_}) b->alloc(K_);
    d_ncp_lists_.alloc((size_t)6 * K_);
    d_ncp_counts_.alloc(8);
    d_ncp_flags_.alloc(K_);
    d_rhsK32_.alloc(K_);
    d_info_.alloc(1);
}

// ---------------------------------------------------------------------------
// Building blocks.
// ---------------------------------------------------------------------------
hipEvent_t Body::next_event() {
    if (ev_used_ == ev_pool_.size()) {
        hipEvent_t e;
        hip_check(hipEventCreate(&e), "hipEventCreate");
        ev_pool_.push_back(e);
    }
    return ev_pool_[ev_used_++];
}

// Sums the recorded kernel durations (milliseconds) and recycles the events.
void Body::collect_timing() {
    hip_check(hipStreamSynchronize(stream_), "timing sync");
    for (auto& p : ev_rows_) { float ms = 0; (void)hipEventElapsedTime(&ms, p.first, p.second); t_spmv_rows_ms += ms; ++n_spmv_rows; }
    for (auto& p : ev_cols_) { float ms = 0; (void)hipEventElapsedTime(&ms, p.first, p.second); t_spmv_cols_ms += ms; ++n_spmv_cols; }
    ev_rows_.clear();
    ev_cols_.clear();
    ev_used_ = 0;
}

// Columns of A^-1 at the partially fixed vertices (three per FP64 product, one
// in each axis slot), then per axis G_d = (A^-1)_{F_d F_d} inverted on the host
// by Cholesky (G_d is a principal submatrix of the SPD A^-1).
void Body::build_partial_dirichlet() {
    const int np = (int)pverts_.size();
    d_pmask_.release();
    if (np == 0) return;
    d_pmask_.upload(pmask_);
    d_pZ_.alloc((size_t)np * nv_);
    std::vector<double> e(dofs_, 0.0);
    for (int p0 = 0; p0 < np; p0 += 3) {
        const int na = std::min(3, np - p0);
        std::fill(e.begin(), e.end(), 0.0);
        for (int a = 0; a < na; ++a) e[3 * pverts_[p0 + a] + a] = 1.0;
        hip_check(hipStreamSynchronize(stream_), "pdir");
        d_r_.upload(e);
        spmv_S<double>(d_r_.p, w64_, y64_, part64_, V64_.p);
        spmv_ST<double>(y64_, xpart64_, V64_.p, nullptr, nullptr, d_tmp_.p);
        launch_pdir_columns(stream_, nv_, p0, na, d_tmp_.p, d_pZ_.p);
    }
    hip_check(hipStreamSynchronize(stream_), "pdir columns");
    // Z rows at the partial vertices: Zpp[q][p] = (A^-1)_{v_p v_q}
    std::vector<double> Zpp;
    {
        DevBuf<int> dpv;
        DevBuf<double> dzz;
        dpv.upload(pverts_);
        dzz.alloc((size_t)np * np);
        launch_pdir_gather(stream_, np, nv_, dpv.p, d_pZ_.p, dzz.p);
        hip_check(hipStreamSynchronize(stream_), "pdir G");
        Zpp = dzz.to_host();
    }
    std::vector<int> map((size_t)3 * np, -1), fv;
    std::vector<double> ginv;
    for (int d = 0; d < 3; ++d) {
        pax_.off[d] = (int)fv.size();
        pax_.goff[d] = (long long)ginv.size();
        std::vector<int> F;
        for (int p = 0; p < np; ++p)
            if (pmask_[3 * pverts_[p] + d]) { map[(size_t)d * np + p] = (int)F.size(); F.push_back(p); }
        const int nf = (int)F.size();
        // Cholesky of G_d, then G_d^-1 column by column
        std::vector<double> L((size_t)nf * nf, 0.0);
        for (int j = 0; j < nf; ++j) {
            double s = Zpp[(size_t)F[j] * np + F[j]];
            for (int k = 0; k < j; ++k) s -= L[(size_t)j * nf + k] * L[(size_t)j * nf + k];
            check(s > 0, "per-DOF Dirichlet: the constrained block of A^-1 is not positive definite.");
            L[(size_t)j * nf + j] = std::sqrt(s);
            for (int i = j + 1; i < nf; ++i) {
                double t = Zpp[(size_t)F[j] * np + F[i]];
                for (int k = 0; k < j; ++k) t -= L[(size_t)i * nf + k] * L[(size_t)j * nf + k];
                L[(size_t)i * nf + j] = t / L[(size_t)j * nf + j];
            }
        }
        std::vector<double> G((size_t)nf * nf);
        std::vector<double> col(nf);
        for (int c = 0; c < nf; ++c) {
            for (int i = 0; i < nf; ++i) {  // L z = e_c
                double t = i == c ? 1.0 : 0.0;
                for (int k = 0; k < i; ++k) t -= L[(size_t)i * nf + k] * col[k];
                col[i] = t / L[(size_t)i * nf + i];
            }
            for (int i = nf - 1; i >= 0; --i) {  // L^T x = z
                double t = col[i];
                for (int k = i + 1; k < nf; ++k) t -= L[(size_t)k * nf + i] * col[k];
                col[i] = t / L[(size_t)i * nf + i];
            }
            for (int i = 0; i < nf; ++i) G[(size_t)i * nf + c] = col[i];
        }
        for (int p : F) fv.push_back(pverts_[p]);
        ginv.insert(ginv.end(), G.begin(), G.end());
    }
    pax_.off[3] = (int)fv.size();
    d_pmap_.upload(map);
    d_pfv_.upload(fv);
    d_pginv_.upload(ginv);
    d_pcoef_.alloc((size_t)3 * np);
}

void Body::partial_dirichlet(double* y, const double* t, const double* o) {
    const int np = (int)pverts_.size();
    if (np == 0) return;
    launch_pdir_coef(stream_, np, pax_, d_pmap_.p, d_pfv_.p, d_pginv_.p, y, t, d_pcoef_.p);
    launch_pdir_apply(stream_, nv_, np, d_pZ_.p, d_pcoef_.p, d_pmask_.p, o, y);
}

void Body::upload_plan() {
    // kernels queued on stream_ may still read the previous plan's buffers
    hip_check(hipStreamSynchronize(stream_), "plan upload");
    d_gu_group_.upload(plan_.gu_group);
    d_gu_chunk_.upload(plan_.gu_chunk);
    d_rbase_.upload(plan_.rbase);
    d_items_.upload(plan_.items);
    d_cu_tile_.upload(plan_.cu_tile);
    d_cu_i0_.upload(plan_.cu_i0);
    d_cu_i1_.upload(plan_.cu_i1);
    d_tile_unit_ptr_.upload(plan_.tile_unit_ptr);
    part32_.alloc(3 * (size_t)plan_.rbase[nv_] + 3); part64_.alloc(3 * (size_t)plan_.rbase[nv_] + 3);
    xpart32_.alloc(192 * plan_.cu_tile.size() + 192); xpart64_.alloc(192 * plan_.cu_tile.size() + 192);
}

void Body::set_launch_shapes(int chunk_w, int unit_items, int local_block) {
    check(ready_, "set_launch_shapes needs the factor built (phd_factor_build).");
    auto in = [](int v, const int* c, int n) { return std::find(c, c + n, v) != c + n; };
    check(in(chunk_w, kTuneChunks, 3), "chunk width must be 512, 1024 or 2048.");
    check(in(unit_items, kTuneUnits, 4), "unit height must be 256, 512, 1024 or 2048.");
    check(in(local_block, kTuneBlocks, 3), "local-step block must be 64, 128 or 256.");
    chunk_w_ = chunk_w;
    unit_items_ = unit_items;
    local_block_ = local_block;
    plan_.build(factor_, unit_items_, chunk_w_);
    upload_plan();
}

Body::TuneResult Body::autotune(double dt, int threads, int reps) {
    check(reps >= 1 && reps <= 1000, "autotune: reps must be in [1, 1000].");
    setup(dt, threads);
    TuneResult res;
    hipEvent_t ea, eb;
    hip_check(hipEventCreate(&ea), "event");
    hip_check(hipEventCreate(&eb), "event");
    auto time_ms = [&](auto&& launch) {
        launch();  // warm-up
        hip_check(hipEventRecord(ea, stream_), "event");
        for (int i = 0; i < reps; ++i) launch();
        hip_check(hipEventRecord(eb, stream_), "event");
        hip_check(hipEventSynchronize(eb), "event");
        float ms = 0;
        hip_check(hipEventElapsedTime(&ms, ea, eb), "event");
        return (double)ms / reps;
    };
    // inputs: the rest positions as the right-hand side of S w, and as q
    d_tmp_.upload(X_);
    d_q_.upload(X_);
    launch_perm_in<float>(stream_, nv_, d_perm_.p, d_tmp_.p, nullptr, w32_.p);
    int best = 0;
    for (int i = 0; i < 3; ++i) {
        plan_.build(factor_, unit_items_, kTuneChunks[i]);
        upload_plan();
        res.rows_ms[i] = time_ms([&] {
            launch_spmv_rows_grp<float>(stream_, (int)plan_.gu_group.size(), d_gu_group_.p, d_gu_chunk_.p, d_fd_.p,
                                        d_base_.p, d_rbase_.p, V32_.p, w32_.p, nv_, part32_.p, kTuneChunks[i]);
            launch_spmv_rows_slots<float>(stream_, nv_, d_rbase_.p, part32_.p, y32_.p);
        });
        if (res.rows_ms[i] < res.rows_ms[best]) best = i;
    }
    chunk_w_ = kTuneChunks[best];
    best = 0;
    for (int i = 0; i < 4; ++i) {
        plan_.build(factor_, kTuneUnits[i], chunk_w_);
        upload_plan();
        res.cols_ms[i] = time_ms([&] {
            launch_spmv_cols<float>(stream_, (int)plan_.cu_tile.size(), d_cu_tile_.p, d_cu_i0_.p, d_cu_i1_.p, d_items_.p,
                                    d_fd_.p, d_base_.p, V32_.p, y32_.p, nv_, xpart32_.p);
            launch_spmv_cols_fin<float>(stream_, nv_, d_tile_unit_ptr_.p, xpart32_.p, d_perm_.p, nullptr, nullptr,
                                        nullptr, d_g_.p);
        });
        if (res.cols_ms[i] < res.cols_ms[best]) best = i;
    }
    unit_items_ = kTuneUnits[best];
    best = 0;
    for (int i = 0; i < 3; ++i) {
        res.local_ms[i] = time_ms([&] { launch_local_step(stream_, mat_, d_q_.p, d_elem_.p, kTuneBlocks[i]); });
        if (res.local_ms[i] < res.local_ms[best]) best = i;
    }
    local_block_ = kTuneBlocks[best];
    plan_.build(factor_, unit_items_, chunk_w_);
    upload_plan();
    (void)hipEventDestroy(ea);
    (void)hipEventDestroy(eb);
    res.chunk_w = chunk_w_;
    res.unit_items = unit_items_;
    res.local_block = local_block_;
    return res;
}

template <class TV>
void Body::spmv_S(const double* r, DevBuf<TV>& w, DevBuf<TV>& y, DevBuf<TV>& part, const TV* V) {
    launch_perm_in<TV>(stream_, nv_, d_perm_.p, r, nullptr, w.p);
    hipEvent_t a = nullptr, b = nullptr;
    if (timing) { a = next_event(); b = next_event(); (void)hipEventRecord(a, stream_); }
    launch_spmv_rows_grp<TV>(stream_, (int)plan_.gu_group.size(), d_gu_group_.p, d_gu_chunk_.p, d_fd_.p, d_base_.p,
                             d_rbase_.p, V, w.p, nv_, part.p, chunk_w_);
    if (timing) { (void)hipEventRecord(b, stream_); ev_rows_.push_back({a, b}); }
    launch_spmv_rows_slots<TV>(stream_, nv_, d_rbase_.p, part.p, y.p);
}

template <class TV>
void Body::spmv_ST(DevBuf<TV>& y, DevBuf<TV>& xpart, const TV* V, const double* add, const double* dval, double* out) {
    hipEvent_t a = nullptr, b = nullptr;
    if (timing) { a = next_event(); b = next_event(); (void)hipEventRecord(a, stream_); }
    launch_spmv_cols<TV>(stream_, (int)plan_.cu_tile.size(), d_cu_tile_.p, d_cu_i0_.p, d_cu_i1_.p, d_items_.p, d_fd_.p,
                         d_base_.p, V, y.p, nv_, xpart.p);
    if (timing) { (void)hipEventRecord(b, stream_); ev_cols_.push_back({a, b}); }
    launch_spmv_cols_fin<TV>(stream_, nv_, d_tile_unit_ptr_.p, xpart.p, d_perm_.p, add, dval ? d_dmask_.p : nullptr, dval, out);
}

template <class TV>
void Body::contact_gather(DevBuf<TV>& y, double* out) {
    const TV* val = std::is_same<TV, float>::value ? (const TV*)d_cp_val32_.p : (const TV*)d_cp_val64_.p;
    launch_contact_gather<TV>(stream_, gather_pieces_, d_cp_row_.p, val, y.p, nv_, d_piece_part_.p, out);
}
template <class TV>
void Body::contact_scatter(const double* f, DevBuf<TV>& y) {
    const TV* val = std::is_same<TV, float>::value ? (const TV*)d_rc_val32_.p : (const TV*)d_rc_val64_.p;
    launch_contact_scatter<TV>(stream_, scatter_pieces_, d_rc_row_.p, d_rc_c_.p, val, f, d_piece_part_.p, nv_, y.p);
}

void Body::local_rhs(const double* q, double inv_h2m) {
    launch_local_step(stream_, mat_, q, d_elem_.p, local_block_, local_fp32_);
    launch_gather_rhs(stream_, nv_, d_inc_ptr_.p, d_inc_idx_.p, d_elem_.p, d_c_dir_.p, d_qtilde_.p, inv_h2m, d_pdrhs_.p, d_b_.p);
}

std::vector<double> Body::dots(const std::vector<std::pair<const double*, const double*>>& pairs) {
    DotArgs a;
    a.m = (int)pairs.size();
    check(a.m <= 32, "too many dot products");
    for (int i = 0; i < a.m; ++i) { a.a[i] = pairs[i].first; a.b[i] = pairs[i].second; }
    launch_dots(stream_, dofs_, a, d_dot_partial_.p, d_dot_out_.p);
    double* pd = reinterpret_cast<double*>(pin_small_);
    hip_check(hipMemcpyAsync(pd, d_dot_out_.p, sizeof(double) * a.m, hipMemcpyDeviceToHost, stream_), "dots D2H");
    hip_check(hipStreamSynchronize(stream_), "dots sync");
    return std::vector<double>(pd, pd + a.m);
}

void Body::kpd_apply(const double* x, double* out, int precision) {
    if (precision == 1) launch_kpd_apply<float>(stream_, mat_, d_W32_.p, x, d_elem_.p);
    else launch_kpd_apply<double>(stream_, mat_, d_W64_.p, x, d_elem_.p);
    launch_gather_vertex(stream_, nv_, d_inc_ptr_.p, d_inc_idx_.p, d_elem_.p, out);
}

// q~ = q + h v + h^2 / m (f_ext + m g) (forward_gpu.cpp:430-436)
void Body::set_inertial_target(const double* q, const double* v, const double* f_ext, double dt) {
    const int n3 = dofs_;
    const double h = dt, mass = element_volume_ * density_, h2m = h * h / mass;
    std::vector<double> fs(n3), qt(n3);
    state_force(q, v, fs.data());
    for (int i = 0; i < n3; ++i) qt[i] = q[i] + h * v[i] + h2m * (f_ext ? f_ext[i] : 0.0) + h2m * fs[i];
    d_qtilde_.upload(qt);
}

WarmStartNet& Body::net() {
    if (!net_) net_.reset(new WarmStartNet());
    return *net_;
}

// Per-vertex inputs of the warm-start network at state (q, v): mean log10
// element stiffness, contact flag (candidate with lambda > 0 in the previous
// frame), obstacle distance for candidates.
void Body::mlp_features(const double* q_dev, const double* v_dev, double dt, bool standardize) {
    const double L = std::cbrt(element_volume_);
    const double mass = element_volume_ * density_, inv_h2m = mass / (dt * dt);
    if (d_adiag_.n != (size_t)nv_) {
        std::vector<float> ad(nv_, 1.f);
        for (int i = 0; i < nv_; ++i)
            for (int p = Ap_[i]; p < Ap_[i + 1]; ++p)
                if (Ai_[p] == i) ad[i] = (float)Ax_[p];
        d_adiag_.upload(ad);
    }
    if (d_logk_.n != (size_t)nv_) {
        std::vector<double> acc(nv_, 0.0);
        std::vector<int> cnt(nv_, 0);
        for (int e = 0; e < ne_; ++e) {
            double w = 0;
            for (const Energy& en : energies_) if (en.member[e]) w += en.elem_w(e);
            const double lw = std::log10(std::max(w, 1e-300));
            for (int k = 0; k < edim_; ++k) { acc[elems_[(size_t)e * edim_ + k]] += lw; cnt[elems_[(size_t)e * edim_ + k]]++; }
        }
        std::vector<float> lk(nv_);
        for (int i = 0; i < nv_; ++i) lk[i] = cnt[i] ? (float)(acc[i] / cnt[i]) : 0.f;
        d_logk_.upload(lk);
        d_Xrest_.upload(X_);
    }
    std::vector<float> cf(nv_, 0.f), cd(nv_, 4.f);
    std::vector<double> qh;
    if (!cand_.empty() || !fb_vertices_.empty()) {
        qh.resize(dofs_);
        hip_check(hipMemcpyAsync(qh.data(), q_dev, sizeof(double) * dofs_, hipMemcpyDeviceToHost, stream_), "q D2H");
        hip_check(hipStreamSynchronize(stream_), "features");
        for (auto& kv : fb_vertices_) {
            double d, nn[3];
            if (boundary_.type == 5) {  // unsigned distance to the opposite surface
                double cl[3], fn[3];
                const double* x = &qh[3 * kv.first];
                d = boundary_.mesh_closest(x, qh.data(), kv.first, cl, fn)
                        ? std::sqrt((x[0] - cl[0]) * (x[0] - cl[0]) + (x[1] - cl[1]) * (x[1] - cl[1]) +
                                    (x[2] - cl[2]) * (x[2] - cl[2])) - boundary_.radius
                        : 4.0 * L;
            } else {
                boundary_.eval(&qh[3 * kv.first], d, nn);
            }
            cd[kv.first] = (float)std::max(-4.0, std::min(4.0, d / L));
        }
        for (size_t i = 0; i < prev_candidates_.size() && i < prev_lambda_.size(); ++i)
            if (prev_lambda_[i] > 0) cf[prev_candidates_[i]] = 1.f;
    }
    d_cflag_.upload(cf);
    d_cdist_.upload(cd);
    if (d_feat_.n < (size_t)nv_ * kMlpInPad) d_feat_.alloc((size_t)nv_ * kMlpInPad);
    WarmStartNet& nt = net();
    d_fmean_.upload(nt.fmean);
    d_fstd_.upload(nt.fstd);
    double g = 0;
    if (gravity_.size() == 3) g = std::sqrt(gravity_[0] * gravity_[0] + gravity_[1] * gravity_[1] + gravity_[2] * gravity_[2]);
    // PD residual at the x + h v guess (d_qtilde_ holds this frame's inertial target)
    if (d_guess_.n < (size_t)2 * dofs_) d_guess_.alloc((size_t)2 * dofs_);
    launch_warm_guess(stream_, nv_, q_dev, v_dev, dt, nullptr, 0.f, d_dmask_.p, d_dval_full_.p, d_guess_.p,
                      d_guess_.p + dofs_);
    local_rhs(d_guess_.p, inv_h2m);
    launch_residual(stream_, nv_, ell_w_, d_Aie_.p, d_Axe_.p, d_b_.p, d_guess_.p, d_dmask_.p, d_r_.p, d_pmask_.p);
    launch_mlp_features(stream_, nv_, q_dev, v_dev, d_Xrest_.p, d_r_.p, d_adiag_.p, d_logk_.p, d_cflag_.p, d_cdist_.p, dt,
                        L, g, d_fmean_.p, d_fstd_.p, standardize ? 1 : 0, d_feat_.p);
}

// Initial iterate of the PD loop from the network: x + h v + tscale * net,
// kept only when its first global step is smaller than that of the plain
// x + h v guess (learned = false: x + h v). d_q_ holds q on entry and the
// chosen iterate on return.
void Body::warm_start(const double* v, double dt, double inv_h2m, bool learned) {
    WarmStartNet& nt = net();
    check(!learned || nt.ready(), "phd_warm_start = 1 needs a trained warm-start network (phd_mlp_init / phd_mlp_train).");
    const int n3 = dofs_;
    d_v_.upload(v, n3);
    if (d_mlpout_.n < (size_t)n3) d_mlpout_.alloc(n3);
    if (learned) {
        mlp_features(d_q_.p, d_v_.p, dt, true);
        nt.forward(stream_, nv_, d_feat_.p, d_mlpout_.p);
    } else {
        hip_check(hipMemsetAsync(d_mlpout_.p, 0, sizeof(float) * n3, stream_), "no net");
    }
    if (d_guess_.n < (size_t)2 * n3) d_guess_.alloc((size_t)2 * n3);
    launch_warm_guess(stream_, nv_, d_q_.p, d_v_.p, dt, d_mlpout_.p, nt.tscale, d_dmask_.p, d_dval_full_.p, d_guess_.p,
                      d_guess_.p + n3);
    if (!learned) {
        hip_check(hipMemcpyAsync(d_q_.p, d_guess_.p, sizeof(double) * n3, hipMemcpyDeviceToDevice, stream_), "x + h v");
        return;
    }
    // size of the first global step from each guess, ||A^-1 (b(g) - A g)||
    // (contact impulses zero): A^-1 damps the high-frequency part of a
    // per-vertex prediction's error, which the first solve removes anyway
    double rn[2];
    for (int g = 0; g < 2; ++g) {
        const double* x = d_guess_.p + (size_t)g * n3;
        local_rhs(x, inv_h2m);
        launch_residual(stream_, nv_, ell_w_, d_Aie_.p, d_Axe_.p, d_b_.p, x, d_dmask_.p, d_r_.p, d_pmask_.p);
        spmv_S<double>(d_r_.p, w64_, y64_, part64_, V64_.p);
        spmv_ST<double>(y64_, xpart64_, V64_.p, nullptr, nullptr, d_tmp_.p);
        rn[g] = std::sqrt(dots({{d_tmp_.p, d_tmp_.p}})[0]);
    }
    const int pick = (std::isfinite(rn[1]) && rn[1] < rn[0]) ? 1 : 0;
    if (pick) ++warm_used; else ++warm_fallback;
    hip_check(hipMemcpyAsync(d_q_.p, d_guess_.p + (size_t)pick * n3, sizeof(double) * n3, hipMemcpyDeviceToDevice, stream_),
              "warm start");
}

// Contact system M dl = rhs on the index set J (already in d_idxJ_, with
// d_om_ / d_E_ / d_nrm_ current): assembled and solved by the blocked
// Cholesky of phd_dense.hip. rhs (host, n) is overwritten with the solution.
// Returns false when a pivot was not positive.
// The compressed contact system on the candidates J (n rows): its right-hand
// side is in d_rhsK_, the solution stays on the device (d_cx32_ / d_cx64_);
// only the pivot flag comes back. False when a pivot was not positive.
bool Body::solve_contact_system(int n, int precision, const int* J, bool plain_w) {
    const double* nrm = plain_w ? nullptr : d_nrm_.p;
    const int np = chol_padded(n);
    launch_clear_flag(stream_, d_info_.p);
    if (precision == 1) {
        if (d_C32_.n < (size_t)chol_elems(n)) d_C32_.alloc((size_t)chol_elems(n));
        if (d_cy32_.n < (size_t)chol_work(n)) { d_cy32_.alloc((size_t)chol_work(n)); d_cx32_.alloc(np); }
        if (h32_) launch_chol_assemble<float, float>(stream_, n, J, d_H32_.p, K_, nrm, d_om_.p, d_E_.p, d_rhsK_.p, d_C32_.p);
        else launch_chol_assemble<float, double>(stream_, n, J, d_H_.p, K_, nrm, d_om_.p, d_E_.p, d_rhsK_.p, d_C32_.p);
        launch_chol_solve<float>(stream_, n, d_C32_.p, d_cy32_.p, d_cx32_.p, d_info_.p);
    } else {
        if (d_C64_.n < (size_t)chol_elems(n)) d_C64_.alloc((size_t)chol_elems(n));
        if (d_cy64_.n < (size_t)chol_work(n)) { d_cy64_.alloc((size_t)chol_work(n)); d_cx64_.alloc(np); }
        launch_chol_assemble<double, double>(stream_, n, J, d_H_.p, K_, nrm, d_om_.p, d_E_.p, d_rhsK_.p, d_C64_.p);
        launch_chol_solve<double>(stream_, n, d_C64_.p, d_cy64_.p, d_cx64_.p, d_info_.p);
    }
    int* flag = reinterpret_cast<int*>(pin_small_);
    hip_check(hipMemcpyAsync(flag, d_info_.p, sizeof(int), hipMemcpyDeviceToHost, stream_), "pivot flag");
    hip_check(hipStreamSynchronize(stream_), "contact solve");
    return flag[0] == 0;
}

// x = A^-1 b with optional FP64 refinement (test entry point; Dirichlet rows are identity).
void Body::apply_ainv(double dt, const double* b, double* x, int precision, int refine) {
    setup(dt, 8);
    d_b_.upload(b, dofs_);
    hip_check(hipMemsetAsync(d_q_.p, 0, sizeof(double) * dofs_, stream_), "zero x");
    hip_check(hipMemcpyAsync(d_r_.p, d_b_.p, sizeof(double) * dofs_, hipMemcpyDeviceToDevice, stream_), "r=b");
    for (int it = 0; it <= refine; ++it) {
        if (it > 0) launch_residual(stream_, nv_, ell_w_, d_Aie_.p, d_Axe_.p, d_b_.p, d_q_.p, nullptr, d_r_.p, d_pmask_.p);
        if (precision == 1) {
            spmv_S<float>(d_r_.p, w32_, y32_, part32_, V32_.p);
            spmv_ST<float>(y32_, xpart32_, V32_.p, d_q_.p, nullptr, d_q_.p);
        } else {
            spmv_S<double>(d_r_.p, w64_, y64_, part64_, V64_.p);
            spmv_ST<double>(y64_, xpart64_, V64_.p, d_q_.p, nullptr, d_q_.p);
        }
        // per-DOF Dirichlet rows are identities: x_F = b_F, the increment is 0 there
        partial_dirichlet(d_q_.p, it == 0 ? nullptr : d_b_.p, d_b_.p);
    }
    hip_check(hipStreamSynchronize(stream_), "apply_ainv");
    d_q_.download(x, dofs_);
}

// x = A^-1 r with the FP64 factor on device vectors (the baseline's global solve).
void Body::ainv64(const double* r, double* x) {
    spmv_S<double>(r, w64_, y64_, part64_, V64_.p);
    spmv_ST<double>(y64_, xpart64_, V64_.p, nullptr, nullptr, x);
}

// Scratch bytes apply_ainv_noalloc needs: w (3 ldw), y (3n), row-chunk partials,
// column-tile partials (factor precision) and the FP64 residual (3n), each
// segment 256-byte aligned.
size_t Body::noalloc_scratch_bytes(int precision) const {
    const size_t tv = precision == 1 ? sizeof(float) : sizeof(double);
    auto al = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t n = (size_t)nv_;
    return al(3 * (size_t)spmv_ldw(nv_) * tv) + al(3 * n * tv) + al((3 * (size_t)plan_.rbase[nv_] + 3) * tv) +
           al((192 * plan_.cu_tile.size() + 192) * tv) + al(3 * n * sizeof(double));
}

template <class TV>
void Body::ainv_passes(const double* b_dev, double* x_dev, int refine, char* scr) {
    auto al = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t n = (size_t)nv_;
    TV* w = reinterpret_cast<TV*>(scr);
    scr += al(3 * (size_t)spmv_ldw(nv_) * sizeof(TV));
    TV* y = reinterpret_cast<TV*>(scr);
    scr += al(3 * n * sizeof(TV));
    TV* part = reinterpret_cast<TV*>(scr);
    scr += al((3 * (size_t)plan_.rbase[nv_] + 3) * sizeof(TV));
    TV* xpart = reinterpret_cast<TV*>(scr);
    scr += al((192 * plan_.cu_tile.size() + 192) * sizeof(TV));
    double* r = reinterpret_cast<double*>(scr);
    const TV* V = std::is_same<TV, float>::value ? (const TV*)V32_.p : (const TV*)V64_.p;
    for (int it = 0; it <= refine; ++it) {
        const double* rhs = b_dev;
        if (it > 0) {
            launch_residual(stream_, nv_, ell_w_, d_Aie_.p, d_Axe_.p, b_dev, x_dev, nullptr, r);
            rhs = r;
        }
        launch_perm_in<TV>(stream_, nv_, d_perm_.p, rhs, nullptr, w);
        launch_spmv_rows_grp<TV>(stream_, (int)plan_.gu_group.size(), d_gu_group_.p, d_gu_chunk_.p, d_fd_.p, d_base_.p,
                                 d_rbase_.p, V, w, nv_, part, chunk_w_);
        launch_spmv_rows_slots<TV>(stream_, nv_, d_rbase_.p, part, y);
        launch_spmv_cols<TV>(stream_, (int)plan_.cu_tile.size(), d_cu_tile_.p, d_cu_i0_.p, d_cu_i1_.p, d_items_.p,
                             d_fd_.p, d_base_.p, V, y, nv_, xpart);
        // the first pass writes x = S^T S b; refinement passes add S^T S r to x in place
        launch_spmv_cols_fin<TV>(stream_, nv_, d_tile_unit_ptr_.p, xpart, d_perm_.p, it == 0 ? nullptr : x_dev, nullptr,
                                 nullptr, x_dev);
    }
}

// apply_ainv on caller-owned device vectors with every intermediate in the
// file-scope scratch array: no device allocation. Needs the factor built;
// calls that overlap on one GPU share the scratch and race.
void Body::apply_ainv_noalloc(const double* b_dev, double* x_dev, int precision, int refine) {
    if (!ready_) throw std::runtime_error("apply_ainv_noalloc: the factor is not built (call phd_factor_build first)");
    if (!b_dev || !x_dev) throw std::runtime_error("apply_ainv_noalloc: null vector");
    if (precision != 0 && precision != 1) throw std::runtime_error("apply_ainv_noalloc: precision must be 0 or 1");
    if (refine < 0) throw std::runtime_error("apply_ainv_noalloc: refine must be >= 0");
    if (b_dev == x_dev) throw std::runtime_error("apply_ainv_noalloc: b and x must not alias");
    if (!pverts_.empty())
        throw std::runtime_error("apply_ainv_noalloc: per-DOF Dirichlet conditions are not supported (whole vertices only)");
    const size_t need = noalloc_scratch_bytes(precision);
    if (need > kNoAllocScratchBytes)
        throw std::runtime_error("apply_ainv_noalloc: this factor needs " + std::to_string(need) +
                                 " bytes of scratch; the maximum is " + std::to_string(kNoAllocScratchBytes));
    hip_check(hipSetDevice(device_), "hipSetDevice");
    char* scr = static_cast<char*>(scratch_);
    if (precision == 1) ainv_passes<float>(b_dev, x_dev, refine, scr);
    else ainv_passes<double>(b_dev, x_dev, refine, scr);
    hip_check(hipGetLastError(), "apply_ainv_noalloc launch");
    hip_check(hipStreamSynchronize(stream_), "apply_ainv_noalloc");
}

// ---------------------------------------------------------------------------
// Type-II Anderson acceleration on device vectors (forward_gpu.cpp:1346-1379,
// backward_gpu.cpp:822-851): the last (x, g), a window of their differences,
// and the Gram matrix of the g differences cached as they arrive.
// ---------------------------------------------------------------------------
struct DevAnderson {
    int window = 5;
    int n = 0;
    DevBuf<double>* Q = nullptr;  // 8 buffers: 0..6 differences of x, 7 the last x
    DevBuf<double>* G = nullptr;  // 8 buffers: 0..6 differences of g, 7 the last g
    bool have_last = false;
    bool replay = false;          // the replay's variant: reg 1e-10 max(tr / m, 1), no restart
    std::deque<int> diffs;        // pool indices of the differences, oldest first
    std::vector<int> free_slots;
    double gram[7][7] = {};       // dG_a . dG_b by pool index, filled as differences arrive
    void init(int w, int n_, DevBuf<double>* q, DevBuf<double>* g) {
        window = std::min(w, 6);
        n = n_;
        Q = q;
        G = g;
        clear();
        for (int i = 0; i < 8; ++i) { Q[i].alloc(n); G[i].alloc(n); }
    }
    void clear() {
        diffs.clear();
        free_slots.clear();
        for (int i = 6; i >= 0; --i) free_slots.push_back(i);
        have_last = false;
    }
};

// One Anderson step (forward_gpu.cpp:1346-1379, backward_gpu.cpp:822-851)
// on the map x -> x + g at x = xprev: records the newest difference pair
// (x - x_last, g - g_last), forms the Gram matrix of the g differences from
// the entries cached as the differences arrived (one new row per call), solves
// the regularized least squares and writes xprev + g - sum gamma_a (dQ_a + dG_a)
// into out. `extra` dot pairs ride along in the same reduction. Returns false
// when nothing was combined: fewer than two iterates, or a restart because
// |gamma| >= 10 (the history is then dropped, the current iterate included).
bool Body::anderson_step(DevAnderson& aa, const double* xprev, const double* g, double* out,
                         const std::vector<std::pair<const double*, const double*>>& extra,
                         std::vector<double>* extra_out) {
    const int n3 = aa.n;
    // the newest difference pair and the new last iterate in one pass
    double *dq = nullptr, *dg = nullptr;
    if (aa.have_last) {
        const int pi = aa.free_slots.back();
        aa.free_slots.pop_back();
        dq = aa.Q[pi].p;
        dg = aa.G[pi].p;
        aa.diffs.push_back(pi);
    }
    launch_aa_push(stream_, n3, xprev, g, aa.Q[7].p, aa.G[7].p, dq, dg);
    aa.have_last = true;
    const int m = (int)aa.diffs.size();
    std::vector<std::pair<const double*, const double*>> pr;
    if (m >= 1) {
        const int nw = aa.diffs.back();
        for (int a = 0; a < m; ++a) pr.push_back({aa.G[aa.diffs[a]].p, aa.G[nw].p});
        for (int a = 0; a < m; ++a) pr.push_back({aa.G[aa.diffs[a]].p, g});
    }
    for (const auto& e : extra) pr.push_back(e);
    if (pr.empty()) return false;
    const std::vector<double> dd = dots(pr);
    if (extra_out) extra_out->assign(dd.begin() + 2 * m, dd.end());
    if (m == 0) return false;
    const int nw = aa.diffs.back();
    for (int a = 0; a < m; ++a) aa.gram[aa.diffs[a]][nw] = aa.gram[nw][aa.diffs[a]] = dd[a];
    std::vector<double> GtG(m * m), rr(m);
    double sq = 0;
    for (int a = 0; a < m; ++a) {
        for (int b = 0; b < m; ++b) GtG[a * m + b] = aa.gram[aa.diffs[a]][aa.diffs[b]];
        rr[a] = dd[m + a];
        sq += GtG[a * m + a];
    }
    const double reg = aa.replay ? 1e-10 * std::max(sq / std::max(m, 1), 1.0) : std::max(1e-6 * sq / std::max(m, 1), 1e-12);
    for (int a = 0; a < m; ++a) GtG[a * m + a] += reg;
    const std::vector<double> gamma = small_ldlt_solve(GtG, m, rr);
    if (!aa.replay && nrm2(gamma) >= 10.0) {
        aa.clear();
        return false;
    }
    std::vector<const double*> pq(m), pg(m);
    for (int a = 0; a < m; ++a) { pq[a] = aa.Q[aa.diffs[a]].p; pg[a] = aa.G[aa.diffs[a]].p; }
    launch_aa_combine(stream_, n3, xprev, g, m, pq.data(), pg.data(), gamma.data(), out);
    while ((int)aa.diffs.size() > aa.window - 1) {
        aa.free_slots.push_back(aa.diffs.front());
        aa.diffs.pop_front();
    }
    return true;
}

// ---------------------------------------------------------------------------
// Forward step.
// ---------------------------------------------------------------------------
void Body::forward(const double* q, const double* v, const double* f_ext, double dt,
                   const std::map<std::string, double>& opt, double* q_next, double* v_next, std::vector<int>& active) {
    forward_impl(q, v, f_ext, dt, opt, q_next, v_next, active, nullptr, nullptr);
}

// The forward PD loop; with replay_active set it is the backward's
// fixed-active-set replay (ReplayAlgPhdFixedActiveSet,
// deformable_projective_dynamics_backward_gpu.cpp:113-418): only the listed
// contacts take part, without friction, from lambda = 0, with the obstacle
// normals re-evaluated at every iterate, r_j = max(W_jj, 1e-10) (no h^2),
// W_ij = H_ij (the reference scales the column solves of row i by n_i on both
// sides, :203-220, so for unit normals and one factor per axis the normal
// products drop out), lambda clamped below only, and Anderson acceleration with
// reg = 1e-10 max(tr / m, 1) and no restart (:381-405). The result goes to
// *replay_out; the tape, the lambda warm start and the frame statistics are
// left as they were.
void Body::forward_impl(const double* q, const double* v, const double* f_ext, double dt,
                        const std::map<std::string, double>& opt, double* q_next, double* v_next,
                        std::vector<int>& active, const std::vector<int>* replay_active, ReplayFrame* replay_out) {
    auto tstart = std::chrono::steady_clock::now();
    const bool rp = replay_active != nullptr;
    const int max_pd_iter = (int)opt_req(opt, "max_pd_iter");
    const double abs_tol = opt_req(opt, "abs_tol");
    const double rel_tol = opt_req(opt, "rel_tol");
    const int verbose = (int)opt_req(opt, "verbose");
    const int threads = (int)opt_req(opt, "thread_ct");
    const int aa_window = (int)opt_or(opt, "aa_window", 5);
    check(aa_window >= 0 && aa_window <= 6, "aa_window must be in [0, 6].");
    check((int)act_.size() == act_dofs_, "Inconsistent actuation size.");
    const double fmu = opt_or(opt, "friction_mu", 0.0);
    const double restitution = opt_or(opt, "restitution", 0.0);
    const int precision = (int)opt_or(opt, "phd_precision", 0.0);
    check(precision == 0 || precision == 1, "phd_precision must be 0 (FP64) or 1 (mixed FP32/FP64).");
    local_fp32_ = opt_or(opt, "phd_local_fp32", 0.0) != 0.0;
    h32_ = precision == 1 && opt_or(opt, "phd_contact_h32", 1.0) != 0.0;
    setup(dt, threads);
    PhaseTimer t_total(this, kFwdTotal);
    const int n3 = dofs_;
    const double h = dt, mass = element_volume_ * density_, h2m = h * h / mass, inv_h2m = mass / (h * h);
    (void)h2m;
    // new-rollout detection for the tape (:437-449)
    if (!rp && !tape_.empty()) {
        const auto& last = tape_.back();
        double dn = 0, a = 0, b = 0;
        for (int i = 0; i < n3; ++i) { dn += (last.q_sol[i] - q[i]) * (last.q_sol[i] - q[i]); a += last.q_sol[i] * last.q_sol[i]; b += q[i] * q[i]; }
        if (!(std::sqrt(dn) <= 1e-8 * std::max(1.0, std::max(std::sqrt(a), std::sqrt(b))))) tape_.clear();
    }
    set_inertial_target(q, v, f_ext, dt);
    d_q_.upload(q, n3);
    const int K = K_;
    const DevBoundary bnd = dev_boundary();
    // per-candidate contact state lives on the device for the whole frame
    std::vector<double> lambda(K, 0.0);
    const bool do_fric = !rp && K > 0 && fmu > 0;
    const unsigned char* enabled = nullptr;
    if (rp) {  // the replay's contacts: a subset of the candidates
        std::unordered_map<int, int> cidx;
        for (int c = 0; c < K; ++c) cidx[cand_[c]] = c;
        std::vector<unsigned char> en(std::max(K, 1), 0);
        for (int node : *replay_active) {
            auto it = cidx.find(node);
            check(it != cidx.end(), "Active contact " + std::to_string(node) + " is not a frictional-boundary vertex.");
            en[it->second] = 1;
        }
        d_enabled_.upload(en);
        enabled = d_enabled_.p;
    }
    const bool mesh = boundary_.type == 5;
    check(!(rp && mesh), "The fixed-active-set replay does not support the mesh boundary in this build.");
    if (K > 0) {
        if (mesh) {
            // frame-start anchors and the candidate filter (:460-509); the
            // filtered-out candidates stay out of every contact set this frame
            d_enabled_.alloc(K);
            launch_contact_frames(stream_, K, bnd, d_cand_.p, d_q_.p, d_nrm_.p, d_t1_.p, d_t2_.p, d_enabled_.p);
            enabled = d_enabled_.p;
        } else {
            launch_contact_frames(stream_, K, bnd, d_cand_.p, d_q_.p, d_nrm_.p, d_t1_.p, d_t2_.p);
        }
        std::vector<double> rprec(K);
        for (int c = 0; c < K; ++c) rprec[c] = std::max((rp ? 1.0 : h * h) * Hdiag_[c], 1e-10);
        if (!rp && !prev_candidates_.empty() && prev_lambda_.size() == prev_candidates_.size()) {
            std::unordered_map<int, int> pidx;
            for (size_t i = 0; i < prev_candidates_.size(); ++i) pidx[prev_candidates_[i]] = (int)i;
            for (int c = 0; c < K; ++c) {
                auto it = pidx.find(cand_[c]);
                if (it != pidx.end()) lambda[c] = prev_lambda_[it->second];
            }
        }
        d_rprec_.upload(rprec);
        d_lambda_.upload(lambda);
        d_lt1_.alloc(K); d_lt2_.alloc(K);
        hip_check(hipMemsetAsync(d_lt1_.p, 0, sizeof(double) * K, stream_), "lt1=0");
        hip_check(hipMemsetAsync(d_lt2_.p, 0, sizeof(double) * K, stream_), "lt2=0");
        hip_check(hipMemsetAsync(d_F_.p, 0, sizeof(double) * 3 * K, stream_), "F=0");
        launch_gather_nodes(stream_, K, d_cand_.p, d_q_.p, d_qframeC_.p);  // vertex positions at the frame start
        if ((int)d_rhsK_.n < K) d_rhsK_.alloc(K);
    }
    const int warm = (int)opt_or(opt, "phd_warm_start", 0.0);
    check(warm >= 0 && warm <= 2, "phd_warm_start must be 0 (q_t), 1 (learned, falls back to q + h v) or 2 (q + h v).");
    if (warm > 0 && !rp) warm_start(v, h, inv_h2m, warm == 1);
    DevAnderson aa;
    aa.init(aa_window, n3, d_hist_q_, d_hist_g_);
    aa.replay = rp;
    hip_check(hipMemsetAsync(d_pdrhs_prev_.p, 0, sizeof(double) * n3, stream_), "prev pd_rhs");
    bool dense64 = precision == 0;
    bool converged = false;
    int iters = 0;
    const int set_masks[4] = {kNcpJ, kNcpP, kNcpF, kNcpG};
    const int fric_masks[2] = {kNcpL, kNcpN};
    int* const lists = d_ncp_lists_.p;  // J, P, F, G lists (K each), then the friction lists
    const int* const listJ = lists;
    const int* const listP = lists + K;
    const int* const listF = lists + 2 * (size_t)K;
    const int* const listG = lists + 3 * (size_t)K;
    for (int iter = 0; iter < max_pd_iter; ++iter) {
        {
            PhaseTimer tl(this, kFwdLocal);
            local_rhs(d_q_.p, inv_h2m);
        }
        hip_check(hipMemcpyAsync(d_qprev_.p, d_q_.p, sizeof(double) * n3, hipMemcpyDeviceToDevice, stream_), "q_prev");
        // residual of the global system with the previous contact impulses
        {
            PhaseTimer ts(this, kFwdSolve);
            launch_residual(stream_, nv_, ell_w_, d_Aie_.p, d_Axe_.p, d_b_.p, d_q_.p, d_dmask_.p, d_r_.p, d_pmask_.p);
            if (K > 0) launch_add_node_force(stream_, K, d_cand_.p, d_F_.p, d_r_.p);
            if (precision == 1) spmv_S<float>(d_r_.p, w32_, y32_, part32_, V32_.p);
            else spmv_S<double>(d_r_.p, w64_, y64_, part64_, V64_.p);
        }
        if (K > 0) {
            PhaseTimer tc(this, kFwdContact);
            std::unique_ptr<PhaseTimer> tcs(new PhaseTimer(this, kFwdCStage));
            // (A^-1 b)_C = q_C + (S^T S r)_C - H F_prev: the candidate rows of S^T y,
            // the candidate positions and their obstacle distances
            if (precision == 1) contact_gather<float>(y32_, d_uC_.p); else contact_gather<double>(y64_, d_uC_.p);
            launch_gather_nodes(stream_, K, d_cand_.p, d_q_.p, d_dF_.p);
            launch_contact_distance(stream_, K, bnd, d_cand_.p, d_q_.p, d_delta_.p, d_nrm_.p);
            if (rp) launch_contact_frames(stream_, K, bnd, d_cand_.p, d_q_.p, d_nrm_.p, d_t1_.p, d_t2_.p);
            tcs.reset();
            std::unique_ptr<PhaseTimer> tcset(new PhaseTimer(this, kFwdCSets));
            // omega, E, h (forward_gpu.cpp:1160-1197) and the index sets. Candidates with
            // omega == 0 (separated, lambda == 0) decouple exactly: their row of
            // Omega W Omega + E is E_c e_c, so the dense work runs on J = {omega != 0};
            // P adds the support of the previous impulses, G that of g's impulses.
            launch_ncp_sets(stream_, K, enabled, d_delta_.p, d_lambda_.p, d_rprec_.p, d_nrm_.p, d_t1_.p, d_t2_.p, d_dF_.p,
                            d_uC_.p, d_F_.p, d_lt1_.p, d_lt2_.p, do_fric ? 1 : 0, d_om_.p, d_E_.p, d_hvec_.p, d_Fg_.p,
                            d_ncp_flags_.p);
            launch_ncp_compact(stream_, K, d_ncp_flags_.p, 4, set_masks, lists, d_ncp_counts_.p, d_posJ_.p);
            hip_check(hipMemcpyAsync(reinterpret_cast<int*>(pin_small_), d_ncp_counts_.p, 4 * sizeof(int), hipMemcpyDeviceToHost, stream_),
                      "set sizes");
            hip_check(hipStreamSynchronize(stream_), "set sizes");
            const int* cnt = reinterpret_cast<const int*>(pin_small_);
            const int nJ = cnt[0], nP = cnt[1], nF = cnt[2], nG = cnt[3];
            tcset.reset();
            std::unique_ptr<PhaseTimer> tcap(new PhaseTimer(this, kFwdCApply));
            // uC -= H F_prev on P; the compressed system's rhs on J with (A^-1 g)_C =
            // (A^-1 b)_C + H F_g, F_g = omega lambda n + lt1 t1 + lt2 t2 (zero sums
            // when a column set is empty)
            if (h32_) {
                launch_h_apply_idx<float>(stream_, nP, listP, nF, listF, d_H32_.p, K, d_F_.p, d_hprev_.p);
                launch_h_apply_idx<float>(stream_, nJ, listJ, nG, listG, d_H32_.p, K, d_Fg_.p, d_tmpK3_.p);
            } else {
                launch_h_apply_idx<double>(stream_, nP, listP, nF, listF, d_H_.p, K, d_F_.p, d_hprev_.p);
                launch_h_apply_idx<double>(stream_, nJ, listJ, nG, listG, d_H_.p, K, d_Fg_.p, d_tmpK3_.p);
            }
            launch_ncp_rhs(stream_, nP, listP, d_hprev_.p, d_posJ_.p, d_tmpK3_.p, d_hvec_.p, d_om_.p, d_nrm_.p, d_uC_.p,
                           d_rhsK_.p);
            tcap.reset();
            int solved_prec = -1;  // precision of the solution in d_cx32_ / d_cx64_
            if (nJ > 0) {
                PhaseTimer td(this, kFwdDense);
                phase[kDenseCalls] += 1;  // counters are kept at every timing level
                phase[kDenseRows] += nJ;
                if (precision == 1 && !dense64) {
                    if (solve_contact_system(nJ, 1, listJ, rp)) solved_prec = 1;
                    else {
                        dense64 = true;  // the FP32 Cholesky broke down: FP64 for the rest of the frame
                        phase[kDenseFp32Fail] += 1;
                    }
                }
                if (solved_prec < 0) {  // FP64 system (also the fallback when the FP32 Cholesky breaks down)
                    check(solve_contact_system(nJ, 0, listJ, rp),
                          "Dense Cholesky of the contact system failed (matrix not positive definite).");
                    solved_prec = 0;
                }
            }
            std::unique_ptr<PhaseTimer> tlam(new PhaseTimer(this, kFwdCApply));
            const double lam_max = rp ? INFINITY : 1e8;
            if (solved_prec == 1)
                launch_ncp_lambda<float>(stream_, K, d_posJ_.p, d_cx32_.p, d_hvec_.p, d_E_.p, d_om_.p, lam_max, d_lambda_.p);
            else
                launch_ncp_lambda<double>(stream_, K, d_posJ_.p, d_cx64_.p, d_hvec_.p, d_E_.p, d_om_.p, lam_max, d_lambda_.p);
            tlam.reset();
            // NCP-integrated Coulomb friction (:1255-1302)
            PhaseTimer tfr(this, kFwdFric);
            if (do_fric) {
                int* flist = lists + 4 * (size_t)K;
                launch_fric_sets(stream_, K, d_lambda_.p, d_om_.p, d_nrm_.p, d_Fn_.p, d_ncp_flags_.p);
                launch_ncp_compact(stream_, K, d_ncp_flags_.p, 2, fric_masks, flist, d_ncp_counts_.p + 4, d_posL_.p);
                if (h32_)
                    launch_h_apply_idx_dev<float>(stream_, K, d_ncp_counts_.p + 4, flist, d_ncp_counts_.p + 5, flist + K,
                                                  d_H32_.p, K, d_Fn_.p, d_tmpK3_.p);
                else
                    launch_h_apply_idx_dev<double>(stream_, K, d_ncp_counts_.p + 4, flist, d_ncp_counts_.p + 5, flist + K,
                                                   d_H_.p, K, d_Fn_.p, d_tmpK3_.p);
                launch_fric_coulomb(stream_, K, d_lambda_.p, d_posL_.p, d_tmpK3_.p, d_uC_.p, d_qframeC_.p, d_nrm_.p,
                                    d_t1_.p, d_t2_.p, d_Hdiag_.p, fmu, d_lt1_.p, d_lt2_.p);
            }
            // new impulses; the S pass carries only the change
            launch_ncp_impulses(stream_, K, d_om_.p, d_lambda_.p, d_nrm_.p, d_lt1_.p, d_lt2_.p, d_t1_.p, d_t2_.p,
                                do_fric ? 1 : 0, d_F_.p, d_dF_.p);
            if (precision == 1) contact_scatter<float>(d_dF_.p, y32_); else contact_scatter<double>(d_dF_.p, y64_);
        }
        {
            PhaseTimer ts(this, kFwdSolve);
            if (precision == 1) spmv_ST<float>(y32_, xpart32_, V32_.p, d_qprev_.p, d_dval_full_.p, d_q_.p);
            else spmv_ST<double>(y64_, xpart64_, V64_.p, d_qprev_.p, d_dval_full_.p, d_q_.p);
            partial_dirichlet(d_q_.p, d_qprev_.p, d_dval_full_.p);
        }
        ++iters;
        // Anderson acceleration
        if (aa_window > 0) {
            PhaseTimer ta(this, kFwdAA);
            launch_axpbypcz(stream_, n3, 1.0, d_q_.p, -1.0, d_qprev_.p, 0.0, nullptr, d_g_.p);
            if (anderson_step(aa, d_qprev_.p, d_g_.p, d_q_.p, {}, nullptr))
                launch_set_dirichlet(stream_, (int)ddof_.size(), d_ddof_.p, d_dval_.p, d_q_.p);
        }
        // dual gate (:1382-1397)
        PhaseTimer tg(this, kFwdGate);
        launch_axpbypcz(stream_, n3, 1.0, d_q_.p, -1.0, d_qprev_.p, 0.0, nullptr, d_g_.p);
        launch_axpbypcz(stream_, n3, 1.0, d_pdrhs_.p, -1.0, d_pdrhs_prev_.p, 0.0, nullptr, d_tmp_.p);
        const std::vector<double> gd = dots({{d_g_.p, d_g_.p}, {d_qprev_.p, d_qprev_.p}, {d_tmp_.p, d_tmp_.p},
                                             {d_pdrhs_prev_.p, d_pdrhs_prev_.p}});
        hip_check(hipMemcpyAsync(d_pdrhs_prev_.p, d_pdrhs_.p, sizeof(double) * n3, hipMemcpyDeviceToDevice, stream_), "prev");
        const double abs_err = std::sqrt(gd[0]), rhs_norm = std::sqrt(gd[1]), rc = std::sqrt(gd[2]), rr = std::sqrt(gd[3]);
        if (verbose > 0) printf("alg_phd PD iter %d  dq=%g  dpd_rhs=%g\n", iter, abs_err, rc);
        if (iter > 0 && abs_err <= rel_tol * rhs_norm + abs_tol && rc <= rel_tol * rr + abs_tol) { converged = true; break; }
    }
    hip_check(hipStreamSynchronize(stream_), "forward");
    std::vector<double> qs(n3);
    d_q_.download(qs.data(), n3);
    std::vector<double> nrm(3 * K), lt1(K, 0.0), lt2(K, 0.0), omega_last(K, 0.0);
    if (K > 0) {
        d_lambda_.download(lambda.data(), K);
        d_lt1_.download(lt1.data(), K);
        d_lt2_.download(lt2.data(), K);
        d_om_.download(omega_last.data(), K);
        d_nrm_.download(nrm.data(), 3 * K);
    }
    if (rp) {
        ReplayFrame& fr = *replay_out;
        fr = ReplayFrame();
        fr.active = *replay_active;
        fr.q_sol = qs;
        std::unordered_map<int, int> cidx;
        for (int c = 0; c < K; ++c) cidx[cand_[c]] = c;
        for (int node : fr.active) {
            const int c = cidx[node];
            fr.lambda.push_back(lambda[c]);
            fr.omega.push_back(omega_last[c]);
            fr.normals.push_back({nrm[3 * c], nrm[3 * c + 1], nrm[3 * c + 2]});
        }
        return;
    }
    if (!converged) fprintf(stderr, "[Warning] alg_phd: PD loop did not converge.\n");
    // restitution post-pass (:1483-1531)
    if (restitution > 0 && K > 0) {
        // mesh: the opposite surface at the last iterate's start (the positions
        // of the last UpdateVertices call), slop 0.1 radius (:1485-1505)
        std::vector<double> qf;
        if (mesh) { qf.resize(n3); d_qprev_.download(qf.data(), n3); }
        const double slop = opt_or(opt, "friction_contact_slop", mesh ? 0.1 * boundary_.radius : 1e-8);
        for (int c = 0; c < K; ++c) {
            if (lambda[c] <= 0) continue;
            const int node = cand_[c];
            double* x = &qs[3 * node];
            double sd, nh[3];
            if (mesh) {
                double cl[3], fn[3];
                if (!boundary_.mesh_closest(x, qf.data(), node, cl, fn)) continue;
                const double* rn = &nrm[3 * c];
                double dp[3] = {x[0] - cl[0], x[1] - cl[1], x[2] - cl[2]};
                sd = dp[0] * rn[0] + dp[1] * rn[1] + dp[2] * rn[2] - boundary_.radius;
                const double dl = std::sqrt(dp[0] * dp[0] + dp[1] * dp[1] + dp[2] * dp[2]);
                for (int d = 0; d < 3; ++d) nh[d] = dl > 1e-12 ? dp[d] / dl : fn[d];
                if (nh[0] * rn[0] + nh[1] * rn[1] + nh[2] * rn[2] < 0) for (int d = 0; d < 3; ++d) nh[d] = -nh[d];
            } else {
                boundary_.eval(x, sd, nh);  // obstacle at the converged position
            }
            const double nn = std::sqrt(nh[0] * nh[0] + nh[1] * nh[1] + nh[2] * nh[2]);
            if (nn <= 1e-12) continue;
            for (int d = 0; d < 3; ++d) nh[d] /= nn;
            if (sd > slop) continue;
            const double vin = v[3 * node] * nh[0] + v[3 * node + 1] * nh[1] + v[3 * node + 2] * nh[2];
            if (vin < 0) {
                double vout = 0;
                for (int d = 0; d < 3; ++d) vout += (x[d] - q[3 * node + d]) / h * nh[d];
                const double dv = -restitution * vin - vout;
                if (dv > 0) for (int d = 0; d < 3; ++d) x[d] += h * dv * nh[d];
            }
        }
    }
    for (int i = 0; i < n3; ++i) { q_next[i] = qs[i]; v_next[i] = (qs[i] - q[i]) / h; }
    active.clear();
    for (int c = 0; c < K; ++c) if (lambda[c] > 0) active.push_back(cand_[c]);
    prev_candidates_ = cand_;
    prev_lambda_ = lambda;
    last_lambda_ = lambda;
    last_lt1_ = lt1;
    last_lt2_ = lt2;
    ReplayFrame fr;
    fr.active = active;
    fr.q_sol = qs;
    std::unordered_map<int, int> cidx;
    for (int c = 0; c < K; ++c) cidx[cand_[c]] = c;
    for (int node : active) {
        const int c = cidx[node];
        fr.lambda.push_back(lambda[c]);
        fr.omega.push_back(omega_last[c]);
        fr.normals.push_back({nrm[3 * c], nrm[3 * c + 1], nrm[3 * c + 2]});
    }
    tape_.push_back(std::move(fr));
    if (tape_.size() > 8192) tape_.erase(tape_.begin(), tape_.begin() + (tape_.size() - 8192));
    StepStats st;
    st.iters = iters;
    st.converged = converged ? 1 : 0;
    st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - tstart).count();
    frame_stats.push_back(st);
}

// ---------------------------------------------------------------------------
// Backward step.
// ---------------------------------------------------------------------------
void Body::backward(const double* q, const double* v, const double* f_ext, double dt, const double* q_next,
                    const double* v_next, const std::vector<int>& active, const double* dl_dq_next, const double* dl_dv_next,
                    const std::map<std::string, double>& opt, double* dl_dq, double* dl_dv, double* dl_df_ext,
                    double* dl_dmat_w, double* dl_dstate_p, double* dl_dw_elem) {
    (void)v_next;
    const int max_pd_iter = (int)opt_req(opt, "max_pd_iter");
    const double abs_tol = opt_req(opt, "abs_tol");
    const double rel_tol = opt_req(opt, "rel_tol");
    opt_req(opt, "verbose");
    const int threads = (int)opt_req(opt, "thread_ct");
    const int precision = (int)opt_or(opt, "phd_precision", 0.0);
    check(precision == 0 || precision == 1, "phd_precision must be 0 (FP64) or 1 (mixed FP32/FP64).");
    check((int)act_.size() == act_dofs_, "Inconsistent actuation size.");
    setup(dt, threads);
    const bool exact = opt_or(opt, "alg_phd_bwd_exact_kpd", 0.0) > 0.5;
    if (exact != kpd_exact_) { kpd_exact_ = exact; build_materials(); }
    const int kfilter = (int)opt_or(opt, "alg_phd_bwd_kpd_filter", 0.0);
    const double trust_rho = opt_or(opt, "alg_phd_bwd_trust_radius", 0.9);
    check(kfilter >= 0 && kfilter <= 3, "alg_phd_bwd_kpd_filter must be 0 (none), 1 (clamp), 2 (abs) or 3 (trust region).");
    check(kfilter != 3 || (trust_rho > 0 && trust_rho < 1), "alg_phd_bwd_trust_radius must lie in (0, 1).");
    const bool state_wb = opt_or(opt, "alg_phd_bwd_state_woodbury", 0.0) > 0.5;
    last_bwd = BackwardStats();
    last_bwd.samples = (double)ne_ * samples_;
    PhaseTimer t_total(this, kBwdTotal);
    const int n3 = dofs_;
    const double h = dt, inv_h = 1 / h, mass = element_volume_ * density_, h2m = h * h / mass, inv_h2m = mass / (h * h);
    for (auto& kv : dirichlet_) check(q_next[kv.first] == kv.second, "Boundary conditions violated.");
    std::vector<double> rhs_qbar(n3);
    for (int i = 0; i < n3; ++i) {
        dl_dq[i] = -dl_dv_next[i] * inv_h;
        dl_dv[i] = 0;
        dl_df_ext[i] = 0;
        rhs_qbar[i] = dl_dq_next[i] + dl_dv_next[i] * inv_h;
    }
    const int nE = (int)energies_.size();
    for (int i = 0; i < nE; ++i) dl_dmat_w[i] = 0;
    for (int k = 0; k < n_state_params_; ++k) dl_dstate_p[k] = 0;
    if (dl_dw_elem) std::fill(dl_dw_elem, dl_dw_elem + (size_t)nE * ne_, 0.0);
    // proximal-map Jacobians at q_next
    d_q_.upload(q_next, n3);
    if (precision == 1) { d_W32_.alloc((size_t)ne_ * samples_ * 81); d_W64_.release(); }
    else { d_W64_.alloc((size_t)ne_ * samples_ * 81); d_W32_.release(); }
    const double* kn[4] = {d_kn_[0].p, d_kn_[1].p, d_kn_[2].p, d_kn_[3].p};
    // Proximal-map Jacobians, filtered per sample. Trust region (mode 3): stage 0
    // is the unfiltered operator; the backbone moves to stage 1 (clamp) and 2
    // (abs), both bounded below by -rho, when its fixed point diverges.
    int stage = kfilter == 3 ? 0 : 1;
    d_nfilt_.alloc(1);
    auto build_kpd = [&]() {
        PhaseTimer t_kpd(this, kBwdKpd);
        double theta = 1.0, lo = -INFINITY;
        int active = kfilter != 0 && !(kfilter == 3 && stage == 0);
        if (kfilter == 1) theta = 0.0;
        if (kfilter == 2) theta = -1.0;
        if (kfilter == 3) { theta = stage == 1 ? 0.0 : -1.0; lo = -trust_rho; }
        hip_check(hipMemsetAsync(d_nfilt_.p, 0, sizeof(int), stream_), "filter count");
        if (precision == 1)
            launch_kpd_build<float>(stream_, mat_, has_cor_ ? d_kc_.p : nullptr, has_vol_ ? d_kv_.p : nullptr, kn, d_q_.p,
                                    active, theta, lo, d_nfilt_.p, d_W32_.p);
        else
            launch_kpd_build<double>(stream_, mat_, has_cor_ ? d_kc_.p : nullptr, has_vol_ ? d_kv_.p : nullptr, kn, d_q_.p,
                                     active, theta, lo, d_nfilt_.p, d_W64_.p);
    };
    build_kpd();
    // tape
    ReplayFrame replay;
    bool from_tape = false;
    if (!tape_.empty()) {
        const auto& tail = tape_.back();
        bool match = tail.active == active && (int)tail.q_sol.size() == n3;
        if (match) {
            double dn = 0, a = 0, b = 0;
            for (int i = 0; i < n3; ++i) { dn += (tail.q_sol[i] - q_next[i]) * (tail.q_sol[i] - q_next[i]); a += tail.q_sol[i] * tail.q_sol[i]; b += q_next[i] * q_next[i]; }
            match = std::sqrt(dn) <= 1e-8 * std::max(1.0, std::max(std::sqrt(a), std::sqrt(b)));
        }
        if (match) { replay = tail; tape_.pop_back(); from_tape = true; }
        else if ((int)tail.q_sol.size() == n3) {  // stale tape (:730-737)
            double dn = 0, a = 0, b = 0;
            for (int i = 0; i < n3; ++i) { dn += (tail.q_sol[i] - q_next[i]) * (tail.q_sol[i] - q_next[i]); a += tail.q_sol[i] * tail.q_sol[i]; b += q_next[i] * q_next[i]; }
            if (std::sqrt(dn) / std::max(1.0, std::max(std::sqrt(a), std::sqrt(b))) > 1e-6) tape_.clear();
        }
    }
    if (!from_tape && !active.empty()) {
        // ReplayAlgPhdFixedActiveSet (:113-418) on the device. The replay reloads
        // d_q_ with a blocking copy, so the K_pd build queued above (which reads
        // q_next from d_q_) must finish first; d_q_ is restored after.
        hip_check(hipStreamSynchronize(stream_), "K_pd before replay");
        std::vector<double> qn(n3), vn(n3);
        std::vector<int> unused;
        const std::vector<double> zeros(n3, 0.0);
        forward_impl(q, v, f_ext ? f_ext : zeros.data(), dt, opt, qn.data(), vn.data(), unused, &active, &replay);
        d_q_.upload(q_next, n3);
    }
    last_bwd.from_tape = from_tape ? 1 : 0;
    const int aa_window = (int)opt_or(opt, "aa_window", 5);
    check(aa_window >= 0 && aa_window <= 6, "aa_window must be in [0, 6].");
    const double clip = opt_or(opt, "alg_phd_bwd_clip_norm", 1e6);
    DevBuf<double> d_rhs, d_s, d_sprev, d_zero;
    d_rhs.upload(rhs_qbar);
    d_s.alloc(n3);
    d_sprev.alloc(n3);
    d_zero.alloc(n3);
    hip_check(hipMemsetAsync(d_zero.p, 0, sizeof(double) * n3, stream_), "zero");

    // safe_clip (backward_gpu.cpp:809-816) on a device vector
    auto safe_clip = [&](double* x) {
        launch_nan_to_zero(stream_, n3, x);
        if (clip > 0) {  // |x| reduced and applied on the device: no host round trip
            DotArgs a;
            a.m = 1;
            a.a[0] = a.b[0] = x;
            launch_dots(stream_, n3, a, d_dot_partial_.p, d_dot_out_.p);
            launch_clip_scale(stream_, n3, clip, d_dot_out_.p, x);
        }
    };
    // Constrained solve data (Woodbury basis); nb == 0 means the plain backbone.
    int nb = 0;
    DevBuf<double> d_G, d_Gx, d_dirs, d_Gs64;
    DevBuf<float> d_Gs32;  // G^-1 as a full symmetric matrix (leading dimension gld)
    int64_t gld = 0;
    DevBuf<int> d_bcidx, d_bcptr, d_bclist;  // basis row -> candidate; candidate -> its basis rows
    std::vector<int> b_c;            // candidate index of each basis row
    std::vector<double> b_dir;       // 3 per basis row
    // one fixed point  s <- A_c^-1 (rhs + K s)  in residual form with AA
    // div (may be null): stop early and set it when the iteration diverges: a
    // non-finite residual, an iterate at the clip norm, or a residual 1e3 x its
    // smallest so far after 10 iterations
    auto fixed_point = [&](int max_it, int* used, bool* conv, bool* div) {
        *conv = false;
        if (div) *div = false;
        double err_min = INFINITY;
        hip_check(hipMemsetAsync(d_s.p, 0, sizeof(double) * n3, stream_), "s=0");
        DevAnderson aa;
        aa.init(aa_window, n3, d_hist_q_, d_hist_g_);
        int it_used = 0;
        for (int i = 0; i < max_it; ++i) {
            hip_check(hipMemcpyAsync(d_sprev.p, d_s.p, sizeof(double) * n3, hipMemcpyDeviceToDevice, stream_), "sprev");
            kpd_apply(d_s.p, d_kz_.p, precision);
            launch_axpbypcz(stream_, n3, 1.0, d_rhs.p, 1.0, d_kz_.p, 0.0, nullptr, d_tmp_.p);
            launch_residual(stream_, nv_, ell_w_, d_Aie_.p, d_Axe_.p, d_tmp_.p, d_s.p, d_dmask_.p, d_r_.p, d_pmask_.p);
            if (precision == 1) spmv_S<float>(d_r_.p, w32_, y32_, part32_, V32_.p);
            else spmv_S<double>(d_r_.p, w64_, y64_, part64_, V64_.p);
            if (nb > 0) {
                // s <- s + A^-1 r - A^-1 J^T G^-1 J (A^-1 r + s) on the device: project the
                // candidate vectors on the basis, apply G^-1 (formed once per frame, FP32
                // copy in mixed mode), expand back to candidate impulses and scatter through
                // S_C. The constraint residual J s rides along, so with an FP32 factor and
                // G^-1 the iterates converge onto J s = 0 instead of drifting off it.
                if (precision == 1) contact_gather<float>(y32_, d_uC_.p); else contact_gather<double>(y64_, d_uC_.p);
                PhaseTimer tbs(this, kBwdBasisSolve);
                launch_gather_nodes(stream_, K_, d_cand_.p, d_sprev.p, d_tmpK3_.p);
                launch_basis_project(stream_, nb, d_bcidx.p, d_dirs.p, d_uC_.p, d_tmpK3_.p, d_rhsK_.p);
                if (precision == 1) launch_gemv_sym<float>(stream_, nb, gld, d_Gs32.p, d_rhsK_.p, d_Gx.p);
                else launch_gemv_sym<double>(stream_, nb, gld, d_Gs64.p, d_rhsK_.p, d_Gx.p);
                launch_basis_expand(stream_, K_, d_bcptr.p, d_bclist.p, d_dirs.p, d_Gx.p, d_dF_.p);
                if (precision == 1) contact_scatter<float>(d_dF_.p, y32_); else contact_scatter<double>(d_dF_.p, y64_);
            }
            if (precision == 1) spmv_ST<float>(y32_, xpart32_, V32_.p, d_sprev.p, d_zero.p, d_s.p);
            else spmv_ST<double>(y64_, xpart64_, V64_.p, d_sprev.p, d_zero.p, d_s.p);
            partial_dirichlet(d_s.p, d_sprev.p, nullptr);
            safe_clip(d_s.p);
            launch_axpbypcz(stream_, n3, 1.0, d_s.p, -1.0, d_sprev.p, 0.0, nullptr, d_g_.p);
            double err;
            if (aa_window > 0) {
                std::vector<double> e2;
                if (anderson_step(aa, d_sprev.p, d_g_.p, d_s.p, {{d_g_.p, d_g_.p}}, &e2)) safe_clip(d_s.p);
                err = std::sqrt(e2[0]);
            } else {
                err = std::sqrt(dots({{d_g_.p, d_g_.p}})[0]);
            }
            it_used = i + 1;
            const double sn = std::sqrt(dots({{d_s.p, d_s.p}})[0]);
            if (i > 0 && err < rel_tol * sn + abs_tol) { *conv = true; break; }
            if (div) {
                err_min = std::min(err_min, err);
                if (!std::isfinite(err) || (clip > 0 && sn >= 0.999 * clip) || (i >= 10 && err > 1e3 * err_min)) {
                    *div = true;
                    break;
                }
            }
        }
        if (used) *used = it_used;
    };
    int bb = 0;
    bool bb_conv = false;
    {
        PhaseTimer tb(this, kBwdBackbone);
        int total = 0;
        for (;;) {
            bool div = false;
            fixed_point(max_pd_iter, &bb, &bb_conv, (kfilter == 3 && stage < 2) ? &div : nullptr);
            total += bb;
            if (!div) break;
            ++stage;
            build_kpd();
        }
        bb = total;
        last_bwd.trust_stage = kfilter == 3 ? stage : 0;
    }
    last_backward_iters = bb;
    std::vector<double> db(n3);
    hip_check(hipStreamSynchronize(stream_), "backbone");
    d_s.download(db.data(), n3);
    {
        int nf = 0;
        d_nfilt_.download(&nf, 1);
        last_bwd.filtered = nf;
    }
    const bool diverged = clip > 0 && nrm2(db) >= clip * 0.999;
    last_bwd.backbone_iters = bb;
    last_bwd.backbone_converged = bb_conv ? 1 : 0;
    last_bwd.clipped = diverged ? 1 : 0;
    last_bwd.db_norm = nrm2(db);
    // Cascade injection (backward_gpu.cpp:878-916) from the backbone adjoint db,
    // or, with alg_phd_bwd_state_woodbury = 1 (no reference counterpart), from
    // the Woodbury-corrected adjoint z: the exact adjoint of the step with its
    // active contacts held (the reference routes only dl_dmat_w through z).
    auto inject = [&](const std::vector<double>& src) {
        std::vector<double> inj(n3);
        for (int i = 0; i < n3; ++i) inj[i] = inv_h2m * src[i];
        const double cap = nrm2(rhs_qbar) + 1.0, in = nrm2(inj);
        if (in > cap) for (double& t : inj) t *= cap / in;
        for (int i = 0; i < n3; ++i) { dl_dq[i] += inj[i]; dl_df_ext[i] += h2m * inj[i]; dl_dv[i] += h * inj[i]; }
        // through h2m f_state(q, v) (backward_gpu.cpp:904-915)
        std::vector<double> dl_df(n3), sq(n3), sv(n3), sp(n_state_params_);
        for (int i = 0; i < n3; ++i) dl_df[i] = h2m * inj[i];
        state_force_backward(q, v, dl_df.data(), sq.data(), sv.data(), sp.data());
        for (int i = 0; i < n3; ++i) { dl_dq[i] += sq[i]; dl_dv[i] += sv[i]; }
        for (int k = 0; k < n_state_params_; ++k) dl_dstate_p[k] += sp[k];
    };
    if (!diverged && !state_wb) inject(db);
    std::vector<double> z = db;
    const bool skip_wb = opt_or(opt, "skip_contact_woodbury", 0.0) > 0.5;
    const int nl = (int)replay.active.size();
    if (!diverged && !skip_wb && nl > 0) {
        const double fmu = opt_or(opt, "friction_mu", 0.0);
        const bool tb = fmu > 0;
        std::unordered_map<int, int> cidx;
        for (int c = 0; c < K_; ++c) cidx[cand_[c]] = c;
        struct BC { int ci, basis; };
        std::vector<BC> full;
        const double tlmin = opt_or(opt, "alg_phd_bwd_tangent_lambda_min", 1e-8);
        std::vector<std::array<double, 3>> T1(nl), T2(nl);
        for (int c = 0; c < nl; ++c) {
            if (!tb) break;
            const auto& n = replay.normals[c];
            int axis = 0;
            for (int d = 1; d < 3; ++d) if (std::fabs(n[d]) < std::fabs(n[axis])) axis = d;
            double e[3] = {0, 0, 0};
            e[axis] = 1;
            const double ne_ = n[0] * e[0] + n[1] * e[1] + n[2] * e[2];
            double a1[3] = {e[0] - ne_ * n[0], e[1] - ne_ * n[1], e[2] - ne_ * n[2]};
            const double an = std::sqrt(a1[0] * a1[0] + a1[1] * a1[1] + a1[2] * a1[2]);
            if (an > 1e-12) for (double& t : a1) t /= an; else { a1[0] = 1; a1[1] = 0; a1[2] = 0; }
            double a2[3] = {n[1] * a1[2] - n[2] * a1[1], n[2] * a1[0] - n[0] * a1[2], n[0] * a1[1] - n[1] * a1[0]};
            const double bn = std::sqrt(a2[0] * a2[0] + a2[1] * a2[1] + a2[2] * a2[2]);
            if (bn > 1e-12) for (double& t : a2) t /= bn; else { a2[0] = 0; a2[1] = 1; a2[2] = 0; }
            for (int d = 0; d < 3; ++d) { T1[c][d] = a1[d]; T2[c][d] = a2[d]; }
        }
        for (int c = 0; c < nl; ++c) {
            if (replay.omega[c] < 0.5) continue;
            full.push_back({c, 0});
            if (tb && replay.lambda[c] > tlmin) { full.push_back({c, 1}); full.push_back({c, 2}); }
        }
        auto dir = [&](const BC& b, double* o) {
            for (int d = 0; d < 3; ++d)
                o[d] = b.basis == 0 ? replay.omega[b.ci] * replay.normals[b.ci][d] : (b.basis == 1 ? T1[b.ci][d] : T2[b.ci][d]);
        };
        const int nbf = (int)full.size();
        if (nbf > 0) {
            std::vector<double> rf(nbf);
            double rmax = 0;
            for (int i = 0; i < nbf; ++i) {
                double dd[3];
                dir(full[i], dd);
                const int node = replay.active[full[i].ci];
                rf[i] = dd[0] * z[3 * node] + dd[1] * z[3 * node + 1] + dd[2] * z[3 * node + 2];
                rmax = std::max(rmax, std::fabs(rf[i]));
            }
            const double gate = std::max(opt_or(opt, "alg_phd_bwd_rhs_prune_abs", 1e-12), opt_or(opt, "alg_phd_bwd_rhs_prune_rel", 1e-3) * rmax);
            std::vector<int> sel;
            for (int i = 0; i < nbf; ++i) if (std::fabs(rf[i]) >= gate) sel.push_back(i);
            const int topk = (int)opt_or(opt, "alg_phd_bwd_basis_topk", 0);
            if (topk > 0 && (int)sel.size() > topk) {
                std::nth_element(sel.begin(), sel.begin() + topk, sel.end(), [&](int a, int b) { return std::fabs(rf[a]) > std::fabs(rf[b]); });
                sel.resize(topk);
            }
            nb = (int)sel.size();
            if (nb > 0) {
                b_c.resize(nb);
                b_dir.resize(3 * nb);
                for (int a = 0; a < nb; ++a) {
                    const BC& b = full[sel[a]];
                    b_c[a] = cidx.at(replay.active[b.ci]);
                    dir(b, &b_dir[3 * a]);
                }
                d_bcidx.upload(b_c);
                d_dirs.upload(b_dir);
                {  // basis rows grouped by candidate, increasing row order within a candidate
                    std::vector<int> cptr(K_ + 1, 0), clist(nb);
                    for (int a = 0; a < nb; ++a) ++cptr[b_c[a] + 1];
                    for (int c = 0; c < K_; ++c) cptr[c + 1] += cptr[c];
                    std::vector<int> fill(cptr.begin(), cptr.end() - 1);
                    for (int a = 0; a < nb; ++a) clist[fill[b_c[a]]++] = a;
                    d_bcptr.upload(cptr);
                    d_bclist.upload(clist);
                }
                std::unique_ptr<PhaseTimer> tbf(new PhaseTimer(this, kBwdBasisFactor));
                phase[kBasisRows] += nb;
                if ((int)d_rhsK_.n < nb) d_rhsK_.alloc(nb);
                gld = ((int64_t)nb + 63) / 64 * 64;
                int info = 0;
                // Mixed precision factors and inverts G in FP32 (the fixed point applies
                // an FP32 copy of G^-1 anyway, and its constraint residual rides along,
                // so the iterates still converge onto J s = 0); FP64 when the FP32
                // factor breaks down or with phd_gram32 = 0.
                bool g32 = precision == 1 && opt_or(opt, "phd_gram32", 1.0) != 0.0;
                if (g32) {
                    DevBuf<float> dGf;
                    dGf.alloc((size_t)nb * nb);
                    launch_assemble_g<float>(stream_, nb, d_bcidx.p, d_dirs.p, d_H_.p, K_, dGf.p);
                    check(rocsolver_spotrf((rocblas_handle)blas_, rocblas_fill_lower, nb, dGf.p, nb, d_info_.p) ==
                              rocblas_status_success, "spotrf (basis) failed");
                    hip_check(hipMemcpyAsync(&info, d_info_.p, sizeof(int), hipMemcpyDeviceToHost, stream_), "info");
                    hip_check(hipStreamSynchronize(stream_), "basis factor");
                    if (info == 0) {
                        check(rocsolver_spotri((rocblas_handle)blas_, rocblas_fill_lower, nb, dGf.p, nb, d_info_.p) ==
                                  rocblas_status_success, "spotri (basis) failed");
                        hip_check(hipMemcpyAsync(&info, d_info_.p, sizeof(int), hipMemcpyDeviceToHost, stream_), "info");
                        hip_check(hipStreamSynchronize(stream_), "basis inverse");
                    }
                    if (info == 0) {
                        d_Gs32.alloc((size_t)(gld * gld));
                        launch_sym_full<float, float>(stream_, nb, gld, dGf.p, d_Gs32.p);
                        hip_check(hipStreamSynchronize(stream_), "basis sym");
                    } else {
                        g32 = false;
                        ++last_bwd.gram_fp64_fallbacks;
                    }
                }
                if (!g32) {
                    d_G.alloc((size_t)nb * nb);
                    launch_assemble_g<double>(stream_, nb, d_bcidx.p, d_dirs.p, d_H_.p, K_, d_G.p);
                    check(rocsolver_dpotrf((rocblas_handle)blas_, rocblas_fill_lower, nb, d_G.p, nb, d_info_.p) ==
                              rocblas_status_success, "dpotrf (basis) failed");
                    hip_check(hipMemcpyAsync(&info, d_info_.p, sizeof(int), hipMemcpyDeviceToHost, stream_), "info");
                    hip_check(hipStreamSynchronize(stream_), "basis factor");
                    check(info == 0, "Contact basis Gram matrix is not positive definite.");
                    check(rocsolver_dpotri((rocblas_handle)blas_, rocblas_fill_lower, nb, d_G.p, nb, d_info_.p) ==
                              rocblas_status_success, "dpotri (basis) failed");
                    hip_check(hipMemcpyAsync(&info, d_info_.p, sizeof(int), hipMemcpyDeviceToHost, stream_), "info");
                    hip_check(hipStreamSynchronize(stream_), "basis inverse");
                    check(info == 0, "Contact basis Gram matrix is singular.");
                    if (precision == 1) {
                        d_Gs32.alloc((size_t)(gld * gld));
                        launch_sym_full<double, float>(stream_, nb, gld, d_G.p, d_Gs32.p);
                    } else {
                        d_Gs64.alloc((size_t)(gld * gld));
                        launch_sym_full<double, double>(stream_, nb, gld, d_G.p, d_Gs64.p);
                    }
                }
                d_Gx.alloc(nb);
                tbf.reset();
                PhaseTimer twb(this, kBwdWoodbury);
                int used = 0;
                bool wconv = false;
                // The reference caps each of its per-basis fixed points at
                // min(max_pd_iter, alg_phd_bwd_wt_max_iter) (:1073-1075); the one
                // constrained fixed point here takes the same cap.
                const int wt_cap = (int)opt_or(opt, "alg_phd_bwd_wt_max_iter", 120);
                fixed_point(std::min(max_pd_iter, std::max(1, wt_cap)), &used, &wconv, nullptr);
                last_bwd.woodbury_iters = used;
                last_bwd.woodbury_converged = wconv ? 1 : 0;
                last_bwd.basis_rows = nb;
                hip_check(hipStreamSynchronize(stream_), "woodbury");
                d_s.download(z.data(), n3);
            }
        }
    }
    if (!diverged && state_wb) inject(z);
    // dl/da and dl/d(act_w) (backward_gpu.cpp:1209-1228): z^T of the actuation
    // force's derivatives, kept only when every dl/da entry is finite
    last_dl_da_.assign(act_dofs_, 0.0);
    last_dl_dact_w_.assign(muscles_.size(), 0.0);
    if (!diverged && act_dofs_ > 0) {
        d_s.upload(z);
        d_gda_.alloc(ne_); d_gdw_.alloc(ne_);
        launch_dact(stream_, mat_, d_q_.p, d_s.p, element_volume_ / samples_, d_gda_.p, d_gdw_.p);
        hip_check(hipStreamSynchronize(stream_), "dact");
        std::vector<double> ga(ne_), gw(ne_), da(act_dofs_, 0.0), dw(muscles_.size(), 0.0);
        d_gda_.download(ga.data(), ne_);
        d_gdw_.download(gw.data(), ne_);
        bool ok = true;
        for (size_t gi = 0; gi < muscles_.size(); ++gi)
            for (int e : muscles_[gi].elems) {
                da[mus_act_[e]] = ga[e];
                dw[gi] += gw[e];
                ok = ok && std::isfinite(ga[e]);
            }
        if (ok) { last_dl_da_ = da; last_dl_dact_w_ = dw; }
    }
    // dl/d(mat_w)
    const bool skip_mat = opt_or(opt, "skip_material_grad", 0.0) > 0.5;
    if (!skip_mat && !diverged) {
        PhaseTimer tm(this, kBwdMatGrad);
        d_s.upload(z);
        d_gcor_.alloc(ne_); d_gvol_.alloc(ne_); d_gnh_.alloc((size_t)std::max(1, n_nh_) * ne_);
        launch_dmatw(stream_, mat_, d_q_.p, d_s.p, element_volume_ / samples_, has_cor_ ? d_gcor_.p : nullptr,
                     has_vol_ ? d_gvol_.p : nullptr, n_nh_ > 0 ? d_gnh_.p : nullptr);
        hip_check(hipStreamSynchronize(stream_), "dmatw");
        std::vector<double> gc(ne_, 0.0), gv(ne_, 0.0), gn((size_t)std::max(1, n_nh_) * ne_, 0.0);
        if (has_cor_) d_gcor_.download(gc.data(), ne_);
        if (has_vol_) d_gvol_.download(gv.data(), ne_);
        if (n_nh_ > 0) d_gnh_.download(gn.data(), (size_t)n_nh_ * ne_);
        std::vector<double> contrib(nE, 0.0);
        bool ok = true;
        int hslot = 0;
        for (int ei = 0; ei < nE; ++ei) {
            const Energy& en = energies_[ei];
            const double* src = en.type == kCorotated ? gc.data() : en.type == kVolume ? gv.data() : &gn[(size_t)(hslot++) * ne_];
            double s = 0;
            for (int e = 0; e < ne_; ++e) if (en.member[e]) s += src[e];
            if (!std::isfinite(s)) ok = false;
            contrib[ei] = s;
        }
        if (ok) {
            for (int ei = 0; ei < nE; ++ei) dl_dmat_w[ei] += contrib[ei];
            if (dl_dw_elem) {
                hslot = 0;
                for (int ei = 0; ei < nE; ++ei) {
                    const Energy& en = energies_[ei];
                    const double* src = en.type == kCorotated ? gc.data() : en.type == kVolume ? gv.data() : &gn[(size_t)(hslot++) * ne_];
                    for (int e = 0; e < ne_; ++e) if (en.member[e]) dl_dw_elem[(size_t)ei * ne_ + e] = src[e];
                }
            }
        }
    }
}

}  // namespace phd
