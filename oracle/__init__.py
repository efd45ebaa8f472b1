"""FP64 CPU oracle for the DiffPhD `pd_eigen_alg_phd` forward/backward step.

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg, where it is the checker (or the timed CPU
baseline), never the product path.

`OracleDeformable` mirrors the method names of the reference's SWIG class
`HexDeformable`/`TetDeformable` (cpp/core/src/py_diff_pd_core.i:49-74,
cpp/core/include/fem/deformable.h:23-141) so that a test can build the same
scene on the oracle and on the product with one helper.
"""
import ctypes
import os
import struct
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None


def build(force=False):
    """Compile oracle.cpp with g++ into oracle/_build/liboracle.so."""
    os.makedirs(os.path.join(_HERE, "_build"), exist_ok=True)
    src = os.path.join(_HERE, "oracle.cpp")
    if force or not os.path.exists(_LIB_PATH) or os.path.getmtime(_LIB_PATH) < os.path.getmtime(src):
        tmp = _LIB_PATH + ".tmp"
        subprocess.check_call(["g++", "-O3", "-std=c++17", "-fopenmp", "-fPIC", "-shared", "-o", tmp, src])
        os.replace(tmp, _LIB_PATH)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        cp = ctypes.c_char_p
        L.orc_create.restype = P
        L.orc_destroy.argtypes = [P]
        L.orc_last_error.restype = cp
        L.orc_last_error.argtypes = [P]
        L.orc_init_mesh.argtypes = [P, ctypes.c_int, ctypes.c_int, dp, ctypes.c_int, ip, ctypes.c_double]
        L.orc_add_pd_energy.argtypes = [P, cp, dp, ctypes.c_int, ip, ctypes.c_int]
        L.orc_set_dirichlet.argtypes = [P, ctypes.c_int, ctypes.c_double]
        L.orc_add_gravity.argtypes = [P, dp]
        L.orc_set_frictional_boundary.argtypes = [P, cp, dp, ctypes.c_int, ip, ctypes.c_int]
        L.orc_set_press_plane.argtypes = [P, ctypes.c_double]
        L.orc_add_planar_contact.argtypes = [P, dp, ctypes.c_int]
        L.orc_add_arc_contact.argtypes = [P, dp, ctypes.c_int]
        L.orc_add_mesh_contact.argtypes = [P, dp, ctypes.c_int]
        L.orc_add_billiard_ball.argtypes = [P, dp, ctypes.c_int]
        L.orc_add_hydrodynamics.argtypes = [P, dp, ctypes.c_int]
        L.orc_num_state_params.argtypes = [P]
        L.orc_state_force.argtypes = [P, dp, dp, dp]
        L.orc_state_force_backward.argtypes = [P, dp, dp, dp, dp, dp, dp]
        L.orc_add_actuation.argtypes = [P, ctypes.c_double, dp, ip, ctypes.c_int]
        L.orc_set_actuation.argtypes = [P, dp, ctypes.c_int]
        L.orc_act_dofs.argtypes = [P]
        L.orc_num_muscles.argtypes = [P]
        L.orc_actuation_grads.argtypes = [P, dp, ctypes.c_int, dp, ctypes.c_int]
        L.orc_backward_baseline.argtypes = [P, dp, dp, dp, ctypes.c_double, dp, ip, ctypes.c_int, dp, dp,
                                            ctypes.POINTER(cp), dp, ctypes.c_int, dp, dp, dp, dp, dp, ip]
        L.orc_forward_baseline.argtypes = [P, cp, dp, dp, dp, ctypes.c_double, ctypes.POINTER(cp), dp, ctypes.c_int,
                                           dp, dp, ip, ctypes.c_int, ip, ip, ctypes.c_int, ip]
        for f in ("orc_dofs", "orc_num_energies", "orc_num_elements", "orc_last_backward_iters"):
            getattr(L, f).argtypes = [P]
        L.orc_forward.argtypes = [P, dp, dp, dp, ctypes.c_double, ctypes.POINTER(cp), dp, ctypes.c_int,
                                  dp, dp, ip, ip, ctypes.c_int]
        L.orc_backward.argtypes = [P, dp, dp, dp, ctypes.c_double, dp, dp, ip, ctypes.c_int, dp, dp,
                                   ctypes.POINTER(cp), dp, ctypes.c_int, dp, dp, dp, dp, dp, dp]
        L.orc_apply_ainv.argtypes = [P, ctypes.c_double, dp, dp]
        L.orc_local_step.argtypes = [P, dp, dp]
        L.orc_local_differential.argtypes = [P, dp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, dp]
        L.orc_last_backward_stats.argtypes = [P, dp, ctypes.c_int]
        L.orc_filter_prox_jacobian.argtypes = [dp, ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.orc_filter_prox_jacobian.restype = ctypes.c_double
        L.orc_jacobi_eigen.argtypes = [ctypes.c_int, dp, dp, dp]
        L.orc_coulomb_step.argtypes = [dp, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, dp]
        L.orc_tri_closest.argtypes = [dp, dp, dp, dp, dp, dp]
        L.orc_boundary_query.argtypes = [P, dp, dp]
        L.orc_frame_stats.argtypes = [P, ip, ip, ctypes.c_int]
        L.orc_reset_frame_stats.argtypes = [P]
        L.orc_set_refinement.argtypes = [P, ctypes.c_int]
        L.orc_last_contact_multipliers.argtypes = [P, dp, dp, dp, ctypes.c_int]
        L.orc_svd3.argtypes = [dp, dp, dp, dp]
        L.orc_project.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, dp, dp]
        L.orc_project_diff.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, dp, dp, dp]
        L.orc_nh_prox.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, dp, dp]
        _lib = L
    return _lib


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _opts(options):
    keys = [k for k in options]
    karr = (ctypes.c_char_p * max(1, len(keys)))(*[k.encode() for k in keys])
    vals, vp = _d([float(options[k]) for k in keys] or [0.0])
    return karr, vals, vp, len(keys)


def read_bin_mesh(path):
    """Reads the reference `.bin` mesh (cpp/core/src/mesh/mesh.cpp:4-32):
    int vdim, int edim, then vertices (int rows, int cols, rows*cols f64 in
    column-major order) and elements (int rows, int cols, i32 column-major)."""
    with open(path, "rb") as f:
        vdim, edim = struct.unpack("ii", f.read(8))
        r, c = struct.unpack("ii", f.read(8))
        V = np.frombuffer(f.read(8 * r * c), dtype=np.float64).reshape(c, r).T
        r2, c2 = struct.unpack("ii", f.read(8))
        E = np.frombuffer(f.read(4 * r2 * c2), dtype=np.int32).reshape(c2, r2).T
    return vdim, edim, np.array(V), np.array(E)


class OracleDeformable:
    """Oracle body with the reference Deformable's method names."""

    def __init__(self, edim=8):
        self._edim = edim
        self._h = lib().orc_create()
        self._dofs = 0

    def __del__(self):
        try:
            lib().orc_destroy(self._h)
        except Exception:
            pass

    def _ok(self, rc):
        if rc != 0:
            raise RuntimeError(lib().orc_last_error(self._h).decode())

    def Initialize(self, bin_file, density, material_type, youngs_modulus, poissons_ratio):
        _, edim, V, E = read_bin_mesh(bin_file)
        self.InitializeFromArrays(V, E, density)

    def InitializeFromArrays(self, vertices, elements, density):
        V = np.asarray(vertices, dtype=np.float64)
        E = np.asarray(elements, dtype=np.int32)
        assert V.shape[0] == 3 and E.shape[0] == self._edim
        Vf, vp = _d(V.T.ravel())
        Ef, ep = _i(E.T.ravel())
        self._ok(lib().orc_init_mesh(self._h, self._edim, V.shape[1], vp, E.shape[1], ep, float(density)))
        self._dofs = 3 * V.shape[1]
        self._ne = E.shape[1]

    def AddPdEnergy(self, energy_type, params, indices):
        p, pp = _d(list(params) or [0.0])
        ix, ip = _i(list(indices) or [0])
        self._ok(lib().orc_add_pd_energy(self._h, energy_type.encode(), pp, len(params), ip, len(indices)))

    def SetDirichletBoundaryCondition(self, dof, val):
        self._ok(lib().orc_set_dirichlet(self._h, int(dof), float(val)))

    def AddStateForce(self, force_type, params):
        p, pp = _d(params)
        if force_type == "gravity":
            self._ok(lib().orc_add_gravity(self._h, pp))
        elif force_type == "planar_contact":
            self._ok(lib().orc_add_planar_contact(self._h, pp, len(params)))
        elif force_type == "arc_contact":
            self._ok(lib().orc_add_arc_contact(self._h, pp, len(params)))
        elif force_type == "mesh_contact":
            self._ok(lib().orc_add_mesh_contact(self._h, pp, len(params)))
        elif force_type == "billiard_ball":
            self._ok(lib().orc_add_billiard_ball(self._h, pp, len(params)))
        elif force_type == "hydrodynamics":
            self._ok(lib().orc_add_hydrodynamics(self._h, pp, len(params)))
        else:
            raise RuntimeError("Unsupported state force: " + str(force_type))

    def SetFrictionalBoundary(self, boundary_type, params, indices):
        p, pp = _d(list(params))
        ix, ip = _i(list(indices) or [0])
        self._ok(lib().orc_set_frictional_boundary(self._h, boundary_type.encode(), pp, len(params), ip, len(indices)))

    def SetPressPlaneHeight(self, z):
        self._ok(lib().orc_set_press_plane(self._h, float(z)))

    def AddActuation(self, stiffness, fiber_direction, indices):
        f, fp = _d(list(fiber_direction))
        ix, ip = _i(list(indices) or [0])
        self._ok(lib().orc_add_actuation(self._h, float(stiffness), fp, ip, len(indices)))

    def _set_actuation(self, a):
        a = np.asarray(list(a), dtype=np.float64)
        buf, bp = _d(a if a.size else np.zeros(1))
        self._ok(lib().orc_set_actuation(self._h, bp, int(a.size)))

    def dofs(self):
        return self._dofs

    def act_dofs(self):
        return lib().orc_act_dofs(self._h)

    def NumOfPdElementEnergies(self):
        return lib().orc_num_energies(self._h)

    def NumOfPdMuscleEnergies(self):
        return lib().orc_num_muscles(self._h)

    def NumOfStateForceParameters(self):
        return lib().orc_num_state_params(self._h)

    def state_force(self, q, v):
        q, qp = _d(q)
        v, vp = _d(v)
        f, fp = _d(np.zeros_like(q))
        self._ok(lib().orc_state_force(self._h, qp, vp, fp))
        return f

    def state_force_backward(self, q, v, dl_df):
        q, qp = _d(q)
        v, vp = _d(v)
        g, gp = _d(dl_df)
        a, ap = _d(np.zeros_like(q))
        b, bp = _d(np.zeros_like(q))
        c, cp = _d(np.zeros(max(1, self.NumOfStateForceParameters())))
        self._ok(lib().orc_state_force_backward(self._h, qp, vp, gp, ap, bp, cp))
        return a, b, c[: self.NumOfStateForceParameters()]

    def PyForward(self, method, q, v, a, f_ext, dt, options, q_next, v_next, active_contact_idx):
        if method not in ("pd_eigen_alg_phd", "pd_eigen_pcg_proj_gpu"):
            if not method.startswith("pd_eigen"):
                raise RuntimeError("Unsupported forward method: " + str(method))
            return self._forward_baseline(method, q, v, a, f_ext, dt, options, q_next, v_next, active_contact_idx)
        n = self._dofs
        qa, qp = _d(q)
        va, vp = _d(v)
        fa, fp = _d(f_ext)
        qn, qnp = _d(np.zeros(n))
        vn, vnp = _d(np.zeros(n))
        act, ap = _i(np.zeros(n // 3 + 1))
        na = ctypes.c_int(0)
        karr, vals, valp, nk = _opts(options)
        self._set_actuation(a)
        self._ok(lib().orc_forward(self._h, qp, vp, fp, float(dt), karr, valp, nk, qnp, vnp, ap, ctypes.byref(na),
                                   n // 3 + 1))
        q_next[:] = list(qn)
        v_next[:] = list(vn)
        active_contact_idx[:] = [int(x) for x in act[: na.value]]

    def _forward_baseline(self, method, q, v, a, f_ext, dt, options, q_next, v_next, active_contact_idx):
        """ForwardProjectiveDynamics (forward.cpp:529-664): active_contact_idx
        holds the previous frame's active set on entry, this frame's on exit."""
        n = self._dofs
        qa, qp = _d(q)
        va, vp = _d(v)
        fa, fp = _d(f_ext)
        qn, qnp = _d(np.zeros(n))
        vn, vnp = _d(np.zeros(n))
        ain, aip = _i(list(active_contact_idx) or [0])
        cap = n // 3 + 1
        aout, aop = _i(np.zeros(cap))
        na = ctypes.c_int(0)
        st, stp = _i(np.zeros(2))
        karr, vals, valp, nk = _opts(options)
        self._set_actuation(a)
        self._ok(lib().orc_forward_baseline(self._h, method.encode(), qp, vp, fp, float(dt), karr, valp, nk, qnp, vnp,
                                            aip, len(active_contact_idx), aop, ctypes.byref(na), cap, stp))
        q_next[:] = list(qn)
        v_next[:] = list(vn)
        active_contact_idx[:] = [int(x) for x in aout[: na.value]]
        self.last_baseline_stats = {"contact_iters": int(st[0]), "pd_iters": int(st[1])}

    def PyBackward(self, method, q, v, a, f_ext, dt, q_next, v_next, active_contact_idx, dl_dq_next, dl_dv_next,
                   options, dl_dq, dl_dv, dl_da, dl_df_ext, dl_dmat_w, dl_dact_w, dl_dstate_p, dl_dw_elem=None):
        if method not in ("pd_eigen_alg_phd", "pd_eigen_pcg_proj_gpu"):
            if not method.startswith("pd_eigen"):
                raise RuntimeError("Unsupported backward method: " + str(method))
            return self._backward_baseline(q, v, a, f_ext, dt, q_next, active_contact_idx, dl_dq_next, dl_dv_next,
                                           options, dl_dq, dl_dv, dl_da, dl_df_ext, dl_dmat_w, dl_dact_w, dl_dstate_p)
        n = self._dofs
        ne = lib().orc_num_energies(self._h)
        nel = lib().orc_num_elements(self._h)
        arrs = [_d(x) for x in (q, v, f_ext, q_next, v_next, dl_dq_next, dl_dv_next)]
        act, ap = _i(list(active_contact_idx) or [0])
        outs = [_d(np.zeros(k)) for k in (n, n, n, max(ne, 1), max(1, self.NumOfStateForceParameters()),
                                          max(ne * nel, 1))]
        karr, vals, valp, nk = _opts(options)
        self._set_actuation(a)
        self._ok(lib().orc_backward(self._h, arrs[0][1], arrs[1][1], arrs[2][1], float(dt), arrs[3][1], arrs[4][1],
                                    ap, len(active_contact_idx), arrs[5][1], arrs[6][1], karr, valp, nk,
                                    *[o[1] for o in outs]))
        na, nm = self.act_dofs(), self.NumOfPdMuscleEnergies()
        ga, gap = _d(np.zeros(max(na, 1)))
        gw, gwp = _d(np.zeros(max(nm, 1)))
        self._ok(lib().orc_actuation_grads(self._h, gap, na, gwp, nm))
        dl_dq[:] = list(outs[0][0])
        dl_dv[:] = list(outs[1][0])
        dl_da[:] = list(ga[:na])
        dl_df_ext[:] = list(outs[2][0])
        dl_dmat_w[:] = list(outs[3][0][:ne])
        dl_dact_w[:] = list(gw[:nm])
        dl_dstate_p[:] = list(outs[4][0][: self.NumOfStateForceParameters()])
        if dl_dw_elem is not None:
            dl_dw_elem[:] = list(outs[5][0][: ne * nel])

    def _backward_baseline(self, q, v, a, f_ext, dt, q_next, active_contact_idx, dl_dq_next, dl_dv_next, options,
                           dl_dq, dl_dv, dl_da, dl_df_ext, dl_dmat_w, dl_dact_w, dl_dstate_p):
        """BackwardProjectiveDynamics for a pd_eigen baseline method
        (deformable_projective_dynamics_backward.cpp:141-679)."""
        n = self._dofs
        ne = lib().orc_num_energies(self._h)
        arrs = [_d(x) for x in (q, v, f_ext, q_next, dl_dq_next, dl_dv_next)]
        act, ap = _i(list(active_contact_idx) or [0])
        outs = [_d(np.zeros(k)) for k in (n, n, n, max(ne, 1), max(1, self.NumOfStateForceParameters()))]
        st, stp = _i(np.zeros(2))
        karr, vals, valp, nk = _opts(options)
        self._set_actuation(a)
        self._ok(lib().orc_backward_baseline(self._h, arrs[0][1], arrs[1][1], arrs[2][1], float(dt), arrs[3][1], ap,
                                             len(active_contact_idx), arrs[4][1], arrs[5][1], karr, valp, nk,
                                             *[o[1] for o in outs], stp))
        na, nm = self.act_dofs(), self.NumOfPdMuscleEnergies()
        ga, gap = _d(np.zeros(max(na, 1)))
        gw, gwp = _d(np.zeros(max(nm, 1)))
        self._ok(lib().orc_actuation_grads(self._h, gap, na, gwp, nm))
        dl_dq[:] = list(outs[0][0])
        dl_dv[:] = list(outs[1][0])
        dl_da[:] = list(ga[:na])
        dl_df_ext[:] = list(outs[2][0])
        dl_dmat_w[:] = list(outs[3][0][:ne])
        dl_dact_w[:] = list(gw[:nm])
        dl_dstate_p[:] = list(outs[4][0][: self.NumOfStateForceParameters()])
        self.last_baseline_bwd_stats = {"iters": int(st[0]), "newton_fallback": int(st[1])}

    # --- test-only helpers -------------------------------------------------
    def apply_ainv(self, dt, b):
        b, bp = _d(b)
        x, xp = _d(np.zeros_like(b))
        self._ok(lib().orc_apply_ainv(self._h, float(dt), bp, xp))
        return x

    def local_step(self, q, a=None):
        if a is not None:
            self._set_actuation(a)
        q, qp = _d(q)
        out, op = _d(np.zeros_like(q))
        self._ok(lib().orc_local_step(self._h, qp, op))
        return out

    def local_differential(self, q, exact=False, kpd_filter=0, trust_radius=0.9, trust_stage=1, a=None):
        if a is not None:
            self._set_actuation(a)
        q, qp = _d(q)
        nd = 3 * self._edim
        out, op = _d(np.zeros(self._ne * nd * nd))
        self._ok(lib().orc_local_differential(self._h, qp, int(exact), int(kpd_filter), float(trust_radius),
                                              int(trust_stage), op))
        return out.reshape(self._ne, nd, nd)

    def frame_stats(self):
        cap = 100000
        it, ip = _i(np.zeros(cap))
        cv, cpp = _i(np.zeros(cap))
        n = lib().orc_frame_stats(self._h, ip, cpp, cap)
        return list(it[:n]), list(cv[:n])

    def reset_frame_stats(self):
        lib().orc_reset_frame_stats(self._h)

    def set_refinement(self, steps):
        """Iterative refinement steps (extended-precision residual) after each LDL^T solve."""
        lib().orc_set_refinement(self._h, int(steps))

    def last_contact_multipliers(self):
        cap = self._dofs // 3 + 1
        a, ap = _d(np.zeros(cap))
        b, bp = _d(np.zeros(cap))
        c, cp = _d(np.zeros(cap))
        n = lib().orc_last_contact_multipliers(self._h, ap, bp, cp, cap)
        return a[:n].copy(), b[:n].copy(), c[:n].copy()

    def last_backward_iters(self):
        return lib().orc_last_backward_iters(self._h)

    def boundary_query(self, x):
        """Frictional boundary at the point x: (distance, normal, t1, t2, press side)."""
        x, xp = _d(np.asarray(x, dtype=np.float64).copy())
        o, op = _d(np.zeros(11))
        self._ok(lib().orc_boundary_query(self._h, xp, op))
        return o[0], o[1:4].copy(), o[4:7].copy(), o[7:10].copy(), int(o[10])

    def last_backward_stats(self):
        out, op = _d(np.zeros(10))
        lib().orc_last_backward_stats(self._h, op, 10)
        keys = ("backbone_iters", "backbone_converged", "clipped", "woodbury_iters", "basis_rows", "filtered",
                "samples", "db_norm", "trust_stage", "from_tape")
        return dict(zip(keys, out.tolist()))


def filter_prox_jacobian(M, aw, theta, lo):
    """Spectral filter of one 9x9 proximal-map block (numpy matrix, row i col j):
    eigenvalues mu of sym(M)/aw above 1 move to 1 + theta (mu - 1), then up to
    lo. Returns (filtered copy, largest eigenvalue shift)."""
    m, mp = _d(np.asarray(M, dtype=np.float64).T.ravel().copy())
    shift = lib().orc_filter_prox_jacobian(mp, float(aw), float(theta), float(lo))
    return m.reshape(9, 9).T.copy(), float(shift)


def coulomb_step(Wtt, u_t, fmu, lam):
    """Per-contact Coulomb step: (lt1, lt2) from the 2x2 tangential Delassus
    block, the tangential displacement (u1, u2), mu and lambda."""
    w, wp = _d(np.asarray(Wtt, dtype=np.float64).ravel().copy())
    o, op = _d(np.zeros(2))
    lib().orc_coulomb_step(wp, float(u_t[0]), float(u_t[1]), float(fmu), float(lam), op)
    return o.copy()


def tri_closest(p, a, b, c, weights=False):
    """Closest point of triangle (a, b, c) to p (the mesh boundary's and the mesh
    contact force's query), and with weights=True also its barycentric weights."""
    arr = [_d(np.asarray(x, dtype=np.float64).copy()) for x in (p, a, b, c)]
    o, op = _d(np.zeros(3))
    w, wp = _d(np.zeros(3))
    lib().orc_tri_closest(arr[0][1], arr[1][1], arr[2][1], arr[3][1], op, wp)
    return (o.copy(), w.copy()) if weights else o.copy()


def jacobi_eigen(A):
    """Eigenvalues and eigenvectors (columns) of a small symmetric matrix."""
    A = np.asarray(A, dtype=np.float64)
    n = A.shape[0]
    a, ap = _d(A.ravel().copy())
    w, wp = _d(np.zeros(n))
    z, zp = _d(np.zeros(n * n))
    lib().orc_jacobi_eigen(n, ap, wp, zp)
    return w.copy(), z.reshape(n, n).copy()


def svd3(F):
    F = np.asarray(F, dtype=np.float64)
    f, fp = _d(F.T.ravel())
    U, up = _d(np.zeros(9))
    s, sp = _d(np.zeros(3))
    V, vp = _d(np.zeros(9))
    lib().orc_svd3(fp, up, sp, vp)
    return U.reshape(3, 3).T.copy(), s.copy(), V.reshape(3, 3).T.copy()


def project(kind, F, k=0.0, mu=0.0, la=0.0):
    t = {"corotated": 0, "volume": 1, "pd_neohookean": 2}[kind]
    f, fp = _d(np.asarray(F, dtype=np.float64).T.ravel())
    o, op = _d(np.zeros(9))
    lib().orc_project(t, k, mu, la, fp, op)
    return o.reshape(3, 3).T.copy()


def project_diff(kind, F, dF, k=0.0, mu=0.0, la=0.0):
    t = {"corotated": 0, "volume": 1, "pd_neohookean": 2}[kind]
    f, fp = _d(np.asarray(F, dtype=np.float64).T.ravel())
    g, gp = _d(np.asarray(dF, dtype=np.float64).T.ravel())
    o, op = _d(np.zeros(9))
    lib().orc_project_diff(t, k, mu, la, fp, gp, op)
    return o.reshape(3, 3).T.copy()


def nh_prox(sig_F, k, mu, la):
    s, sp = _d(sig_F)
    o, op = _d(np.zeros(3))
    lib().orc_nh_prox(k, mu, la, sp, op)
    return o.copy()
