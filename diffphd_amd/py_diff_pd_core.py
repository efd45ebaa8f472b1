"""Python module with the names of the reference's SWIG module `py_diff_pd_core`
(cpp/core/src/py_diff_pd_core.i:1-83) for the `pd_eigen_alg_phd` path.

`HexDeformable` / `TetDeformable` keep the reference's method names and
argument meaning (cpp/core/include/fem/deformable.h:23-141): outputs are
caller-allocated `StdRealVector`s that the call resizes and fills
(deformable.cpp:157-187), errors are raised as RuntimeError. The work is done
by libphd.so through the C ABI of include/phd.h.
"""
import ctypes

import numpy as np

from . import _lib
from .hex_mesh import write_bin

METHODS = ("pd_eigen_alg_phd", "pd_eigen_pcg_proj_gpu")


class StdRealVector(list):
    """std::vector<real> stand-in: a list the callee fills in place. Built from
    a size (zeros) or from a sequence, like the SWIG vector constructors."""

    def __init__(self, n=0):
        if np.ndim(n) == 0:
            super().__init__([0.0] * int(n))
        else:
            super().__init__(float(x) for x in np.asarray(n, dtype=np.float64).ravel())


class StdIntVector(list):
    def __init__(self, n=0):
        if np.ndim(n) == 0:
            super().__init__([0] * int(n))
        else:
            super().__init__(int(x) for x in np.asarray(n).ravel())


StdMap = dict


def _read_bin(path):
    import struct
    with open(path, "rb") as f:
        vdim, edim = struct.unpack("ii", f.read(8))
        r, c = struct.unpack("ii", f.read(8))
        V = np.frombuffer(f.read(8 * r * c), dtype=np.float64).reshape(c, r).T
        r2, c2 = struct.unpack("ii", f.read(8))
        E = np.frombuffer(f.read(4 * r2 * c2), dtype=np.int32).reshape(c2, r2).T
    return vdim, edim, np.array(V), np.array(E)


class _Mesh:
    """Mesh<3, edim> (cpp/core/src/mesh/mesh.cpp)."""
    _edim = 8

    def __init__(self):
        self._V = np.zeros((3, 0))
        self._E = np.zeros((self._edim, 0), dtype=np.int32)

    def Initialize(self, bin_file_name):
        vdim, edim, V, E = _read_bin(bin_file_name)
        if vdim != 3 or edim != self._edim:
            raise RuntimeError("Corrupted mesh file: " + str(bin_file_name))
        self._V, self._E = V, E

    def InitializeFromArrays(self, V, E):
        self._V, self._E = np.asarray(V, dtype=np.float64), np.asarray(E, dtype=np.int32)

    def NumOfVertices(self):
        return self._V.shape[1]

    def NumOfElements(self):
        return self._E.shape[1]

    def py_vertex(self, i):
        return list(self._V[:, i])

    def py_vertices(self):
        return list(self._V.T.ravel())

    def py_element(self, i):
        return list(int(x) for x in self._E[:, i])

    def SaveToFile(self, path):
        write_bin(path, self._V, self._E)


class HexMesh3d(_Mesh):
    _edim = 8


class TetMesh3d(_Mesh):
    _edim = 4


_last_ctx = [None]


class _Deformable:
    _edim = 8

    def __init__(self, device=0):
        L = _lib.lib()
        self._ctx = L.phd_ctx_create(int(device))
        if not self._ctx:
            raise RuntimeError("phd_ctx_create failed: " + L.phd_last_error(None).decode())
        self._dofs = 0
        self._ne = 0
        _last_ctx[0] = self

    def __del__(self):
        try:
            if self._ctx:
                _lib.lib().phd_ctx_destroy(self._ctx)
                self._ctx = None
        except Exception:
            pass

    def _ok(self, rc):
        if rc != 0:
            raise RuntimeError(_lib.lib().phd_last_error(self._ctx).decode())

    # --- Deformable::Initialize (deformable.cpp:17-43) -----------------------
    def Initialize(self, binary_file_name, density, material_type, youngs_modulus, poissons_ratio):
        _, edim, V, E = _read_bin(binary_file_name)
        if edim != self._edim:
            raise RuntimeError("Corrupted mesh file: " + str(binary_file_name))
        self._init_arrays(V, E, density, material_type)

    def InitializeFromArrays(self, vertices, elements, density, material_type="none", youngs_modulus=0.0,
                             poissons_ratio=0.0):
        self._init_arrays(np.asarray(vertices), np.asarray(elements), density, material_type)

    def _init_arrays(self, V, E, density, material_type):
        if material_type != "none":
            raise RuntimeError("Only material type 'none' is supported on the pd_eigen_alg_phd path.")
        V = np.asarray(V, dtype=np.float64)
        E = np.asarray(E, dtype=np.int32)
        Vf, vp = _lib.dptr(V.T.ravel())
        Ef, ep = _lib.iptr(E.T.ravel())
        self._ok(_lib.lib().phd_init_mesh(self._ctx, self._edim, V.shape[1], vp, E.shape[1], ep, float(density)))
        self._dofs = 3 * V.shape[1]
        self._ne = E.shape[1]
        self._V, self._E = V, E

    def AddPdEnergy(self, energy_type, params, indices):
        p, pp = _lib.dptr(list(params) or [0.0])
        ix, ip = _lib.iptr(list(indices) or [0])
        self._ok(_lib.lib().phd_add_pd_energy(self._ctx, energy_type.encode(), pp, len(params), ip, len(indices)))

    def SetDirichletBoundaryCondition(self, dof, val):
        self._ok(_lib.lib().phd_set_dirichlet(self._ctx, int(dof), float(val)))

    def AddStateForce(self, force_type, params):
        p, pp = _lib.dptr(list(params))
        self._ok(_lib.lib().phd_add_state_force(self._ctx, force_type.encode(), pp, len(params)))

    def SetFrictionalBoundary(self, boundary_type, params, indices):
        p, pp = _lib.dptr(list(params))
        ix, ip = _lib.iptr(list(indices) or [0])
        self._ok(_lib.lib().phd_set_frictional_boundary(self._ctx, boundary_type.encode(), pp, len(params), ip,
                                                        len(indices)))

    def SetPressPlaneHeight(self, z):
        self._ok(_lib.lib().phd_set_press_plane(self._ctx, float(z)))

    def dofs(self):
        return self._dofs

    def AddActuation(self, stiffness, fiber_direction, indices):
        f, fp = _lib.dptr(list(fiber_direction))
        ix, ip = _lib.iptr(list(indices) or [0])
        self._ok(_lib.lib().phd_add_actuation(self._ctx, float(stiffness), fp, ip, len(indices)))

    def _set_actuation(self, a):
        aa = np.asarray(list(a), dtype=np.float64)
        buf, bp = _lib.dptr(aa if aa.size else np.zeros(1))
        self._ok(_lib.lib().phd_set_actuation(self._ctx, bp, int(aa.size)))

    def act_dofs(self):
        return _lib.lib().phd_act_dofs(self._ctx)

    def NumOfPdElementEnergies(self):
        return _lib.lib().phd_num_pd_element_energies(self._ctx)

    def NumOfPdMuscleEnergies(self):
        return _lib.lib().phd_num_pd_muscle_energies(self._ctx)

    def NumOfStateForceParameters(self):
        return _lib.lib().phd_num_state_force_parameters(self._ctx)

    def mesh(self):
        m = HexMesh3d() if self._edim == 8 else TetMesh3d()
        m.InitializeFromArrays(self._V, self._E)
        return m

    # --- PyForward / PyBackward (deformable.cpp:157-187) ---------------------
    def PyForward(self, method, q, v, a, f_ext, dt, options, q_next, v_next, active_contact_idx):
        if method not in METHODS:
            if str(method).startswith("pd_eigen"):
                return self._forward_baseline(method, q, v, a, f_ext, dt, options, q_next, v_next, active_contact_idx)
            raise RuntimeError("Unsupported forward method: " + str(method))
        n = self._dofs
        qa, qp = _lib.dptr(q)
        va, vp = _lib.dptr(v)
        fa, fp = _lib.dptr(f_ext)
        if qa.size != n or va.size != n or fa.size != n:
            raise RuntimeError("Inconsistent state size.")
        qn, qnp = _lib.dptr(np.zeros(n))
        vn, vnp = _lib.dptr(np.zeros(n))
        cap = n // 3 + 1
        act, ap = _lib.iptr(np.zeros(cap))
        import ctypes
        na = ctypes.c_int(cap)
        karr, vals, valp, nk = _lib.options(options)
        self._set_actuation(a)
        self._ok(_lib.lib().phd_step_forward(self._ctx, qp, vp, fp, float(dt), karr, valp, nk, qnp, vnp, ap,
                                             ctypes.byref(na)))
        q_next[:] = qn.tolist()
        v_next[:] = vn.tolist()
        active_contact_idx[:] = [int(x) for x in act[: na.value]]
        _last_ctx[0] = self

    def _forward_baseline(self, method, q, v, a, f_ext, dt, options, q_next, v_next, active_contact_idx):
        """The DiffPD baseline forward (ForwardProjectiveDynamics with a pd_eigen
        method, deformable_projective_dynamics_forward.cpp:529-664):
        active_contact_idx holds the previous frame's active set on entry, this
        frame's on exit."""
        n = self._dofs
        qa, qp = _lib.dptr(q)
        va, vp = _lib.dptr(v)
        fa, fp = _lib.dptr(f_ext)
        if qa.size != n or va.size != n or fa.size != n:
            raise RuntimeError("Inconsistent state size.")
        qn, qnp = _lib.dptr(np.zeros(n))
        vn, vnp = _lib.dptr(np.zeros(n))
        ain, aip = _lib.iptr(list(active_contact_idx) or [0])
        cap = n // 3 + 1
        aout, aop = _lib.iptr(np.zeros(cap))
        na = ctypes.c_int(cap)
        st, stp = _lib.iptr(np.zeros(3))
        karr, vals, valp, nk = _lib.options(options)
        self._set_actuation(a)
        self._ok(_lib.lib().phd_step_forward_baseline(self._ctx, str(method).encode(), qp, vp, fp, float(dt), karr, valp,
                                                      nk, qnp, vnp, aip, len(active_contact_idx), aop, ctypes.byref(na),
                                                      stp))
        q_next[:] = qn.tolist()
        v_next[:] = vn.tolist()
        active_contact_idx[:] = [int(x) for x in aout[: na.value]]
        self.last_baseline_stats = {"contact_iters": int(st[0]), "pd_iters": int(st[1]), "ls_steps": int(st[2])}
        _last_ctx[0] = self

    def PyBackward(self, method, q, v, a, f_ext, dt, q_next, v_next, active_contact_idx, dl_dq_next, dl_dv_next,
                   options, dl_dq, dl_dv, dl_da, dl_df_ext, dl_dmat_w, dl_dact_w, dl_dstate_p, dl_dw_elem=None):
        if method not in METHODS:
            if str(method).startswith("pd_eigen"):
                return self._backward_baseline(q, v, a, f_ext, dt, q_next, active_contact_idx, dl_dq_next, dl_dv_next,
                                               options, dl_dq, dl_dv, dl_da, dl_df_ext, dl_dmat_w, dl_dact_w,
                                               dl_dstate_p)
            raise RuntimeError("Unsupported backward method: " + str(method))
        n = self._dofs
        ins = [_lib.dptr(x) for x in (q, v, f_ext, q_next, v_next, dl_dq_next, dl_dv_next)]
        act, ap = _lib.iptr(list(active_contact_idx) or [0])
        nE = self.NumOfPdElementEnergies()
        outs = [_lib.dptr(np.zeros(k)) for k in (n, n, n, max(nE, 1), max(1, self.NumOfStateForceParameters()),
                                                 max(nE * self._ne, 1))]
        karr, vals, valp, nk = _lib.options(options)
        self._set_actuation(a)
        self._ok(_lib.lib().phd_step_backward(self._ctx, ins[0][1], ins[1][1], ins[2][1], float(dt), ins[3][1],
                                              ins[4][1], ap, len(active_contact_idx), ins[5][1], ins[6][1], karr,
                                              valp, nk, *[o[1] for o in outs]))
        na, nm = self.act_dofs(), self.NumOfPdMuscleEnergies()
        ga, gap = _lib.dptr(np.zeros(max(na, 1)))
        gw, gwp = _lib.dptr(np.zeros(max(nm, 1)))
        self._ok(_lib.lib().phd_actuation_gradients(self._ctx, gap, na, gwp, nm))
        dl_dq[:] = outs[0][0].tolist()
        dl_dv[:] = outs[1][0].tolist()
        dl_da[:] = ga[:na].tolist()
        dl_df_ext[:] = outs[2][0].tolist()
        dl_dmat_w[:] = outs[3][0][:nE].tolist()
        dl_dact_w[:] = gw[:nm].tolist()
        dl_dstate_p[:] = outs[4][0][: self.NumOfStateForceParameters()].tolist()
        if dl_dw_elem is not None:
            dl_dw_elem[:] = outs[5][0][: nE * self._ne].tolist()
        _last_ctx[0] = self

    def _backward_baseline(self, q, v, a, f_ext, dt, q_next, active_contact_idx, dl_dq_next, dl_dv_next, options,
                           dl_dq, dl_dv, dl_da, dl_df_ext, dl_dmat_w, dl_dact_w, dl_dstate_p):
        """BackwardProjectiveDynamics for a pd_eigen baseline method
        (deformable_projective_dynamics_backward.cpp:141-679)."""
        n = self._dofs
        ins = [_lib.dptr(x) for x in (q, v, f_ext, q_next, dl_dq_next, dl_dv_next)]
        act, ap = _lib.iptr(list(active_contact_idx) or [0])
        nE = self.NumOfPdElementEnergies()
        outs = [_lib.dptr(np.zeros(k)) for k in (n, n, n, max(nE, 1), max(1, self.NumOfStateForceParameters()))]
        st, stp = _lib.iptr(np.zeros(1))
        karr, vals, valp, nk = _lib.options(options)
        self._set_actuation(a)
        self._ok(_lib.lib().phd_step_backward_baseline(self._ctx, ins[0][1], ins[1][1], ins[2][1], float(dt), ins[3][1],
                                                       ap, len(active_contact_idx), ins[4][1], ins[5][1], karr, valp,
                                                       nk, *[o[1] for o in outs], stp))
        na, nm = self.act_dofs(), self.NumOfPdMuscleEnergies()
        ga, gap = _lib.dptr(np.zeros(max(na, 1)))
        gw, gwp = _lib.dptr(np.zeros(max(nm, 1)))
        self._ok(_lib.lib().phd_actuation_gradients(self._ctx, gap, na, gwp, nm))
        dl_dq[:] = outs[0][0].tolist()
        dl_dv[:] = outs[1][0].tolist()
        dl_da[:] = ga[:na].tolist()
        dl_df_ext[:] = outs[2][0].tolist()
        dl_dmat_w[:] = outs[3][0][:nE].tolist()
        dl_dact_w[:] = gw[:nm].tolist()
        dl_dstate_p[:] = outs[4][0][: self.NumOfStateForceParameters()].tolist()
        self.last_baseline_bwd_stats = {"iters": int(st[0])}
        _last_ctx[0] = self

    # --- device tensors (DLPack handoff) ----------------------------------
    def _dev(self, t, name):
        import torch
        t = torch.from_dlpack(t) if not isinstance(t, torch.Tensor) else t
        if not t.is_cuda or t.dtype != torch.float64 or not t.is_contiguous() or t.numel() != self._dofs:
            raise RuntimeError(f"{name}: expected a contiguous float64 device tensor of {self._dofs} entries")
        return t

    def forward_device(self, method, q, v, a, f_ext, dt, options):
        """PyForward on device tensors (torch, or anything torch.from_dlpack
        takes): the tensors' memory goes to the library by pointer, q_next and
        v_next come back as new tensors on the same device. Returns (q_next,
        v_next, active_contact_idx)."""
        import torch
        if method not in METHODS:
            raise RuntimeError("Unsupported forward method: " + str(method))
        q, v, f_ext = self._dev(q, "q"), self._dev(v, "v"), self._dev(f_ext, "f_ext")
        qn, vn = torch.empty_like(q), torch.empty_like(q)
        cap = max(1, self._dofs // 3 + 1)
        act, ap = _lib.iptr(np.zeros(cap))
        na = ctypes.c_int(cap)
        karr, vals, valp, nk = _lib.options(options)
        self._set_actuation(a.detach().cpu().numpy() if hasattr(a, "detach") else a)
        torch.cuda.synchronize(q.device)
        self._ok(_lib.lib().phd_step_forward_device(self._ctx, q.data_ptr(), v.data_ptr(), f_ext.data_ptr(), float(dt),
                                                    karr, valp, nk, qn.data_ptr(), vn.data_ptr(), ap, ctypes.byref(na)))
        _last_ctx[0] = self
        return qn, vn, [int(x) for x in act[: na.value]]

    def backward_device(self, method, q, v, a, f_ext, dt, q_next, v_next, active_contact_idx, dl_dq_next, dl_dv_next,
                        options):
        """PyBackward on device tensors. Returns a dict of dl_dq, dl_dv,
        dl_df_ext (tensors on the inputs' device) and dl_da, dl_dmat_w,
        dl_dact_w, dl_dstate_p (numpy arrays)."""
        import torch
        if method not in METHODS:
            raise RuntimeError("Unsupported backward method: " + str(method))
        ins = [self._dev(x, name) for x, name in ((q, "q"), (v, "v"), (f_ext, "f_ext"), (q_next, "q_next"),
                                                  (v_next, "v_next"), (dl_dq_next, "dl_dq_next"),
                                                  (dl_dv_next, "dl_dv_next"))]
        outs = [torch.empty_like(ins[0]) for _ in range(3)]
        nE = self.NumOfPdElementEnergies()
        host = [_lib.dptr(np.zeros(k)) for k in (max(nE, 1), max(1, self.NumOfStateForceParameters()),
                                                 max(nE * self._ne, 1))]
        act, ap = _lib.iptr(list(active_contact_idx) or [0])
        karr, vals, valp, nk = _lib.options(options)
        self._set_actuation(a.detach().cpu().numpy() if hasattr(a, "detach") else a)
        torch.cuda.synchronize(ins[0].device)
        self._ok(_lib.lib().phd_step_backward_device(
            self._ctx, ins[0].data_ptr(), ins[1].data_ptr(), ins[2].data_ptr(), float(dt), ins[3].data_ptr(),
            ins[4].data_ptr(), ap, len(active_contact_idx), ins[5].data_ptr(), ins[6].data_ptr(), karr, valp, nk,
            outs[0].data_ptr(), outs[1].data_ptr(), outs[2].data_ptr(), *[h[1] for h in host]))
        na, nm = self.act_dofs(), self.NumOfPdMuscleEnergies()
        ga, gap = _lib.dptr(np.zeros(max(na, 1)))
        gw, gwp = _lib.dptr(np.zeros(max(nm, 1)))
        self._ok(_lib.lib().phd_actuation_gradients(self._ctx, gap, na, gwp, nm))
        _last_ctx[0] = self
        return {"dl_dq": outs[0], "dl_dv": outs[1], "dl_df_ext": outs[2], "dl_da": ga[:na].copy(),
                "dl_dmat_w": host[0][0][:nE].copy(), "dl_dact_w": gw[:nm].copy(),
                "dl_dstate_p": host[1][0][: self.NumOfStateForceParameters()].copy()}

    # --- extensions used by tests and the benchmark --------------------------
    def factor_build(self, dt, threads=8):
        self._ok(_lib.lib().phd_factor_build(self._ctx, float(dt), int(threads)))

    def apply_ainv(self, dt, b, precision=0, refine=0):
        ba, bp = _lib.dptr(b)
        x, xp = _lib.dptr(np.zeros_like(ba))
        self._ok(_lib.lib().phd_apply_ainv(self._ctx, float(dt), bp, xp, int(precision), int(refine)))
        return x

    def apply_ainv_noalloc(self, b_dev_ptr, x_dev_ptr, precision=0, refine=0):
        """x = A^-1 b on device vectors (raw device addresses of 3N doubles);
        allocates nothing on the device. The factor must already be built."""
        self._ok(_lib.lib().phd_apply_ainv_noalloc(self._ctx, ctypes.c_void_p(int(b_dev_ptr)),
                                                   ctypes.c_void_p(int(x_dev_ptr)), int(precision), int(refine)))

    def factor_info(self):
        out, op = _lib.dptr(np.zeros(15))
        self._ok(_lib.lib().phd_factor_info(self._ctx, op, 15))
        keys = ["nnz_S", "v_len", "max_depth", "order_s", "factor_s", "setup_s", "row_units", "col_units", "nnz_L", "K",
                "contact_entries", "contact_rows", "chunk_w", "unit_items", "local_block"]
        return dict(zip(keys, out.tolist()))

    def autotune(self, dt, threads=8, reps=10):
        """Times the launch-shape candidates on this body and keeps the fastest
        (phd_autotune); returns the choice and the per-candidate milliseconds."""
        out, op = _lib.dptr(np.zeros(13))
        self._ok(_lib.lib().phd_autotune(self._ctx, float(dt), int(threads), int(reps), op, 13))
        return {"chunk_w": int(out[0]), "unit_items": int(out[1]), "local_block": int(out[2]),
                "rows_ms": dict(zip((512, 1024, 2048), out[3:6].tolist())),
                "cols_ms": dict(zip((256, 512, 1024, 2048), out[6:10].tolist())),
                "local_ms": dict(zip((64, 128, 256), out[10:13].tolist()))}

    def set_launch_shapes(self, chunk_w, unit_items, local_block):
        self._ok(_lib.lib().phd_set_launch_shapes(self._ctx, int(chunk_w), int(unit_items), int(local_block)))

    def frame_stats(self):
        L = _lib.lib()
        n = L.phd_frame_stats(self._ctx, None, None, None, 0)
        it, ip = _lib.iptr(np.zeros(max(n, 1)))
        cv, cp = _lib.iptr(np.zeros(max(n, 1)))
        sc, sp = _lib.dptr(np.zeros(max(n, 1)))
        L.phd_frame_stats(self._ctx, ip, cp, sp, n)
        return it[:n].tolist(), cv[:n].tolist(), sc[:n].tolist()

    def reset_frame_stats(self):
        _lib.lib().phd_reset_frame_stats(self._ctx)

    def reset_contact_history(self):
        self._ok(_lib.lib().phd_reset_contact_history(self._ctx))

    def last_backward_iters(self):
        return _lib.lib().phd_last_backward_iters(self._ctx)

    def last_backward_stats(self):
        out, op = _lib.dptr(np.zeros(12))
        _lib.lib().phd_last_backward_stats(self._ctx, op, 12)
        keys = ("backbone_iters", "backbone_converged", "clipped", "woodbury_iters", "woodbury_converged",
                "basis_rows", "filtered", "samples", "db_norm", "trust_stage", "from_tape", "gram_fp64_fallbacks")
        return dict(zip(keys, out.tolist()))

    def last_contact_multipliers(self):
        cap = self._dofs // 3 + 1
        a, ap = _lib.dptr(np.zeros(cap))
        b, bp = _lib.dptr(np.zeros(cap))
        c, cp = _lib.dptr(np.zeros(cap))
        n = _lib.lib().phd_last_contact_multipliers(self._ctx, ap, bp, cp, cap)
        return a[:n].copy(), b[:n].copy(), c[:n].copy()

    # ---- learned warm start (phd_warm_start option) ----
    def mlp_init(self, seed=0):
        self._ok(_lib.lib().phd_mlp_init(self._ctx, int(seed)))

    def mlp_set_normalization(self, fmean, fstd, tscale):
        a, ap = _lib.fptr(fmean)
        b, bp = _lib.fptr(fstd)
        self._ok(_lib.lib().phd_mlp_set_normalization(self._ctx, ap, bp, float(tscale)))

    def mlp_params(self):
        out, op = _lib.fptr(np.zeros(_lib.lib().phd_mlp_num_params()))
        self._ok(_lib.lib().phd_mlp_get_params(self._ctx, op))
        return out.copy()

    def mlp_set_params(self, p):
        a, ap = _lib.fptr(p)
        self._ok(_lib.lib().phd_mlp_set_params(self._ctx, ap))

    def mlp_features(self, q, v, dt, f_ext=None):
        """Raw per-vertex features (n x 16) of state (q, v) for this body."""
        qa, qp = _lib.dptr(q)
        va, vp = _lib.dptr(v)
        fa, fp = _lib.dptr(np.zeros(self._dofs) if f_ext is None else f_ext)
        out, op = _lib.fptr(np.zeros((self._dofs // 3, 16)))
        self._ok(_lib.lib().phd_mlp_features(self._ctx, qp, vp, fp, float(dt), op))
        return out.copy()

    def mlp_forward(self, feat):
        f, fp = _lib.fptr(feat)
        n = f.shape[0]
        out, op = _lib.fptr(np.zeros((n, 3)))
        self._ok(_lib.lib().phd_mlp_forward(self._ctx, n, fp, op))
        return out.copy()

    def mlp_loss_grad(self, feat, target):
        f, fp = _lib.fptr(feat)
        t, tp = _lib.fptr(target)
        g, gp = _lib.fptr(np.zeros(_lib.lib().phd_mlp_num_params()))
        loss = ctypes.c_double(0.0)
        self._ok(_lib.lib().phd_mlp_loss_grad(self._ctx, f.shape[0], fp, tp, gp, ctypes.byref(loss)))
        return loss.value, g.copy()

    def mlp_train(self, feat, target, steps, batch, lr=1e-3, seed=0):
        f, fp = _lib.fptr(feat)
        t, tp = _lib.fptr(target)
        losses, lp = _lib.fptr(np.zeros(max(steps, 1)))
        self._ok(_lib.lib().phd_mlp_train(self._ctx, f.shape[0], fp, tp, int(steps), int(batch), float(lr), int(seed), lp))
        return losses[:steps].copy()

    def warm_start_stats(self, reset=True):
        u, f = ctypes.c_int(0), ctypes.c_int(0)
        self._ok(_lib.lib().phd_warm_start_stats(self._ctx, ctypes.byref(u), ctypes.byref(f), int(reset)))
        return u.value, f.value

    def copy_probe(self, nbytes, reps=10):
        out, op = _lib.dptr(np.zeros(1))
        self._ok(_lib.lib().phd_copy_probe(self._ctx, int(nbytes), int(reps), op))
        return float(out[0])

    def timing_enable(self, on=True):
        self._ok(_lib.lib().phd_timing_enable(self._ctx, int(on)))

    PHASES = ("fwd_total", "fwd_local", "fwd_solve", "fwd_contact", "fwd_dense", "fwd_aa", "fwd_gate",
              "bwd_total", "bwd_kpd", "bwd_backbone", "bwd_basis_factor", "bwd_woodbury", "bwd_basis_solve",
              "bwd_matgrad", "dense_calls", "dense_rows", "dense_fp32_fail", "basis_rows", "fwd_c_stage",
              "fwd_c_sets", "fwd_c_apply", "fwd_friction")

    def phase_times(self, reset=True):
        """Wall-clock split (s) recorded with timing_enable(2); the counters
        (dense_calls, dense_rows, dense_fp32_fail, basis_rows) are kept at
        every timing level."""
        out, op = _lib.dptr(np.zeros(len(self.PHASES)))
        n = _lib.lib().phd_phase_times(self._ctx, op, len(self.PHASES), int(reset))
        return dict(zip(self.PHASES[:n], out[:n].tolist()))

    def timing_read(self, reset=True):
        out, op = _lib.dptr(np.zeros(4))
        self._ok(_lib.lib().phd_timing_read(self._ctx, op, int(reset)))
        return {"rows_ms": out[0], "rows_launches": int(out[1]), "cols_ms": out[2], "cols_launches": int(out[3])}


class HexDeformable(_Deformable):
    """Deformable<3, 8>."""
    _edim = 8


class TetDeformable(_Deformable):
    """Deformable<3, 4>."""
    _edim = 4


# --- alg_phd_gpu_memory.h:9-27 (answers for the most recently used body) -------
def AlgPhdSMemoryBytes():
    d = _last_ctx[0]
    return int(_lib.lib().phd_s_memory_bytes(d._ctx)) if d is not None else 0


def AlgPhdSNnz():
    d = _last_ctx[0]
    return int(_lib.lib().phd_s_nnz(d._ctx)) if d is not None else 0


def AlgPhdCachedN():
    d = _last_ctx[0]
    return int(_lib.lib().phd_cached_n(d._ctx)) if d is not None else -1


def AlgPhdResetFrameStats():
    d = _last_ctx[0]
    if d is not None:
        d.reset_frame_stats()


def AlgPhdFrameIters():
    d = _last_ctx[0]
    return d.frame_stats()[0] if d is not None else []


def AlgPhdFrameConverged():
    d = _last_ctx[0]
    return d.frame_stats()[1] if d is not None else []
