"""Rollout harness: forward over all frames, then the reverse adjoint loop.

Restates EnvBase.simulate (python/py_diff_pd/env/env_base.py:163-427) for the
`pd_eigen_alg_phd` path: per-frame PyForward, loss at the last frame, reverse
PyBackward with the per-frame adjoint clip `_clip_adj` (default 1e3,
env_base.py:368-378) and `dl_dmat_w` summed over frames. Works with any body
exposing the reference Deformable's method names.
"""
import time

import numpy as np

from .py_diff_pd_core import StdRealVector


def _clip_adj(g, clip):
    if clip is None:
        return g
    g = np.nan_to_num(g, nan=0.0, posinf=clip, neginf=-clip)
    n = np.linalg.norm(g)
    if n > clip:
        g = g * (clip / n)
    return g


def simulate(body, dt, frames, method, opt, q0, v0, loss_and_grad, f_ext=None, require_grad=True,
             plane_schedule=None, grad_clip_norm=1e3, bwd_opt=None, a=None):
    """Returns (loss, grad, info); grad = (dl_dq0, dl_dv0, dl_df_ext per frame).

    loss_and_grad(q_T, v_T) -> (loss, dl_dq_T, dl_dv_T). plane_schedule: press
    plane height per frame (prescribed obstacle motion), or None. a: actuation
    levels, one vector for every frame or one row per frame (act_dofs entries);
    info then holds dl_da per frame and dl_dact_w summed over frames."""
    dofs = body.dofs()
    na = body.act_dofs()
    if a is None:
        acts_in = np.zeros((frames, na))
    else:
        a = np.asarray(a, dtype=np.float64)
        acts_in = np.tile(a, (frames, 1)) if a.ndim == 1 else a
        if acts_in.shape != (frames, na):
            raise RuntimeError("Inconsistent actuation size.")
    f_ext = np.zeros(dofs) if f_ext is None else np.asarray(f_ext, dtype=np.float64)
    bwd_opt = opt if bwd_opt is None else bwd_opt
    q = [np.asarray(q0, dtype=np.float64).copy()]
    v = [np.asarray(v0, dtype=np.float64).copy()]
    active = [[]]
    t_fwd = []
    for i in range(frames):
        t0 = time.time()
        if plane_schedule is not None:
            body.SetPressPlaneHeight(plane_schedule[i])
        # the previous frame's active set goes in (the baseline forward starts its
        # active-set loop from it, env_base.py:263-265); this frame's comes out
        qn, vn, act = StdRealVector(dofs), StdRealVector(dofs), list(active[-1])
        body.PyForward(method, q[-1], v[-1], list(acts_in[i]), f_ext, dt, opt, qn, vn, act)
        q.append(np.asarray(qn))
        v.append(np.asarray(vn))
        active.append(list(act))
        t_fwd.append(time.time() - t0)
    loss, dl_dq, dl_dv = loss_and_grad(q[-1], v[-1])
    info = {"q": q, "v": v, "active_contact_indices": active, "forward_time": float(np.sum(t_fwd)),
            "forward_time_per_frame": np.asarray(t_fwd)}
    if not require_grad:
        return loss, None, info
    t0 = time.time()
    nE = body.NumOfPdElementEnergies()
    dl_dmat_w = np.zeros(nE)
    dl_dw_elem = None
    dl_df = np.zeros((frames, dofs))
    dl_da = np.zeros((frames, na))
    dl_dact_w = np.zeros(body.NumOfPdMuscleEnergies())
    dl_dstate_p = np.zeros(body.NumOfStateForceParameters())
    dq_next, dv_next = np.asarray(dl_dq, dtype=np.float64).copy(), np.asarray(dl_dv, dtype=np.float64).copy()
    for i in reversed(range(frames)):
        if plane_schedule is not None:
            body.SetPressPlaneHeight(plane_schedule[i])
        o = [StdRealVector(0) for _ in range(8)]
        body.PyBackward(method, q[i], v[i], list(acts_in[i]), f_ext, dt, q[i + 1], v[i + 1], active[i + 1], dq_next,
                        dv_next, bwd_opt, o[0], o[1], o[2], o[3], o[4], o[5], o[6], o[7])
        dq_next = _clip_adj(np.asarray(o[0]), grad_clip_norm)
        dv_next = _clip_adj(np.asarray(o[1]), grad_clip_norm)
        dl_df[i] = _clip_adj(np.asarray(o[3]), grad_clip_norm)
        dl_dmat_w += np.asarray(o[4])
        if na:
            dl_da[i] = np.asarray(o[2])
        if len(o[5]):
            dl_dact_w += np.asarray(o[5])
        if len(o[6]):
            dl_dstate_p += np.asarray(o[6])
        w = np.asarray(o[7])
        dl_dw_elem = w.copy() if dl_dw_elem is None else dl_dw_elem + w
    info["backward_time"] = time.time() - t0
    info["dl_dmat_w"] = dl_dmat_w
    info["dl_da"] = dl_da
    info["dl_dact_w"] = dl_dact_w
    info["dl_dstate_p"] = dl_dstate_p
    info["dl_dw_elem"] = None if dl_dw_elem is None else dl_dw_elem.reshape(nE, -1)
    return loss, (dq_next, dv_next, dl_df), info
