"""torch autograd wrapper of one time step (python/py_diff_pd/common/sim.py:9-85).

Same names and argument order as the reference's `SimFunction` / `Sim`. Two
defects of the reference wrapper are fixed:
  * its backward passes an empty active-contact set and omits the last two
    outputs of `PyBackward` (`dl_dact_w`, `dl_dstate_p`); here the active set
    returned by the forward is kept and handed back, and the full output list
    is passed;
  * tensors on a GPU device are moved to host memory for the call (the solver
    takes host arrays) and the results are returned on the caller's device and
    dtype.
The per-energy material gradient `dl_dmat_w` of the last backward is kept in
`Sim.last_dl_dmat_w` (the reference wrapper drops it).
"""
import numpy as np
import torch
import torch.autograd as autograd
import torch.nn as nn

from .py_diff_pd_core import StdIntVector, StdRealVector


def _host(t):
    return t.detach().to("cpu", torch.float64).numpy()


def _like(arr, ref):
    return torch.as_tensor(np.asarray(arr, dtype=np.float64)).to(device=ref.device, dtype=ref.dtype)


class SimFunction(autograd.Function):

    @staticmethod
    def forward(ctx, deformable, dofs, act_dofs, method, q, v, a, f_ext, dt, option):
        ctx.deformable = deformable
        ctx.dofs = dofs
        ctx.act_dofs = act_dofs
        ctx.method = method
        ctx.dt = dt
        ctx.option = option
        q_next = StdRealVector(dofs)
        v_next = StdRealVector(dofs)
        active = StdIntVector(0)
        deformable.PyForward(method, _host(q), _host(v), _host(a), _host(f_ext), dt, option, q_next, v_next, active)
        ctx.active = list(active)
        qn = _like(q_next, q)
        vn = _like(v_next, v)
        ctx.save_for_backward(q, v, a, f_ext, qn, vn)
        return qn, vn

    @staticmethod
    def backward(ctx, dl_dq_next, dl_dv_next):
        q, v, a, f_ext, q_next, v_next = ctx.saved_tensors
        dofs = ctx.dofs
        if dl_dq_next is None:
            dl_dq_next = torch.zeros_like(q_next)
        if dl_dv_next is None:
            dl_dv_next = torch.zeros_like(v_next)
        dl_dq, dl_dv, dl_da, dl_df_ext = (StdRealVector(dofs), StdRealVector(dofs), StdRealVector(ctx.act_dofs),
                                          StdRealVector(dofs))
        dl_dmat_w, dl_dact_w, dl_dstate_p = StdRealVector(0), StdRealVector(0), StdRealVector(0)
        ctx.deformable.PyBackward(ctx.method, _host(q), _host(v), _host(a), _host(f_ext), ctx.dt, _host(q_next),
                                  _host(v_next), StdIntVector(ctx.active), _host(dl_dq_next), _host(dl_dv_next),
                                  ctx.option, dl_dq, dl_dv, dl_da, dl_df_ext, dl_dmat_w, dl_dact_w, dl_dstate_p)
        Sim.last_dl_dmat_w = np.asarray(dl_dmat_w, dtype=np.float64)
        da = _like(dl_da, a) if len(dl_da) == a.numel() else torch.zeros_like(a)
        return (None, None, None, None, _like(dl_dq, q), _like(dl_dv, v), da, _like(dl_df_ext, f_ext), None, None)


class Sim(nn.Module):
    last_dl_dmat_w = None

    def __init__(self, deformable):
        super().__init__()
        self.deformable = deformable

    def forward(self, dofs, act_dofs, method, q, v, a, f_ext, dt, option):
        return SimFunction.apply(self.deformable, dofs, act_dofs, method, q, v, a, f_ext, dt, option)
