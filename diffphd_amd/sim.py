"""torch autograd wrapper of one time step (python/py_diff_pd/common/sim.py:9-85).

Same names and argument order as the reference's `SimFunction` / `Sim`. Two
defects of the reference wrapper are fixed:
  * its backward passes an empty active-contact set and omits the last two
    outputs of `PyBackward` (`dl_dact_w`, `dl_dstate_p`); here the active set
    returned by the forward is kept and handed back, and the full output list
    is passed;
  * GPU tensors reach the HIP body without a host round-trip through numpy or
    lists: float64 device tensors go to the library by pointer (DLPack
    handoff, `forward_device` / `backward_device`) and the results are new
    tensors on the same device. Other tensors (CPU, another dtype, or a body
    without the device entry points) are copied to host arrays for the call
    and come back on the caller's device and dtype.
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


def _device_path(deformable, *tensors):
    return hasattr(deformable, "forward_device") and all(
        t.is_cuda and t.dtype == torch.float64 and t.device == tensors[0].device for t in tensors)


class SimFunction(autograd.Function):

    @staticmethod
    def forward(ctx, deformable, dofs, act_dofs, method, q, v, a, f_ext, dt, option):
        ctx.deformable = deformable
        ctx.dofs = dofs
        ctx.act_dofs = act_dofs
        ctx.method = method
        ctx.dt = dt
        ctx.option = option
        if _device_path(deformable, q, v, f_ext):
            qn, vn, active = deformable.forward_device(method, q.detach().contiguous(), v.detach().contiguous(), a,
                                                       f_ext.detach().contiguous(), dt, option)
            ctx.active = list(active)
        else:
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
        if _device_path(ctx.deformable, q, v, f_ext, dl_dq_next, dl_dv_next):
            out = ctx.deformable.backward_device(ctx.method, q.contiguous(), v.contiguous(), a, f_ext.contiguous(),
                                                 ctx.dt, q_next, v_next, ctx.active, dl_dq_next.contiguous(),
                                                 dl_dv_next.contiguous(), ctx.option)
            Sim.last_dl_dmat_w = out["dl_dmat_w"]
            da = _like(out["dl_da"], a) if out["dl_da"].size == a.numel() else torch.zeros_like(a)
            return (None, None, None, None, out["dl_dq"], out["dl_dv"], da, out["dl_df_ext"], None, None)
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
