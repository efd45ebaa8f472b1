"""Batched rollouts of independent bodies (many small, uneven meshes).

Each body owns its HIP stream, factor and buffers, and the host loop of a step
releases the GIL inside the C ABI call, so rollouts run concurrently from a
thread pool and their kernels overlap on the GPU. Results do not depend on the
batching: each body's arithmetic is the same as in a sequential rollout.
"""
from concurrent.futures import ThreadPoolExecutor

from . import scenes
from .env import simulate


def _rollout(body, spec, frames, opt, require_grad):
    g = spec["loss_g"]
    loss, grad, info = simulate(body, spec["dt"], frames, "pd_eigen_alg_phd", opt, spec["q0"], spec["v0"],
                                lambda q, v: (float(g @ q), g.copy(), 0 * v), require_grad=require_grad)
    return {"loss": loss, "q_T": info["q"][-1], "grad_v0": None if grad is None else grad[1],
            "dl_dmat_w": info.get("dl_dmat_w")}


def rollout_batch(specs, make_body, frames=None, options=None, workers=8, require_grad=True):
    """Forward (and, with require_grad, backward) rollouts of every spec.

    make_body: a callable returning a fresh body (e.g. lambda: HexDeformable(0)).
    The bodies are built here, in the calling thread; the rollouts then run on
    `workers` threads. Returns one dict per spec (loss, final q, dl/dv0,
    dl/dmat_w), in order."""
    bodies = [scenes.apply(s, make_body()) for s in specs]

    def run(i):
        spec = specs[i]
        opt = dict(spec["options"], **(options or {}))
        return _rollout(bodies[i], spec, frames or spec["frames"], opt, require_grad)

    if workers <= 1:
        return [run(i) for i in range(len(specs))]
    with ThreadPoolExecutor(max_workers=workers) as ex:
        return list(ex.map(run, range(len(specs))))
