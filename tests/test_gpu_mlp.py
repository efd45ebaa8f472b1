"""Learned warm start (phd_mlp.hip): MFMA lane layout, forward and gradient
against the CPU restatement (oracle/mlp.py), training on the GPU, and what
the trained network gains over the plain x + h v start."""
import numpy as np
import pytest

import oracle.mlp as M
from diffphd_amd import _lib, scenes, warm_start
from helpers import product_body, rel_err, rollout

pytestmark = pytest.mark.gpu


def test_mfma_bf16_lane_layout():
    rng = np.random.default_rng(0)
    A = M.bf16(rng.standard_normal((16, 16))).astype(np.float32)
    B = M.bf16(rng.standard_normal((16, 16))).astype(np.float32)
    a, ap = _lib.fptr(A)
    b, bp = _lib.fptr(B)
    d, dp = _lib.fptr(np.zeros((16, 16)))
    assert _lib.lib().phd_mlp_mfma_probe(ap, bp, dp) == 0
    ref = A.astype(np.float64) @ B.astype(np.float64)
    print("mfma", np.abs(d - ref).max())
    assert np.abs(d - ref).max() <= 1e-5 * np.abs(ref).max()


@pytest.fixture(scope="module")
def body():
    spec = scenes.cantilever()
    b = product_body(spec)
    b.mlp_init(1)
    return spec, b


def test_mlp_forward_and_gradient_match_oracle(body):
    _, b = body
    rng = np.random.default_rng(1)
    p = b.mlp_params().astype(np.float64)
    n = 203  # not a multiple of the 16-row tiles
    X = np.zeros((n, 16), np.float32)
    X[:, :13] = rng.standard_normal((n, 13))
    T = rng.standard_normal((n, 3)).astype(np.float32)
    out = b.mlp_forward(X)
    ref = M.forward(p, X.astype(np.float64))
    print("forward", rel_err(out, ref))
    assert rel_err(out, ref) < 1e-5
    loss, g = b.mlp_loss_grad(X, T)
    lref, gref = M.loss_grad(p, X.astype(np.float64), T.astype(np.float64))
    print("loss", loss, lref, "grad", rel_err(g, gref))
    assert abs(loss - lref) < 1e-5 * lref
    assert rel_err(g, gref) < 1e-5
    # padded rows / columns carry no gradient
    G = M.unpack(g)
    assert np.all(G["W1"][13:] == 0) and np.all(G["W4"][:, 3:] == 0) and np.all(G["b4"][3:] == 0)


def test_warm_start_training_and_gain(body):
    """Train on three rollouts of the cantilever (random uniform initial
    velocities) on the GPU, then measure on a held-out rollout what the
    network's start gains over x + h v (start error, PD iterations; printed:
    with three training rollouts the network does not generalize beyond
    x + h v). Asserted: training lowers the loss, the learned start is not
    worse than x + h v by more than half, and the converged answer is the
    same with the learned start, with x + h v, and with the reference's q_t."""
    spec, b = body
    rng = np.random.default_rng(5)
    opt = dict(spec["options"], rel_tol=1e-9)
    frames = 10
    v0s = [np.tile(rng.uniform(-0.3, 0.3, 3), spec["q0"].size // 3) for _ in range(4)]
    feats, targets, _ = warm_start.collect(b, spec, frames, opt, v0s[:3])
    losses = warm_start.train(b, feats, targets, steps=800, batch=4096, lr=1e-3, seed=0)
    print("training loss", losses[:3], losses[-3:])
    assert np.mean(losses[-20:]) < 0.3 * np.mean(losses[:20])
    fh, th, it_plain = warm_start.collect(b, spec, frames, opt, v0s[3:])
    mean, std, tscale = warm_start.normalization(feats, targets)
    pred = b.mlp_forward(warm_start.standardize(fh, mean, std)) * tscale
    e_net, e_xhv = np.linalg.norm(th - pred), np.linalg.norm(th)
    print("held-out start error: network", e_net, "x + h v", e_xhv)
    assert e_net < 1.5 * e_xhv
    s = dict(spec, v0=v0s[3])
    b.reset_frame_stats()
    b.warm_start_stats()
    q1, _, _ = rollout(b, s, frames, dict(opt, phd_warm_start=1))
    it_net = b.frame_stats()[0]
    used, fell = b.warm_start_stats()
    b.reset_frame_stats()
    q2, _, _ = rollout(b, s, frames, dict(opt, phd_warm_start=2))
    it_xhv = b.frame_stats()[0]
    print("PD iterations: q_t", sum(it_plain), "x + h v", sum(it_xhv), "learned", sum(it_net), "used/fallback", used, fell)
    q0, _, _ = rollout(b, s, frames, opt)
    assert rel_err(q1[-1], q0[-1]) < 1e-6 and rel_err(q2[-1], q0[-1]) < 1e-6
