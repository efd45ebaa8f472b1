"""Learned warm start on the press scene: a per-vertex network trained on the
GPU from plain-solver rollouts of reduced presses (their own stiffness fields
and grounds, seeded random initial velocities), evaluated on a held-out press
rolled out from rest at the 1e-9 dual gate. The network's initial iterate is
compared with q + h v and with the reference's q_t."""
import numpy as np
import pytest

from diffphd_amd import scenes, warm_start as W
from helpers import product_body, rel_err

pytestmark = pytest.mark.gpu


def test_press_learned_start_gain():
    cells = (16, 16, 8)
    scene = lambda s: scenes.press(cells=cells, fmu=0.0, seed=s)
    opt = dict(scene(0)["options"], rel_tol=1e-9, max_pd_iter=3000, phd_precision=0)
    F, T, its = W.press_training_set(product_body, scene, seeds=[1, 2, 3, 4], rollouts=2, frames=8, opt=opt)
    print("training samples", F.shape[0], "plain PD iterations", sum(its))
    spec = scene(0)
    body = product_body(spec)
    body.mlp_init(0)
    losses = W.train(body, F, T, steps=2000, batch=4096, lr=1e-3, seed=0)
    print("training loss", np.mean(losses[:20]), "->", np.mean(losses[-20:]))
    assert np.mean(losses[-20:]) < 0.1 * np.mean(losses[:20])
    norm = W.normalization(F, T)
    frames = 10
    fh, th, _, sh = W.collect(body, spec, frames, opt, [spec["v0"]], spec["plane_schedule"], with_step=True)
    err = W.start_errors(body, norm, fh, th, sh)
    print("held-out distance of the initial iterate to the converged state", err)
    # the gain: the learned iterate is much closer than q + h v and q_t
    assert err["learned"] < 0.75 * err["q_plus_hv"] and err["q_plus_hv"] < err["q_t"]
    res = W.compare_starts(body, spec, frames, opt)
    for mode, name in ((0, "q_t"), (2, "q + h v"), (1, "learned")):
        r = res[mode]
        print(f"{name}: PD iterations {sum(r['iters'])} {r['iters']} converged {sum(r['converged'])}/{frames} "
              f"used {r['used']} fallback {r['fallback']} {r['seconds']:.2f} s")
        assert sum(r["converged"]) == frames
    assert res[1]["used"] + res[1]["fallback"] == frames
    # every start reaches the same state to the solve's tolerance
    assert rel_err(res[1]["q"], res[0]["q"]) < 1e-5 and rel_err(res[2]["q"], res[0]["q"]) < 1e-5
