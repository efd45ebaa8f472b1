"""Training data and training of the learned warm start (phd_warm_start = 1).

Data: plain-solver rollouts; per frame, the raw features of the start state
(computed by the body itself, right before that frame's forward, so the
contact flags are those the solver will see) and the raw target
q_next - q - h v. Normalization: per-feature mean / standard deviation and
the RMS of the targets."""
import numpy as np

from .py_diff_pd_core import StdRealVector


def collect(body, spec, frames, opt, initial_velocities, plane_schedule=None):
    feats, targets, iters = [], [], []
    dofs = body.dofs()
    f_ext = np.zeros(dofs)
    h = spec["dt"]
    for v0 in initial_velocities:
        q, v = np.asarray(spec["q0"], float), np.asarray(v0, float)
        body.reset_frame_stats()
        for i in range(frames):
            if plane_schedule is not None:
                body.SetPressPlaneHeight(plane_schedule[i])
            feats.append(body.mlp_features(q, v, h, f_ext))
            qn, vn, act = StdRealVector(dofs), StdRealVector(dofs), []
            body.PyForward("pd_eigen_alg_phd", q, v, [], f_ext, h, opt, qn, vn, act)
            qn, vn = np.asarray(qn), np.asarray(vn)
            targets.append((qn - q - h * v).reshape(-1, 3))
            q, v = qn, vn
        iters.extend(body.frame_stats()[0])
    return np.concatenate(feats), np.concatenate(targets), iters


def normalization(feats, targets):
    f = feats[:, :13].astype(np.float64)
    mean = f.mean(0)
    std = f.std(0)
    std[std < 1e-6] = 1.0
    tscale = float(np.sqrt(np.mean(targets.astype(np.float64) ** 2))) or 1.0
    return mean.astype(np.float32), std.astype(np.float32), tscale


def standardize(feats, mean, std):
    out = np.zeros_like(feats, dtype=np.float32)
    out[:, :13] = (feats[:, :13] - mean) / std
    return out


def train(body, feats, targets, steps=2000, batch=4096, lr=1e-3, seed=0):
    """Initializes (if needed), normalizes and trains the body's network;
    returns the per-step minibatch losses."""
    mean, std, tscale = normalization(feats, targets)
    body.mlp_set_normalization(mean, std, tscale)
    return body.mlp_train(standardize(feats, mean, std), targets / tscale, steps, batch, lr, seed)
