"""Training data and training of the learned warm start (phd_warm_start = 1).

Data: plain-solver rollouts; per frame, the raw features of the start state
(computed by the body itself, right before that frame's forward, so the
contact flags are those the solver will see) and the raw target
q_next - q - h v. Normalization: per-feature mean / standard deviation and
the RMS of the targets."""
import numpy as np

from .py_diff_pd_core import StdRealVector


def collect(body, spec, frames, opt, initial_velocities, plane_schedule=None, with_step=False):
    """Raw features and targets of `frames` plain-solver frames per initial
    velocity, the PD iterations per frame and (with_step) q_next - q."""
    feats, targets, iters, steps = [], [], [], []
    dofs = body.dofs()
    f_ext = np.zeros(dofs)
    h = spec["dt"]
    for v0 in initial_velocities:
        q, v = np.asarray(spec["q0"], float), np.asarray(v0, float)
        body.reset_frame_stats()
        if hasattr(body, "reset_contact_history"):
            body.reset_contact_history()
        for i in range(frames):
            if plane_schedule is not None:
                body.SetPressPlaneHeight(plane_schedule[i])
            feats.append(body.mlp_features(q, v, h, f_ext))
            qn, vn, act = StdRealVector(dofs), StdRealVector(dofs), []
            body.PyForward("pd_eigen_alg_phd", q, v, [], f_ext, h, opt, qn, vn, act)
            qn, vn = np.asarray(qn), np.asarray(vn)
            targets.append((qn - q - h * v).reshape(-1, 3))
            steps.append((qn - q).reshape(-1, 3))
            q, v = qn, vn
        iters.extend(body.frame_stats()[0])
    if with_step:
        return np.concatenate(feats), np.concatenate(targets), iters, np.concatenate(steps)
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


def transfer(src, dst, feats, targets):
    """Copies src's trained network and the normalization of (feats, targets) to dst."""
    mean, std, tscale = normalization(feats, targets)
    dst.mlp_init(0)
    dst.mlp_set_normalization(mean, std, tscale)
    dst.mlp_set_params(src.mlp_params())


def press_training_set(make_body, scene, seeds, rollouts, frames, opt, v_scale=0.05, seed=0):
    """Plain-solver training data on the press: for each scene seed (its own
    stiffness field and ground), `rollouts` rollouts of `frames` frames from
    seeded random initial velocities (a uniform translation of magnitude up to
    v_scale plus per-DOF noise of v_scale / 5). Returns the pooled raw
    features, targets and the plain solver's PD iterations per frame."""
    rng = np.random.default_rng(seed)
    F, T, its = [], [], []
    for s in seeds:
        spec = scene(s)
        body = make_body(spec)
        n = spec["q0"].size
        v0s = [np.tile(rng.uniform(-v_scale, v_scale, 3), n // 3) + 0.2 * v_scale * rng.standard_normal(n)
               for _ in range(rollouts)]
        f, t, it = collect(body, spec, frames, opt, v0s, spec.get("plane_schedule"))
        F.append(f)
        T.append(t)
        its.extend(it)
    return np.concatenate(F), np.concatenate(T), its


def compare_starts(body, spec, frames, opt, v0=None):
    """One rollout per initial iterate (phd_warm_start 0 = q_t, the reference's;
    2 = q + h v; 1 = the learned start with its fallback to q + h v) from the
    same state: per-frame PD iterations, converged flags, wall seconds, the
    learned start's used / fallback counts, and the final state of each."""
    import time
    from .py_diff_pd_core import StdRealVector
    dofs = body.dofs()
    f_ext = np.zeros(dofs)
    out = {}
    for mode in (0, 2, 1):
        q = np.asarray(spec["q0"], float)
        v = np.asarray(spec["v0"] if v0 is None else v0, float)
        body.reset_frame_stats()
        body.reset_contact_history()
        body.warm_start_stats(reset=True)
        t0 = time.perf_counter()
        for i in range(frames):
            if spec.get("plane_schedule") is not None:
                body.SetPressPlaneHeight(spec["plane_schedule"][i])
            qn, vn, act = StdRealVector(dofs), StdRealVector(dofs), []
            body.PyForward("pd_eigen_alg_phd", q, v, [], f_ext, spec["dt"], dict(opt, phd_warm_start=mode), qn, vn, act)
            q, v = np.asarray(qn), np.asarray(vn)
        secs = time.perf_counter() - t0
        it, conv, _ = body.frame_stats()
        used, fell = body.warm_start_stats(reset=True)
        out[mode] = {"iters": [int(x) for x in it], "converged": [int(x) for x in conv], "seconds": secs,
                     "used": used, "fallback": fell, "q": q}
    return out


def start_errors(body, norm, feats, targets, q_step):
    """Distance of each initial iterate to the converged next state over the
    given frames: the learned start (network prediction), q + h v (target
    zero) and q_t (q_step = q_next - q per frame). feats are raw features,
    norm = normalization() of the training set."""
    mean, std, tscale = norm
    pred = body.mlp_forward(standardize(feats, mean, std)) * tscale
    return {"learned": float(np.linalg.norm(targets - pred)), "q_plus_hv": float(np.linalg.norm(targets)),
            "q_t": float(np.linalg.norm(q_step))}
