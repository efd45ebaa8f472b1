"""Benchmark: forward+backward PD steps/s on the 131k-hex contact-rich press.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--cells 64,64,32]
                    [--precision mixed|fp64] [--no-fp64] [--no-cpu-baseline]
                    [--converging F] [--no-warm-start]

One step = one forward PD time step plus its adjoint step on the 64x64x32
hex press scene (131,072 elements, 139,425 vertices, ~418k DOFs; corotated +
volume with a 10x log-space random stiffness field; heightfield ground and a
descending press plane, friction 0.3). Inputs are resident in HBM; the timed
region covers exactly K forward frames and the K matching backward frames.
The FP32/FP64 "mixed" solver is the measured path; the plain FP64 path is
timed beside it for the speedup figure. The "converging" object adds a
time-to-tolerance leg (the frictionless press from rest, where the forward
loop reaches the 1e-9 dual gate) and the learned initial iterate on it. Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=4)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--cells", type=str, default="64,64,32")
    p.add_argument("--precision", choices=["mixed", "fp64"], default="mixed")
    p.add_argument("--no-fp64", action="store_true", help="skip timing the plain FP64 path")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--threads", type=int, default=min(16, os.cpu_count() or 8))
    p.add_argument("--start-frame", type=int, default=24,
                   help="forward-only frames run before the warmup to reach the pressing phase (untimed)")
    p.add_argument("--prep-iters", type=int, default=100, help="PD iteration cap of the untimed preparation frames")
    p.add_argument("--phases", action="store_true",
                   help="synchronize at phase boundaries and report the wall-clock split (perturbs the timing)")
    p.add_argument("--converging", type=int, default=3,
                   help="frames of the time-to-tolerance leg (frictionless press from rest at the 1e-9 dual gate, "
                        "FP64 and mixed); 0 skips it")
    p.add_argument("--converging-cap", type=int, default=3000, help="PD iteration cap of the time-to-tolerance leg")
    p.add_argument("--no-autotune", action="store_true",
                   help="keep the default launch shapes instead of timing the candidates after the factorization")
    p.add_argument("--no-tet", action="store_true",
                   help="skip the heterogeneous Neo-Hookean tet block leg (40^3 cubes, 320k tets)")
    p.add_argument("--no-napkin", action="store_true",
                   help="skip the napkin leg (alg_phd and the DiffPD baseline against the published answers)")
    p.add_argument("--no-warm-start", action="store_true",
                   help="skip training the learned warm start on plain-solver press rollouts and comparing it with "
                        "q_t and q + h v on the time-to-tolerance leg")
    p.add_argument("--profile-spmv", type=int, default=0,
                   help="only time N global solves (FP32 and FP64 factor) and a 1 GiB copy, for profiling")
    return p.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# Backward options of the measured path: per-element stiffness in K_pd and the
# trust region on the proximal-map spectrum (the reference's mean-stiffness K_pd
# makes the backbone adjoint diverge to the clip norm on this scene).
BWD_OPTS = {"alg_phd_bwd_exact_kpd": 1, "alg_phd_bwd_kpd_filter": 3}


def run_steps(body, spec, opt, start_q, start_v, frame0, K, bwd_opt=None):
    """K forward frames from (start_q, start_v) then the K backward frames.
    Returns states, active sets, the gradients at the window start and the
    per-frame backward statistics (newest frame first)."""
    from diffphd_amd.py_diff_pd_core import StdRealVector
    dofs = body.dofs()
    f = np.zeros(dofs)
    q, v, acts = [start_q], [start_v], [[]]
    for i in range(K):
        body.SetPressPlaneHeight(spec["plane_schedule"][frame0 + i])
        qn, vn, act = StdRealVector(dofs), StdRealVector(dofs), []
        body.PyForward("pd_eigen_alg_phd", q[-1], v[-1], [], f, spec["dt"], opt, qn, vn, act)
        q.append(np.asarray(qn))
        v.append(np.asarray(vn))
        acts.append(list(act))
        log(f"forward frame {frame0 + i}: active contacts {len(act)}")
    n = q[-1].size
    dq = 2.0 * (q[-1] - spec["target"]) / n  # loss = mean squared distance to the target
    dv = np.zeros(n)
    bopt = dict(opt, **(bwd_opt or {}))
    bstats = []
    for i in reversed(range(K)):
        body.SetPressPlaneHeight(spec["plane_schedule"][frame0 + i])
        o = [StdRealVector(0) for _ in range(8)]
        body.PyBackward("pd_eigen_alg_phd", q[i], v[i], [], f, spec["dt"], q[i + 1], v[i + 1], acts[i + 1], dq, dv,
                        bopt, *o)
        dq, dv = np.asarray(o[0]), np.asarray(o[1])
        st = body.last_backward_stats() if hasattr(body, "last_backward_stats") else {}
        st["frame"] = frame0 + i
        bstats.append(st)
        log(f"backward frame {frame0 + i}: |dl_dq| {np.linalg.norm(dq):.3e} "
            + " ".join(f"{k} {st[k]:g}" for k in ("backbone_iters", "woodbury_iters", "trust_stage", "clipped",
                                                   "filtered", "basis_rows") if k in st))
    return q, v, acts, dq, dv, bstats


def prepare_state(spec, args):
    """Forward-only frames 0..start_frame-1 (PD iterations capped at
    --prep-iters) on the mixed body: the plane reaches the block's top face
    after ~8 frames, so the timed window lies in the pressing phase with the
    top face and the ground bumps in contact. Returns (q, v) at start_frame."""
    from diffphd_amd import scenes
    from diffphd_amd.py_diff_pd_core import HexDeformable, StdRealVector
    body = scenes.apply(spec, HexDeformable(0))
    opt = dict(spec["options"], phd_precision=1, thread_ct=args.threads, max_pd_iter=args.prep_iters)
    q, v = spec["q0"], spec["v0"]
    f = np.zeros(q.size)
    for i in range(args.start_frame):
        body.SetPressPlaneHeight(spec["plane_schedule"][i])
        qn, vn, act = StdRealVector(q.size), StdRealVector(q.size), []
        body.PyForward("pd_eigen_alg_phd", q, v, [], f, spec["dt"], opt, qn, vn, act)
        q, v = np.asarray(qn), np.asarray(vn)
        log(f"prep frame {i}: active contacts {len(act)}")
    return q, v


def time_path(spec, precision, K, W, threads, q0, v0, frame0, phases=False, autotune=True):
    from diffphd_amd import scenes
    from diffphd_amd.py_diff_pd_core import HexDeformable
    body = scenes.apply(spec, HexDeformable(0))
    opt = dict(spec["options"], phd_precision=1 if precision == "mixed" else 0, thread_ct=threads)
    log(f"{precision}: building the factor")
    t0 = time.perf_counter()
    body.factor_build(spec["dt"], threads)
    t_setup = time.perf_counter() - t0
    tune = None
    if autotune:
        t0 = time.perf_counter()
        tune = body.autotune(spec["dt"], threads, reps=5)
        tune["seconds"] = time.perf_counter() - t0
        log(f"{precision}: autotune {tune}")
    log(f"{precision}: setup {t_setup:.1f}s {body.factor_info()}")
    q, v = [q0], [v0]
    if W > 0:
        q, v, _, _, _, _ = run_steps(body, spec, opt, q0, v0, frame0, W, BWD_OPTS)
    body.reset_frame_stats()
    body.timing_enable(2 if phases else 1)
    body.timing_read(reset=True)
    body.phase_times(reset=True)
    t0 = time.perf_counter()
    qq, vv, acts, dq0, dv0, bstats = run_steps(body, spec, opt, q[-1], v[-1], frame0 + W, K, BWD_OPTS)
    elapsed = time.perf_counter() - t0
    tim = body.timing_read(reset=True)
    ph = body.phase_times(reset=True)
    if phases:
        log(f"{precision} phases: " + json.dumps({k: round(v, 4) for k, v in ph.items()}))
    iters, conv, _ = body.frame_stats()
    info = body.factor_info()
    return {
        "body": body, "elapsed": elapsed, "setup_s": t_setup, "timing": tim, "fwd_iters": iters, "fwd_conv": conv,
        "nnz": info["nnz_S"], "max_depth": info["max_depth"], "factor_s": info["factor_s"], "autotune": tune,
        "active": [len(a) for a in acts[1:]], "q": qq, "dl_dq": dq0, "dl_dv": dv0, "bwd": bstats,
        "dense": {"solves": int(ph.get("dense_calls", 0)), "mean_rows": ph.get("dense_rows", 0) / max(ph.get("dense_calls", 0), 1),
                  "fp32_breakdowns_to_fp64": int(ph.get("dense_fp32_fail", 0))},
    }


def converging_leg(args, cells):
    """Time to tolerance: the press without friction from rest (frames 0 ..
    F-1: the block settles on the ground, the plane is still above it), where
    the forward PD loop reaches the 1e-9 dual gate, forward + backward in FP64
    and in mixed precision from the same state. Setup (factorization) is
    outside the timed region."""
    from diffphd_amd import scenes
    from diffphd_amd.py_diff_pd_core import HexDeformable
    F = args.converging
    spec = scenes.press(cells=cells, fmu=0.0, frames=max(100, F))
    out = {"workload": f"press {cells[0]}x{cells[1]}x{cells[2]} hex, friction 0, frames 0-{F - 1} from rest, "
                       f"fwd+bwd, dual gate 1e-9 (cap {args.converging_cap})"}
    states = {}
    for name, prec in (("fp64", 0), ("mixed", 1)):
        body = scenes.apply(spec, HexDeformable(0))
        body.factor_build(spec["dt"], args.threads)
        opt = dict(spec["options"], rel_tol=1e-9, max_pd_iter=args.converging_cap, phd_precision=prec,
                   thread_ct=args.threads)
        body.reset_frame_stats()
        t0 = time.perf_counter()
        q, v, acts, dq, dv, bstats = run_steps(body, spec, opt, spec["q0"], spec["v0"], 0, F, BWD_OPTS)
        secs = time.perf_counter() - t0
        iters, conv, _ = body.frame_stats()
        rows, clipped, bwd_iters = bwd_summary(bstats)
        out[name] = {"steps_per_s": F / secs, "seconds": secs, "fwd_pd_iters": [int(x) for x in iters],
                     "fwd_frames_converged": int(np.sum(conv)), "bwd_iters": bwd_iters,
                     "bwd_converged": int(sum(r["backbone_converged"] for r in rows)), "clipped_bwd_frames": clipped}
        states[name] = (q[-1], dq, dv)
        log(f"converging leg {name}: {F / secs:.3f} steps/s, forward iterations {list(iters)}")
        del body

    def dev(a, b):
        return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
    out["speedup_mixed_vs_fp64"] = out["mixed"]["steps_per_s"] / out["fp64"]["steps_per_s"]
    out["rel_dev_mixed_vs_fp64"] = {k: dev(a, b) for k, a, b in zip(("q_end", "dl_dq_start", "dl_dv_start"),
                                                                     states["mixed"], states["fp64"])}
    return out


def warm_start_leg(args, cells):
    """The learned warm start (per-vertex MLP 13 -> 3 x 128 -> 3, BF16 weights
    on MFMA) on the time-to-tolerance press in mixed precision:
    trained on the GPU from plain-solver rollouts of two other presses (their
    own stiffness fields and grounds, seeded random initial velocities), then
    the same frames from rest with q_t, q + h v and the learned start."""
    from diffphd_amd import scenes, warm_start as W
    from diffphd_amd.py_diff_pd_core import HexDeformable
    F = args.converging
    scene = lambda s: scenes.press(cells=cells, fmu=0.0, frames=max(100, F), seed=s)
    make = lambda spec: scenes.apply(spec, HexDeformable(0))
    spec = scene(0)
    opt = dict(spec["options"], rel_tol=1e-9, max_pd_iter=args.converging_cap, phd_precision=1,
               thread_ct=args.threads)
    t0 = time.perf_counter()
    feats, targets, plain_its = W.press_training_set(make, scene, seeds=[1, 2], rollouts=2, frames=F, opt=opt)
    t_data = time.perf_counter() - t0
    body = make(spec)
    body.mlp_init(0)
    t0 = time.perf_counter()
    losses = W.train(body, feats, targets, steps=2000, batch=4096, lr=1e-3, seed=0)
    t_train = time.perf_counter() - t0
    norm = W.normalization(feats, targets)
    fh, th, _, sh = W.collect(body, spec, F, opt, [spec["v0"]], spec["plane_schedule"], with_step=True)
    err = W.start_errors(body, norm, fh, th, sh)
    res = W.compare_starts(body, spec, F, opt)
    out = {"training": {"samples": int(feats.shape[0]), "frames": len(plain_its), "data_s": t_data,
                        "train_s": t_train, "adam_steps": 2000, "loss_first": float(np.mean(losses[:20])),
                        "loss_last": float(np.mean(losses[-20:]))},
           "start_distance_to_converged": err}
    for mode, name in ((0, "q_t"), (2, "q_plus_hv"), (1, "learned")):
        r = res[mode]
        out[name] = {"steps_per_s_fwd": F / r["seconds"], "fwd_pd_iters": r["iters"],
                     "fwd_frames_converged": int(sum(r["converged"])), "used": r["used"], "fallback": r["fallback"]}
        log(f"warm start {name}: {F / r['seconds']:.3f} fwd steps/s, iterations {r['iters']}")
    out["learned"]["fallback_rate"] = res[1]["fallback"] / max(F, 1)
    return out


def tet_leg(args, n=40, prep=14, K=2):
    """The heterogeneous Neo-Hookean tet block (40^3 cubes split in 5 tets:
    68,921 vertices, 320k tets; 100x inclusions, friction 0.5, dt 5e-3) landing
    on the ground: `prep` untimed frames until the bottom touches, then K timed
    fwd+bwd frames in mixed and in FP64 from the same state, with the deviation
    of the end state and of the window's gradients."""
    from diffphd_amd import scenes
    from diffphd_amd.py_diff_pd_core import StdRealVector, TetDeformable
    spec = scenes.tet_block(n=n)
    out = {"workload": f"tet block {n}^3 cubes x 5 tets ({spec['E'].shape[1]} tets, {spec['V'].shape[1]} vertices), "
                       "pd_neohookean, friction 0.5, dt 5e-3"}
    res = {}
    for name, prec in (("mixed", 1), ("fp64", 0)):
        body = scenes.apply(spec, TetDeformable(0))
        opt = dict(spec["options"], phd_precision=prec, thread_ct=args.threads)
        body.factor_build(spec["dt"], args.threads)
        ndof = body.dofs()
        q, v, act = np.asarray(spec["q0"], float), np.asarray(spec["v0"], float), []
        for _ in range(prep):
            qn, vn, a = StdRealVector(ndof), StdRealVector(ndof), list(act)
            body.PyForward("pd_eigen_alg_phd", q, v, [], np.zeros(ndof), spec["dt"], dict(opt, max_pd_iter=200), qn, vn, a)
            q, v, act = np.asarray(qn), np.asarray(vn), list(a)
        qs, vs, acts = [q], [v], [act]
        body.reset_frame_stats()
        t0 = time.time()
        for _ in range(K):
            qn, vn, a = StdRealVector(ndof), StdRealVector(ndof), list(acts[-1])
            body.PyForward("pd_eigen_alg_phd", qs[-1], vs[-1], [], np.zeros(ndof), spec["dt"], opt, qn, vn, a)
            qs.append(np.asarray(qn))
            vs.append(np.asarray(vn))
            acts.append(list(a))
        iters, conv = body.frame_stats()[:2]
        g = np.random.default_rng(0).standard_normal(ndof)
        dq, dv = g.copy(), np.zeros(ndof)
        bst = []
        for i in reversed(range(K)):
            o = [StdRealVector(0) for _ in range(8)]
            body.PyBackward("pd_eigen_alg_phd", qs[i], vs[i], [], np.zeros(ndof), spec["dt"], qs[i + 1], vs[i + 1],
                            acts[i + 1], dq, dv, dict(opt, **BWD_OPTS), *o)
            dq, dv = np.asarray(o[0]), np.asarray(o[1])
            bst.append(body.last_backward_stats())
        sec = time.time() - t0
        res[name] = (qs[-1], dq, dv)
        out[name] = {"steps_per_s": K / sec, "fwd_pd_iters": [int(x) for x in iters[-K:]],
                     "fwd_frames_converged": int(np.sum(conv[-K:])), "active_contacts": len(acts[-1]),
                     "bwd_backbone_iters": [int(b["backbone_iters"]) for b in bst],
                     "clipped_bwd_frames": int(sum(b["clipped"] for b in bst))}
        del body

    def dev(a, b):
        return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
    out["speedup_vs_fp64"] = out["mixed"]["steps_per_s"] / out["fp64"]["steps_per_s"]
    out["rel_dev_mixed_vs_fp64"] = {"q_end": dev(res["mixed"][0], res["fp64"][0]),
                                    "dl_dq_start": dev(res["mixed"][1], res["fp64"][1]),
                                    "dl_dv_start": dev(res["mixed"][2], res["fp64"][2])}
    return out


def napkin_leg():
    """The reference's installation check (python/example/napkin_3d_test.py):
    the heterogeneous napkin on a sphere, 125 frames, forward with
    pd_eigen_alg_phd and with the DiffPD baseline pd_eigen_pcg on the HIP body,
    against the published summary (README.md:214-217)."""
    from diffphd_amd import scenes
    from diffphd_amd.env import simulate
    from diffphd_amd.py_diff_pd_core import HexDeformable
    spec = scenes.napkin()
    baseline_opt = {"max_pd_iter": 5000, "max_ls_iter": 10, "abs_tol": 1e-9, "rel_tol": 1e-4, "verbose": 0,
                    "thread_ct": 8, "use_bfgs": 1, "use_acc": 1, "bfgs_history_size": 10, "project_newton": 0}
    out, zs = {}, {}
    for method, opt in (("pd_eigen_alg_phd", spec["options"]), ("pd_eigen_pcg", baseline_opt)):
        body = scenes.apply(spec, HexDeformable(0))
        t0 = time.time()
        _, _, info = simulate(body, spec["dt"], spec["frames"], method, opt, spec["q0"], spec["v0"],
                              lambda q, v: (0.0, 0 * q, 0 * v), require_grad=False)
        sec = time.time() - t0
        z = np.array(info["q"]).reshape(len(info["q"]), -1, 3)[:, :, 2]
        zs[method] = z
        out[method] = {"forward_s": sec, "max_z_std_after_landing": float(z[spec["frames"] - 35:].std(axis=1).max())}
    out["max_abs_z_drift"] = float(np.abs(zs["pd_eigen_alg_phd"] - zs["pd_eigen_pcg"]).max())
    out["published"] = {"pd_eigen_alg_phd": 0.06671, "pd_eigen_pcg": 0.06569, "max_abs_z_drift": 1.7370e-02}
    return out


def bwd_summary(bstats):
    """Per-frame backward record (oldest frame first) and totals."""
    rows = [{"frame": int(s["frame"]), "backbone_iters": int(s["backbone_iters"]),
             "woodbury_iters": int(s["woodbury_iters"]), "backbone_converged": int(s["backbone_converged"]),
             "trust_stage": int(s["trust_stage"]), "clipped": int(s["clipped"]), "basis_rows": int(s["basis_rows"]),
             "filtered_samples": int(s["filtered"])} for s in reversed(bstats)]
    return rows, sum(r["clipped"] for r in rows), sum(r["backbone_iters"] + r["woodbury_iters"] for r in rows)


def _cpu_step(args, cells):
    """One fwd+bwd step (frame 0) of the press generator at `cells` on the FP64
    oracle, after its factorization; returns (seconds, vertices, PD iterations,
    factorization seconds, candidates)."""
    import oracle
    from diffphd_amd import scenes
    spec = scenes.press(cells=cells, frames=max(100, args.start_frame + args.warmup + args.steps))
    orc = scenes.apply(spec, oracle.OracleDeformable(8))
    opt = dict(spec["options"], thread_ct=args.threads)
    t0 = time.perf_counter()
    orc.apply_ainv(spec["dt"], np.zeros(orc.dofs()))  # LDL^T factorization
    t_factor = time.perf_counter() - t0
    t0 = time.perf_counter()
    run_steps(orc, spec, opt, spec["q0"], spec["v0"], 0, 1, BWD_OPTS)
    t_step = time.perf_counter() - t0
    iters = orc.frame_stats()[0]
    return t_step, spec["q0"].size // 3, int(np.sum(iters)), t_factor, len(spec["boundary"][2])


def cpu_baseline(args, cells_list=((8, 8, 4), (16, 16, 8)), target_cells=(64, 64, 32)):
    """FP64 oracle (CPU restatement of the reference solver: per-axis sparse
    LDL^T, Delassus columns by triangular solves, host dense NCP and Anderson)
    timed on the host cores for one forward+backward step of the same press
    generator at two reduced grids. The 64x64x32 oracle needs minutes for its
    LDL^T alone and ~3K triangular solves per frame for the Delassus columns
    (K ~ 8.4k candidates), beyond a bounded sample. `value` is the larger
    grid's measured steps/s; the power law through the two measured sizes
    gives the extrapolated 64x64x32 figure, reported beside it."""
    import oracle
    os.environ["OMP_NUM_THREADS"] = str(args.threads)
    oracle.build()
    pts = []
    for cells in cells_list:
        t, nv, it, tf, K = _cpu_step(args, cells)
        pts.append({"cells": list(cells), "vertices": nv, "candidates": K, "seconds_per_step": t,
                    "fwd_pd_iters": it, "factor_s": tf})
        log(f"cpu baseline {cells}: {t:.2f} s/step ({it} PD iterations, factor {tf:.2f} s)")
    (a, b) = pts[-2], pts[-1]
    alpha = np.log(b["seconds_per_step"] / a["seconds_per_step"]) / np.log(b["vertices"] / a["vertices"])
    nv_t = (target_cells[0] + 1) * (target_cells[1] + 1) * (target_cells[2] + 1)
    t_extra = b["seconds_per_step"] * (nv_t / b["vertices"]) ** alpha
    return {"value": 1.0 / b["seconds_per_step"], "unit": "steps/s", "cores": args.threads, "kind": "port",
            "sample": (f"oracle fwd+bwd step 0 of the press generator (same options, factorization excluded) at "
                       + " and ".join(f"{p['cells'][0]}x{p['cells'][1]}x{p['cells'][2]} ({p['vertices']} vertices, "
                                      f"{p['seconds_per_step']:.2f} s)" for p in pts)
                       + f"; value = the larger grid. Power law t ~ N^{alpha:.2f} through the two sizes extrapolates "
                       f"to {t_extra:.3g} s/step at {target_cells[0]}x{target_cells[1]}x{target_cells[2]} "
                       f"({nv_t} vertices)"),
            "points": pts, "power_law_exponent": float(alpha),
            "extrapolated_steps_per_s_at_bench_grid": float(1.0 / t_extra)}


def profile_spmv(spec, args):
    """N applications of A^-1 = S^T S with each factor precision plus a plain copy."""
    from diffphd_amd import scenes
    from diffphd_amd.py_diff_pd_core import HexDeformable
    body = scenes.apply(spec, HexDeformable(0))
    body.factor_build(spec["dt"], args.threads)
    info = body.factor_info()
    log(f"factor {info}")
    b = np.random.default_rng(0).standard_normal(body.dofs())
    res = {"nnz_S": info["nnz_S"]}
    for prec, name in ((1, "fp32"), (0, "fp64")):
        body.apply_ainv(spec["dt"], b, precision=prec)
        body.timing_enable(True)
        body.timing_read(reset=True)
        for _ in range(args.profile_spmv):
            body.apply_ainv(spec["dt"], b, precision=prec)
        t = body.timing_read(reset=True)
        body.timing_enable(False)
        bv = 4 if prec == 1 else 8
        for k in ("rows", "cols"):
            ms = t[k + "_ms"] / max(t[k + "_launches"], 1)
            res[f"{name}_{k}_ms"] = ms
            res[f"{name}_{k}_GBs"] = info["nnz_S"] * bv / (ms * 1e-3) / 1e9
        log(f"{name}: {res}")
    ms = body.copy_probe(1 << 30, 20)
    res["copy_1GiB_ms"] = ms
    res["copy_GBs"] = 2 * (1 << 30) / (ms * 1e-3) / 1e9
    print(json.dumps(res))


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: --gpus {args.gpus} is not supported "
              f"(the solver has no multi-GPU path; it runs one body per GPU).", file=sys.stderr)
        sys.exit(2)
    from diffphd_amd import scenes
    cells = tuple(int(c) for c in args.cells.split(","))
    spec = scenes.press(cells=cells, frames=max(100, args.start_frame + args.warmup + args.steps))
    if args.profile_spmv > 0:
        profile_spmv(spec, args)
        return
    q0, v0 = prepare_state(spec, args) if args.start_frame > 0 else (spec["q0"], spec["v0"])
    res = time_path(spec, args.precision, args.steps, args.warmup, args.threads, q0, v0, args.start_frame, args.phases,
                    not args.no_autotune)
    K = args.steps
    steps_per_s = K / res["elapsed"]
    nnz = res["nnz"]
    bv = 4 if args.precision == "mixed" else 8
    tim = res["timing"]
    rows_avg = tim["rows_ms"] / max(tim["rows_launches"], 1)
    cols_avg = tim["cols_ms"] / max(tim["cols_launches"], 1)
    dom = "spmv_cols" if tim["cols_ms"] >= tim["rows_ms"] else "spmv_rows"
    dom_avg = cols_avg if dom == "spmv_cols" else rows_avg
    bytes_per_launch = nnz * bv
    achieved = bytes_per_launch / (dom_avg * 1e-3) / 1e9 if dom_avg > 0 else 0.0
    n_dof = spec["q0"].size
    fwd_iters = int(np.sum(res["fwd_iters"]))
    bwd_rows, clipped, bwd_iters = bwd_summary(res["bwd"])
    # HBM traffic per launch of the dominant kernel from the committed FETCH_SIZE /
    # WRITE_SIZE passes (calibrated on the copy kernel), for the workload they cover
    traffic = None
    pmc = os.path.join(ROOT, "profiles", f"spmv_press_{cells[0]}x{cells[1]}x{cells[2]}_pmc_traffic.json")
    if os.path.exists(pmc):
        kern = ("k_spmv_rows_grp" if dom == "spmv_rows" else "k_spmv_cols") + ("<float>" if bv == 4 else "<double>")
        traffic = json.load(open(pmc))["kernels"].get(kern, {}).get("traffic_bytes")
    copy_ms = res["body"].copy_probe(1 << 30, 10)
    out = {
        "metric": "fwd+bwd steps/s @131k-hex contact scene, mixed/learned vs plain FP64 PD",
        "value": steps_per_s,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * res["elapsed"] / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if args.precision == "mixed" else "f64",
        "data": "synthetic (seeded press scene: GRF stiffness field, sinusoid heightfield)",
        "config": {"workload": f"press {cells[0]}x{cells[1]}x{cells[2]} hex, fwd+bwd PD step, friction 0.3",
                   "elements": int(np.prod(cells)), "dofs": int(n_dof), "precision": args.precision,
                   "warm_start": "q_t (the reference's initial iterate; the learned start is measured on the "
                                 "time-to-tolerance leg, converging.warm_start)",
                   "backward": "per-element K_pd, trust region on the proximal-map spectrum (stage 0 = unfiltered)",
                   "timed_frames": [args.start_frame + args.warmup, args.start_frame + args.warmup + K - 1],
                   "active_contacts": res["active"], "fwd_pd_iters_timed": fwd_iters,
                   "fwd_frames_converged": int(np.sum(res["fwd_conv"])), "max_pd_iter": spec["options"]["max_pd_iter"],
                   "bwd_iters_timed": bwd_iters, "clipped_bwd_frames": clipped, "bwd_frames": bwd_rows,
                   "dense_contact_solves": res["dense"],
                   "setup_s": res["setup_s"], "factor_nnz": nnz, "factor_max_depth": res["max_depth"],
                   "autotune": res["autotune"]},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": "profiles/spmv_press_*_pmc_traffic.md" if traffic else None,
                     "copy_kernel_GBs": 2 * (1 << 30) / (copy_ms * 1e-3) / 1e9,
                     "bytes_per_launch": bytes_per_launch, "avg_launch_ms": dom_avg,
                     "spmv_rows_avg_ms": rows_avg, "spmv_cols_avg_ms": cols_avg,
                     "launches": {"rows": tim["rows_launches"], "cols": tim["cols_launches"]}},
        # 3N x (forward PD iterations + backbone and Woodbury adjoint iterations) per second
        "dof_iters_per_s": n_dof * (fwd_iters + bwd_iters) / res["elapsed"],
    }
    any_clipped = clipped > 0
    if not args.no_fp64 and args.precision == "mixed":
        res.pop("body", None)
        r64 = time_path(spec, "fp64", K, args.warmup, args.threads, q0, v0, args.start_frame, args.phases,
                        not args.no_autotune)
        sps64 = K / r64["elapsed"]
        r64.pop("body", None)
        rows64, clipped64, bwd64 = bwd_summary(r64["bwd"])
        any_clipped |= clipped64 > 0

        def dev(a, b):
            return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
        out["fp64_steps_per_s"] = sps64
        out["speedup_vs_fp64"] = steps_per_s / sps64
        out["fp64"] = {"frames": out["config"]["timed_frames"], "fwd_pd_iters_timed": int(np.sum(r64["fwd_iters"])),
                       "dense_contact_solves": r64["dense"],
                       "fwd_frames_converged": int(np.sum(r64["fwd_conv"])), "bwd_iters_timed": bwd64,
                       "clipped_bwd_frames": clipped64, "bwd_frames": rows64}
        # same start state, same frames: mixed against FP64 at the end of the
        # forward window (q) and at its start (the gradients of the window loss)
        out["rel_dev_vs_fp64"] = {"q_end": dev(res["q"][-1], r64["q"][-1]),
                                  "dl_dq_start": dev(res["dl_dq"], r64["dl_dq"]),
                                  "dl_dv_start": dev(res["dl_dv"], r64["dl_dv"])}
    if args.converging > 0:
        res.pop("body", None)
        out["converging"] = converging_leg(args, cells)
        any_clipped |= out["converging"]["fp64"]["clipped_bwd_frames"] + out["converging"]["mixed"]["clipped_bwd_frames"] > 0
        if not args.no_warm_start:
            try:
                out["converging"]["warm_start"] = warm_start_leg(args, cells)
            except Exception as e:  # reported in the line; the timed legs above stand
                log(f"warm start leg failed: {e!r}")
                out["converging"]["warm_start"] = {"error": repr(e)}
    if not args.no_tet:
        res.pop("body", None)
        out["tet_block"] = tet_leg(args)
    if not args.no_napkin:
        res.pop("body", None)
        out["napkin"] = napkin_leg()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args, target_cells=cells)
    print(json.dumps(out))
    if any_clipped:
        log("a timed backward frame reached the clip norm (its gradients were cut): see config.bwd_frames")
        sys.exit(3)


if __name__ == "__main__":
    main()
