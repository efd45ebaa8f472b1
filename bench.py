"""Benchmark: forward+backward PD steps/s on the 131k-hex contact-rich press.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--cells 64,64,32]
                    [--precision mixed|fp64] [--no-fp64] [--no-cpu-baseline]

One step = one forward PD time step plus its adjoint step on the 64x64x32
hex press scene (131,072 elements, 139,425 vertices, ~418k DOFs; corotated +
volume with a 10x log-space random stiffness field; heightfield ground and a
descending press plane, friction 0.3). Inputs are resident in HBM; the timed
region covers exactly K forward frames and the K matching backward frames.
The FP32/FP64 "mixed" solver is the measured path; the plain FP64 path is
timed beside it for the speedup figure. Prints one JSON line.
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
    p.add_argument("--profile-spmv", type=int, default=0,
                   help="only time N global solves (FP32 and FP64 factor) and a 1 GiB copy, for profiling")
    return p.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def run_steps(body, spec, opt, start_q, start_v, frame0, K):
    """K forward frames from (start_q, start_v) then the K backward frames."""
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
    for i in reversed(range(K)):
        body.SetPressPlaneHeight(spec["plane_schedule"][frame0 + i])
        o = [StdRealVector(0) for _ in range(8)]
        body.PyBackward("pd_eigen_alg_phd", q[i], v[i], [], f, spec["dt"], q[i + 1], v[i + 1], acts[i + 1], dq, dv,
                        opt, *o)
        dq, dv = np.asarray(o[0]), np.asarray(o[1])
        log(f"backward frame {frame0 + i}: |dl_dq| {np.linalg.norm(dq):.3e}")
    return q, v, acts


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


def time_path(spec, precision, K, W, threads, q0, v0, frame0, phases=False):
    from diffphd_amd import scenes
    from diffphd_amd.py_diff_pd_core import HexDeformable
    body = scenes.apply(spec, HexDeformable(0))
    opt = dict(spec["options"], phd_precision=1 if precision == "mixed" else 0, thread_ct=threads)
    log(f"{precision}: building the factor")
    t0 = time.perf_counter()
    body.factor_build(spec["dt"], threads)
    t_setup = time.perf_counter() - t0
    log(f"{precision}: setup {t_setup:.1f}s {body.factor_info()}")
    q, v = [q0], [v0]
    if W > 0:
        q, v, _ = run_steps(body, spec, opt, q0, v0, frame0, W)
    body.reset_frame_stats()
    body.timing_enable(2 if phases else 1)
    body.timing_read(reset=True)
    body.phase_times(reset=True)
    t0 = time.perf_counter()
    qq, vv, acts = run_steps(body, spec, opt, q[-1], v[-1], frame0 + W, K)
    elapsed = time.perf_counter() - t0
    tim = body.timing_read(reset=True)
    if phases:
        log(f"{precision} phases: " + json.dumps({k: round(v, 4) for k, v in body.phase_times(reset=True).items()}))
    iters = body.frame_stats()[0]
    info = body.factor_info()
    return {
        "body": body, "elapsed": elapsed, "setup_s": t_setup, "timing": tim, "fwd_iters": iters,
        "nnz": info["nnz_S"], "max_depth": info["max_depth"], "factor_s": info["factor_s"],
        "active": [len(a) for a in acts[1:]], "q": qq, "bwd_iters": body.last_backward_iters(),
    }


def cpu_baseline(args, cells=(16, 16, 8)):
    """FP64 oracle (CPU restatement of the reference solver: per-axis sparse
    LDL^T, Delassus columns by triangular solves, host dense NCP and Anderson)
    timed on the host cores for one forward+backward step of the same press
    generator at a reduced grid. The full 64x64x32 oracle needs minutes for its
    LDL^T alone and ~3K triangular solves per frame for the Delassus columns
    (K ~ 8.4k candidates), far beyond a bounded sample; the reduced grid runs
    the identical algorithm and options, so its steps/s bounds the CPU rate at
    full size from above."""
    import oracle
    from diffphd_amd import scenes
    os.environ["OMP_NUM_THREADS"] = str(args.threads)
    oracle.build()
    spec = scenes.press(cells=cells, frames=max(100, args.start_frame + args.warmup + args.steps))
    orc = scenes.apply(spec, oracle.OracleDeformable(8))
    opt = dict(spec["options"], thread_ct=args.threads)
    t0 = time.perf_counter()
    orc.apply_ainv(spec["dt"], np.zeros(orc.dofs()))  # LDL^T factorization
    t_factor = time.perf_counter() - t0
    t0 = time.perf_counter()
    run_steps(orc, spec, opt, spec["q0"], spec["v0"], 0, 1)
    t_step = time.perf_counter() - t0
    iters = orc.frame_stats()[0]
    nv = spec["q0"].size // 3
    return {"value": 1.0 / t_step, "unit": "steps/s", "cores": args.threads, "kind": "port",
            "sample": (f"oracle fwd+bwd step 0 of the press generator at {cells[0]}x{cells[1]}x{cells[2]} cells "
                       f"({nv} vertices, {len(spec['boundary'][2])} contact candidates, same options; "
                       f"{int(np.sum(iters))} forward PD iterations) in {t_step:.1f} s, factorization "
                       f"{t_factor:.2f} s excluded; an upper bound on the CPU steps/s at 64x64x32")}


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
    res = time_path(spec, args.precision, args.steps, args.warmup, args.threads, q0, v0, args.start_frame, args.phases)
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
                   "timed_frames": [args.start_frame + args.warmup, args.start_frame + args.warmup + K - 1],
                   "active_contacts": res["active"], "fwd_pd_iters_timed": fwd_iters,
                   "setup_s": res["setup_s"], "factor_nnz": nnz, "factor_max_depth": res["max_depth"]},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": "profiles/spmv_press_*_pmc_traffic.md" if traffic else None,
                     "copy_kernel_GBs": 2 * (1 << 30) / (copy_ms * 1e-3) / 1e9,
                     "bytes_per_launch": bytes_per_launch, "avg_launch_ms": dom_avg,
                     "spmv_rows_avg_ms": rows_avg, "spmv_cols_avg_ms": cols_avg,
                     "launches": {"rows": tim["rows_launches"], "cols": tim["cols_launches"]}},
        "dof_iters_per_s": n_dof * fwd_iters / res["elapsed"],
    }
    if not args.no_fp64 and args.precision == "mixed":
        del res["body"]
        K64 = max(1, K // 2)
        r64 = time_path(spec, "fp64", K64, args.warmup, args.threads, q0, v0, args.start_frame, args.phases)
        sps64 = K64 / r64["elapsed"]
        out["fp64_steps_per_s"] = sps64
        out["speedup_vs_fp64"] = steps_per_s / sps64
        dq = res["q"][K64] - r64["q"][K64]
        out["rel_dev_q_vs_fp64"] = float(np.linalg.norm(dq) / np.linalg.norm(r64["q"][K64]))
        out["fp64_fwd_pd_iters_timed"] = int(np.sum(r64["fwd_iters"]))
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
