"""Benchmark: bundle-adjustment LM iterations/s, mixed precision vs plain FP64.

Workload (default ``ba200``): 200 cameras on a radius-10 circle, 100,000
points, 3..10 observations per point (~650k observations), 5% of them replaced
by uniform outliers in [0,1000]^2 px with Huber(1 px) - the mixed-precision
form of the at-scale problem.  Inputs are seeded and resident in HBM before
timing starts.

A *step* is one LM iteration.  ``--steps K`` LM iterations are timed (all
tolerances 0, so exactly K iterations run) after ``--warmup W`` untimed ones:

* mixed  : FP32 Jacobian / J^T J / Schur assembly, FP64 Cholesky of the
           1800x1800 reduced camera system, FP64 cost and gain ratio, 2
           refinement steps, device-resident LM state, graph-captured loop;
* plain  : the same kernels in FP64 with the host-driven loop (one host
           synchronisation per iteration), DENSE_SCHUR and ITERATIVE_SCHUR.

``value`` is the mixed form's LM iterations/s.  Time-to-tolerance (default
tolerances) is read from per-iteration device timestamps.  The roofline entry
covers the kernel group that takes the most time per iteration, timed with HIP
events on the stream the library launches on.  The CPU baseline times the
float64 oracle (numpy, one thread) on two LM iterations of the same problem.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path[:0] = [str(ROOT), str(ROOT / "ceres-torch-cuda_amd")]

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def make_device(data):
    import torch

    import ceres_mi355x as tc
    from ceres_mi355x.ba import BADevice

    loss = tc.HuberLoss(data.loss[1]) if data.loss[0] == "huber" else None
    return BADevice(torch.tensor(data.cameras), torch.tensor(data.points), data.obs_camera, data.obs_point,
                    data.observations, loss_native=tc.losses.native_loss(loss))


def run_solve(dev, data, opts):
    import torch

    dev.set_parameters(torch.tensor(data.cameras), torch.tensor(data.points))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rep, rec = dev.run(opts)
    torch.cuda.synchronize()
    return rep, rec, time.perf_counter() - t0


def timed_iters(dev, data, *, precision, loop, linear, steps, warmup, refine=2):
    import ceres_mi355x as tc
    from ceres_mi355x.ba import lm_options

    base = tc.SolverOptions(function_tolerance=0.0, gradient_tolerance=0.0, parameter_tolerance=0.0,
                            max_num_refinement_iterations=refine if precision == "mixed" else 0)
    base.max_num_iterations = max(1, warmup)
    run_solve(dev, data, lm_options(base, precision=precision, loop=loop, linear=linear))
    base.max_num_iterations = steps
    opts = lm_options(base, precision=precision, loop=loop, linear=linear)
    run_solve(dev, data, opts)  # builds the plan / graph for this option set
    rep, rec, wall = run_solve(dev, data, opts)
    iters = int(rep.num_records) - 1
    return dict(iters=iters, wall_s=wall, iters_per_s=iters / wall, final_cost=rep.final_cost,
                device_s=rep.device_seconds, lin_iters=[int(x) for x in rec[:, 10]]), rep


def time_to_tol(dev, data, *, precision, loop, linear, max_iters, target):
    import ceres_mi355x as tc
    from ceres_mi355x.ba import lm_options

    o = tc.SolverOptions(max_num_iterations=max_iters, max_num_refinement_iterations=2 if precision == "mixed" else 0)
    opts = lm_options(o, precision=precision, loop=loop, linear=linear)
    run_solve(dev, data, opts)
    rep, rec, wall = run_solve(dev, data, opts)
    hit = None
    if target is not None:
        for row in rec:
            if row[1] <= target:
                hit = float(row[11])
                break
    return rep, rec, wall, hit


def profile_kernels(dev, data, *, precision, linear, iters):
    import ctypes

    import numpy as np
    import torch

    import ceres_mi355x as tc
    from ceres_mi355x import _native as nat
    from ceres_mi355x.ba import lm_options

    o = tc.SolverOptions(max_num_iterations=iters + 2, function_tolerance=0.0, gradient_tolerance=0.0,
                         parameter_tolerance=0.0, max_num_refinement_iterations=2 if precision == "mixed" else 0)
    opts = lm_options(o, precision=precision, loop="host", linear=linear)
    run_solve(dev, data, opts)
    handle = dev._plans[tuple(getattr(opts, f) for f, _ in nat.ct_lm_options._fields_)][0]
    dev.set_parameters(torch.tensor(data.cameras), torch.tensor(data.points))
    ms = (ctypes.c_double * len(nat.KERNELS))()
    cnt = (ctypes.c_int32 * len(nat.KERNELS))()
    nat.check(nat.load().ct_ba_plan_profile(handle, iters, ms, cnt), "ct_ba_plan_profile")
    return {k: (ms[i] / max(cnt[i], 1), cnt[i]) for i, k in enumerate(nat.KERNELS)}


def algorithmic_work(kernel: str, data, precision: str):
    """Compulsory bytes (or flops) of one launch of a kernel group.

    Every input is read once and every output written once; gathered re-reads
    are not counted.  s = bytes per assembly scalar (4 mixed, 8 FP64).
    """
    C, P, N = data.num_cameras, data.num_points, data.num_observations
    s = 4 if precision == "mixed" else 8
    n = 9 * C
    if kernel == "eval_points":
        reads = N * (16 + 4) + P * (24 + 4) + C * 72
        writes = N * (18 * s + 6 * s + 6 * s + 2 * s + 16) + P * (6 * s + 3 * s + 24 + 24)
        return "hbm", reads + writes
    if kernel == "camera_blocks":
        import numpy as np

        counts = np.bincount(data.obs_point, minlength=P)
        terms = int(np.sum(counts * (counts + 1) // 2))
        reads = N * (18 * s + 6 * s + 2 * s + 16 + 4) + terms * 8 + (C * (C + 1) // 2 + 1) * 4
        writes = n * n * 8 + n * 8 * 3
        return "hbm", reads + writes
    if kernel == "candidate":
        reads = N * (18 * s + 6 * s + 16 + 16 + 4) + P * (6 * s + 3 * s + 24 + 4) + C * 72 + n * 8
        writes = P * (24 + 24) + n * 8
        return "hbm", reads + writes
    if kernel == "rcs_solve":
        return "mfma", n ** 3 / 3.0 + 2.0 * n * n  # FP64 flops of potrf + 2 triangular solves
    return None, None


def cpu_baseline(data, iterations=2):
    """Float64 oracle (numpy) on the same problem, one thread: LM iterations/s."""
    from threadpoolctl import threadpool_limits

    from oracle.ba import BAModel
    from oracle.lm import OracleOptions, trust_region_solve

    with threadpool_limits(limits=1):
        model = BAModel(data)
        opts = OracleOptions(max_num_iterations=iterations, function_tolerance=0.0, gradient_tolerance=0.0,
                             parameter_tolerance=0.0)
        t0 = time.perf_counter()
        s, _ = trust_region_solve(model, opts)
        el = time.perf_counter() - t0
    return dict(value=iterations / el, unit="LM iterations/s", cores=1, kind="port",
                sample=f"{iterations} LM iterations of the float64 numpy oracle (block-sparse Schur, DENSE_SCHUR) on "
                       f"the same problem, {el:.1f} s, BLAS limited to 1 thread")


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="ba200", choices=["ba200", "ba16"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (multi-GPU sharding of the reduced "
              "camera system is a design note, not a built path)", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from oracle import generators as gen

    assert torch.cuda.is_available(), "bench.py needs a ROCm GPU"
    name = "ba200_huber" if args.workload == "ba200" else "ba16_huber"
    t0 = time.perf_counter()
    data = gen.ba_config(name)
    dev = make_device(data)
    log(f"[bench] {name}: C={data.num_cameras} P={data.num_points} N={data.num_observations} "
        f"setup {time.perf_counter() - t0:.1f}s")

    out = {}
    mixed, _ = timed_iters(dev, data, precision="mixed", loop="device", linear="dense_schur", steps=args.steps,
                           warmup=args.warmup)
    log(f"[bench] mixed device-loop: {mixed['iters_per_s']:.1f} it/s ({mixed['iters']} its in {mixed['wall_s']*1e3:.1f} ms)")
    plain, _ = timed_iters(dev, data, precision="fp64", loop="host", linear="dense_schur", steps=args.steps,
                           warmup=args.warmup)
    log(f"[bench] plain fp64 host-loop dense: {plain['iters_per_s']:.1f} it/s")
    plain_pcg, _ = timed_iters(dev, data, precision="fp64", loop="host", linear="iterative_schur", steps=args.steps,
                               warmup=args.warmup)
    log(f"[bench] plain fp64 host-loop PCG: {plain_pcg['iters_per_s']:.1f} it/s")
    plain_dev, _ = timed_iters(dev, data, precision="fp64", loop="device", linear="dense_schur", steps=args.steps,
                               warmup=args.warmup)
    log(f"[bench] fp64 device-loop dense: {plain_dev['iters_per_s']:.1f} it/s")

    # time to tolerance against the plain FP64 final cost (reference default tolerances)
    rep64, rec64, wall64, _ = time_to_tol(dev, data, precision="fp64", loop="host", linear="dense_schur",
                                          max_iters=50, target=None)
    cams64, pts64 = dev.cameras.clone(), dev.points.clone()
    target = rep64.final_cost * (1.0 + 1e-6)
    hit64 = next((float(r[11]) for r in rec64 if r[1] <= target), None)
    repm, recm, wallm, hitm = time_to_tol(dev, data, precision="mixed", loop="device", linear="dense_schur",
                                          max_iters=50, target=target)
    camsm, ptsm = dev.cameras.clone(), dev.points.clone()
    rel_cost = abs(repm.final_cost - rep64.final_cost) / rep64.final_cost
    param_rms = float(torch.sqrt(torch.mean(torch.cat([(camsm - cams64).reshape(-1), (ptsm - pts64).reshape(-1)]) ** 2)))
    log(f"[bench] time-to-tol fp64 {hit64} s, mixed {hitm} s; final cost rel err {rel_cost:.2e}; param rms {param_rms:.2e}")

    roofline = None
    kernels = None
    if not args.no_profile:
        kernels = profile_kernels(dev, data, precision="mixed", linear="dense_schur", iters=min(args.steps, 10))
        log("[bench] mixed kernel groups (ms/launch): " + ", ".join(f"{k}={v[0]:.3f}" for k, v in kernels.items()))
        dominant = max((k for k in kernels if algorithmic_work(k, data, "mixed")[0]), key=lambda k: kernels[k][0])
        bound, work = algorithmic_work(dominant, data, "mixed")
        sec = kernels[dominant][0] * 1e-3
        if bound == "hbm":
            ach = work / sec / 1e9
            roofline = dict(kernel=dominant, bound="hbm", achieved=ach, peak=HBM_PEAK_GBS, unit="GB/s",
                            frac=ach / HBM_PEAK_GBS, traffic=None, algorithmic_bytes=work)
        else:
            ach = work / sec / 1e12
            roofline = dict(kernel=dominant, bound="mfma", achieved=ach, peak=FP64_PEAK_TFLOPS, unit="TFLOP/s",
                            frac=ach / FP64_PEAK_TFLOPS, traffic=None, algorithmic_flops=work)

    cpu = None if args.no_cpu else cpu_baseline(data)
    if cpu:
        log(f"[bench] cpu oracle: {cpu['value']:.4f} it/s")

    value = mixed["iters_per_s"]
    line = {
        "metric": "BA LM iters/s + time-to-tol (mixed vs FP64); served solves/s, p50/p99 latency",
        "value": value,
        "unit": "LM iterations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 / value,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64",
        "data": "synthetic (seeded generator, inputs resident in HBM)",
        "config": {
            "workload": f"{name}: {data.num_cameras} cameras, {data.num_points} points, {data.num_observations} "
                        "observations, 5% uniform outliers, Huber(1 px), DENSE_SCHUR",
            "mixed": "FP32 J/JtJ/Schur, FP64 Cholesky+cost+gain ratio, 2 refinement steps, device loop (graphs)",
            "plain": "FP64 everywhere, host-driven loop (one sync per iteration)",
        },
        "plain_fp64_iters_per_s": plain["iters_per_s"],
        "plain_fp64_pcg_iters_per_s": plain_pcg["iters_per_s"],
        "fp64_device_loop_iters_per_s": plain_dev["iters_per_s"],
        "mixed_vs_plain_fp64": value / plain["iters_per_s"],
        "mixed_vs_plain_fp64_pcg": value / plain_pcg["iters_per_s"],
        "time_to_tol_s": {"mixed": hitm, "fp64": hit64, "target": "cost <= (1+1e-6) x plain-FP64 final cost"},
        "mixed_final_cost_rel_err_vs_fp64": rel_cost,
        "mixed_param_rms_err_vs_fp64": param_rms,
        "kernel_ms": None if kernels is None else {k: v[0] for k, v in kernels.items()},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    dev.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
