"""Benchmark: bundle-adjustment LM iterations/s and time-to-tolerance, plain
FP64 vs mixed precision, on one MI355X.

Workload (default ``ba200``): the at-scale synthetic bundle adjustment - 200
cameras on a radius-10 circle, 100,000 points drawn in [-2,2]^3, each seen by
3..10 random cameras (~650k observations), trivial loss.  Inputs are seeded
and resident in HBM before timing starts.

A *step* is one LM iteration.  ``--warmup W`` untimed iterations run first;
then exactly ``--steps K`` LM iterations (all tolerances 0) are timed between
two device synchronisations.  Once converged, LM ends by itself when the trust
region collapses; the K iterations are then back-to-back solves from the same
seeded start.  ``value`` is the plain FP64 form with
ITERATIVE_SCHUR (explicit reduced camera system, Schur-Jacobi PCG, eta 0.1)
and the LM loop kept on the device (graph-captured iterations).  Beside it:

* plain FP64 with DENSE_SCHUR (FP64 Cholesky of the 1800x1800 reduced camera
  system), device loop and host-driven loop;
* the mixed form: FP32 Jacobian / J^T J / Schur assembly, FP64 Cholesky, FP64
  cost and gain ratio, 2 FP64 refinement steps;
* time-to-tolerance: both forms run to convergence (all tolerances 0); the
  device time until the cost falls to (1 + 1e-6) x the converged plain FP64
  final cost, from per-iteration device timestamps; final-cost and parameter
  RMS error of the mixed form vs FP64;
* the same comparison on the 16-camera Huber problem with 5% outliers.

The roofline entry covers the kernel with the largest share of an iteration,
timed with HIP events on the stream the library launches on.  The CPU
baseline times the float64 oracle (numpy) on one LM iteration of the whole
problem with the same linear solver, BLAS on the host's cores (at most 16);
beside it, the oracle's restatement of the reference's dense algebra on the
16-camera workload.
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
    """Exactly ``steps`` LM iterations (all tolerances 0) from the seeded start.

    LM stops by itself once the trust region collapses after convergence; the
    problem is then restarted from the seeded start until ``steps`` iterations
    have run.  The solve lengths are learned in an untimed pass that also
    builds every plan and graph; the timed pass repeats them.  A restart's
    initial evaluation is timed but not counted.
    """
    import torch

    import ceres_mi355x as tc
    from ceres_mi355x.ba import lm_options

    def opts_for(n):
        o = tc.SolverOptions(max_num_iterations=n, function_tolerance=0.0, gradient_tolerance=0.0,
                             parameter_tolerance=0.0, preconditioner_type=tc.PreconditionerType.SCHUR_JACOBI,
                             max_num_refinement_iterations=refine if precision == "mixed" else 0)
        return lm_options(o, precision=precision, loop=loop, linear=linear)

    run_solve(dev, data, opts_for(max(1, warmup)))
    cams0 = torch.tensor(data.cameras, device=dev.device)
    pts0 = torch.tensor(data.points, device=dev.device)
    # The timed pass must repeat the learned solve lengths.  The same option
    # set from the same start gives the same trajectory (repeated solves agree
    # bit for bit); if a pass still differs, it is learned and timed again and
    # the retry is reported.
    for attempt in range(3):
        seq, done = [], 0
        while done < steps:
            o = opts_for(steps - done)
            rep, _, _ = run_solve(dev, data, o)
            n = int(rep.num_records) - 1
            if n <= 0:
                raise RuntimeError("LM made no iteration")
            seq.append(o)
            done += n
        lin, accepted, total = [], 0, 0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for o in seq:
            dev.set_parameters(cams0, pts0)
            rep, rec = dev.run(o)
            total += int(rep.num_records) - 1
            lin.extend(rec[1:, 10].tolist())
            accepted += int(rec[1:, 9].sum())
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        if total == steps:
            break
        log(f"[bench] timed pass ran {total} iterations where the learned solves ran {steps}; learning again")
    else:
        raise RuntimeError(f"timed pass ran {total} iterations, expected {steps}")
    return dict(iters=total, wall_s=wall, iters_per_s=total / wall, solves=len(seq), accepted_steps=accepted,
                mean_linear_iterations=float(sum(lin) / max(len(lin), 1)), final_cost=rep.final_cost,
                timed_pass_retries=attempt), rep


def time_to_tol(dev, data, *, precision, loop, linear, max_iters, target):
    """Run to convergence (all tolerances 0); device timestamp of the first iteration with cost <= target."""
    import ceres_mi355x as tc
    from ceres_mi355x.ba import lm_options

    o = tc.SolverOptions(max_num_iterations=max_iters, function_tolerance=0.0, gradient_tolerance=0.0,
                         parameter_tolerance=0.0, max_num_refinement_iterations=2 if precision == "mixed" else 0,
                         preconditioner_type=tc.PreconditionerType.SCHUR_JACOBI)
    opts = lm_options(o, precision=precision, loop=loop, linear=linear)
    run_solve(dev, data, opts)
    rep, rec, wall = run_solve(dev, data, opts)
    hit = None
    if target is not None:
        for row in rec:
            if row[1] <= target:
                hit = (float(row[11]), int(row[0]))
                break
    return rep, rec, wall, hit


def profile_kernels(dev, data, *, precision, linear, iters):
    """Per-kernel-group device time (HIP events on the library's stream), host-driven loop.

    Returns ({group: (ms per launch, launches)}, mean linear-solver iterations
    per timed LM iteration).  The host-driven loop launches no kernel that
    returns early (no evaluation-only step groups, nothing after convergence),
    so a rocprofv3 kernel trace of the same run averages the same launches.
    """
    import ctypes

    import torch

    import ceres_mi355x as tc
    from ceres_mi355x import _native as nat
    from ceres_mi355x.ba import lm_options

    o = tc.SolverOptions(max_num_iterations=iters, function_tolerance=0.0, gradient_tolerance=0.0,
                         parameter_tolerance=0.0, max_num_refinement_iterations=2 if precision == "mixed" else 0,
                         preconditioner_type=tc.PreconditionerType.SCHUR_JACOBI)
    opts = lm_options(o, precision=precision, loop="host", linear=linear)
    run_solve(dev, data, opts)  # builds the plan
    handle = dev._plans[tuple(getattr(opts, f) for f, _ in nat.ct_lm_options._fields_)][0]
    dev.set_parameters(torch.tensor(data.cameras), torch.tensor(data.points))
    ms = (ctypes.c_double * len(nat.KERNELS))()
    cnt = (ctypes.c_int32 * len(nat.KERNELS))()
    nat.check(nat.load().ct_ba_plan_profile(handle, iters, ms, cnt), "ct_ba_plan_profile")
    # linear-solver iterations of the same LM iterations, from a recorded run of the same plan
    _, rec, _ = run_solve(dev, data, opts)
    lin = rec[1:, 10]
    mean_lin = float(lin.mean()) if lin.size else 0.0
    return {k: (ms[i] / max(cnt[i], 1), cnt[i]) for i, k in enumerate(nat.KERNELS)}, mean_lin


def num_items(data, max_work=224, diagonal_only=False):
    """Camera-block work items (the library splits each block's terms / observations into <= max_work pieces);
    with the implicit PCG only the diagonal blocks are assembled.  Returns (items, observation items, terms read,
    blocks)."""
    import numpy as np

    from ceres_mi355x.ba import build_tables

    t = build_tables(data.obs_camera, data.obs_point, data.num_cameras, data.num_points)
    terms = np.diff(t.block_offsets)
    cam_obs = np.diff(t.camera_offsets)
    diag = t.block_cameras[:, 0] == t.block_cameras[:, 1]
    if diagonal_only:
        terms = terms[diag]
    items = int(np.sum(-(-terms // max_work)))
    items1 = int(np.sum(-(-cam_obs[t.block_cameras[diag, 0]] // max_work)))
    return items + items1, items1, int(np.sum(terms)), int(t.block_cameras.shape[0])


def implicit_pcg(data, precision: str, linear: str) -> bool:
    """Whether the library applies S from the records (its ITERATIVE_SCHUR default): plain FP64, at most 300
    cameras, and point groups (no point seen by more than 64 cameras)."""
    return (linear == "iterative_schur" and precision == "fp64" and 0 < data.num_cameras <= 300
            and fused_evaluation(data))


def fused_evaluation(data) -> bool:
    """The library fuses the observation and point evaluation when no point has more than 64 observations."""
    import numpy as np

    return int(np.max(np.bincount(np.asarray(data.obs_point)))) <= 64 if len(data.obs_point) else False


def algorithmic_work(kernel: str, data, precision: str, *, linear_iterations=0.0, items=None, implicit=False):
    """Compulsory HBM bytes (or FP64 flops) of one launch of a kernel group.

    Every input is read once and every output written once; gathered re-reads
    are not counted.  s = bytes per assembly scalar (4 mixed, 8 FP64).
    ``implicit``: the PCG applies S from the evaluation's records (k_pcg_implicit).
    """
    C, P, N = data.num_cameras, data.num_points, data.num_observations
    s = 4 if precision == "mixed" else 8
    n = 9 * C
    if kernel == "eval_obs":
        if fused_evaluation(data):
            # one kernel: observation Jacobians and the point elimination, records written once
            reads = N * (4 + 4 + 16) + 4 * (P + 1) + C * 72 + P * 24
            writes = N * (16 + 24 * s + 6 * s + 2 * s) + P * (24 + 9 * s + 24)
            return "hbm", reads + writes
        reads = N * (4 + 4 + 16) + C * 72 + P * 24
        writes = N * (16 + 24 * s)
        return "hbm", reads + writes
    if kernel == "eval_points":
        reads = 4 * (P + 1) + N * (6 * s + 16)
        writes = P * (24 + 9 * s + 24) + N * 8 * s
        return "hbm", reads + writes
    if kernel == "camera_items":
        n_items, n_items1, T, B = items
        reads = 16 * n_items + N * (4 + 18 * s + 2 * s + 16 + 6 * s) + 8 * T
        writes = n_items * 81 * s + n_items1 * 27 * 8
        return "hbm", reads + writes
    if kernel == "camera_reduce":
        n_items, n_items1, T, B = items
        reads = n_items * 81 * s + n_items1 * 27 * 8 + 8 * (B + 1) + 16 * n_items
        # the implicit PCG needs S's diagonal blocks only
        writes = (C * 81 * 8 if implicit else n * n * 8) + 3 * n * 8 + 16 * C
        return "hbm", reads + writes
    if kernel == "rcs_solve" and implicit:
        # setup: the C diagonal 9x9 blocks of S read, their inverses written, x cleared, rhs read;
        # each CG iteration: phase 1 reads every observation's JP record (24 doubles) and indices and
        # writes its y (9 doubles); phase 2 reads y, camera_obs, the damping and writes S p
        setup = C * 81 * 8 * 2 + 4 * C + 2 * n * 8
        per_iteration = N * (24 * 8 + 4 + 4 + 9 * 8 + 4 + 9 * 8) + 4 * (P + 1) + n * 8 * 2 + 4 * (C + 1)
        return "hbm", setup + linear_iterations * per_iteration
    if kernel == "rcs_solve":
        if linear_iterations > 0:  # PCG: S read once per iteration, plus its diagonal blocks once
            return "hbm", linear_iterations * (n * n * 8 + 8 * n * 8) + n * 9 * 8 + n * 8
        return "mfma", 2.0 * n ** 3 / 3.0 + 4.0 * n * n  # FP64 flops: factor + V = L^-1, two products with V
    if kernel == "step":
        if fused_evaluation(data):
            # one kernel: cameras' and points' step, back-substitution and the candidate cost
            reads = (8 * n + 72 * C + 24 * P + 4 * (P + 1) + P * 9 * s
                     + N * (4 + 4 + 24 * s + 6 * s + 16 + 16))
            writes = 72 * C + 24 * P + 24 * P
            return "hbm", reads + writes
        reads = 8 * n + 72 * C + 24 * P + 4 * (P + 1) + N * (4 + 18 * s + 6 * s) + P * 9 * s
        writes = 72 * C + 24 * P + 24 * P
        return "hbm", reads + writes
    if kernel == "candidate":
        reads = N * (8 + 18 * s + 6 * s + 16 + 16) + 24 * P + 8 * n + 72 * C + 24 * P
        writes = 0
        return "hbm", reads + writes
    return None, None


KERNEL_SYMBOL = {"eval_obs": "k_eval_fused", "eval_points": "k_eval_points", "camera_items": "k_camera_items",
                 "camera_reduce": "k_camera_reduce", "rcs_solve": "k_pcg_implicit", "step": "k_step_candidate",
                 "candidate": "k_candidate_obs"}
COUNTER_SUMMARY = Path(__file__).resolve().parent / "profiles" / "ba200_fp64_pcg_hbm_counters.md"


def measured_traffic(group, name, precision, linear):
    """HBM bytes per dispatch of a kernel group from the committed rocprofv3 counter summary.

    The summary holds FETCH_SIZE and WRITE_SIZE per dispatch for the ba200 FP64
    ITERATIVE_SCHUR run, scaled by what the counters read for a 256 MiB copy
    in the same run.  Other workloads and forms have no summary: None.
    """
    if name != "ba200" or precision != "fp64" or linear != "iterative_schur" or not COUNTER_SUMMARY.exists():
        return None
    sym = KERNEL_SYMBOL.get(group)
    for row in COUNTER_SUMMARY.read_text().splitlines():
        cells = [c.strip() for c in row.strip("|").split("|")]
        if len(cells) >= 5 and cells[0].strip("`").split("<")[0] == sym:
            return float(cells[4])
    return None


def roofline_entry(kernels, data, precision, mean_lin, items, linear="iterative_schur", workload="ba200"):
    """Roofline of the kernel group with the largest time per launch (lm_begin/lm_end/refine excluded)."""
    cands = {k: v for k, v in kernels.items() if k not in ("lm_begin", "lm_end", "refine") and v[1] > 0}
    dominant = max(cands, key=lambda k: cands[k][0])
    bound, work = algorithmic_work(dominant, data, precision, linear_iterations=mean_lin, items=items,
                                   implicit=implicit_pcg(data, precision, linear))
    sec = kernels[dominant][0] * 1e-3
    if bound == "hbm":
        ach = work / sec / 1e9
        traffic = measured_traffic(dominant, workload, precision, linear)
        out = dict(kernel=dominant, bound="hbm", achieved=ach, peak=HBM_PEAK_GBS, unit="GB/s",
                   frac=ach / HBM_PEAK_GBS, traffic=traffic, algorithmic_bytes=work, ms_per_launch=kernels[dominant][0],
                   launches=kernels[dominant][1], linear_iterations_per_launch=mean_lin if dominant == "rcs_solve" else None)
        if traffic:
            out["traffic_source"] = ("profiles/ba200_fp64_pcg_hbm_counters.md: rocprofv3 FETCH_SIZE x2 + WRITE_SIZE x1 "
                                     "(scales from a 256 MiB copy in the same run), bytes per dispatch")
            out["achieved_on_traffic"] = traffic / sec / 1e9
        return out
    ach = work / sec / 1e12
    return dict(kernel=dominant, bound="mfma", achieved=ach, peak=FP64_PEAK_TFLOPS, unit="TFLOP/s",
                frac=ach / FP64_PEAK_TFLOPS, traffic=None, algorithmic_flops=work, ms_per_launch=kernels[dominant][0],
                launches=kernels[dominant][1])


def dense_cholesky_entry(dev, data, iters=5):
    """Matrix-core rate of the dense FP64 RCS solve (DENSE_SCHUR) on this workload.

    One RCS solve factors S = L L^T on 64x64 tiles with MFMA FP64 tile
    products, forms V = L^-1 in the same launch and solves with two products
    with V: about n^3/3 + n^3/3 + 2 n^2 flops for n = 9 x cameras.  The time is
    the HIP-event time of that launch group per LM iteration.  Its critical
    path is the sequential chain of 64-column diagonal tiles, so the rate sits
    far below the peak at this size.
    """
    kernels, _ = profile_kernels(dev, data, precision="fp64", linear="dense_schur", iters=iters)
    n = 9 * data.num_cameras
    ms = kernels["rcs_solve"][0]
    flops = 2.0 * n ** 3 / 3.0 + 4.0 * n * n
    ach = flops / (ms * 1e-3) / 1e12
    return {"n": n, "ms_per_solve": ms, "algorithmic_flops": flops, "achieved": ach, "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": ach / FP64_PEAK_TFLOPS,
            "counters": "profiles/ba200_cholesky_mfma_counters.md"}


def host_threads() -> int:
    """Threads the CPU baselines use: the host cores this process may run on, at most 16 (one GPU's share)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        n = os.cpu_count() or 1
    return max(1, min(n, int(os.environ.get("OMP_NUM_THREADS", n) or n), 16))


def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:  # pragma: no cover
        pass
    return "unknown"


def cpu_baseline(data, iterations=1):
    """Float64 oracle (numpy, ITERATIVE_SCHUR) on the whole problem, on the host's cores.

    One LM iteration of the full workload (the evaluation at the start, the
    Schur-Jacobi PCG step and the candidate evaluation), as the GPU counts
    iterations of back-to-back solves.  BLAS (the reduced camera system's
    products and factorisations) runs on ``host_threads()`` threads; the
    per-observation and per-point array arithmetic of numpy is single-threaded.
    """
    from threadpoolctl import threadpool_limits

    from oracle.ba import BAModel
    from oracle.lm import OracleOptions, trust_region_solve

    threads = host_threads()
    with threadpool_limits(limits=threads):
        model = BAModel(data, linear_solver="iterative_schur")
        opts = OracleOptions(max_num_iterations=iterations, function_tolerance=0.0, gradient_tolerance=0.0,
                             parameter_tolerance=0.0, preconditioner_type="schur_jacobi")
        t0 = time.perf_counter()
        trust_region_solve(model, opts)
        el = time.perf_counter() - t0
    return dict(value=iterations / el, unit="LM iterations/s", cores=threads, kind="port",
                sample=f"{iterations} LM iteration(s) of the float64 numpy oracle (block-sparse Schur, Schur-Jacobi "
                       f"PCG) on the whole workload ({data.num_cameras} cameras / {data.num_points} points / "
                       f"{data.num_observations} observations) in {el:.1f} s, evaluation at the start included; "
                       f"BLAS on {threads} threads of {os.cpu_count()} ({_cpu_model()}), numpy array arithmetic "
                       "single-threaded")


def cpu_dense_faithful(name="ba16", iterations=1):
    """The reference's own dense algebra, restated by the oracle, on the 16-camera workload.

    Dense Jacobian (16,000 x 6,144), H = J^T J + D, LU of the eliminated point
    block, LU of the Schur complement (linear.py:667-699), the same LM rules;
    one LM iteration on the host's cores.
    """
    from threadpoolctl import threadpool_limits

    from oracle import generators as gen
    from oracle.ba import BAModel
    from oracle.lm import OracleOptions, trust_region_solve

    data = gen.ba_config(name)
    threads = host_threads()
    with threadpool_limits(limits=threads):
        model = BAModel(data, linear_solver="dense_schur", dense_faithful=True)
        opts = OracleOptions(max_num_iterations=iterations, function_tolerance=0.0, gradient_tolerance=0.0,
                             parameter_tolerance=0.0)
        t0 = time.perf_counter()
        trust_region_solve(model, opts)
        el = time.perf_counter() - t0
    return dict(value=iterations / el, unit="LM iterations/s", cores=threads, kind="port",
                sample=f"{iterations} LM iteration(s) of the oracle's dense restatement of the reference's DENSE_SCHUR "
                       f"algebra on {name} ({data.num_cameras} cameras, {data.num_points} points, "
                       f"{data.num_observations} observations) in {el:.1f} s on {threads} BLAS threads "
                       f"({_cpu_model()})")


def similarity_aligned_rms(cams_a, pts_a, cams_b, pts_b) -> dict:
    """RMS differences of two reconstructions after the 7-dof similarity that best maps a's points onto b's.

    Bundle adjustment fixes its solution only up to rotation, translation and
    scale, so two solves of the same problem may differ by such a transform.
    The transform is the least-squares similarity of the point sets (Umeyama,
    1991); point positions and camera centres c = -R(aa)^T t are compared
    after it is applied to a.
    """
    import numpy as np
    from scipy.spatial.transform import Rotation

    A = np.asarray(pts_a, dtype=np.float64)
    B = np.asarray(pts_b, dtype=np.float64)
    ma, mb = A.mean(0), B.mean(0)
    Ac, Bc = A - ma, B - mb
    U, S, Vt = np.linalg.svd(Bc.T @ Ac / len(A))
    d = np.sign(np.linalg.det(U @ Vt))
    D = np.diag([1.0, 1.0, d])
    R = U @ D @ Vt
    scale = float(np.trace(np.diag(S) @ D) / (Ac * Ac).sum(1).mean())
    t = mb - scale * R @ ma
    pts_err = np.sqrt(np.mean(np.sum((scale * A @ R.T + t - B) ** 2, axis=1)))

    def centres(cams):
        cams = np.asarray(cams, dtype=np.float64)
        rot = Rotation.from_rotvec(cams[:, :3]).as_matrix()
        return -np.einsum("cji,cj->ci", rot, cams[:, 3:6])

    ca, cb = centres(cams_a), centres(cams_b)
    cam_err = np.sqrt(np.mean(np.sum((scale * ca @ R.T + t - cb) ** 2, axis=1)))
    return {"points_rms": float(pts_err), "camera_centres_rms": float(cam_err), "scale": scale,
            "points_rms_unaligned": float(np.sqrt(np.mean(np.sum((A - B) ** 2, axis=1))))}


def mixed_vs_fp64(dev, data, *, max_iters):
    """Time-to-tolerance and final-state errors of the mixed form vs plain FP64 (DENSE_SCHUR, device loop).

    The FP64 solve is run to convergence (all tolerances 0, it ends at the
    minimum trust-region radius); target = its final cost x (1 + 1e-6).  Both
    forms report the device time (from the start of the solve) of the first
    iteration whose cost is at or below the target.
    """
    import torch

    rep64, rec64, _, _ = time_to_tol(dev, data, precision="fp64", loop="device", linear="dense_schur",
                                     max_iters=max_iters, target=None)
    cams64, pts64 = dev.cameras.clone(), dev.points.clone()
    target = rep64.final_cost * (1.0 + 1e-6)
    hit64 = next(((float(r[11]), int(r[0])) for r in rec64 if r[1] <= target), None)
    repm, _, _, hitm = time_to_tol(dev, data, precision="mixed", loop="device", linear="dense_schur",
                                   max_iters=max_iters, target=target)
    camsm, ptsm = dev.cameras.clone(), dev.points.clone()
    diff = torch.cat([(camsm - cams64).reshape(-1), (ptsm - pts64).reshape(-1)])
    aligned = similarity_aligned_rms(camsm.cpu().numpy(), ptsm.cpu().numpy(), cams64.cpu().numpy(),
                                     pts64.cpu().numpy())
    return {
        "target": "cost <= (1+1e-6) x converged plain-FP64 final cost (all tolerances 0)",
        "fp64_s": hit64 and hit64[0], "fp64_iteration": hit64 and hit64[1],
        "mixed_s": hitm and hitm[0], "mixed_iteration": hitm and hitm[1],
        "fp64_final_cost": rep64.final_cost, "mixed_final_cost": repm.final_cost,
        "mixed_final_cost_rel_err": abs(repm.final_cost - rep64.final_cost) / rep64.final_cost,
        "mixed_param_rms_err_unaligned": float(torch.sqrt(torch.mean(diff * diff))),
        "mixed_vs_fp64_after_similarity_alignment": aligned,
        "note": "bundle adjustment fixes its solution up to a similarity; the aligned errors remove it",
    }


def bench_curve_batch(num_problems=4096, iterations=30, replays=5, cpu_problems=256):
    """Batched uneven curve fits (y = exp(m x + c), Cauchy(0.5), 30 LM iterations) in one persistent launch.

    The launch shape is autotuned; the solve (start values reset + batched LM)
    is captured in a HIP graph and replayed; time per batch is the mean over
    the replays, bracketed by device synchronisations.  Also reports the FP64
    form and a one-thread oracle sample.

    Roofline entry: a sweep reads a problem's samples once (8 B per sample in
    the mixed form's FP32 storage, 16 B in FP64); per problem there is one
    sweep per evaluation at x (num_records) and one per candidate cost
    (num_residual_evaluations), plus the mixed form's initial FP64 cost.  The
    samples of a problem stay in cache between its sweeps, so the achieved
    figure is the rate at which the sweeps consume samples, against HBM peak.
    """
    import numpy as np
    import torch

    import ceres_mi355x as tc
    from ceres_mi355x.autotune import curve_schedule
    from ceres_mi355x.ba import lm_options
    from ceres_mi355x.curve import CurveBatch
    from oracle import generators as gen

    data = gen.make_curve_batch(num_problems=num_problems, seed=3)
    batch = CurveBatch(data.offsets, data.x, data.y, loss=tc.CauchyLoss(0.5))
    sizes = np.diff(data.offsets)
    out = {"workload": f"{num_problems} curve fits y = exp(m x + c), N ~ U[50, 2000] samples "
                       f"({int(data.offsets[-1])} in total), Cauchy(0.5), {iterations} LM iterations max, "
                       "reference default tolerances, start m = c = 0"}
    final = {}
    for precision in ("mixed", "fp64"):
        opts = lm_options(tc.SolverOptions(max_num_iterations=iterations), precision=precision)
        res = batch.solve(opts, records=False)  # tunes the launch shape on first use
        final[precision] = res.reports["final_cost"].copy()
        (threads, groups, waves), order = curve_schedule(batch, opts)  # the cached choice
        start = torch.zeros(num_problems, 2, dtype=torch.float64, device="cuda")
        params = start.clone()
        graph = torch.cuda.CUDAGraph()
        stream = torch.cuda.Stream()
        shape = (threads, groups, waves)
        with torch.cuda.stream(stream):
            batch.launch(opts, params, schedule=shape, order=order)
            torch.cuda.synchronize()
            with torch.cuda.graph(graph, stream=stream):
                params.zero_()
                params.add_(start)
                batch.launch(opts, params, schedule=shape, order=order)
        graph.replay()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(replays):
            graph.replay()
        torch.cuda.synchronize()
        sec = (time.perf_counter() - t0) / replays
        iters = int(np.sum(res.reports["num_records"]) - num_problems)
        conv = float(np.mean(res.reports["termination"] == 0))
        sweeps = res.reports["num_records"].astype(np.int64) + res.reports["num_residual_evaluations"]
        if precision == "mixed":
            sweeps = sweeps + 1
        sample_bytes = 8 if precision == "mixed" else 16
        alg = float(np.sum(sweeps * sizes)) * sample_bytes
        out[precision] = {"seconds_per_batch": sec, "solves_per_s": num_problems / sec,
                          "lm_iterations_per_s": iters / sec, "lm_iterations": iters, "converged_fraction": conv,
                          "schedule": {"threads": threads, "groups": groups, "wave_problems": waves, "order": order},
                          "median_abs_err_m_vs_truth": float(np.median(np.abs(params[:, 0].cpu().numpy()
                                                                              - data.true_m))),
                          "roofline": {"bound": "hbm", "achieved": alg / sec / 1e9, "peak": HBM_PEAK_GBS,
                                       "unit": "GB/s", "frac": alg / sec / 1e9 / HBM_PEAK_GBS, "traffic": None,
                                       "algorithmic_bytes": alg, "sample_sweeps": float(np.sum(sweeps * sizes)),
                                       "bytes_per_sample_sweep": sample_bytes}}
        log(f"[bench] curve batch {precision}: {sec * 1e3:.3f} ms/batch, {num_problems / sec:.0f} solves/s, "
            f"schedule {threads}x{groups} waves {waves} {order}")
    # speed of the mixed form over the FP64 form, and how far its final costs (of the FP32-rounded samples,
    # evaluated in FP64) are from the FP64 form's
    out["mixed_vs_fp64"] = out["fp64"]["seconds_per_batch"] / out["mixed"]["seconds_per_batch"]
    out["mixed_final_cost_max_rel_diff"] = float(np.max(np.abs(final["mixed"] - final["fp64"])
                                                        / np.maximum(np.abs(final["fp64"]), 1e-300)))
    # one-thread oracle on the first problems of the batch
    from threadpoolctl import threadpool_limits

    from oracle.lm import OracleOptions, trust_region_solve
    from oracle.small_models import CurveModel

    with threadpool_limits(limits=1):
        t0 = time.perf_counter()
        for i in range(cpu_problems):
            lo, hi = data.offsets[i], data.offsets[i + 1]
            trust_region_solve(CurveModel(data.x[lo:hi], data.y[lo:hi]), OracleOptions(max_num_iterations=iterations))
        el = time.perf_counter() - t0
    frac = float(data.offsets[cpu_problems]) / float(data.offsets[-1])
    out["cpu_baseline"] = {"value": num_problems * frac / el, "unit": "solves/s",
                           "cores": 1, "kind": "port",
                           "sample": f"float64 numpy oracle (DENSE_QR step) on the first {cpu_problems} problems "
                                     f"({int(data.offsets[cpu_problems])} samples) in {el:.2f} s; scaled to the "
                                     "batch's sample mix"}
    return out


def bench_ba_batch(num_problems=101, seed=4, repeats=3, plain_sample=30):
    """Batched small bundle adjustments (the serving mix's share: 2-20 cameras, 100-5000 points, 2-6
    observations per point, reference default options, DENSE_SCHUR) in one persistent launch
    (ct_ba_batch_solve), timed with HIP events on the launching stream; the host table build
    (ct_ba_batch_tables + uploads) is reported beside it.  The plain per-problem GPU baseline solves the first
    `plain_sample` problems one after another through per-problem plans (host-driven loop, as a caller without
    batching would)."""
    import numpy as np
    import torch

    import ceres_mi355x as tc
    from ceres_mi355x.ba import BADevice, lm_options
    from ceres_mi355x.ba_batch import BABatch, BAProblemInput
    from oracle import generators as gen

    rng = np.random.default_rng(seed)
    datas = []
    for _ in range(num_problems):
        C = int(rng.integers(2, 21))
        P = int(rng.integers(100, 5001))
        datas.append(gen.make_bundle_adjustment(num_cameras=C, num_points=P, obs_per_point=(2, min(6, C)),
                                                seed=int(rng.integers(1 << 30))))
    inputs = [BAProblemInput(d.cameras, d.points, d.obs_camera, d.obs_point, d.observations) for d in datas]
    opts = lm_options(tc.SolverOptions())
    BABatch(inputs[:2]).solve(opts)  # load the code object
    builds = []
    for _ in range(repeats):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        batch = BABatch(inputs)
        torch.cuda.synchronize()
        builds.append(time.perf_counter() - t0)
    s = torch.cuda.current_stream()
    results = torch.zeros(num_problems * 64, dtype=torch.uint8, device="cuda")
    times = []
    for _ in range(repeats):
        cams, pts = batch.cameras.clone(), batch.points.clone()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        batch.launch(opts, cams, pts, results)
        e1.record(s)
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1) * 1e-3)
    reps, _, _, _ = batch.solve(opts)
    iters = int(sum(int(r["num_records"]) - 1 for r in reps))
    dev_s = np.array([float(r["device_seconds"]) for r in reps])
    launch_s = min(times)
    plain = 0.0
    for d in datas[:plain_sample]:
        dev = BADevice(torch.as_tensor(d.cameras, dtype=torch.float64), torch.as_tensor(d.points, dtype=torch.float64),
                       d.obs_camera, d.obs_point, d.observations)
        dev.run(lm_options(tc.SolverOptions(), loop="host"))  # plan creation and first solve untimed
        dev.set_parameters(torch.as_tensor(d.cameras, dtype=torch.float64),
                           torch.as_tensor(d.points, dtype=torch.float64))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        dev.run(lm_options(tc.SolverOptions(), loop="host"))
        torch.cuda.synchronize()
        plain += time.perf_counter() - t0
        dev.close()
    plain_rate = plain_sample / plain
    out = {
        "workload": f"{num_problems} bundle adjustments (2-20 cameras, 100-5000 points, 2-6 obs/point; "
                    f"{sum(d.num_observations for d in datas)} observations in total), reference default options, "
                    "DENSE_SCHUR, FP64, one persistent launch",
        "launch_ms": launch_s * 1e3,
        "solves_per_s": num_problems / launch_s,
        "lm_iterations": iters,
        "lm_iterations_per_s": iters / launch_s,
        "host_table_build_ms": min(builds) * 1e3,
        "per_problem_device_ms": {"median": float(np.median(dev_s) * 1e3), "max": float(dev_s.max() * 1e3)},
        "plain_per_problem_plans_solves_per_s": plain_rate,
        "plain_sample": f"first {plain_sample} problems, per-problem plans (host loop), plan creation untimed",
        "batched_vs_plain": (num_problems / launch_s) / plain_rate,
    }
    log(f"[bench] batched BA: {out['solves_per_s']:.0f} solves/s in one launch ({out['launch_ms']:.1f} ms, "
        f"host tables {out['host_table_build_ms']:.1f} ms); per-problem plans {plain_rate:.0f} solves/s")
    return out


def bench_block_schur(rows=409600, eliminate=512, remain=128, repeats=3):
    """The reference's dense Schur benchmark shape through the registered ``block_schur`` backend.

    Reference: ``benchmarking.py:131-172`` (``schur_benchmark``): A [409600 x 640]
    Gaussian, x_true Gaussian, b = A x_true, ``solve_linear_system(A, b,
    DENSE_SCHUR, num_eliminate=512)`` and the residual norm |A x - b|, one
    warmup and the median of three runs; published 0.5114 s on CUDA
    (``docs/PERFORMANCE_SUMMARY.md:61,73``).  Same shape and distribution here,
    drawn on the device from a seeded generator.  The solve forms H = A^T A and
    A^T b from row chunks of A on the matrix cores (k_gram_partials), reduces
    the chunks in a fixed order and factors H (n = 640) with the tiled FP64
    Cholesky; FP64 throughout.  The Gram flops are m n (n + 1) + 2 m n (lower
    triangle of H and A^T b); the solve-only time bounds the Gram kernel's rate
    from below (the kernel alone is in the rocprofv3 summary).
    """
    import ctypes

    import torch

    from ceres_mi355x import _native as nat
    from ceres_mi355x import linear as L

    cols = eliminate + remain
    gen_ = torch.Generator(device="cuda").manual_seed(11)
    A = torch.randn((rows, cols), generator=gen_, dtype=torch.float64, device="cuda")
    x_true = torch.randn((cols,), generator=gen_, dtype=torch.float64, device="cuda")
    b = A @ x_true
    L.register_rocm_backends()

    def run():
        res = L.call_optional_backend("block_schur", A, b, num_eliminate=eliminate)
        return res, torch.linalg.norm(A @ res.x - b)

    def timed(fn):
        torch.cuda.synchronize()
        t = time.perf_counter()
        out = fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t, out

    timed(run)  # warmup
    full = sorted(timed(run)[0] for _ in range(repeats))
    # solve alone on the device (HIP events on the stream the library launches on)
    lib = nat.load()
    nbytes = ctypes.c_size_t(0)
    nat.check(lib.ct_dense_workspace_bytes(cols, ctypes.byref(nbytes)), "ct_dense_workspace_bytes")
    ws = torch.empty(int(nbytes.value), dtype=torch.uint8, device="cuda")
    x = torch.empty(cols, dtype=torch.float64, device="cuda")
    solve_ms = []
    for _ in range(repeats + 1):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        rc = lib.ct_block_schur_solve_f64(nat.ptr(A), rows, cols, cols, nat.ptr(b), eliminate, None, nat.ptr(x),
                                          nat.ptr(ws), ws.numel(), nat.current_stream_handle())
        e1.record()
        nat.check(rc, "ct_block_schur_solve_f64")
        torch.cuda.synchronize()
        solve_ms.append(e0.elapsed_time(e1))
    solve_ms = sorted(solve_ms[1:])
    res, resid = run()
    flops = rows * cols * (cols + 1) + 2.0 * rows * cols
    med_solve = solve_ms[len(solve_ms) // 2]
    ach = flops / (med_solve * 1e-3) / 1e12
    out = {
        "workload": f"schur/dense/{rows}x{cols}/elim{eliminate} (A Gaussian, b = A x_true, FP64)",
        "median_s": full[len(full) // 2], "min_s": full[0], "max_s": full[-1], "repeats": repeats,
        "timed": "registered block_schur backend + residual norm |A x - b|, as the reference's run()",
        "reference_cuda_median_s": 0.5113517,
        "speedup_vs_reference": 0.5113517 / full[len(full) // 2],
        "residual_norm": float(resid), "solution_max_abs_error": float((res.x - x_true).abs().max()),
        "solve_only_ms": med_solve, "gram_flops": flops,
        "gram_rate_lower_bound": {"achieved": ach, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                                  "frac": ach / FP64_PEAK_TFLOPS},
    }
    del A, b
    torch.cuda.empty_cache()
    return out


def bench_serving(num_requests=300, rate=40.0, seed=4, slots=8, saturated_slots=(8, 16), baseline_sample=60):
    """Served request stream of 60% curve fits (as the batched workload, Cauchy(0.5)), 30% small bundle
    adjustments (2-20 cameras, 100-5000 points, 2-6 observations per point) and 10% SE2 pose graphs (100-2000
    nodes); reference default solver options.

    Two runs of the same requests: Poisson arrivals at `rate`/s (latency at a light load), and all requests
    arriving at once (saturation: the server's sustained solves/s, with p50/p99 latency at that load), for each
    slot count in `saturated_slots`.  Latency is arrival to result.  The plain per-problem GPU throughput solves a
    sample of the mix alone, one request after another; `served_vs_plain` is the best saturated rate over it."""
    import numpy as np

    import ceres_mi355x as tc
    from ceres_mi355x.serving import BARequest, CurveRequest, PoseRequest, SolveServer
    from oracle import generators as gen

    rng = np.random.default_rng(seed)
    kinds = rng.choice(3, size=num_requests, p=[0.6, 0.3, 0.1])
    reqs = []
    for k in kinds:
        s = int(rng.integers(1 << 30))
        if k == 0:
            cb = gen.make_curve_batch(num_problems=1, seed=s)
            reqs.append(CurveRequest(cb.x, cb.y, loss=tc.CauchyLoss(0.5)))
        elif k == 1:
            C = int(rng.integers(2, 21))
            P = int(rng.integers(100, 5001))
            d = gen.make_bundle_adjustment(num_cameras=C, num_points=P, obs_per_point=(2, min(6, C)), seed=s)
            reqs.append(BARequest(d.cameras, d.points, d.obs_camera, d.obs_point, d.observations))
        else:
            d = gen.make_pose_graph(num_nodes=int(rng.integers(100, 2001)), seed=s)
            reqs.append(PoseRequest(d.poses, d.edges, d.measurements))
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=num_requests))

    def run(server, at):
        t0 = time.perf_counter()
        futures = []
        for r, a in zip(reqs, at):
            delay = t0 + a - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            futures.append(server.submit(r))
        results = [f.result() for f in futures]
        wall = max(r.finished for r in results) - t0
        lat = np.array([r.latency for r in results])
        by_kind = {}
        for kind in ("curve", "ba", "pose2d"):
            lk = np.array([r.latency for r in results if r.kind == kind])
            if lk.size:
                by_kind[kind] = {"count": int(lk.size), "p50_ms": float(np.percentile(lk, 50) * 1e3),
                                 "p99_ms": float(np.percentile(lk, 99) * 1e3)}
        return {"solves_per_s": num_requests / wall, "p50_ms": float(np.percentile(lat, 50) * 1e3),
                "p99_ms": float(np.percentile(lat, 99) * 1e3), "by_kind": by_kind}

    def warm(server):
        # compile paths, caching allocator: the served paths and the plain per-problem paths
        for r in (reqs[int(np.argmax(kinds == 0))], reqs[int(np.argmax(kinds == 1))], reqs[int(np.argmax(kinds == 2))]):
            server.solve_alone(r)
            server.solve_plain(r)

    server = SolveServer(tc.SolverOptions(), num_slots=slots)
    warm(server)
    light = run(server, arrivals)
    # plain per-problem GPU throughput: a sample of the same mix, each request alone, one after another
    sample = reqs[:baseline_sample]
    t1 = time.perf_counter()
    for r in sample:
        server.solve_plain(r)
    alone_s = (time.perf_counter() - t1) / len(sample)
    server.close()
    saturated = {}
    for ns in saturated_slots:
        server = SolveServer(tc.SolverOptions(), num_slots=ns)
        warm(server)
        saturated[f"{ns}_slots"] = run(server, np.zeros(num_requests))
        server.close()
    best = max(saturated, key=lambda k: saturated[k]["solves_per_s"])
    plain = 1.0 / alone_s
    out = {
        "workload": f"{num_requests} requests: 60% curve fits (N ~ U[50,2000], Cauchy(0.5)), 30% bundle adjustments "
                    "(2-20 cameras, 100-5000 points, 2-6 obs/point), 10% SE2 pose graphs (100-2000 nodes); reference "
                    "default options; continuous batching of curve fits and of bundle adjustments (one persistent launch "
                    "per batch), pose graphs on stream slots",
        "light_load": {"offered_rate_per_s": rate, "slots": slots, **light},
        "saturated": saturated,
        "solves_per_s": saturated[best]["solves_per_s"],
        "p50_ms": saturated[best]["p50_ms"],
        "p99_ms": saturated[best]["p99_ms"],
        "saturated_config": best,
        "plain_per_problem_solves_per_s": plain,
        "plain_sample": f"first {len(sample)} requests of the mix, solved alone one after another",
        "served_vs_plain": saturated[best]["solves_per_s"] / plain,
    }
    log(f"[bench] serving light load ({rate:g}/s): {light['solves_per_s']:.1f} solves/s, p50 {light['p50_ms']:.1f} ms, "
        f"p99 {light['p99_ms']:.1f} ms; saturated: "
        + ", ".join(f"{k} {v['solves_per_s']:.1f}/s p50 {v['p50_ms']:.0f} ms p99 {v['p99_ms']:.0f} ms"
                    for k, v in saturated.items())
        + f"; plain {plain:.1f} solves/s")
    return out


def profile_pose_graph(num_nodes: int, solves: int = 3) -> dict:
    """Per-solve time of one SE2 pose graph of the serving mix's largest size (seeded; LM with the reference
    example's default options, host-driven loop so every kernel launches), and the size of the system the plan
    factors densely per iteration."""
    import torch

    import ceres_mi355x as tc
    from ceres_mi355x.ba import lm_options
    from ceres_mi355x.pose2d import PoseGraphDevice
    from oracle import generators as gen

    data = gen.make_pose_graph(num_nodes=num_nodes, seed=7)
    dev = PoseGraphDevice(data.poses, data.edges, data.measurements)
    opts = lm_options(tc.SolverOptions(), precision="fp64", loop="host")
    out = {"workload": f"SE2 pose graph, {num_nodes} nodes, {len(data.edges)} edges (chain plus "
                       f"{len(data.edges) - num_nodes + 1} random loop closures)",
           "unknowns": 3 * (num_nodes - 1), "factored_unknowns": dev.system_size(opts)}
    dev.run(opts)  # plan, first touch
    times, iters = [], []
    for _ in range(solves):
        dev.set_poses(data.poses)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rep, _ = dev.run(opts)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        iters.append(rep.num_records - 1)
    out.update(ms_per_solve=1e3 * min(times), lm_iterations=iters[-1], final_cost=rep.final_cost)
    dev.close()
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="ba200", choices=["ba200", "ba16", "pose2000", "ba_batch", "block_schur"],
                    help="pose2000: time and profile one 2000-node SE2 pose graph solve; ba_batch: the serving mix's "
                         "101 bundle adjustments in one batched launch; block_schur: the dense Schur solve of a "
                         "409600 x 640 system (no JSON line)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--only-value", action="store_true", help="time the value path only")
    ap.add_argument("--profile-only", action="store_true",
                    help="only the per-kernel profile of the value path (the run a rocprofv3 trace is taken of)")
    ap.add_argument("--profile-linear", default="iterative_schur", choices=["iterative_schur", "dense_schur"],
                    help="linear solver of the profiled path")
    ap.add_argument("--profile-precision", default="fp64", choices=["fp64", "mixed"])
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (sharding the points across GPUs "
              "with an all-reduce of the reduced camera system is a design note, not a built path)",
              file=sys.stderr)
        return 2

    import torch

    from oracle import generators as gen

    assert torch.cuda.is_available(), "bench.py needs a ROCm GPU"
    name = args.workload
    if name == "pose2000":
        log(f"[bench] pose graph: {profile_pose_graph(2000)}")
        return 0
    if name == "ba_batch":
        log(f"[bench] batched bundle adjustment: {bench_ba_batch()}")
        return 0
    if name == "block_schur":
        log(f"[bench] dense block Schur: {bench_block_schur()}")
        return 0
    t0 = time.perf_counter()
    data = gen.ba_config(name)
    dev = make_device(data)
    log(f"[bench] {name}: C={data.num_cameras} P={data.num_points} N={data.num_observations} "
        f"setup {time.perf_counter() - t0:.1f}s")

    runs = {}

    def timed(label, **kw):
        r, _ = timed_iters(dev, data, steps=args.steps, warmup=args.warmup, **kw)
        runs[label] = r
        log(f"[bench] {label}: {r['iters_per_s']:.1f} it/s ({r['iters']} its, {r['solves']} solves, "
            f"{r['accepted_steps']} accepted, {r['mean_linear_iterations']:.2f} linear its/it, "
            f"{r['wall_s'] * 1e3:.1f} ms)")
        return r

    extra = {}
    if not args.profile_only:
        value_run = timed("fp64_iterative_schur_device_loop", precision="fp64", loop="device",
                          linear="iterative_schur")
    if not (args.only_value or args.profile_only):
        timed("fp64_iterative_schur_host_loop", precision="fp64", loop="host", linear="iterative_schur")
        timed("fp64_dense_schur_device_loop", precision="fp64", loop="device", linear="dense_schur")
        timed("fp64_dense_schur_host_loop", precision="fp64", loop="host", linear="dense_schur")
        timed("mixed_dense_schur_device_loop", precision="mixed", loop="device", linear="dense_schur")
        extra["mixed_vs_fp64_dense_iters_per_s"] = (runs["mixed_dense_schur_device_loop"]["iters_per_s"] /
                                                    runs["fp64_dense_schur_device_loop"]["iters_per_s"])
        extra["time_to_tol"] = mixed_vs_fp64(dev, data, max_iters=500)
        log(f"[bench] time-to-tol {extra['time_to_tol']}")
        small = gen.ba_config("ba16_huber")
        sdev = make_device(small)
        extra["time_to_tol_ba16_huber"] = mixed_vs_fp64(sdev, small, max_iters=2000)
        log(f"[bench] ba16_huber time-to-tol {extra['time_to_tol_ba16_huber']}")
        sdev.close()
        extra["curve_batch"] = bench_curve_batch()
        extra["ba_batch"] = bench_ba_batch()
        extra["serving"] = bench_serving()
        extra["block_schur"] = bench_block_schur()
        log(f"[bench] dense block Schur: {extra['block_schur']}")

    roofline = None
    kernels = None
    per_kernel = None
    if not args.no_profile:
        prec, lin_solver = args.profile_precision, args.profile_linear
        implicit = implicit_pcg(data, prec, lin_solver)
        items = num_items(data, diagonal_only=implicit)
        kernels, mean_lin = profile_kernels(dev, data, precision=prec, linear=lin_solver, iters=min(args.steps, 10))
        if lin_solver != "iterative_schur":
            mean_lin = 0.0
        log(f"[bench] {prec} {lin_solver} kernel groups (ms/launch): "
            + ", ".join(f"{k}={v[0]:.4f}" for k, v in kernels.items()) + f"; mean linear iterations {mean_lin:.2f}")
        roofline = roofline_entry(kernels, data, prec, mean_lin, items, linear=lin_solver, workload=name)
        per_kernel = {}
        for k, (ms, cnt) in kernels.items():
            bound, work = algorithmic_work(k, data, prec, linear_iterations=mean_lin, items=items, implicit=implicit)
            if bound == "hbm" and cnt:
                per_kernel[k] = {"ms": ms, "GB/s": work / (ms * 1e-3) / 1e9, "algorithmic_bytes": work}
        if not (args.only_value or args.profile_only):
            extra["dense_cholesky"] = dense_cholesky_entry(dev, data)
            log(f"[bench] dense RCS factorisation + inverse + solve: {extra['dense_cholesky']}")
        if args.profile_only:
            # calibration for hardware byte counters: an elementwise copy of exactly 256 MiB read + 256 MiB written
            src = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda")
            dst = torch.empty_like(src)
            torch.mul(src, 1.0, out=dst)
            torch.cuda.synchronize()
            print(json.dumps({"roofline": roofline, "kernel_ms": {k: v[0] for k, v in kernels.items()},
                              "per_kernel_hbm": per_kernel, "calibration_copy_bytes": 256 * 1024 * 1024}),
                  flush=True)
            dev.close()
            return 0

    cpu = None if args.no_cpu else cpu_baseline(data)
    if cpu:
        log(f"[bench] cpu oracle: {cpu['value']:.4f} it/s ({cpu['sample']})")
        if not (args.only_value or args.profile_only):
            extra["cpu_dense_faithful_ba16"] = cpu_dense_faithful("ba16")
            log(f"[bench] cpu dense-faithful ba16: {extra['cpu_dense_faithful_ba16']}")

    value = value_run["iters_per_s"]
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
        "dtype": "f64",
        "data": "synthetic (seeded generator, inputs resident in HBM)",
        "config": {
            "workload": f"{name}: {data.num_cameras} cameras, {data.num_points} points, "
                        f"{data.num_observations} observations, trivial loss, LM with ITERATIVE_SCHUR "
                        "(Schur-Jacobi PCG, eta 0.1), all tolerances 0; K iterations = back-to-back solves from "
                        "the seeded start, restarted when LM terminates",
        },
        "iters_per_s": {k: v["iters_per_s"] for k, v in runs.items()},
        "runs": runs,
        **extra,
        "kernel_ms": None if kernels is None else {k: v[0] for k, v in kernels.items()},
        "per_kernel_hbm": per_kernel,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    dev.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
