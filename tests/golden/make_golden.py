"""Write golden LM trajectories by running the reference solver (build container only).

The reference's per-iteration trust-region trajectory (costs, gain ratios,
radii, accept/reject decisions, final parameters) on bundle-adjustment,
curve-fit and pose-graph problems is pinned by no known-answer test, closed
form or fixture that the reference ships, so this script runs the reference
itself (``samionb/ceres-torch-cuda``, imported from ``/root/reference/src``)
on inputs made by ``oracle/generators.py`` and records its outputs as JSON
fixtures next to this file.  The fixtures are data only; the reference is
never needed to read them.

Usage (from the repository root, in a container that holds /root/reference):

    python tests/golden/make_golden.py [case ...]
"""
from __future__ import annotations

import hashlib
import json
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
REFERENCE_SRC = Path("/root/reference/src")
OUT = Path(__file__).resolve().parent


def _import_reference():
    sys.path.insert(0, str(REFERENCE_SRC))
    import torch
    import ceres_torch as tc

    return torch, tc


def array_digest(*arrays: np.ndarray) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def _iterations(summary):
    out = []
    for it in summary.iterations:
        out.append(
            dict(
                iteration=it.iteration,
                cost=it.cost,
                cost_change=it.cost_change,
                relative_decrease=it.relative_decrease,
                trust_region_radius=it.trust_region_radius,
                step_is_valid=bool(it.step_is_valid),
                step_is_successful=bool(it.step_is_successful),
                step_norm=it.step_norm,
                gradient_max_norm=it.gradient_max_norm,
                gradient_norm=it.gradient_norm,
                linear_solver_iterations=it.linear_solver_iterations,
            )
        )
    return out


def _summary_dict(summary):
    return dict(
        initial_cost=summary.initial_cost,
        final_cost=summary.final_cost,
        termination_type=summary.termination_type.value,
        message=summary.message,
        num_successful_steps=summary.num_successful_steps,
        num_unsuccessful_steps=summary.num_unsuccessful_steps,
        num_residual_evaluations=summary.num_residual_evaluations,
        num_jacobian_evaluations=summary.num_jacobian_evaluations,
        num_linear_solves=summary.num_linear_solves,
        iterations=_iterations(summary),
    )


def _loss(tc, loss):
    if loss is None or loss[0] == "trivial":
        return None
    if loss[0] == "huber":
        return tc.HuberLoss(loss[1])
    if loss[0] == "cauchy":
        return tc.CauchyLoss(loss[1])
    raise ValueError(loss)


def run_ba(name: str, config: str, solver: str, tolerances: str, max_iters: int, store_inputs: bool,
           strategy: str = "levenberg_marquardt", dogleg_type: str = "traditional", constant_cameras=(),
           initial_radius: float | None = None, preconditioner: str = "schur_jacobi"):
    torch, tc = _import_reference()
    from oracle.generators import ba_config

    d = ba_config(config)
    cams = torch.tensor(d.cameras, dtype=torch.float64)
    pts = torch.tensor(d.points, dtype=torch.float64)
    obs = torch.tensor(d.observations, dtype=torch.float64)

    def snavely(camera, point):
        p = tc.angle_axis_rotate_point(camera[:3], point) + camera[3:6]
        xp = -p[0] / p[2]
        yp = -p[1] / p[2]
        r2 = xp * xp + yp * yp
        distortion = 1.0 + r2 * (camera[7] + camera[8] * r2)
        return camera[6] * distortion * torch.stack([xp, yp])

    problem = tc.Problem()
    cam_blocks = [cams[i] for i in range(cams.shape[0])]
    pt_blocks = [pts[i] for i in range(pts.shape[0])]
    for c in cam_blocks:
        problem.AddParameterBlock(c)
    for i in constant_cameras:
        problem.SetParameterBlockConstant(cam_blocks[i])
    for p in pt_blocks:
        problem.AddParameterBlock(p)
        problem.SetParameterBlockOrderingGroup(p, 0)
    loss = d.loss
    for k in range(obs.shape[0]):
        o = obs[k]
        problem.AddResidualBlock(
            tc.AutoDiffCostFunction(lambda camera, point, o=o: snavely(camera, point) - o, [9, 3], 2),
            _loss(tc, loss),
            [cam_blocks[int(d.obs_camera[k])], pt_blocks[int(d.obs_point[k])]],
        )
    kwargs = dict(max_num_iterations=max_iters)
    if initial_radius is not None:
        kwargs["initial_trust_region_radius"] = initial_radius
    if strategy == "dogleg":
        kwargs["trust_region_strategy_type"] = tc.TrustRegionStrategyType.DOGLEG
        kwargs["dogleg_type"] = (tc.DoglegType.SUBSPACE_DOGLEG if dogleg_type == "subspace"
                                 else tc.DoglegType.TRADITIONAL_DOGLEG)
    if tolerances == "zero":
        kwargs.update(function_tolerance=0.0, gradient_tolerance=0.0, parameter_tolerance=0.0)
    if solver == "dense_schur":
        kwargs["linear_solver_type"] = tc.LinearSolverType.DENSE_SCHUR
    else:
        kwargs["linear_solver_type"] = tc.LinearSolverType.ITERATIVE_SCHUR
        kwargs["preconditioner_type"] = tc.PreconditionerType[preconditioner.upper()]
    t0 = time.perf_counter()
    summary = tc.solve(tc.SolverOptions(**kwargs), problem)
    elapsed = time.perf_counter() - t0
    out = dict(
        case=name,
        kind="ba",
        config=config,
        linear_solver=solver,
        tolerances=tolerances,
        max_num_iterations=max_iters,
        loss=list(loss) if loss else None,
        trust_region_strategy=strategy,
        dogleg_type=dogleg_type,
        constant_cameras=list(constant_cameras),
        initial_trust_region_radius=initial_radius,
        preconditioner=preconditioner if solver == "iterative_schur" else None,
        input_sha256=array_digest(d.cameras, d.points, d.observations, d.obs_camera, d.obs_point),
        reference_seconds=elapsed,
        summary=_summary_dict(summary),
        final_cameras=cams.numpy().tolist() if store_inputs else None,
        final_points=pts.numpy().tolist() if store_inputs else None,
    )
    if not store_inputs:
        np.savez_compressed(OUT / f"{name}_final.npz", cameras=cams.numpy(), points=pts.numpy())
    return out


def run_curve(name: str, seed: int, indices: list[int], max_iters: int):
    torch, tc = _import_reference()
    from oracle.generators import make_curve_batch

    batch = make_curve_batch(num_problems=64, seed=seed)
    cases = []
    for idx in indices:
        lo, hi = int(batch.offsets[idx]), int(batch.offsets[idx + 1])
        xs = torch.tensor(batch.x[lo:hi], dtype=torch.float64)
        ys = torch.tensor(batch.y[lo:hi], dtype=torch.float64)
        m = torch.tensor([0.0], dtype=torch.float64)
        c = torch.tensor([0.0], dtype=torch.float64)
        problem = tc.Problem()
        for x, y in zip(xs, ys):
            problem.add_residual_block(
                tc.AutoDiffCostFunction(lambda m, c, x=x, y=y: y - torch.exp(m[0] * x + c[0]), [1, 1]),
                tc.CauchyLoss(0.5),
                [m, c],
            )
        t0 = time.perf_counter()
        summary = tc.solve(tc.SolverOptions(max_num_iterations=max_iters), problem)
        cases.append(
            dict(
                index=idx,
                n=hi - lo,
                reference_seconds=time.perf_counter() - t0,
                summary=_summary_dict(summary),
                final=[float(m.item()), float(c.item())],
            )
        )
    return dict(
        case=name,
        kind="curve",
        seed=seed,
        num_problems=64,
        max_num_iterations=max_iters,
        loss=["cauchy", 0.5],
        input_sha256=array_digest(batch.offsets, batch.x, batch.y),
        problems=cases,
    )


def run_pose(name: str, num_nodes: int, seed: int, max_iters: int):
    torch, tc = _import_reference()
    from oracle.generators import make_pose_graph

    g = make_pose_graph(num_nodes=num_nodes, seed=seed)

    def normalize_angle(theta):
        return torch.atan2(torch.sin(theta), torch.cos(theta))

    def relative_pose(a, b):
        ca = torch.cos(a[2])
        sa = torch.sin(a[2])
        delta = b[:2] - a[:2]
        local = torch.stack([ca * delta[0] + sa * delta[1], -sa * delta[0] + ca * delta[1]])
        return torch.cat([local, normalize_angle(b[2] - a[2]).reshape(1)])

    poses = [torch.tensor(p, dtype=torch.float64) for p in g.poses]
    problem = tc.Problem()
    for p in poses:
        problem.AddParameterBlock(p)
    problem.SetParameterBlockConstant(poses[0])
    for (a, b), meas in zip(g.edges, g.measurements):
        mt = torch.tensor(meas, dtype=torch.float64)
        problem.AddResidualBlock(
            tc.AutoDiffCostFunction(lambda pa, pb, m=mt: relative_pose(pa, pb) - m, [3, 3], 3),
            None,
            [poses[int(a)], poses[int(b)]],
        )
    t0 = time.perf_counter()
    summary = tc.solve(
        tc.SolverOptions(max_num_iterations=max_iters, linear_solver_type=tc.LinearSolverType.DENSE_QR), problem
    )
    return dict(
        case=name,
        kind="pose2d",
        num_nodes=num_nodes,
        seed=seed,
        max_num_iterations=max_iters,
        input_sha256=array_digest(g.poses, g.edges, g.measurements),
        reference_seconds=time.perf_counter() - t0,
        summary=_summary_dict(summary),
        final_poses=torch.stack(poses).numpy().tolist(),
    )


CASES = {
    "ba4_dense_schur_zero_tol": lambda: run_ba("ba4_dense_schur_zero_tol", "ba4", "dense_schur", "zero", 25, True),
    "ba4_dense_schur_default_tol": lambda: run_ba("ba4_dense_schur_default_tol", "ba4", "dense_schur", "default", 50, True),
    "ba4_iterative_schur_zero_tol": lambda: run_ba("ba4_iterative_schur_zero_tol", "ba4", "iterative_schur", "zero", 25, True),
    "ba6_huber_dense_schur": lambda: run_ba("ba6_huber_dense_schur", "ba6_huber", "dense_schur", "default", 50, True),
    "curve_cauchy": lambda: run_curve("curve_cauchy", 3, [0, 5, 11, 17, 23, 31], 30),
    "pose2d_40": lambda: run_pose("pose2d_40", 40, 4, 50),
    "ba4_iterative_jacobi_zero_tol": lambda: run_ba("ba4_iterative_jacobi_zero_tol", "ba4", "iterative_schur", "zero",
                                                    25, True, preconditioner="jacobi"),
    "ba4_iterative_identity_zero_tol": lambda: run_ba("ba4_iterative_identity_zero_tol", "ba4", "iterative_schur",
                                                      "zero", 25, True, preconditioner="identity"),
    "ba4_dogleg_zero_tol": lambda: run_ba("ba4_dogleg_zero_tol", "ba4", "dense_schur", "zero", 25, True,
                                          strategy="dogleg", constant_cameras=(0, 1), initial_radius=0.05),
    "ba4_subspace_dogleg_zero_tol": lambda: run_ba("ba4_subspace_dogleg_zero_tol", "ba4", "dense_schur", "zero", 25,
                                                   True, strategy="dogleg", dogleg_type="subspace",
                                                   constant_cameras=(0, 1), initial_radius=0.05),
    "ba4_dogleg_default_tol": lambda: run_ba("ba4_dogleg_default_tol", "ba4", "dense_schur", "default", 50, True,
                                             strategy="dogleg", constant_cameras=(0, 1), initial_radius=0.05),
    "ba16_dense_schur_zero_tol": lambda: run_ba("ba16_dense_schur_zero_tol", "ba16", "dense_schur", "zero", 25, False),
}


def main(argv: list[str]) -> None:
    names = argv or [n for n in CASES if not n.startswith("ba16")]
    for name in names:
        t0 = time.perf_counter()
        data = CASES[name]()
        (OUT / f"{name}.json").write_text(json.dumps(data, indent=1))
        print(f"{name}: {time.perf_counter() - t0:.1f}s", flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])
