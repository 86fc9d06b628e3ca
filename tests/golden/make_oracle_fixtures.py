"""Write LM trajectories of the float64 oracle at sizes too large to run in a test.

The reference's dense algebra cannot run the 200-camera bundle adjustment
(its dense Jacobian would be 3.1 TB) and a 2000-node pose graph takes minutes
per solve in the oracle's dense DENSE_QR restatement, so the GPU parity tests
at those sizes compare against trajectories this script records by running
``oracle/`` (the block-sparse float64 restatement, itself pinned by the
reference-recorded trajectories in this folder) on inputs from
``oracle/generators.py``.  It never imports the reference.

Usage (from the repository root):

    python tests/golden/make_oracle_fixtures.py [case ...]
"""
from __future__ import annotations

import json
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path[:0] = [str(ROOT), str(ROOT / "tests")]
OUT = Path(__file__).resolve().parent

from conftest import array_digest  # noqa: E402
from oracle import ba as oba  # noqa: E402
from oracle import generators as gen  # noqa: E402
from oracle.lm import OracleOptions, trust_region_solve  # noqa: E402
from oracle.small_models import PoseGraphModel  # noqa: E402


def _summary(s) -> dict:
    return dict(initial_cost=s.initial_cost, final_cost=s.final_cost, termination_type=s.termination_type,
                message=s.message, num_successful_steps=s.num_successful_steps,
                num_unsuccessful_steps=s.num_unsuccessful_steps,
                num_residual_evaluations=s.num_residual_evaluations,
                num_jacobian_evaluations=s.num_jacobian_evaluations, num_linear_solves=s.num_linear_solves,
                iterations=s.iterations)


def run_ba(name: str, config: str, solver: str, max_iters: int, tolerances: str = "zero",
           preconditioner: str = "schur_jacobi", keep_iterations: int | None = None) -> dict:
    d = gen.ba_config(config)
    o = OracleOptions(max_num_iterations=max_iters, preconditioner_type=preconditioner)
    if tolerances == "zero":
        o.function_tolerance = o.gradient_tolerance = o.parameter_tolerance = 0.0
    t0 = time.perf_counter()
    s, (cams, pts) = trust_region_solve(oba.BAModel(d, linear_solver=solver), o)
    summary = _summary(s)
    summary["num_iterations"] = len(s.iterations)
    if keep_iterations is not None:
        summary["iterations"] = summary["iterations"][:keep_iterations]
    return dict(case=name, kind="ba", source="oracle", config=config, linear_solver=solver, tolerances=tolerances,
                max_num_iterations=max_iters, loss=list(d.loss) if d.loss else None,
                preconditioner=preconditioner if solver == "iterative_schur" else None,
                input_sha256=array_digest(d.cameras, d.points, d.observations, d.obs_camera, d.obs_point),
                oracle_seconds=time.perf_counter() - t0, summary=summary,
                final_points_head=np.asarray(pts)[:200].tolist(), final_cameras=np.asarray(cams).tolist())


def run_pose(name: str, num_nodes: int, seed: int, max_iters: int, loop_fraction: float = 0.10) -> dict:
    g = gen.make_pose_graph(num_nodes=num_nodes, seed=seed, loop_fraction=loop_fraction)
    t0 = time.perf_counter()
    s, x = trust_region_solve(PoseGraphModel(g.poses, g.edges, g.measurements), OracleOptions(max_num_iterations=max_iters))
    return dict(case=name, kind="pose2d", source="oracle", num_nodes=num_nodes, seed=seed,
                loop_fraction=loop_fraction, max_num_iterations=max_iters,
                input_sha256=array_digest(g.poses, g.edges, g.measurements),
                oracle_seconds=time.perf_counter() - t0, summary=_summary(s),
                final_poses_head=np.asarray(x)[:300].tolist())


COVARIANCE_BLOCKS = [(("c", 2), ("c", 2)), (("c", 5), ("c", 11)), (("c", 15), ("c", 15)), (("p", 0), ("p", 0)),
                     (("p", 17), ("p", 1999)), (("p", 500), ("c", 9)), (("c", 3), ("p", 1234)), (("p", 42), ("p", 43))]


def run_covariance(name: str, config: str, constant_cameras=(0, 1)) -> dict:
    """Covariance blocks of the dense SVD form (covariance.py:168-223) at the generator's start, gauge fixed
    by constant cameras."""
    from oracle.covariance import dense_covariance

    d = gen.ba_config(config)
    active = np.ones(d.num_cameras, bool)
    active[list(constant_cameras)] = False
    model = oba.BAModel(d, camera_active=active)
    t0 = time.perf_counter()
    ev, _ = model.evaluate(model.x0(), True)
    J = model.dense_jacobian(ev)
    cov, summary = dense_covariance(J)
    ccol = -np.ones(d.num_cameras, np.int64)
    ccol[active] = 3 * d.num_points + 9 * np.arange(int(active.sum()))

    def cols(kind, i):
        return np.arange(3 * i, 3 * i + 3) if kind == "p" else np.arange(ccol[i], ccol[i] + 9)

    blocks = [dict(a=list(a), b=list(b), value=cov[np.ix_(cols(*a), cols(*b))].tolist()) for a, b in COVARIANCE_BLOCKS]
    return dict(case=name, kind="ba_covariance", source="oracle", config=config,
                constant_cameras=list(constant_cameras),
                input_sha256=array_digest(d.cameras, d.points, d.observations, d.obs_camera, d.obs_point),
                oracle_seconds=time.perf_counter() - t0, summary=summary, blocks=blocks)


CASES = {
    # the 200-camera workload: first 5 LM iterations, both linear solvers
    "oracle_ba200_iterative_schur_5": lambda: run_ba("oracle_ba200_iterative_schur_5", "ba200", "iterative_schur", 5),
    "oracle_ba200_dense_schur_5": lambda: run_ba("oracle_ba200_dense_schur_5", "ba200", "dense_schur", 5),
    # the 16-camera Huber problem with 5% outliers, FP64, run until LM stops by itself
    "oracle_ba16_huber_converged": lambda: run_ba("oracle_ba16_huber_converged", "ba16_huber", "dense_schur", 2000,
                                                  keep_iterations=40),
    # the largest pose graph of the served mix: 2000 SE2 nodes, 10 iterations
    "oracle_pose2000_10": lambda: run_pose("oracle_pose2000_10", 2000, 21, 10),
    # covariance blocks of the 16-camera problem (dense SVD of its 16,000 x 6,126 Jacobian)
    "oracle_ba16_covariance": lambda: run_covariance("oracle_ba16_covariance", "ba16"),
}


def main(argv: list[str]) -> None:
    for name in argv or list(CASES):
        t0 = time.perf_counter()
        data = CASES[name]()
        (OUT / f"{name}.json").write_text(json.dumps(data, indent=0, separators=(",", ":")))
        print(f"{name}: {time.perf_counter() - t0:.1f}s", flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])
