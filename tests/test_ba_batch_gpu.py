"""Batched small bundle adjustments (one persistent launch, one workgroup per problem).

Each problem's per-iteration records are checked against the reference's own
recorded trajectories (ba4 / ba6-Huber / ba16 goldens) and against the float64
oracle's trajectory for the served mix's sizes (2-20 cameras, up to 2,000
points), to the plain FP64 contract: accept/reject sequence and costs to 1e-9
relative until the cost has converged, final cost to 1e-10 (1e-7 for the
Huber/outlier case, as in test_ba_gpu.py).  A problem's result is bit-identical
whether it is solved alone or in the batch.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import ceres_mi355x as tc
from ceres_mi355x import _native as nat
from ceres_mi355x.ba import lm_options
from ceres_mi355x.ba_batch import BABatch, BAProblemInput
from ceres_mi355x.solver import _MESSAGES, _TERMINATION
from oracle import generators as gen
from oracle.ba import BAModel
from oracle.lm import OracleOptions, trust_region_solve

from conftest import load_golden
from test_oracle import assert_trajectory_matches

pytestmark = pytest.mark.gpu

LOSS_NATIVE = {"trivial": lambda a: ("trivial", 0.0, 0.0, 1.0), "huber": lambda a: ("huber", a, 0.0, 1.0),
               "cauchy": lambda a: ("cauchy", a, 0.0, 1.0)}


def as_input(data) -> BAProblemInput:
    loss = data.loss if data.loss else ("trivial",)
    native = LOSS_NATIVE[loss[0]](loss[1] if len(loss) > 1 else 0.0)
    return BAProblemInput(data.cameras, data.points, data.obs_camera, data.obs_point, data.observations, native)


class _Summary:
    """Per-problem records in the shape assert_trajectory_matches reads."""

    def __init__(self, rep, rec):
        n = int(rep["num_records"])
        self.iterations = [dict(iteration=int(r[nat.REC["iteration"]]), cost=float(r[nat.REC["cost"]]),
                                step_is_successful=bool(r[nat.REC["step_is_successful"]]),
                                trust_region_radius=float(r[nat.REC["trust_region_radius"]]),
                                linear_solver_iterations=int(r[nat.REC["linear_solver_iterations"]]))
                           for r in rec[:n]]
        self.final_cost = float(rep["final_cost"])
        self.initial_cost = float(rep["initial_cost"])


def _options(max_iter, zero_tol):
    o = tc.SolverOptions(max_num_iterations=max_iter, linear_solver_type=tc.LinearSolverType.DENSE_SCHUR)
    if zero_tol:
        o.function_tolerance = o.gradient_tolerance = o.parameter_tolerance = 0.0
    return o


def test_batched_ba_matches_reference_trajectories():
    cases = ["ba4_dense_schur_zero_tol", "ba6_huber_dense_schur", "ba16_dense_schur_zero_tol",
             "ba4_dense_schur_default_tol"]
    goldens = [load_golden(c) for c in cases]
    print("batched cases: " + ", ".join(cases), flush=True)
    # one launch per option set: grouped by (max iterations, tolerances)
    groups = {}
    for g, case in zip(goldens, cases):
        groups.setdefault((g["max_num_iterations"], g["tolerances"]), []).append((case, g))
    for (max_iter, tol), members in groups.items():
        batch = BABatch([as_input(gen.ba_config(g["config"])) for _, g in members])
        opts = lm_options(_options(max_iter, tol == "zero"))
        reps, recs, _, _ = batch.solve(opts, with_records=True)
        for k, (case, g) in enumerate(members):
            s = _Summary(reps[k], recs[k])
            final_rtol = 1e-7 if "huber" in case else 1e-10
            assert_trajectory_matches(s, g["summary"], final_rtol=final_rtol)
            if tol == "default":
                assert _TERMINATION[int(reps[k]["termination"])].value == g["summary"]["termination_type"]
                assert _MESSAGES[int(reps[k]["reason"])] == g["summary"]["message"]
                assert int(reps[k]["num_records"]) == len(g["summary"]["iterations"])
                assert int(reps[k]["num_successful_steps"]) == g["summary"]["num_successful_steps"]


def _mix(seed, count):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        C = int(rng.integers(2, 21))
        P = int(rng.integers(100, 1200))
        out.append(gen.make_bundle_adjustment(num_cameras=C, num_points=P, obs_per_point=(2, min(6, C)),
                                              seed=int(rng.integers(1 << 30))))
    return out


def test_batched_ba_matches_oracle_on_the_served_sizes():
    datas = _mix(7, 6)
    datas.append(gen.make_bundle_adjustment(num_cameras=20, num_points=2000, obs_per_point=(2, 6), seed=11,
                                            outlier_fraction=0.05, loss=("huber", 1.0)))
    datas.append(gen.make_bundle_adjustment(num_cameras=21, num_points=300, obs_per_point=(2, 6), seed=12))
    datas.append(gen.make_bundle_adjustment(num_cameras=2, num_points=60, obs_per_point=(2, 2), seed=13))
    print("batched oracle mix: " + ", ".join(f"C={d.num_cameras} P={d.num_points}" for d in datas), flush=True)
    batch = BABatch([as_input(d) for d in datas])
    sopts = tc.SolverOptions(linear_solver_type=tc.LinearSolverType.DENSE_SCHUR)
    reps, recs, cams, pts = batch.solve(lm_options(sopts), with_records=True)
    for k, d in enumerate(datas):
        ref, _ = trust_region_solve(BAModel(d), OracleOptions(max_num_iterations=sopts.max_num_iterations))
        golden = dict(iterations=ref.iterations, final_cost=ref.final_cost, initial_cost=ref.initial_cost)
        rtol = 1e-7 if d.loss and d.loss[0] == "huber" else 1e-10
        assert_trajectory_matches(_Summary(reps[k], recs[k]), golden, final_rtol=rtol)
        assert int(reps[k]["termination"]) in (0, 1)
        c, p = batch.split(cams, pts, k)
        assert np.all(np.isfinite(c)) and np.all(np.isfinite(p))


def test_batched_ba_alone_equals_in_the_batch():
    datas = _mix(3, 5)
    sopts = tc.SolverOptions(linear_solver_type=tc.LinearSolverType.DENSE_SCHUR, max_num_iterations=20)
    opts = lm_options(sopts)
    batch = BABatch([as_input(d) for d in datas])
    reps, _, cams, pts = batch.solve(opts)
    for k, d in enumerate(datas):
        one = BABatch([as_input(d)])
        r1, _, c1, p1 = one.solve(opts)
        assert r1[0]["final_cost"] == reps[k]["final_cost"], k
        assert r1[0]["num_records"] == reps[k]["num_records"], k
        c, p = batch.split(cams, pts, k)
        np.testing.assert_array_equal(c1.reshape(-1, 9), c)
        np.testing.assert_array_equal(p1.reshape(-1, 3), p)


def test_batched_ba_matches_the_per_problem_plan():
    """Same problem through the per-problem DENSE_SCHUR plan and through the batch: same trajectory."""
    d = gen.make_bundle_adjustment(num_cameras=12, num_points=900, obs_per_point=(2, 6), seed=5)
    sopts = tc.SolverOptions(linear_solver_type=tc.LinearSolverType.DENSE_SCHUR)
    cams = torch.tensor(d.cameras, dtype=torch.float64, device="cuda")
    pts = torch.tensor(d.points, dtype=torch.float64, device="cuda")
    problem = tc.Problem()
    problem.add_bundle_adjustment_observations(cams, pts, d.obs_camera, d.obs_point, d.observations)
    summary = tc.solve(sopts, problem)
    reps, recs, _, _ = BABatch([as_input(d)]).solve(lm_options(sopts), with_records=True)
    golden = dict(iterations=[vars(i) for i in summary.iterations], final_cost=summary.final_cost,
                  initial_cost=summary.initial_cost)
    assert_trajectory_matches(_Summary(reps[0], recs[0]), golden, final_rtol=1e-10)
    assert int(reps[0]["num_records"]) == len(summary.iterations)


def test_batched_ba_refuses_what_it_does_not_cover():
    d = gen.ba_config("ba4")
    batch = BABatch([as_input(d)])
    o = lm_options(tc.SolverOptions(linear_solver_type=tc.LinearSolverType.ITERATIVE_SCHUR), linear="iterative_schur")
    with pytest.raises(nat.NativeLibraryError):
        batch.solve(o)
    with pytest.raises(ValueError):
        BABatch([BAProblemInput(np.zeros((22, 9)), np.zeros((1, 3)), [0], [0], [[0.0, 0.0]])])
