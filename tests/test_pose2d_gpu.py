"""SE2 pose graphs on the MI355X vs the reference's recorded trajectory and the oracle."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import ceres_mi355x as tc
from ceres_mi355x.ba import lm_options
from ceres_mi355x.pose2d import PoseGraphDevice
from ceres_mi355x.solver import set_device_loop
from conftest import load_golden
from oracle import generators as gen
from oracle.lm import OracleOptions, trust_region_solve
from oracle.small_models import PoseGraphModel
from test_oracle import assert_trajectory_matches

pytestmark = pytest.mark.gpu


def _problem(data, loss=None):
    poses = [torch.tensor(p, dtype=torch.float64) for p in data.poses]
    problem = tc.Problem()
    for (a, b), m in zip(data.edges, data.measurements):
        problem.AddResidualBlock(tc.Pose2DRelativeError(*m), loss, [poses[a], poses[b]])
    problem.SetParameterBlockConstant(poses[0])
    return problem, poses


@pytest.mark.parametrize("loop", ["host", "device"])
def test_pose_graph_matches_reference_trajectory(loop):
    g = load_golden("pose2d_40")
    data = gen.make_pose_graph(num_nodes=g["num_nodes"], seed=g["seed"])
    problem, poses = _problem(data)
    set_device_loop(loop)
    try:
        summary = tc.solve(tc.SolverOptions(max_num_iterations=g["max_num_iterations"]), problem)
    finally:
        set_device_loop("device")
    assert_trajectory_matches(summary, g["summary"])
    assert summary.termination_type == tc.TerminationType.CONVERGENCE
    np.testing.assert_allclose(torch.stack(poses).numpy(), np.asarray(g["final_poses"]), rtol=1e-7, atol=1e-8)


def test_pose_graph_with_robust_loss_matches_oracle():
    data = gen.make_pose_graph(num_nodes=120, seed=9)
    loss = tc.HuberLoss(0.1)
    problem, poses = _problem(data, loss)
    opts = tc.SolverOptions(max_num_iterations=30)
    summary = tc.solve(opts, problem)
    ref, x = trust_region_solve(PoseGraphModel(data.poses, data.edges, data.measurements, loss=("huber", 0.1)),
                                OracleOptions(max_num_iterations=30))
    assert_trajectory_matches(summary, {"iterations": ref.iterations, "final_cost": ref.final_cost,
                                        "initial_cost": ref.initial_cost})
    np.testing.assert_allclose(torch.stack(poses).numpy(), x, rtol=1e-7, atol=1e-8)


def _separator_unknowns(data, fixed=0):
    """3 x the nodes touched by a non-chain edge, other than the constant node."""
    touched = {int(v) for a, b in data.edges if abs(int(a) - int(b)) != 1 for v in (a, b)}
    touched.discard(fixed)
    return 3 * len(touched)


def test_pose_graph_large_uses_blocked_cholesky():
    data = gen.make_pose_graph(num_nodes=400, seed=12)  # 1197 unknowns, 234 of them on the separator
    dev = PoseGraphDevice(data.poses, data.edges, data.measurements)
    opts = lm_options(tc.SolverOptions(max_num_iterations=20), precision="fp64", loop="device")
    assert dev.system_size(opts) == _separator_unknowns(data) == 234  # above 192: the tiled factorisation
    rep, rec = dev.run(opts)
    ref, x = trust_region_solve(PoseGraphModel(data.poses, data.edges, data.measurements),
                                OracleOptions(max_num_iterations=20))
    assert rep.final_cost == pytest.approx(ref.final_cost, rel=1e-9)
    np.testing.assert_allclose(dev.poses.cpu().numpy(), x, rtol=1e-6, atol=1e-7)
    err = np.abs(dev.poses.cpu().numpy()[:, :2] - data.true_poses[:, :2])
    assert np.median(err) < 1.0


@pytest.mark.parametrize("loop_fraction,form", [(0.0, "chain"), (3.0, "dense")])
def test_pose_graph_chain_and_dense_forms_match_oracle(loop_fraction, form):
    """A pure chain has an empty separator (every node is eliminated along the chain, nothing is factored
    densely); with loop closures on every node no node lies in a run and the plan keeps the dense form."""
    print(f"pose graph form={form}", flush=True)
    data = gen.make_pose_graph(num_nodes=60, seed=31, loop_fraction=loop_fraction)
    dev = PoseGraphDevice(data.poses, data.edges, data.measurements)
    opts = lm_options(tc.SolverOptions(max_num_iterations=20), precision="fp64", loop="device")
    assert dev.system_size(opts) == (0 if form == "chain" else 3 * 59)
    dev.close()
    problem, poses = _problem(data)
    summary = tc.solve(tc.SolverOptions(max_num_iterations=20), problem)
    ref, x = trust_region_solve(PoseGraphModel(data.poses, data.edges, data.measurements),
                                OracleOptions(max_num_iterations=20))
    assert_trajectory_matches(summary, {"iterations": ref.iterations, "final_cost": ref.final_cost,
                                        "initial_cost": ref.initial_cost})
    np.testing.assert_allclose(torch.stack(poses).numpy(), x, rtol=1e-6, atol=1e-7)


def _graph_with_runs(num_nodes, extra_edges, seed):
    """A chain plus the given non-chain edges, measured from the true poses with the generator's noise."""
    base = gen.make_pose_graph(num_nodes=num_nodes, seed=seed, loop_fraction=0.0)
    rng = np.random.default_rng(seed + 1)
    extra = np.asarray(extra_edges, dtype=np.int32).reshape(-1, 2)
    meas = gen._relative_pose(base.true_poses[extra[:, 0]], base.true_poses[extra[:, 1]])
    meas[:, :2] += rng.normal(0.0, 0.05, size=(len(extra), 2))
    meas[:, 2] += rng.normal(0.0, 0.01, size=len(extra))
    return gen.PoseGraphData(poses=base.poses, edges=np.concatenate([base.edges, extra]),
                             measurements=np.concatenate([base.measurements, meas]), true_poses=base.true_poses)


@pytest.mark.parametrize("loop", ["host", "device"])
@pytest.mark.parametrize("layout", ["chain200", "runs_64_65_68"])
def test_pose_runs_longer_than_one_lds_chunk_match_oracle(layout, loop):
    """Run elimination and back-substitution walk a run 64 nodes at a time, carrying the factor and
    the solve across chunk boundaries.  A 200-node chain is one run of 199 nodes (4 chunks); loop
    closures (0, 65) and (65, 131) cut a 200-node chain into runs of exactly 64, 65 and 68 nodes."""
    print(f"pose layout={layout} loop={loop}", flush=True)
    if layout == "chain200":
        data = gen.make_pose_graph(num_nodes=200, seed=51, loop_fraction=0.0)
        unknowns = 0
    else:
        data = _graph_with_runs(200, [(0, 65), (65, 131)], seed=52)
        unknowns = 3 * 2  # nodes 65 and 131 (node 0 is constant)
    dev = PoseGraphDevice(data.poses, data.edges, data.measurements)
    assert dev.system_size(lm_options(tc.SolverOptions(max_num_iterations=1), precision="fp64", loop="device")) == unknowns
    dev.close()
    problem, poses = _problem(data)
    set_device_loop(loop)
    try:
        summary = tc.solve(tc.SolverOptions(max_num_iterations=20), problem)
    finally:
        set_device_loop("device")
    ref, x = trust_region_solve(PoseGraphModel(data.poses, data.edges, data.measurements),
                                OracleOptions(max_num_iterations=20))
    assert_trajectory_matches(summary, {"iterations": ref.iterations, "final_cost": ref.final_cost,
                                        "initial_cost": ref.initial_cost})
    np.testing.assert_allclose(torch.stack(poses).numpy(), x, rtol=1e-6, atol=1e-7)


@pytest.mark.parametrize("loop", ["host", "device"])
def test_pose2000_matches_oracle_trajectory(loop):
    """The largest pose graph of the served mix (2,000 SE2 nodes, odometry plus 200 loop closures),
    first 10 iterations against the oracle's (tests/golden/make_oracle_fixtures.py)."""
    from conftest import array_digest

    print(f"pose2000 loop={loop}", flush=True)
    g = load_golden("oracle_pose2000_10")
    data = gen.make_pose_graph(num_nodes=g["num_nodes"], seed=g["seed"], loop_fraction=g["loop_fraction"])
    assert g["input_sha256"] == array_digest(data.poses, data.edges, data.measurements)
    problem, poses = _problem(data)
    set_device_loop(loop)
    try:
        summary = tc.solve(tc.SolverOptions(max_num_iterations=g["max_num_iterations"]), problem)
    finally:
        set_device_loop("device")
    assert_trajectory_matches(summary, g["summary"])
    np.testing.assert_allclose(torch.stack(poses)[:300].numpy(), np.asarray(g["final_poses_head"]), rtol=1e-6,
                               atol=1e-7)
