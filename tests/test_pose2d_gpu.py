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
