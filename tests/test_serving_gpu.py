"""Continuous-batching server: results in the mix equal results alone; admission respects the HBM budget."""
from __future__ import annotations

import numpy as np
import pytest

import ceres_mi355x as tc
from ceres_mi355x.serving import BARequest, CurveRequest, PoseRequest, SolveServer, request_bytes
from oracle import generators as gen
from oracle.ba import BAModel
from oracle.lm import OracleOptions, trust_region_solve

pytestmark = pytest.mark.gpu


def _mix(seed=21):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(24):
        cb = gen.make_curve_batch(num_problems=1, seed=int(rng.integers(1 << 30)))
        reqs.append(CurveRequest(cb.x, cb.y, loss=tc.CauchyLoss(0.5)))
    for C, P in ((3, 150), (8, 600), (12, 900), (20, 400), (24, 300)):
        d = gen.make_bundle_adjustment(num_cameras=C, num_points=P, obs_per_point=(2, min(6, C)),
                                       seed=int(rng.integers(1 << 30)))
        reqs.append(BARequest(d.cameras, d.points, d.obs_camera, d.obs_point, d.observations))
    for n in (100, 260):
        d = gen.make_pose_graph(num_nodes=n, seed=int(rng.integers(1 << 30)))
        reqs.append(PoseRequest(d.poses, d.edges, d.measurements))
    order = rng.permutation(len(reqs))
    return [reqs[i] for i in order]


def test_results_in_the_mix_equal_results_alone():
    reqs = _mix()
    with SolveServer(tc.SolverOptions(max_num_iterations=30), num_slots=3) as server:
        futures = [server.submit(r) for r in reqs]
        mixed = [f.result(timeout=120) for f in futures]
        alone = [server.solve_alone(r) for r in reqs]
    for k, (m, a) in enumerate(zip(mixed, alone)):
        assert m.kind == a.kind
        # deterministic kernels, per-request plans, per-problem batched arithmetic: identical results
        assert m.final_cost == a.final_cost, (k, m.kind)
        np.testing.assert_array_equal(m.params, a.params)
        assert m.termination == a.termination and m.iterations == a.iterations
        assert m.final_cost <= m.initial_cost
        assert m.latency >= 0.0


def test_admission_keeps_within_budget():
    reqs = _mix(seed=5)[:10]
    budget = max(request_bytes(r) for r in reqs)  # at most one large request at a time
    with SolveServer(tc.SolverOptions(max_num_iterations=10), hbm_budget_bytes=budget, num_slots=4) as server:
        futures = [server.submit(r) for r in reqs]
        results = [f.result(timeout=120) for f in futures]
        assert server.in_use == 0
    assert all(r.finished >= r.admitted >= r.arrival for r in results)
    with pytest.raises(ValueError):
        SolveServer(hbm_budget_bytes=1).submit(reqs[0])


def test_served_bundle_adjustments_match_the_oracle():
    """Batched BA requests in a served mix: each final cost equals the float64 oracle's (1e-10 relative),
    including a 24-camera request, which takes a per-request plan on a slot."""
    rng = np.random.default_rng(9)
    datas, reqs = [], []
    for C, P in ((2, 120), (6, 500), (15, 1500), (21, 800), (24, 300)):
        d = gen.make_bundle_adjustment(num_cameras=C, num_points=P, obs_per_point=(2, min(6, C)),
                                       seed=int(rng.integers(1 << 30)))
        datas.append(d)
        reqs.append(BARequest(d.cameras, d.points, d.obs_camera, d.obs_point, d.observations))
    for i in range(6):
        cb = gen.make_curve_batch(num_problems=1, seed=100 + i)
        reqs.append(CurveRequest(cb.x, cb.y, loss=tc.CauchyLoss(0.5)))
    with SolveServer(tc.SolverOptions(), num_slots=2) as server:
        results = [f.result(timeout=120) for f in [server.submit(r) for r in reqs]]
    for d, res in zip(datas, results):
        ref, _ = trust_region_solve(BAModel(d), OracleOptions())
        assert res.kind == "ba"
        assert res.final_cost == pytest.approx(ref.final_cost, rel=1e-10), d.num_cameras
        assert res.iterations == len(ref.iterations) - 1, d.num_cameras
