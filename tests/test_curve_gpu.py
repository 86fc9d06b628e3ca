"""Batched curve fits on the MI355X vs the reference's recorded trajectories and the oracle."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

import ceres_mi355x as tc
from ceres_mi355x import _native as nat
from ceres_mi355x.ba import lm_options
from ceres_mi355x.curve import CurveBatch
from ceres_mi355x.solver import _records_to_iterations
from conftest import load_golden
from oracle import generators as gen
from test_oracle import assert_trajectory_matches

pytestmark = pytest.mark.gpu


class _Summary:
    def __init__(self, report, records):
        n = int(report["num_records"])
        self.iterations = [vars(it) for it in _records_to_iterations(records[:n], 0.1)]
        self.final_cost = float(report["final_cost"])
        self.initial_cost = float(report["initial_cost"])


def _golden_batch():
    g = load_golden("curve_cauchy")
    data = gen.make_curve_batch(num_problems=g["num_problems"], seed=g["seed"])
    batch = CurveBatch(data.offsets, data.x, data.y, loss=tc.CauchyLoss(0.5))
    opts = lm_options(tc.SolverOptions(max_num_iterations=g["max_num_iterations"]), precision="fp64")
    return g, data, batch, opts


def _waves(batch, limit):
    """Problems of at most ``limit`` samples: solved one per wave at the end of the largest-first order."""
    return int(np.sum(batch.sizes <= limit))


@pytest.mark.parametrize("schedule,order,wave_limit", [((64, 0), "index", 0), ((256, 0), "largest", 0),
                                                       ((512, 64), "index", 0), ((256, 0), "largest", 1000),
                                                       ((128, 8), "largest", 10 ** 6)])
def test_curve_batch_matches_reference_trajectories(schedule, order, wave_limit):
    g, data, batch, opts = _golden_batch()
    if wave_limit:
        schedule = schedule + (_waves(batch, wave_limit),)
        assert 0 < schedule[2] <= batch.num_problems
    res = batch.solve(opts, schedule=schedule, order=order, records=True)
    params = res.params.cpu().numpy()
    for case in g["problems"]:
        i = case["index"]
        s = _Summary(res.reports[i], res.records[i])
        assert_trajectory_matches(s, case["summary"])
        assert int(res.reports[i]["termination"]) == 0  # CONVERGENCE, as the reference
        np.testing.assert_allclose(params[i], case["final"], rtol=1e-7, atol=1e-9)


def test_curve_problem_api_matches_reference():
    g, data, _, _ = _golden_batch()
    case = g["problems"][3]
    i = case["index"]
    lo, hi = data.offsets[i], data.offsets[i + 1]
    m = torch.zeros(1, dtype=torch.float64)
    c = torch.zeros(1, dtype=torch.float64)
    problem = tc.Problem()
    for xv, yv in zip(data.x[lo:hi], data.y[lo:hi]):
        problem.AddResidualBlock(tc.ExpCurveResidual(xv, yv), tc.CauchyLoss(0.5), [m, c])
    summary = tc.solve(tc.SolverOptions(max_num_iterations=g["max_num_iterations"]), problem)
    assert_trajectory_matches(summary, case["summary"])
    assert summary.termination_type == tc.TerminationType.CONVERGENCE
    np.testing.assert_allclose([float(m[0]), float(c[0])], case["final"], rtol=1e-7, atol=1e-9)


def test_curve_results_do_not_depend_on_queue_order():
    _, _, batch, opts = _golden_batch()
    a = batch.solve(opts, schedule=(128, 32), order="index")
    b = batch.solve(opts, schedule=(128, 512), order="largest")
    assert torch.equal(a.params, b.params)  # same workgroup size: bit-identical per problem
    c = batch.solve(opts, schedule=(512, 0), order="index")
    np.testing.assert_allclose(c.reports["final_cost"], a.reports["final_cost"], rtol=1e-12)
    # a problem solved by one wave gives the same bits whatever the workgroup size and count
    w = batch.num_problems
    d = batch.solve(opts, schedule=(128, 16, w), order="largest")
    e = batch.solve(opts, schedule=(512, 4, w), order="index")
    assert torch.equal(d.params, e.params)
    np.testing.assert_allclose(d.reports["final_cost"], a.reports["final_cost"], rtol=1e-12)


def test_curve_mixed_precision_reaches_fp64_cost():
    """The mixed form (FP32 samples, FP32 evaluation, FP64 costs) against the FP64 form and the FP64 oracle."""
    from oracle.lm import OracleOptions, trust_region_solve
    from oracle.small_models import CurveModel

    data = gen.make_curve_batch(num_problems=256, seed=7)
    batch = CurveBatch(data.offsets, data.x, data.y, loss=tc.CauchyLoss(0.5))
    x_r = data.x.astype(np.float32).astype(np.float64)
    y_r = data.y.astype(np.float32).astype(np.float64)
    rounded = CurveBatch(data.offsets, x_r, y_r, loss=tc.CauchyLoss(0.5))
    base = tc.SolverOptions(max_num_iterations=200, function_tolerance=1e-13, gradient_tolerance=0.0,
                            parameter_tolerance=0.0)
    r32 = batch.solve(lm_options(base, precision="mixed"), schedule=(256, 0, 128), order="largest")
    r64 = rounded.solve(lm_options(base, precision="fp64"), schedule=(256, 0), order="largest")
    assert np.all(r32.reports["termination"] == 0)
    # the same FP32-rounded problem: the FP64 costs of both forms agree far below the FP32 evaluation error
    np.testing.assert_allclose(r32.reports["final_cost"], r64.reports["final_cost"], rtol=1e-9)
    np.testing.assert_allclose(r32.params.cpu().numpy(), r64.params.cpu().numpy(), rtol=1e-5, atol=1e-7)
    # against the FP64 oracle on the original samples (the rounding of the samples moves the optimum slightly)
    opts = OracleOptions(max_num_iterations=200, function_tolerance=1e-13, gradient_tolerance=0.0,
                         parameter_tolerance=0.0)
    for i in range(0, 256, 51):
        lo, hi = data.offsets[i], data.offsets[i + 1]
        ref, _ = trust_region_solve(CurveModel(data.x[lo:hi], data.y[lo:hi]), opts)
        assert abs(r32.reports["final_cost"][i] - ref.final_cost) <= 1e-6 * ref.final_cost, i
    # the fit recovers the generating parameters
    err_m = np.abs(r32.params[:, 0].cpu().numpy() - data.true_m)
    assert np.median(err_m) < 0.02


def test_curve_mixed_form_requires_fp32_samples():
    g, data, batch, _ = _golden_batch()
    b = batch._struct(None)
    b.x32 = None
    opts = lm_options(tc.SolverOptions(max_num_iterations=5), precision="mixed")
    params = torch.zeros(batch.num_problems, 2, dtype=torch.float64, device="cuda")
    sched = nat.ct_batch_schedule(64, 0, 0)
    rc = nat.load().ct_curve_solve(ctypes.byref(b), ctypes.byref(opts), ctypes.byref(sched), nat.ptr(params), None,
                                   None, nat.ptr(batch._workspace), batch._workspace.numel(),
                                   nat.current_stream_handle())
    assert rc == nat.CT_ERR_ARGUMENT
    assert b"x32" in nat.load().ct_last_error()


def test_curve_graph_capture_replays_match_eager():
    _, _, batch, opts = _golden_batch()
    start = torch.zeros(batch.num_problems, 2, dtype=torch.float64, device="cuda")
    eager = start.clone()
    shape = (256, 2048)
    batch.launch(opts, eager, schedule=shape, order="largest")
    torch.cuda.synchronize()
    params = start.clone()
    graph = torch.cuda.CUDAGraph()
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        batch.launch(opts, params, schedule=shape, order="largest")  # warm-up outside the capture
        torch.cuda.synchronize()
        with torch.cuda.graph(graph, stream=stream):
            params.zero_()  # fill kernels, not memset nodes
            params.add_(start)
            batch.launch(opts, params, schedule=shape, order="largest")
    for k in range(3):
        params.fill_(-1.0)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(params, eager), k
        _ = torch.randn(1000, device="cuda").sum().item()  # eager work between replays


def test_curve_malformed_ranges_are_rejected():
    x = torch.linspace(0, 1, 10, dtype=torch.float64, device="cuda")
    y = torch.exp(0.3 * x)
    off = torch.tensor([0, 6, 4, 10], dtype=torch.int64, device="cuda")  # problem 1 has end < start
    b = nat.ct_curve_batch()
    b.num_problems, b.num_samples = 3, 10
    b.offsets, b.x, b.y, b.order = nat.ptr(off), nat.ptr(x), nat.ptr(y), None
    b.loss = nat.ct_loss(0, 0.0, 0.0, 1.0)
    lib = nat.load()
    bad = ctypes.c_int32(-1)
    assert lib.ct_curve_validate(ctypes.byref(b), ctypes.byref(bad)) == nat.CT_ERR_ARGUMENT
    assert bad.value == 1
    # solved without validation, the malformed problem ends with INVALID_INPUT; the others converge
    opts = lm_options(tc.SolverOptions(max_num_iterations=20), precision="fp64")
    params = torch.zeros(3, 2, dtype=torch.float64, device="cuda")
    res = torch.zeros(3 * ctypes.sizeof(nat.ct_lm_report), dtype=torch.uint8, device="cuda")
    ws = torch.zeros(256, dtype=torch.uint8, device="cuda")
    sched = nat.ct_batch_schedule(64, 0)
    rc = lib.ct_curve_solve(ctypes.byref(b), ctypes.byref(opts), ctypes.byref(sched), nat.ptr(params), nat.ptr(res),
                            None, nat.ptr(ws), 256, nat.current_stream_handle())
    assert rc == 0
    torch.cuda.synchronize()
    rep = np.frombuffer(res.cpu().numpy().tobytes(), dtype=nat.REPORT_DTYPE)
    assert rep["reason"][1] == 8 and rep["termination"][1] == 2
    assert rep["termination"][0] == 0 and rep["termination"][2] == 0
