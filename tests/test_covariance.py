"""Host-side covariance API (no GPU): options, block requests, refusals."""
from __future__ import annotations

import pytest
import torch

import ceres_mi355x as tc


def test_covariance_options_validate():
    # tests/test_covariance_interpolation.py:133-137
    with pytest.raises(ValueError, match="null_space_rank"):
        tc.CovarianceOptions(null_space_rank=-2).validate()
    with pytest.raises(ValueError, match="min_reciprocal_condition_number"):
        tc.CovarianceOptions(min_reciprocal_condition_number=-1.0).validate()


def _ba_problem():
    cams = torch.zeros(2, 9, dtype=torch.float64)
    pts = torch.zeros(2, 3, dtype=torch.float64)
    problem = tc.Problem()
    problem.add_bundle_adjustment_observations(cams, pts, [0, 1, 0, 1], [0, 0, 1, 1], torch.zeros(4, 2))
    return problem


def test_covariance_rejects_duplicate_block_requests():
    # tests/test_covariance_interpolation.py:47-59
    problem = _ba_problem()
    a, b = problem.parameter_blocks[0], problem.parameter_blocks[2]
    with pytest.raises(ValueError, match="duplicate"):
        tc.Covariance().Compute([(a, b), (b, a)], problem)
    with pytest.raises(ValueError, match="duplicates"):
        tc.Covariance().Compute([a, a], problem)


def test_covariance_refuses_what_the_device_path_does_not_cover():
    problem = _ba_problem()
    a = problem.parameter_blocks[0]
    with pytest.raises(NotImplementedError, match="null_space_rank"):
        tc.Covariance(tc.CovarianceOptions(null_space_rank=-1)).Compute([(a, a)], problem)
    other = tc.Problem()
    x = torch.zeros(1, dtype=torch.float64)
    other.AddResidualBlock(tc.AutoDiffCostFunction(lambda x: 2.0 * x - 1.0, [1]), None, [x])
    with pytest.raises(NotImplementedError):
        tc.Covariance().Compute([(x, x)], other)


def test_covariance_with_no_blocks_succeeds_without_a_gpu():
    problem = _ba_problem()
    cov = tc.Covariance()
    assert cov.Compute([], problem)
    assert cov.summary.message == "No covariance blocks requested."
    assert cov.summary.ambient_dimension == 24 and cov.summary.tangent_dimension == 24
    assert "Covariance Report" in cov.summary.BriefReport() and "Requested blocks: 0" in cov.summary.FullReport()
