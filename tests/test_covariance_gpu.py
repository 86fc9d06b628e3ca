"""Covariance of bundle-adjustment estimates on the MI355X against the oracle's
restatement of the reference's dense SVD covariance (covariance.py:86-223).

The gauge of bundle adjustment (rotation, translation, scale) is fixed by
holding two cameras constant, so J has full rank and the reference's
(J^T J)^+ is (J^T J)^-1.  Tolerance: 1e-7 relative to the largest entry of
each block (the device forms S^-1 from an FP64 Cholesky, the oracle from an
SVD of the dense Jacobian; both lose about cond(J^T J) x eps).
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import ceres_mi355x as tc
from oracle import ba as oba
from oracle import generators as gen
from oracle.covariance import dense_covariance

pytestmark = pytest.mark.gpu


def _problem(data, constant_cameras=(0, 1), loss=None):
    cams = torch.tensor(data.cameras, dtype=torch.float64)
    pts = torch.tensor(data.points, dtype=torch.float64)
    problem = tc.Problem()
    problem.add_bundle_adjustment_observations(cams, pts, data.obs_camera, data.obs_point, data.observations,
                                               loss=loss)
    for i in constant_cameras:
        problem.SetParameterBlockConstant(problem.parameter_blocks[i])
    return problem, cams, pts


def _oracle(data, constant_cameras=(0, 1), apply_loss=True):
    active = np.ones(data.num_cameras, bool)
    active[list(constant_cameras)] = False
    d = data
    if not apply_loss:
        import copy

        d = copy.copy(data)
        d.loss = ("trivial",)
    model = oba.BAModel(d, camera_active=active)
    ev, _ = model.evaluate(model.x0(), True)
    J = model.dense_jacobian(ev)
    cov, summary = dense_covariance(J)
    npa = data.num_points
    ccol = -np.ones(data.num_cameras, np.int64)
    ccol[active] = 3 * npa + 9 * np.arange(int(active.sum()))

    def cols(kind, i):
        return np.arange(3 * i, 3 * i + 3) if kind == "p" else np.arange(ccol[i], ccol[i] + 9)

    return cov, summary, cols


def _check(cov_dev, problem, cov, cols, pairs, C):
    for (ka, ia), (kb, ib) in pairs:
        a = problem.parameter_blocks[ia if ka == "c" else C + ia]
        b = problem.parameter_blocks[ib if kb == "c" else C + ib]
        got = cov_dev.GetCovarianceBlockInTangentSpace(a, b).numpy()
        ref = cov[np.ix_(cols(ka, ia), cols(kb, ib))]
        scale = np.abs(ref).max()
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-7 * scale, err_msg=str(((ka, ia), (kb, ib))))


@pytest.mark.parametrize("config,pairs", [
    ("ba4", [(("c", 2), ("c", 2)), (("c", 2), ("c", 3)), (("p", 0), ("p", 0)), (("p", 5), ("p", 17)),
             (("p", 7), ("c", 3)), (("c", 2), ("p", 9))]),
    ("ba6_huber", [(("c", 4), ("c", 5)), (("p", 3), ("p", 3)), (("p", 11), ("c", 2)), (("p", 40), ("p", 41))]),
])
def test_covariance_blocks_match_the_dense_svd_oracle(config, pairs):
    print(f"covariance {config}", flush=True)
    data = gen.ba_config(config)
    loss = tc.HuberLoss(data.loss[1]) if data.loss[0] == "huber" else None
    problem, _, _ = _problem(data, loss=loss)
    C = data.num_cameras
    blocks = []
    for (ka, ia), (kb, ib) in pairs:
        blocks.append((problem.parameter_blocks[ia if ka == "c" else C + ia],
                       problem.parameter_blocks[ib if kb == "c" else C + ib]))
    cov_dev = tc.Covariance()
    assert cov_dev.Compute(blocks, problem), cov_dev.summary.message
    cov, summary, cols = _oracle(data)
    assert summary["success"]
    assert cov_dev.summary.rank == summary["rank"] and cov_dev.summary.nullity == 0
    assert cov_dev.summary.num_rows == summary["num_rows"] and cov_dev.summary.num_columns == summary["num_columns"]
    _check(cov_dev, problem, cov, cols, pairs, C)
    # transposed requests and the ambient form of a constant camera
    a, b = blocks[0]
    torch.testing.assert_close(cov_dev.GetCovarianceBlock(b, a), cov_dev.GetCovarianceBlock(a, b).T)
    const = problem.parameter_blocks[0]
    cov2 = tc.Covariance()
    assert cov2.Compute([(const, const), (const, a)], problem)
    assert torch.equal(cov2.GetCovarianceBlock(const, const), torch.zeros(9, 9, dtype=torch.float64))
    assert cov2.GetCovarianceBlockInTangentSpace(const, a).shape == (0, a.tangent_size)


def test_covariance_tiled_reduced_system_and_loss_option():
    """24 cameras (9C = 216 > 192: the tiled factorisation) and apply_loss_function=False."""
    data = gen.make_bundle_adjustment(num_cameras=24, num_points=300, obs_per_point=(3, 6), seed=61,
                                      outlier_fraction=0.05, loss=("huber", 1.0))
    problem, _, _ = _problem(data, loss=tc.HuberLoss(1.0))
    C = data.num_cameras
    pairs = [(("c", 5), ("c", 23)), (("c", 12), ("c", 12)), (("p", 4), ("p", 4)), (("p", 4), ("p", 250)),
             (("p", 99), ("c", 7))]
    blocks = [(problem.parameter_blocks[ia if ka == "c" else C + ia],
               problem.parameter_blocks[ib if kb == "c" else C + ib]) for (ka, ia), (kb, ib) in pairs]
    for apply_loss in (True, False):
        cov_dev = tc.Covariance(tc.CovarianceOptions(apply_loss_function=apply_loss))
        assert cov_dev.Compute(blocks, problem), cov_dev.summary.message
        cov, summary, cols = _oracle(data, apply_loss=apply_loss)
        assert summary["success"]
        _check(cov_dev, problem, cov, cols, pairs, C)


def test_covariance_of_a_free_gauge_is_rank_deficient_like_the_oracle():
    """No constant camera: J^T J has the 7-dimensional similarity null space; the reference's SVD reports
    rank = columns - 7 and fails; the device's pivot test finds the same 7 null directions."""
    data = gen.ba_config("ba4")
    problem, _, _ = _problem(data, constant_cameras=())
    cov_dev = tc.Covariance()
    a = problem.parameter_blocks[1]
    assert not cov_dev.Compute([(a, a)], problem)
    _, summary, _ = _oracle(data, constant_cameras=())
    assert not summary["success"] and summary["nullity"] == 7
    assert cov_dev.summary.message == summary["message"] == "Jacobian is rank deficient."
    assert cov_dev.summary.nullity == 7 and cov_dev.summary.rank == summary["rank"]
    with pytest.raises(RuntimeError):
        cov_dev.GetCovarianceBlock(a, a)


def test_covariance_sparse_qr_matches_dense_svd():
    data = gen.ba_config("ba4")
    problem, _, _ = _problem(data)
    blocks = [problem.parameter_blocks[2], problem.parameter_blocks[4 + 3]]
    svd = tc.Covariance(tc.CovarianceOptions(algorithm_type=tc.CovarianceAlgorithmType.DENSE_SVD))
    qr = tc.Covariance(tc.CovarianceOptions(algorithm_type=tc.CovarianceAlgorithmType.SPARSE_QR))
    assert svd.Compute(blocks, problem) and qr.Compute(blocks, problem)
    assert qr.summary.message == "Computed covariance with QR."
    torch.testing.assert_close(qr.GetCovarianceMatrix(blocks), svd.GetCovarianceMatrix(blocks), rtol=0, atol=0)
