"""Dense covariance of a least-squares problem: numpy restatement of the
reference's ``Covariance.compute`` for a given Jacobian.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows ``src/ceres_torch/covariance.py``:

* ``compute``                     :86-166 (the J^T J spectrum and rank tests)
* ``_compute_svd_covariance``     :168-223 (DENSE_SVD: V diag(1/s^2) V^T over the kept singular values)
* ``_compute_qr_covariance``      :225-276 (SPARSE_QR: R^-1 R^-T of the reduced QR)
* ``_svd_threshold`` / ``_qr_threshold`` :374-382
* ``_set_summary_from_singular_values``  :384-422
* ``_eigenvalue_ratios_from_singular_values`` :449-453

The summary is a dict with the reference's ``CovarianceSummary`` field names.
"""
from __future__ import annotations

import numpy as np

_EPS = float(np.finfo(np.float64).eps)
_TINY = float(np.finfo(np.float64).tiny)


def eigenvalue_ratios(s: np.ndarray) -> np.ndarray:
    """covariance.py:449-453: (s / max s)^2."""
    if s.size == 0:
        return s
    return (s / max(float(s.max()), _TINY)) ** 2


def _threshold(values: np.ndarray, shape, column_pivot_threshold: float) -> float:
    """covariance.py:374-382 (the same rule for singular values and |diag R|)."""
    if column_pivot_threshold >= 0:
        return float(column_pivot_threshold)
    return 20.0 * (shape[0] + shape[1]) * _EPS * float(values.max())


def _summary(s: np.ndarray, shape, success: bool, message: str, column_pivot_threshold: float,
             keep: np.ndarray | None = None) -> dict:
    """covariance.py:384-422."""
    if s.size == 0:
        rank, max_sigma, min_retained, rcond = 0, 0.0, 0.0, 0.0
    else:
        if keep is None:
            keep = s > _threshold(s, shape, column_pivot_threshold)
        rank = int(np.sum(keep))
        max_sigma = float(s.max())
        retained = s[keep]
        min_retained = float(retained.min()) if retained.size else 0.0
        rcond = float((retained.min() / max_sigma) ** 2) if retained.size and max_sigma > 0.0 else 0.0
    return dict(success=success, message=message, rank=rank, nullity=max(0, shape[1] - rank), num_rows=shape[0],
                num_columns=shape[1], max_singular_value=max_sigma, min_retained_singular_value=min_retained,
                reciprocal_condition_number=rcond)


def dense_covariance(J: np.ndarray, *, algorithm: str = "dense_svd", column_pivot_threshold: float = -1.0,
                     min_reciprocal_condition_number: float = 1e-14, null_space_rank: int = 0):
    """Tangent covariance [n, n] (None when the reference leaves none) and summary."""
    J = np.asarray(J, dtype=np.float64)
    shape = J.shape
    if shape[1] == 0:
        return np.zeros((0, 0)), _summary(np.zeros(0), shape, True, "No active tangent parameters.",
                                          column_pivot_threshold)
    if algorithm == "sparse_qr":
        if shape[0] < shape[1]:
            out = _summary(np.zeros(0), shape, False,
                           "Sparse QR covariance requires rows >= columns and cannot use null_space_rank.",
                           column_pivot_threshold)
            out.update(rank=shape[0], nullity=shape[1] - shape[0])
            return None, out
        R = np.linalg.qr(J, mode="r")
        diag = np.abs(np.diagonal(R))
        rank = int(np.sum(diag > _threshold(diag, shape, column_pivot_threshold)))
        s = np.linalg.svd(J, compute_uv=False)
        if rank < shape[1]:
            return None, _summary(s, shape, False,
                                  "Sparse QR detected a rank deficient Jacobian and cannot use null_space_rank.",
                                  column_pivot_threshold)
        if np.any(eigenvalue_ratios(s) < min_reciprocal_condition_number) and null_space_rank == 0:
            return None, _summary(s, shape, False, "Sparse QR spectrum is below min_reciprocal_condition_number.",
                                  column_pivot_threshold)
        from scipy.linalg import solve_triangular

        R_inv = solve_triangular(R, np.eye(R.shape[1]), lower=False)
        return R_inv @ R_inv.T, _summary(s, shape, True, "Computed covariance with QR.", column_pivot_threshold)
    _, s, Vh = np.linalg.svd(J, full_matrices=False)
    keep = s > _threshold(s, shape, column_pivot_threshold)
    ratios = eigenvalue_ratios(s)
    if null_space_rank == -1:
        keep = keep & (ratios >= min_reciprocal_condition_number)
    elif null_space_rank > 0:
        if null_space_rank > s.size:
            return Vh.T @ np.diag(np.zeros_like(s)) @ Vh, _summary(
                s, shape, False, "Requested null_space_rank exceeds the number of singular values.",
                column_pivot_threshold, keep)
        keep = keep.copy()
        keep[-null_space_rank:] = False
    inv_s2 = np.where(keep, 1.0 / (s * s), 0.0)
    cov = (Vh.T * inv_s2) @ Vh
    rank = int(np.sum(keep))
    if null_space_rank == 0 and rank < shape[1]:
        return cov, _summary(s, shape, False, "Jacobian is rank deficient.", column_pivot_threshold, keep)
    if np.any(ratios[keep] < min_reciprocal_condition_number):
        return cov, _summary(s, shape, False,
                             "Retained covariance spectrum is below min_reciprocal_condition_number.",
                             column_pivot_threshold, keep)
    return cov, _summary(s, shape, True, "Computed covariance with dense SVD.", column_pivot_threshold, keep)
