"""Plugin-registry backends (dense damped normal equations) and the standalone Cholesky on the MI355X.

Reference semantics (cuda_backends.py:215-307): x solves (A^T A + diag(d)) x = A^T b.
The oracle here is the textbook identity with the damped least-squares problem
min |A x - b|^2 + sum d x^2, solved by numpy lstsq on [A; sqrt(d)] in float64.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

from ceres_mi355x import _native as nat
from ceres_mi355x import linear as L

pytestmark = pytest.mark.gpu


def _damped_lstsq(A, b, d):
    aug = np.vstack([A, np.diag(np.sqrt(d))]) if d is not None else A
    rhs = np.concatenate([b, np.zeros(A.shape[1])]) if d is not None else b
    return np.linalg.lstsq(aug, rhs, rcond=None)[0]


@pytest.mark.parametrize("m,n", [(40, 7), (300, 144), (900, 257), (2000, 520)])
@pytest.mark.parametrize("damped", [False, True])
def test_normal_equations_backend_matches_damped_least_squares(m, n, damped):
    rng = np.random.default_rng(m + n)
    A = rng.normal(size=(m, n))
    b = rng.normal(size=m)
    d = rng.uniform(0.1, 2.0, size=n) if damped else None
    ref = _damped_lstsq(A, b, d)
    At = torch.tensor(A, device="cuda")
    res = L.rocm_normal_equations(At, torch.tensor(b, device="cuda"),
                                  damping=None if d is None else torch.tensor(d, device="cuda"))
    np.testing.assert_allclose(res.x.cpu().numpy(), ref, rtol=1e-9, atol=1e-11)
    assert res.summary.success and res.summary.num_iterations == 1
    # the reported residual is that of the damped least-squares problem
    expect = np.sqrt(np.sum((A @ ref - b) ** 2) + (0.0 if d is None else np.sum(d * ref * ref)))
    assert res.summary.residual_norm == pytest.approx(expect, rel=1e-9)


@pytest.mark.parametrize("m,n,lda_pad", [(5000, 70, 0), (40000, 200, 3), (51200, 640, 0), (3, 130, 1)])
def test_gram_over_row_chunks_matches_the_normal_equations(m, n, lda_pad):
    """H = A^T A and A^T b are summed over row chunks of A (fewer than eight
    chunks, whole groups of eight, the 640-column block-Schur shape, fewer rows
    than a stage); A may sit in a wider row-major buffer (lda > n).  The oracle
    is numpy's solve of the damped normal equations in float64."""
    print(f"gram chunks m={m} n={n} lda_pad={lda_pad}", flush=True)
    rng = np.random.default_rng(m + n)
    A = rng.normal(size=(m, n))
    b = rng.normal(size=m)
    d = rng.uniform(0.5, 2.0, size=n)
    ref = np.linalg.solve(A.T @ A + np.diag(d), A.T @ b)
    lda = n + lda_pad
    buf = np.zeros((m, lda))
    buf[:, :n] = A
    lib = nat.load()
    nbytes = ctypes.c_size_t(0)
    assert lib.ct_dense_workspace_bytes(n, ctypes.byref(nbytes)) == 0
    ws = torch.empty(int(nbytes.value), dtype=torch.uint8, device="cuda")
    At, bt, dt = (torch.tensor(v, device="cuda") for v in (buf, b, d))
    x = torch.full((n,), float("nan"), dtype=torch.float64, device="cuda")
    rc = lib.ct_normal_equations_solve_f64(nat.ptr(At), m, n, lda, nat.ptr(bt), nat.ptr(dt), nat.ptr(x), nat.ptr(ws),
                                           ws.numel(), nat.current_stream_handle())
    assert rc == 0, lib.ct_last_error().decode()
    np.testing.assert_allclose(x.cpu().numpy(), ref, rtol=1e-9, atol=1e-11)
    # the same call again on the same workspace: the fixed-order reduce makes it bit-identical
    x2 = torch.empty_like(x)
    rc = lib.ct_normal_equations_solve_f64(nat.ptr(At), m, n, lda, nat.ptr(bt), nat.ptr(dt), nat.ptr(x2), nat.ptr(ws),
                                           ws.numel(), nat.current_stream_handle())
    assert rc == 0
    torch.cuda.synchronize()
    assert torch.equal(x, x2)


def test_block_schur_backend_same_solution_and_float32():
    rng = np.random.default_rng(5)
    A = rng.normal(size=(120, 30))
    b = rng.normal(size=120)
    d = rng.uniform(0.5, 1.0, size=30)
    ref = _damped_lstsq(A, b, d)
    r64 = L.rocm_block_schur(torch.tensor(A, device="cuda"), torch.tensor(b, device="cuda"),
                             damping=torch.tensor(d, device="cuda"), num_eliminate=12)
    np.testing.assert_allclose(r64.x.cpu().numpy(), ref, rtol=1e-9, atol=1e-11)
    r32 = L.rocm_block_schur(torch.tensor(A, device="cuda", dtype=torch.float32),
                             torch.tensor(b, device="cuda", dtype=torch.float32),
                             damping=torch.tensor(d, device="cuda", dtype=torch.float32), num_eliminate=12)
    assert r32.x.dtype == torch.float32
    np.testing.assert_allclose(r32.x.cpu().numpy(), ref, rtol=1e-4, atol=1e-5)


def test_registry_contract_declines_host_tensors_and_dispatches():
    L.clear_optional_backends()
    names = L.register_rocm_backends()
    assert set(names) == set(L.ROCM_BACKEND_NAMES)
    A = torch.randn(50, 8, dtype=torch.float64)
    b = torch.randn(50, dtype=torch.float64)
    assert L.call_optional_backend("block_schur", A, b, num_eliminate=3) is None  # CPU tensors: declined
    res = L.call_optional_backend("sparse_normal_cholesky", A.cuda(), b.cuda(), damping=None, num_eliminate=0,
                                  min_iterations=0, max_iterations=500)
    ref = np.linalg.lstsq(A.numpy(), b.numpy(), rcond=None)[0]
    np.testing.assert_allclose(res.x.cpu().numpy(), ref, rtol=1e-9, atol=1e-11)
    L.unregister_rocm_backends()
    assert all(L.get_optional_backend(n) is None for n in L.ROCM_BACKEND_NAMES)


def test_backends_in_a_foreign_registry_solve_on_the_device():
    # through a dispatcher with the reference's contract and its own classes:
    # a device solve comes back as that registry's LinearSolverResult
    import standin_registry as R

    L.register_rocm_backends(R)
    try:
        rng = np.random.default_rng(7)
        A = rng.normal(size=(60, 10))
        b = rng.normal(size=60)
        d = rng.uniform(0.1, 1.0, size=10)
        for name in L.ROCM_BACKEND_NAMES:
            res = R.try_optional_linear_backend(name, torch.tensor(A, device="cuda"), torch.tensor(b, device="cuda"),
                                                damping=torch.tensor(d, device="cuda"), num_eliminate=4,
                                                solver_type=None, block_sizes=None)
            assert isinstance(res, R.LinearSolverResult) and isinstance(res.summary, R.LinearSolverSummary)
            assert res.summary.success
            ref = np.linalg.solve(A.T @ A + np.diag(d), A.T @ b)
            np.testing.assert_allclose(res.x.cpu().numpy(), ref, rtol=1e-9, atol=1e-11)
            assert res.summary.residual_norm > 0.0
        # host tensors decline to None, so the reference falls back to its torch path
        assert R.try_optional_linear_backend("block_schur", torch.tensor(A), torch.tensor(b), num_eliminate=4) is None
    finally:
        L.unregister_rocm_backends(R)


# n <= 192: one-workgroup factorisation (tile boundaries at multiples of 16, the
# right-hand-side row in a tile of its own at n = 176 and 192); above: tiled
@pytest.mark.parametrize("n", [1, 9, 16, 17, 144, 175, 176, 191, 192, 193, 256, 257, 300, 320, 1000, 1800])
def test_standalone_potrf_potrs(n):
    rng = np.random.default_rng(n)
    M = rng.normal(size=(n, n))
    S = M @ M.T + n * np.eye(n)
    rhs = rng.normal(size=n)
    lib = nat.load()
    nbytes = ctypes.c_size_t(0)
    assert lib.ct_rcs_workspace_bytes(n, ctypes.byref(nbytes)) == 0
    ws = torch.zeros(int(nbytes.value), dtype=torch.uint8, device="cuda")
    St = torch.tensor(S, device="cuda")
    assert lib.ct_rcs_potrf_f64(nat.ptr(St), n, nat.ptr(ws), ws.numel(), nat.current_stream_handle()) == 0
    Lf = np.tril(St.cpu().numpy())
    np.testing.assert_allclose(Lf, np.linalg.cholesky(S), rtol=1e-10, atol=1e-10)
    x = torch.empty(n, dtype=torch.float64, device="cuda")
    for k in range(2):  # two solves with one factor
        r = torch.tensor(rhs * (k + 1), device="cuda")
        assert lib.ct_rcs_potrs_f64(nat.ptr(St), n, nat.ptr(r), nat.ptr(x), nat.ptr(ws), ws.numel(),
                                    nat.current_stream_handle()) == 0
        np.testing.assert_allclose(x.cpu().numpy(), np.linalg.solve(S, rhs * (k + 1)), rtol=1e-9, atol=1e-12)


def test_potrf_reports_not_positive_definite():
    n = 40
    S = np.eye(n)
    S[17, 17] = -1.0
    lib = nat.load()
    nbytes = ctypes.c_size_t(0)
    lib.ct_rcs_workspace_bytes(n, ctypes.byref(nbytes))
    ws = torch.zeros(int(nbytes.value), dtype=torch.uint8, device="cuda")
    St = torch.tensor(S, device="cuda")
    assert lib.ct_rcs_potrf_f64(nat.ptr(St), n, nat.ptr(ws), ws.numel(), nat.current_stream_handle()) == \
        nat.CT_ERR_NOT_SPD


def _factor(S, ws=None):
    """ct_rcs_potrf_f64 of a host SPD matrix; returns (status, factor tensor, workspace)."""
    n = S.shape[0]
    lib = nat.load()
    if ws is None:
        nbytes = ctypes.c_size_t(0)
        assert lib.ct_rcs_workspace_bytes(n, ctypes.byref(nbytes)) == 0
        ws = torch.zeros(int(nbytes.value), dtype=torch.uint8, device="cuda")
    St = torch.tensor(S, device="cuda")
    rc = lib.ct_rcs_potrf_f64(nat.ptr(St), n, nat.ptr(ws), ws.numel(), nat.current_stream_handle())
    return rc, St, ws


def _solve(St, ws, rhs):
    n = St.shape[0]
    x = torch.empty(n, dtype=torch.float64, device="cuda")
    r = torch.tensor(rhs, device="cuda")
    assert nat.load().ct_rcs_potrs_f64(nat.ptr(St), n, nat.ptr(r), nat.ptr(x), nat.ptr(ws), ws.numel(),
                                       nat.current_stream_handle()) == 0
    return x.cpu().numpy()


def test_potrf_small_reports_not_positive_definite_in_a_later_panel():
    # n <= 192: the bad pivot sits in the seventh 16-column panel
    n = 150
    S = np.eye(n)
    S[100, 100] = -1.0
    rc, _, _ = _factor(S)
    assert rc == nat.CT_ERR_NOT_SPD


def test_small_factor_ill_conditioned_and_repeatable():
    # one-workgroup path: condition number 1e8, repeated factorisations bit-identical
    rng = np.random.default_rng(6)
    n = 150
    Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    S = (Q * np.logspace(0.0, 8.0, n)) @ Q.T
    S = 0.5 * (S + S.T)
    b = rng.normal(size=n)
    rc, St, ws = _factor(S)
    assert rc == 0
    x = _solve(St, ws, b)
    backward = np.linalg.norm(S @ x - b) / (np.linalg.norm(S, 2) * np.linalg.norm(x) + np.linalg.norm(b))
    assert backward < 1e-12
    ref = np.linalg.solve(S, b)
    assert np.linalg.norm(x - ref) / np.linalg.norm(ref) < 1e-6
    rc2, St2, ws2 = _factor(S, ws)
    assert rc2 == 0
    np.testing.assert_array_equal(np.tril(St2.cpu().numpy()), np.tril(St.cpu().numpy()))
    np.testing.assert_array_equal(_solve(St2, ws2, b), x)


def test_potrf_tiled_reports_not_positive_definite():
    # n > 192 takes the tiled factorisation; the bad pivot sits in the fourth 64-column tile
    n = 300
    S = np.eye(n)
    S[200, 200] = -1.0
    rc, _, _ = _factor(S)
    assert rc == nat.CT_ERR_NOT_SPD


def test_refactor_on_one_workspace_is_bit_identical():
    # consecutive factorisations reuse the workspace's tile flags (a new epoch each launch)
    rng = np.random.default_rng(11)
    n = 700
    mats, rhss = [], []
    for _ in range(3):
        M = rng.normal(size=(n, n))
        mats.append(M @ M.T + n * np.eye(n))
        rhss.append(rng.normal(size=n))
    fresh = []
    for S, b in zip(mats, rhss):
        rc, St, ws = _factor(S)
        assert rc == 0
        fresh.append((np.tril(St.cpu().numpy()), _solve(St, ws, b)))
    ws = None
    for (S, b), (Lf, xf) in zip(zip(mats, rhss), fresh):
        rc, St, ws = _factor(S, ws)
        assert rc == 0
        np.testing.assert_array_equal(np.tril(St.cpu().numpy()), Lf)
        np.testing.assert_array_equal(_solve(St, ws, b), xf)
        np.testing.assert_allclose(_solve(St, ws, b), np.linalg.solve(S, b), rtol=1e-9, atol=1e-12)


def test_inverse_based_solve_on_ill_conditioned_system():
    # solves with the factor go through V = L^-1; condition number 1e8
    rng = np.random.default_rng(5)
    n = 500
    Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    S = (Q * np.logspace(0.0, 8.0, n)) @ Q.T
    S = 0.5 * (S + S.T)
    b = rng.normal(size=n)
    rc, St, ws = _factor(S)
    assert rc == 0
    x = _solve(St, ws, b)
    backward = np.linalg.norm(S @ x - b) / (np.linalg.norm(S, 2) * np.linalg.norm(x) + np.linalg.norm(b))
    assert backward < 1e-12
    ref = np.linalg.solve(S, b)
    assert np.linalg.norm(x - ref) / np.linalg.norm(ref) < 1e-6
