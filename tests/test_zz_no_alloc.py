"""The no-allocation dense solve against the workspace form, through the C-ABI.

``ct_normal_equations_solve_f64_noalloc`` runs the same kernels as
``ct_normal_equations_solve_f64`` with its workspace in a static device array,
so on the same inputs the two must agree bit for bit.  Two calls in a row, the
second smaller, show that what the first call left in the shared array does
not leak into the second (the tiled factorisation's control words and the
state word must start at zero on every call).
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

from ceres_mi355x import _native as nat

pytestmark = pytest.mark.gpu


def _with_workspace(lib, A, b, d, x):
    m, n = A.shape
    nbytes = ctypes.c_size_t(0)
    assert lib.ct_dense_workspace_bytes(n, ctypes.byref(nbytes)) == 0
    ws = torch.empty(int(nbytes.value), dtype=torch.uint8, device="cuda")
    rc = lib.ct_normal_equations_solve_f64(nat.ptr(A), m, n, n, nat.ptr(b), nat.ptr(d), nat.ptr(x), nat.ptr(ws),
                                           ws.numel(), nat.current_stream_handle())
    assert rc == 0, lib.ct_last_error().decode()


def _no_alloc(lib, A, b, d, x):
    m, n = A.shape
    rc = lib.ct_normal_equations_solve_f64_noalloc(nat.ptr(A), m, n, n, nat.ptr(b), nat.ptr(d), nat.ptr(x),
                                                   nat.current_stream_handle())
    assert rc == 0, lib.ct_last_error().decode()


def _case(n, seed):
    rng = np.random.default_rng(seed)
    m = n + 37
    A = torch.tensor(rng.normal(size=(m, n)), device="cuda")
    b = torch.tensor(rng.normal(size=m), device="cuda")
    d = torch.tensor(rng.uniform(0.1, 1.0, size=n), device="cuda")
    return A, b, d


@pytest.mark.parametrize("sizes", [(400, 300), (512, 100), (150, 40)])
def test_no_alloc_solve_matches_the_workspace_solve(sizes):
    # tiled (n > 192) then smaller tiled / one-workgroup factorisations
    print(f"no-alloc sizes={sizes}", flush=True)
    lib = nat.load()
    for k, n in enumerate(sizes):
        A, b, d = _case(n, 100 + k + n)
        x_ref = torch.empty(n, dtype=torch.float64, device="cuda")
        x_new = torch.full((n,), float("nan"), dtype=torch.float64, device="cuda")
        _with_workspace(lib, A, b, d, x_ref)
        _no_alloc(lib, A, b, d, x_new)
        torch.cuda.synchronize()
        assert torch.equal(x_ref, x_new), (n, float((x_ref - x_new).abs().max()))
        An, bn, dn = A.cpu().numpy(), b.cpu().numpy(), d.cpu().numpy()
        exact = np.linalg.solve(An.T @ An + np.diag(dn), An.T @ bn)
        np.testing.assert_allclose(x_new.cpu().numpy(), exact, rtol=1e-9, atol=1e-11)


def test_no_alloc_solve_rejects_sizes_beyond_its_limit():
    lib = nat.load()
    n = 513
    A, b, d = _case(8, 1)  # never read: the size check comes first
    x = torch.empty(8, dtype=torch.float64, device="cuda")
    rc = lib.ct_normal_equations_solve_f64_noalloc(nat.ptr(A), n + 1, n, n, nat.ptr(b), nat.ptr(d), nat.ptr(x),
                                                   nat.current_stream_handle())
    assert rc == nat.CT_ERR_ARGUMENT
    assert "no-allocation limit" in lib.ct_last_error().decode()
