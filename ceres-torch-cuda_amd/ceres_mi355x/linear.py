"""Linear-solver plugin registry and the MI355X backends behind its names.

The reference lets accelerated solvers plug into ``solve_linear_system``
under fixed names through ``register_optional_backend(name, fn)``
(``linear.py:84-103``); a backend is ``fn(A, b, *, damping, num_eliminate,
**options) -> LinearSolverResult`` for a dense ``A [m, n]``, and declines a
call by raising ``OptionalBackendUnavailable`` (``linear.py:1311-1328``).  Its
CUDA tier registers ``sparse_normal_cholesky``, ``sparse_cholesky``,
``sparse_schur`` and ``block_schur`` (``cuda_backends.py:158-307``), each
solving the damped normal equations ``(A^T A + diag(damping)) x = A^T b``.

Here the same names (plus ``dense_schur``) are served by the HIP library:
``ct_normal_equations_solve_f64`` forms ``A^T A`` with a tiled Gram kernel and
solves it with the FP64 Cholesky of the reduced-camera-system solver.  The
Schur-complement names return the same solution (eliminating unknowns of an
SPD system does not change its solution).  The functions plug into the
reference's own registry too; they then raise its exception class and return
its result class::

    import ceres_torch.linear as ref_linear
    from ceres_mi355x.linear import register_rocm_backends
    register_rocm_backends(ref_linear)
"""
from __future__ import annotations

import ctypes
import sys
from dataclasses import dataclass
from typing import Any, Callable, Optional

import torch

from . import _native as nat


class OptionalBackendUnavailable(RuntimeError):
    """A backend declines a call (linear.py:84-85)."""


@dataclass
class LinearSolverSummary:
    """linear.py:12-17."""

    residual_norm: float
    num_iterations: int
    success: bool
    message: str = ""


@dataclass
class LinearSolverResult:
    """linear.py:20-23."""

    x: torch.Tensor
    summary: LinearSolverSummary


OptionalBackend = Callable[..., "torch.Tensor | LinearSolverResult"]
_optional_backends: dict[str, OptionalBackend] = {}


def register_optional_backend(name: str, backend: OptionalBackend) -> None:
    _optional_backends[name] = backend


def unregister_optional_backend(name: str) -> None:
    _optional_backends.pop(name, None)


def clear_optional_backends() -> None:
    _optional_backends.clear()


def get_optional_backend(name: str) -> Optional[OptionalBackend]:
    return _optional_backends.get(name)


ROCM_BACKEND_NAMES = ("sparse_normal_cholesky", "sparse_cholesky", "sparse_schur", "block_schur", "dense_schur")


def _require_device_tensor(t: torch.Tensor, name: str) -> None:
    """cuda_backends.py:322-326: decline host tensors and non-float dtypes."""
    if t.device.type != "cuda":
        raise OptionalBackendUnavailable(f"{name} must be a GPU tensor for the MI355X backend.")
    if t.dtype not in (torch.float32, torch.float64):
        raise OptionalBackendUnavailable(f"{name} must be float32 or float64 for the MI355X backend.")


def _summary(A, b, x, damping, message: str) -> LinearSolverSummary:
    """Residual of the damped least-squares problem, as cuda_backends.py:336-354 reports it."""
    residual = A @ x - b.reshape(-1).to(dtype=A.dtype, device=A.device)
    sq = torch.sum(residual * residual)
    if damping is not None:
        d = damping.reshape(-1).to(dtype=A.dtype, device=A.device)
        sq = sq + torch.sum(torch.clamp(d, min=0.0) * x * x)
    return LinearSolverSummary(float(torch.sqrt(torch.clamp(sq, min=0.0))), 1, True, message)


def _solve(A: torch.Tensor, b: torch.Tensor, damping: Optional[torch.Tensor], num_eliminate: Optional[int],
           message: str) -> LinearSolverResult:
    _require_device_tensor(A, "A")
    _require_device_tensor(b, "b")
    if A.ndim != 2:
        raise ValueError("A must be a 2D matrix")
    m, n = int(A.shape[0]), int(A.shape[1])
    b = b.reshape(-1)
    if b.numel() != m:
        raise ValueError("b must have one entry per row of A")
    # the kernels compute in FP64; FP32 inputs are widened and the result narrowed back
    A64 = A.detach().to(torch.float64).contiguous()
    b64 = b.detach().to(torch.float64).contiguous()
    d64 = None
    if damping is not None:
        _require_device_tensor(damping, "damping")
        d64 = damping.detach().reshape(-1).to(device=A.device, dtype=torch.float64).contiguous()
        if d64.numel() != n:
            raise ValueError("damping must have one entry per column of A")
    x = torch.empty(n, dtype=torch.float64, device=A.device)
    lib = nat.load()
    nbytes = ctypes.c_size_t(0)
    nat.check(lib.ct_dense_workspace_bytes(n, ctypes.byref(nbytes)), "ct_dense_workspace_bytes")
    ws = torch.empty(int(nbytes.value), dtype=torch.uint8, device=A.device)
    if num_eliminate is None:
        rc = lib.ct_normal_equations_solve_f64(nat.ptr(A64), m, n, n, nat.ptr(b64), nat.ptr(d64), nat.ptr(x),
                                               nat.ptr(ws), ws.numel(), nat.current_stream_handle())
    else:
        rc = lib.ct_block_schur_solve_f64(nat.ptr(A64), m, n, n, nat.ptr(b64), int(num_eliminate), nat.ptr(d64),
                                          nat.ptr(x), nat.ptr(ws), ws.numel(), nat.current_stream_handle())
    if rc == nat.CT_ERR_NOT_SPD:
        # the reference's torch.linalg.solve would still return (LU); this backend declines instead
        raise OptionalBackendUnavailable(nat.load().ct_last_error().decode())
    nat.check(rc, "dense normal-equations solve")
    x = x.to(A.dtype)
    return LinearSolverResult(x, _summary(A, b, x, damping, message))


def rocm_normal_equations(A: torch.Tensor, b: torch.Tensor, *, damping: Optional[torch.Tensor] = None,
                          **_: Any) -> LinearSolverResult:
    """Backend for ``sparse_normal_cholesky`` / ``sparse_cholesky`` (cuda_backends.py:215-229, 267-282)."""
    return _solve(A, b, damping, None, "mi355x normal equations (HIP Gram + FP64 Cholesky)")


def rocm_block_schur(A: torch.Tensor, b: torch.Tensor, *, damping: Optional[torch.Tensor] = None,
                     num_eliminate: int = 0, **_: Any) -> LinearSolverResult:
    """Backend for ``sparse_schur`` / ``block_schur`` / ``dense_schur`` (cuda_backends.py:232-264, 285-306)."""
    return _solve(A, b, damping, int(num_eliminate), "mi355x block schur (HIP Gram + FP64 Cholesky)")


def _adapt_to_registry(fn: OptionalBackend, registry: Any, name: str) -> OptionalBackend:
    """Wrap a backend so that it speaks a foreign registry's types.

    The reference's dispatcher (``linear.py:1311-1328``) falls back only on its
    own ``OptionalBackendUnavailable`` and accepts a ``LinearSolverResult``
    only when it is an instance of its own class; anything else it passes to
    ``_summarize_solution`` as the solution tensor.  So a decline is re-raised
    as the owner's exception and a result is rebuilt from the owner's
    ``LinearSolverResult`` / ``LinearSolverSummary``.
    """
    owner_unavailable = registry.OptionalBackendUnavailable
    owner_result = registry.LinearSolverResult
    owner_summary = registry.LinearSolverSummary

    def backend(A: torch.Tensor, b: torch.Tensor, **kwargs: Any):
        try:
            res = fn(A, b, **kwargs)
        except OptionalBackendUnavailable as exc:
            raise owner_unavailable(str(exc)) from exc
        s = res.summary
        return owner_result(res.x, owner_summary(s.residual_norm, s.num_iterations, s.success, s.message))

    backend.__name__ = backend.__qualname__ = f"{fn.__name__}_for_{name}"
    backend.__doc__ = fn.__doc__
    return backend


def register_rocm_backends(registry: Any = None, *, overwrite: bool = True) -> tuple[str, ...]:
    """Register the MI355X backends under the reference's plugin names (cuda_backends.py:158-207).

    ``registry`` is the module that owns the registry: this module by default,
    or the reference's ``ceres_torch.linear``.  It must provide
    ``register_optional_backend``, ``get_optional_backend``,
    ``OptionalBackendUnavailable``, ``LinearSolverResult`` and
    ``LinearSolverSummary``; for a foreign registry the backends raise its
    exception class and return its result class.  Returns the names registered.
    """
    nat.load()
    backends = {
        "sparse_normal_cholesky": rocm_normal_equations,
        "sparse_cholesky": rocm_normal_equations,
        "sparse_schur": rocm_block_schur,
        "block_schur": rocm_block_schur,
        "dense_schur": rocm_block_schur,
    }
    own = sys.modules[__name__]
    if registry is None:
        registry = own
    for attr in ("register_optional_backend", "get_optional_backend", "OptionalBackendUnavailable",
                 "LinearSolverResult", "LinearSolverSummary"):
        if not hasattr(registry, attr):
            raise TypeError(f"registry {registry!r} has no attribute {attr!r}")
    done = []
    for name, fn in backends.items():
        if overwrite or registry.get_optional_backend(name) is None:
            registry.register_optional_backend(name, fn if registry is own else _adapt_to_registry(fn, registry, name))
            done.append(name)
    return tuple(done)


def unregister_rocm_backends(registry: Any = None) -> None:
    """Remove the names :func:`register_rocm_backends` registers from ``registry``."""
    unregister = (registry or sys.modules[__name__]).unregister_optional_backend
    for name in ROCM_BACKEND_NAMES:
        unregister(name)


def call_optional_backend(name: str, A: torch.Tensor, b: torch.Tensor, **kwargs: Any) -> Optional[LinearSolverResult]:
    """The reference's dispatch contract for one name (linear.py:1311-1328): None when absent or declined."""
    backend = get_optional_backend(name)
    if backend is None:
        return None
    try:
        raw = backend(A, b, **kwargs)
    except TypeError:
        raw = backend(A, b)
    except OptionalBackendUnavailable:
        return None
    if isinstance(raw, LinearSolverResult):
        return raw
    return LinearSolverResult(raw, _summary(A, b, raw, kwargs.get("damping"), f"optional {name} backend"))
