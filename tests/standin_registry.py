"""A stand-in for the reference's ``ceres_torch.linear`` registry.

It has its own ``OptionalBackendUnavailable``, ``LinearSolverSummary`` and
``LinearSolverResult`` classes, its own registry dict, and a dispatcher with
the reference's try/except/isinstance contract (``linear.py:84-103`` for the
registry, ``linear.py:1311-1328`` for ``_try_optional_linear_backend``,
``linear.py:1373-1383`` for ``_summarize_solution``).  Tests plug the MI355X
backends into it to show they interoperate with a registry whose classes are
not the package's own.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Any, Callable, Optional

import torch


@dataclass
class LinearSolverSummary:
    residual_norm: float
    num_iterations: int
    success: bool
    message: str = ""


@dataclass
class LinearSolverResult:
    x: torch.Tensor
    summary: LinearSolverSummary


class OptionalBackendUnavailable(RuntimeError):
    pass


_backends: dict[str, Callable[..., Any]] = {}


def register_optional_backend(name: str, backend: Callable[..., Any]) -> None:
    _backends[name] = backend


def unregister_optional_backend(name: str) -> None:
    _backends.pop(name, None)


def get_optional_backend(name: str) -> Optional[Callable[..., Any]]:
    return _backends.get(name)


def summarize_solution(A: torch.Tensor, b: torch.Tensor, x: torch.Tensor, iterations: int, message: str):
    # a non-tensor x fails here exactly as it does in the reference
    residual_norm = float(torch.linalg.norm(A @ x.reshape(-1) - b.reshape(-1)).detach().cpu())
    return LinearSolverResult(x.reshape(-1), LinearSolverSummary(residual_norm, iterations, True, message))


def try_optional_linear_backend(name: str, A: torch.Tensor, b: torch.Tensor, **kwargs: Any):
    backend = get_optional_backend(name)
    if backend is None:
        return None
    try:
        raw_result = backend(A, b, **kwargs)
    except TypeError:
        raw_result = backend(A, b)
    except OptionalBackendUnavailable:
        return None
    if isinstance(raw_result, LinearSolverResult):
        return raw_result
    return summarize_solution(A, b, raw_result, 1, f"optional {name} backend")
