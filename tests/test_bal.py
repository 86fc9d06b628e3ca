"""BAL file reader (examples/simple_bundle_adjuster.py:9-70) on the CPU; the
tiny-BAL solve of tests/test_examples.py:151-165 runs in tests/test_ba_gpu.py."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from ceres_mi355x.bal import BALProblem, build_problem
from oracle import ba as oba


def tiny_bal_text():
    """The reference example's tiny problem (simple_bundle_adjuster.py:142-176): 2 cameras, 3 points."""
    true_cams = np.array([[0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 800.0, 0.0, 0.0],
                          [0.0, 0.02, 0.0, -0.4, 0.0, 0.05, 800.0, 0.0, 0.0]])
    true_pts = np.array([[0.2, -0.1, 4.0], [-0.3, 0.15, 4.5], [0.1, 0.2, 5.0]])
    init_pts = true_pts + np.array([[0.04, -0.03, 0.15], [-0.05, 0.02, -0.12], [0.03, 0.04, 0.10]])
    rows = []
    for c in range(2):
        for p in range(3):
            x, y = oba.project(true_cams[c][None], true_pts[p][None])[0]
            rows.append(f"{c} {p} {x:.17g} {y:.17g}")
    lines = ["2 3 6", *rows, *(f"{v:.17g}" for v in true_cams.reshape(-1)), *(f"{v:.17g}" for v in init_pts.reshape(-1))]
    return "\n".join(lines) + "\n", true_cams, true_pts, init_pts


def test_reads_the_reference_layout():
    text, cams, _, init_pts = tiny_bal_text()
    bal = BALProblem.from_text(text)
    assert (bal.num_cameras, bal.num_points, bal.num_observations) == (2, 3, 6)
    assert bal.camera_indices == [0, 0, 0, 1, 1, 1]
    assert bal.point_indices == [0, 1, 2, 0, 1, 2]
    np.testing.assert_array_equal(bal.cameras.numpy(), cams)
    np.testing.assert_array_equal(bal.points.numpy(), init_pts)
    assert bal.cameras.dtype == torch.float64


def test_round_trip_is_exact(tmp_path):
    rng = np.random.default_rng(4)
    C, P, N = 5, 40, 120
    text = "\n".join([f"{C} {P} {N}",
                      *(f"{rng.integers(C)} {rng.integers(P)} {rng.normal():.17g} {rng.normal():.17g}" for _ in range(N)),
                      *(f"{v:.17g}" for v in rng.normal(size=9 * C + 3 * P))]) + "\n"
    path = tmp_path / "p.bal"
    path.write_text(text)
    a = BALProblem.from_file(path)
    b = BALProblem.from_text(a.to_text())
    assert a.camera_indices == b.camera_indices and a.point_indices == b.point_indices
    for x, y in ((a.cameras, b.cameras), (a.points, b.points), (a.observations, b.observations)):
        assert torch.equal(x, y)


@pytest.mark.parametrize("text,msg", [
    ("1 1 1\n0 0 1.0 2.0\n" + "0 " * 12 + "7\n", "trailing values"),
    ("1 1 1\n0 0 1.0 2.0\n" + "0 " * 11, "ends early"),
    ("1 1 1\n0 3 1.0 2.0\n" + "0 " * 12, "missing camera or point"),
    ("1 1 1\n0 0.5 1.0 2.0\n" + "0 " * 12, "integers"),
    ("1 x 1\n", "three integers"),
])
def test_malformed_input_is_rejected(text, msg):
    with pytest.raises(ValueError, match=msg):
        BALProblem.from_text(text)


def test_build_problem_registers_cameras_then_points():
    text, *_ = tiny_bal_text()
    bal = BALProblem.from_text(text)
    problem, cams, pts = build_problem(bal, fix_cameras=True)
    assert problem.num_parameter_blocks() == 5
    assert all(problem.parameter_blocks[i].constant for i in range(2))
    assert not any(problem.parameter_blocks[i].constant for i in range(2, 5))
    assert all(problem.parameter_blocks[i].ordering_group == 0 for i in range(2, 5))
    assert problem.num_residuals() == 12
    assert cams is bal.cameras and pts is bal.points
