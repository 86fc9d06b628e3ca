"""CPU-side checks of the C ABI and the host layer (no GPU compute)."""
from __future__ import annotations

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

import ceres_mi355x as tc
from ceres_mi355x import _native as nat
from ceres_mi355x.ba import build_tables

HEADER = Path(__file__).resolve().parents[1] / "include" / "ceres_mi355x.h"


def declared_functions() -> list[str]:
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ct_[a-z0-9_]+)\s*\(", text)))


def test_library_loads_and_exports_every_declared_symbol():
    lib = nat.load()
    names = declared_functions()
    assert names, "header declares no functions"
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(nat.exported_symbols())
    assert lib.ct_abi_version() == nat.ABI_VERSION
    header_version = re.search(r"#define CT_ABI_VERSION (\d+)", HEADER.read_text()).group(1)
    assert int(header_version) == nat.ABI_VERSION


def _problem_struct(C=3, P=4, N=8):
    p = nat.ct_ba_problem()
    p.num_cameras, p.num_points, p.num_observations = C, P, N
    p.num_blocks = C * (C + 1) // 2
    p.num_terms = 10
    fake = ctypes.c_void_p(0x1000)
    for f in ("obs_camera", "obs_point", "point_offsets", "observations", "camera_offsets", "camera_obs", "block_cameras",
              "block_offsets", "block_terms"):
        setattr(p, f, fake)
    p.loss = nat.ct_loss(0, 0.0, 0.0, 1.0)
    return p


def test_workspace_query_and_argument_errors():
    lib = nat.load()
    opts = tc.solver.lm_options(tc.SolverOptions(max_num_iterations=10))
    nbytes = ctypes.c_size_t(0)
    assert lib.ct_ba_workspace_bytes(ctypes.byref(_problem_struct()), ctypes.byref(opts), ctypes.byref(nbytes)) == 0
    assert nbytes.value > 0
    bad = _problem_struct()
    bad.num_points = -1
    assert lib.ct_ba_workspace_bytes(ctypes.byref(bad), ctypes.byref(opts), ctypes.byref(nbytes)) == nat.CT_ERR_ARGUMENT
    assert b"negative" in lib.ct_last_error()
    bad = _problem_struct()
    bad.loss = nat.ct_loss(42, 0.0, 0.0, 1.0)
    assert lib.ct_ba_workspace_bytes(ctypes.byref(bad), ctypes.byref(opts), ctypes.byref(nbytes)) == nat.CT_ERR_ARGUMENT
    assert b"loss" in lib.ct_last_error()
    opts.linear_solver = 7
    assert lib.ct_ba_workspace_bytes(ctypes.byref(_problem_struct()), ctypes.byref(opts), ctypes.byref(nbytes)) == nat.CT_ERR_ARGUMENT
    opts.linear_solver = 0
    opts.trust_region_strategy = 9
    assert lib.ct_ba_workspace_bytes(ctypes.byref(_problem_struct()), ctypes.byref(opts), ctypes.byref(nbytes)) == nat.CT_ERR_ARGUMENT
    assert b"strategy" in lib.ct_last_error()
    opts.trust_region_strategy = 1  # DOGLEG sizes the Cholesky workspace even for ITERATIVE_SCHUR
    opts.linear_solver = 1
    assert lib.ct_ba_workspace_bytes(ctypes.byref(_problem_struct()), ctypes.byref(opts), ctypes.byref(nbytes)) == 0


def test_plan_create_rejects_small_workspace():
    lib = nat.load()
    opts = tc.solver.lm_options(tc.SolverOptions(max_num_iterations=10))
    handle = ctypes.c_void_p()
    st = lib.ct_ba_plan_create(ctypes.byref(_problem_struct()), ctypes.byref(opts), ctypes.c_void_p(0x10),
                               ctypes.c_void_p(0x10), ctypes.c_void_p(0x10), 16, None, ctypes.byref(handle))
    assert st == nat.CT_ERR_WORKSPACE
    assert b"workspace too small" in lib.ct_last_error()


def test_pair_tables_match_brute_force():
    rng = np.random.default_rng(0)
    C, P = 5, 30
    oc, op = [], []
    for p in range(P):
        cams = rng.choice(C, size=rng.integers(1, 5), replace=False)
        oc += list(cams)
        op += [p] * len(cams)
    perm = rng.permutation(len(oc))
    oc, op = np.array(oc)[perm], np.array(op)[perm]
    t = build_tables(oc, op, C, P)
    cam = t.obs_camera
    pt = np.asarray(op)[t.order]
    assert np.all(np.diff(pt) >= 0)
    # every block (i <= j) holds exactly the observation pairs of common points
    for b, (i, j) in enumerate(t.block_cameras):
        terms = t.block_terms[t.block_offsets[b]: t.block_offsets[b + 1]]
        want = sorted((a, c) for a in range(len(cam)) for c in range(len(cam))
                      if pt[a] == pt[c] and cam[a] == i and cam[c] == j)
        assert sorted(map(tuple, terms)) == want
    for i in range(C):
        obs = t.camera_obs[t.camera_offsets[i]: t.camera_offsets[i + 1]]
        assert sorted(obs) == sorted(np.nonzero(cam == i)[0])


def test_options_validation_messages_match_reference():
    # tests/test_problem_solver.py:151-173 of the reference
    with pytest.raises(ValueError, match="DOGLEG only supports exact factorization based linear solvers"):
        tc.SolverOptions(trust_region_strategy_type=tc.TrustRegionStrategyType.DOGLEG,
                         linear_solver_type=tc.LinearSolverType.ITERATIVE_SCHUR).validate()
    with pytest.raises(ValueError, match="use_mixed_precision_solves cannot be used with DENSE_QR"):
        tc.SolverOptions(use_mixed_precision_solves=True).validate()
    with pytest.raises(ValueError, match="max_num_iterations must be >= 0"):
        tc.SolverOptions(max_num_iterations=-1).validate()
    with pytest.raises(ValueError, match="trust region radii must satisfy min <= initial <= max"):
        tc.SolverOptions(initial_trust_region_radius=1e20).validate()
    tc.SolverOptions(linear_solver_type=tc.LinearSolverType.DENSE_SCHUR, use_mixed_precision_solves=True,
                     max_num_refinement_iterations=2).validate()


def test_summary_reports_keep_reference_fields():
    s = tc.SolverSummary()
    assert not s.IsSolutionUsable()
    assert "Termination: FAILURE" in s.BriefReport()
    full = s.FullReport()
    for key in ("Jacobian evaluation time", "Linear solver time", "Preprocessor time", "Preconditioner",
                "Line search interpolation", "Inner iteration steps"):
        assert key in full


def test_solve_refuses_unsupported_problems_without_fallback():
    import torch

    x = torch.tensor([0.5], dtype=torch.float64)
    problem = tc.Problem()
    problem.AddResidualBlock(tc.AutoDiffCostFunction(lambda x: 10.0 - x, [1], 1), None, [x])
    with pytest.raises((NotImplementedError, tc.NativeLibraryError)):
        tc.solve(tc.SolverOptions(), problem)
    # DOGLEG runs on the device for bundle adjustment only
    with pytest.raises(NotImplementedError, match="DOGLEG"):
        tc.solve(tc.SolverOptions(trust_region_strategy_type=tc.TrustRegionStrategyType.DOGLEG), problem)


def test_losses_native_descriptors_match_torch_values():
    import torch

    s = torch.tensor([0.3, 4.0], dtype=torch.float64)
    from oracle.losses import loss_rho

    for loss in [tc.TrivialLoss(), tc.HuberLoss(1.0), tc.CauchyLoss(0.5), tc.SoftLOneLoss(0.7), tc.ArctanLoss(2.0),
                 tc.TolerantLoss(1.0, 0.5), tc.TukeyLoss(3.0), tc.ScaledLoss(tc.CauchyLoss(0.5), 2.0)]:
        mine = [t.numpy() for t in loss.evaluate(s)]
        ref = loss_rho(tc.losses.oracle_loss(loss), s.numpy())
        for a, b in zip(mine, ref):
            np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-15)


def test_kernel_group_names_follow_header_enum():
    text = HEADER.read_text()
    groups = re.findall(r"CT_KERNEL_([A-Z_]+) = (\d+)", text)
    names = [name.lower() for name, idx in sorted(groups, key=lambda g: int(g[1])) if name != "COUNT"]
    assert names == nat.KERNELS
    assert int(dict(groups)["COUNT"]) == len(nat.KERNELS)


def test_registry_declines_host_tensors_like_the_reference():
    import torch

    from ceres_mi355x import linear as L

    L.clear_optional_backends()
    assert set(L.register_rocm_backends()) == set(L.ROCM_BACKEND_NAMES)
    A = torch.randn(20, 4, dtype=torch.float64)
    b = torch.randn(20, dtype=torch.float64)
    for name in L.ROCM_BACKEND_NAMES:
        assert L.call_optional_backend(name, A, b, damping=None, num_eliminate=1) is None
    with pytest.raises(L.OptionalBackendUnavailable):
        L.rocm_normal_equations(A.to(torch.int64), b)
    # overwrite=False keeps an existing registration
    L.register_optional_backend("block_schur", lambda A, b, **kw: "mine")
    L.register_rocm_backends(overwrite=False)
    assert L.get_optional_backend("block_schur")(A, b) == "mine"
    L.unregister_rocm_backends()
    assert all(L.get_optional_backend(n) is None for n in L.ROCM_BACKEND_NAMES)


def test_backends_in_a_foreign_registry_decline_with_its_exception():
    # the reference's dispatcher catches only its own OptionalBackendUnavailable
    # (linear.py:1311-1328): host tensors must come back as None through it
    import torch

    import standin_registry as R
    from ceres_mi355x import linear as L

    names = L.register_rocm_backends(R)
    try:
        assert set(names) == set(L.ROCM_BACKEND_NAMES)
        A = torch.randn(20, 4, dtype=torch.float64)
        b = torch.randn(20, dtype=torch.float64)
        for name in L.ROCM_BACKEND_NAMES:
            assert R.try_optional_linear_backend(name, A, b, damping=None, num_eliminate=1,
                                                 solver_type=None, block_sizes=None) is None
        with pytest.raises(R.OptionalBackendUnavailable):
            R.get_optional_backend("block_schur")(A.to(torch.int64), b, num_eliminate=1)
        # nothing registered in the package's own registry
        assert all(L.get_optional_backend(n) is None for n in L.ROCM_BACKEND_NAMES)
        # overwrite=False keeps the owner's existing registration
        R.register_optional_backend("block_schur", lambda A, b, **kw: "theirs")
        assert "block_schur" not in L.register_rocm_backends(R, overwrite=False)
        assert R.get_optional_backend("block_schur")(A, b) == "theirs"
    finally:
        L.unregister_rocm_backends(R)
    assert all(R.get_optional_backend(n) is None for n in L.ROCM_BACKEND_NAMES)
    with pytest.raises(TypeError):
        L.register_rocm_backends(object())


def test_crs_matrix_round_trip_and_layout():
    # problem.py:15-41 of the reference: nonzeros row by row, rows = row starts
    import torch

    m = torch.tensor([[0.0, 2.0, 0.0], [0.0, 0.0, 0.0], [1.5, 0.0, -3.0]], dtype=torch.float64)
    crs = tc.CRSMatrix.from_dense(m)
    assert (crs.num_rows, crs.num_cols) == (3, 3)
    assert crs.rows == [0, 1, 1, 3]
    assert crs.cols == [1, 0, 2]
    assert crs.values == [2.0, 1.5, -3.0]
    assert torch.equal(crs.to_dense(), m)
    empty = tc.CRSMatrix.from_dense(torch.zeros(2, 4, dtype=torch.float64))
    assert empty.rows == [0, 0, 0] and empty.cols == [] and torch.equal(empty.to_dense(), torch.zeros(2, 4, dtype=torch.float64))


def test_iterative_schur_needs_the_points_as_the_eliminated_group():
    # solver.py:714-741: the reference eliminates the first ordering group; without groups
    # ITERATIVE_SCHUR is CG on the normal equations (linear.py:248-257), which the device path refuses
    import torch

    from ceres_mi355x.solver import _BAExtract, _check_elimination, _eliminated_blocks

    cams = torch.zeros(2, 9, dtype=torch.float64)
    pts = torch.zeros(3, 3, dtype=torch.float64)
    oc, op = np.array([0, 1, 0, 1, 0, 1]), np.array([0, 0, 1, 1, 2, 2])
    obs = np.zeros((6, 2))
    opts = tc.SolverOptions(linear_solver_type=tc.LinearSolverType.ITERATIVE_SCHUR,
                            preconditioner_type=tc.PreconditionerType.SCHUR_JACOBI)
    problem = tc.Problem()
    problem.add_bundle_adjustment_observations(cams, pts, oc, op, obs)
    ex = _BAExtract(problem)
    assert _eliminated_blocks(problem) == {id(b) for b in ex.pt_blocks}
    _check_elimination(problem, ex, opts)
    problem2 = tc.Problem()
    problem2.add_bundle_adjustment_observations(cams, pts, oc, op, obs, point_ordering_group=None)
    assert _eliminated_blocks(problem2) == set()
    with pytest.raises(NotImplementedError, match="ordering group"):
        _check_elimination(problem2, _BAExtract(problem2), opts)
    problem2.SetParameterBlockOrderingGroup(problem2.parameter_blocks[0], 0)  # a camera first
    with pytest.raises(NotImplementedError, match="ordering group"):
        _check_elimination(problem2, _BAExtract(problem2), opts)
    # exact solvers take any order
    _check_elimination(problem2, _BAExtract(problem2), tc.SolverOptions(linear_solver_type=tc.LinearSolverType.DENSE_SCHUR))


def test_pose2d_system_size_query_and_workspace_bounds():
    """ct_pose2d_plan_system_size rejects null arguments without touching the device; the workspace
    query bounds the separator tables by the node and pair counts (it grows with the pairs)."""
    lib = nat.load()
    n = ctypes.c_int32(-7)
    assert lib.ct_pose2d_plan_system_size(None, ctypes.byref(n)) == nat.CT_ERR_ARGUMENT
    assert lib.ct_pose2d_plan_system_size(None, None) == nat.CT_ERR_ARGUMENT
    assert n.value == -7
    opts = tc.solver.lm_options(tc.SolverOptions(max_num_iterations=10))
    sizes = []
    for pairs in (10, 100):
        p = nat.ct_pose2d_problem()
        p.num_nodes, p.num_edges, p.num_pairs, p.fixed_node = 50, 0, pairs, 0  # sizes only: no tables needed
        nbytes = ctypes.c_size_t(0)
        assert lib.ct_pose2d_workspace_bytes(ctypes.byref(p), ctypes.byref(opts), ctypes.byref(nbytes)) == 0
        sizes.append(nbytes.value)
    assert sizes[1] > sizes[0] > 0
