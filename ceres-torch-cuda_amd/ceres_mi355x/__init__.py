"""ceres_mi355x: MI355X-native trust-region solver core with the ceres_torch API.

``import ceres_mi355x as tc`` gives the names code written against the
reference's ``ceres_torch`` (``src/ceres_torch/__init__.py``) uses on the
least-squares path: ``Problem``, cost and loss functions, manifolds,
``SolverOptions``, ``solve``, ``SolverSummary`` and the linear-solver plugin
registry.  The compute path is the HIP library ``libceres_mi355x.so``.
"""
from .costs import (
    AnalyticCostFunction,
    AutoDiffCostFunction,
    CallableCostFunction,
    CostFunction,
    DynamicAutoDiffCostFunction,
    ExpCurveResidual,
    Pose2DRelativeError,
    QuaternionReprojectionError,
    SnavelyReprojectionError,
    quaternion_reprojection,
    relative_pose,
    snavely_reprojection,
)
from .losses import (
    ArctanLoss,
    CauchyLoss,
    ComposedLoss,
    HuberLoss,
    LossFunction,
    LossFunctionWrapper,
    ScaledLoss,
    SoftLOneLoss,
    TolerantLoss,
    TrivialLoss,
    TukeyLoss,
)
from .manifolds import EigenQuaternionManifold, EuclideanManifold, Manifold, QuaternionManifold, SubsetManifold
from .problem import CRSMatrix, EvaluateOptions, EvaluationResult, ParameterBlock, Problem, ResidualBlock
from .bal import BALProblem
from .covariance import Covariance, CovarianceOptions, CovarianceSummary
from .rotation import (
    angle_axis_rotate_point,
    angle_axis_to_quaternion,
    convert_ceres_to_eigen_quaternion,
    convert_eigen_to_ceres_quaternion,
    normalize_quaternion,
    quaternion_conjugate,
    quaternion_product,
    quaternion_rotate_point,
    unit_quaternion_rotate_point,
)
from .solver import set_device_loop, solve
from .types import (
    CallbackReturnType,
    CovarianceAlgorithmType,
    DenseLinearAlgebraLibraryType,
    DoglegType,
    IterationSummary,
    LinearSolverOrderingType,
    LinearSolverType,
    LineSearchDirectionType,
    LineSearchInterpolationType,
    LineSearchType,
    LoggingType,
    MinimizerType,
    NonlinearConjugateGradientType,
    PreconditionerType,
    SolverOptions,
    SolverSummary,
    SparseLinearAlgebraLibraryType,
    TerminationType,
    TrustRegionStrategyType,
    VisibilityClusteringType,
)
from ._native import NativeLibraryError

Solve = solve

__all__ = [name for name in dir() if not name.startswith("_")]
