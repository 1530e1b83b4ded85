"""mycel_amd -- MI355X (gfx950) implementation of Mycel.Core's matrix-free Hex8 elasticity path.

The compute path is the HIP library libmycel_hip.so (C ABI in include/mycel_hip.h). This package is
the host-side mirror of the reference's managed API (Mycel.Core.Fem / Mycel.Core.Optimization)
over that ABI. There is no CPU fallback: importing the API without the built library raises.
"""
from ._lib import (MATVEC_AUTO, MATVEC_ELEMENT, MATVEC_NODE_GATHER, MATVEC_REFERENCE_ORDER, PRECISION_FP64,
                   PRECISION_FP64_ORDERED, PRECISION_MIXED_BF16,
                   PRECISION_MIXED_FP32, PRECISION_MIXED_GMG, ArgumentError,
                   ArgumentNullError, ArgumentOutOfRangeError, DeviceError, InvalidOperationError, MycelError,
                   OperationCanceledError, OutOfMemoryError)
from .fem import (FeMaterial, FemOperator, FeModel, GridSpec, Hex8Element, JacobiPreconditioner, LoadCase,
                  MultigridPreconditioner, PcgResult, PcgSolver, StaticAnalysis, StaticAnalysisOptions, StaticAnalysisResult, SurfacePatch)
from .optimization import (DensityFilter, IterationSnapshot, OcOptimizer, OptimizerOptions, SensitivityAnalysis,
                           TopologyOptimizationResult, TopologyOptimizer)

__all__ = [
    "PRECISION_FP64", "PRECISION_MIXED_FP32", "PRECISION_MIXED_BF16", "PRECISION_MIXED_GMG",
    "PRECISION_FP64_ORDERED", "MATVEC_AUTO", "MATVEC_NODE_GATHER", "MATVEC_ELEMENT", "MATVEC_REFERENCE_ORDER",
    "ArgumentError", "ArgumentNullError", "ArgumentOutOfRangeError", "DeviceError", "InvalidOperationError",
    "MycelError", "OperationCanceledError", "OutOfMemoryError",
    "FeMaterial", "FemOperator", "FeModel", "GridSpec", "Hex8Element", "JacobiPreconditioner", "LoadCase",
    "MultigridPreconditioner",
    "PcgResult", "PcgSolver", "StaticAnalysis", "StaticAnalysisOptions", "StaticAnalysisResult", "SurfacePatch",
    "DensityFilter", "IterationSnapshot", "OcOptimizer", "OptimizerOptions", "SensitivityAnalysis",
    "TopologyOptimizationResult", "TopologyOptimizer",
]
