"""helmholtz_amd: MI355X (gfx950) path for Helmholtz.jl's Helmholtz operator and
shifted-Laplacian multigrid solve, behind the reference's own API names.

The compute runs in libhelmholtz_amd.so (hand-written HIP, C ABI in
include/helmholtz_amd.h); this package is the host-side mirror of the
reference interface.  Importing it does not touch the GPU.
"""
from ._lib import HelmholtzError, LIB_PATH, lib  # noqa: F401
from .helmholtz import (  # noqa: F401
    GetHelmholtzOperator, GetHelmholtzOperatorHO, GetHelmholtzShiftOP, HelmholtzOperator, HelmholtzParam, RegularMesh,
    getHelmholtzFun,
    getABL, getAcousticPointSource, getMaximalFrequency, getMidPointSrc, getRegularMesh,
    getShiftedHelmholtzParam, getSommerfeldBC, getTopPointSrc, loc2cs, loc2cs3D, point_sources,
)
from .solver import (  # noqa: F401
    MGparam, ShiftedLaplacianMultigridSolver, clear, copySolver, getMGparam,
    getShiftedLaplacianMultigridSolver, hierarchy_levels, kernel_time, precondition, solveLinearSystem,
    time_kernel,
)
