"""mercury_amd -- MI355X-native batched mass-spring step (Mercury's hot path).

The compute path is libmercury_hip.so (HIP kernels for gfx950 behind the C-ABI
in include/mercury_hip.h).  This package is its Python host view: the
reference's Simulator API, robot builders and the evaluate loop.
"""
from ._lib import (MH_FITNESS_DISPLACEMENT, MH_FITNESS_REFERENCE, MH_KERNEL_AUTO,
                   MH_KERNEL_LDS, MH_KERNEL_MASS, MH_KERNEL_MASS_GATHER, MH_KERNEL_MASS_STREAM,
                   MH_KERNEL_STREAM,
                   MH_SOLVER_FORCE, MH_SOLVER_XPBD_REF, MhError, load)
from .materials import MaterialSet, workload_materials
from .robots import Element, voxel_build, voxel_population
from .simulator import ElementTracker, SimConfig, Simulator, pack

__all__ = ["MaterialSet", "workload_materials", "Element", "voxel_build", "voxel_population",
           "SimConfig", "Simulator", "ElementTracker", "pack", "load", "MhError",
           "MH_SOLVER_XPBD_REF", "MH_SOLVER_FORCE", "MH_KERNEL_AUTO", "MH_KERNEL_LDS",
           "MH_KERNEL_STREAM", "MH_KERNEL_MASS", "MH_KERNEL_MASS_GATHER", "MH_KERNEL_MASS_STREAM",
           "MH_FITNESS_REFERENCE",
           "MH_FITNESS_DISPLACEMENT"]
