"""MI355X-native (gfx950) drop-in for warp_fdtd's uniform-grid FDTD path.

Python API of warp_fdtd (solver, detectors, adjoint, material, driven_uniform,
periodic, volume) over hand-written HIP kernels reached through the C ABI of
libwfdtd.so (include/wfdtd.h).  Device arrays are torch tensors on ROCm.
"""

from . import (_lib, adjoint, array, autotune, batch, cases, convergence, design, detectors,
               driven_uniform, guide_observer, guide_problem, guide_validation, kernels,
               fno, material, modes, periodic, solver, surrogate, volume)

__all__ = ["_lib", "adjoint", "array", "autotune", "batch", "cases", "convergence", "design",
           "detectors", "driven_uniform", "guide_observer", "guide_problem", "guide_validation",
           "fno", "kernels", "material", "modes", "periodic", "solver", "surrogate", "volume"]
