"""nuclearmpm_amd: an MI355X-native MLS-MPM engine with the API of jparr721/NuclearMPM.

The step (p2g -> grid update -> g2p, src/nclr.h:80-84) runs as hand-written HIP kernels for
gfx950 in libnclr_mpm.so behind a C ABI (include/nclr_mpm.h).  This package is the Python
mirror of the reference's C++ API plus the scene generators used by the benchmarks.
"""
from ._lib import EngineError, EngineMissing, LIB_PATH, lib
from .scenes import Scene, XorShift128, cube, linspaced
from .sim import MaterialModel, MPMBatch, MPMSimulation, Particle, Particles, SceneSpec, diag

__all__ = [
    "EngineError",
    "EngineMissing",
    "LIB_PATH",
    "MaterialModel",
    "MPMBatch",
    "MPMSimulation",
    "SceneSpec",
    "Particle",
    "Particles",
    "Scene",
    "XorShift128",
    "cube",
    "diag",
    "lib",
    "linspaced",
]
