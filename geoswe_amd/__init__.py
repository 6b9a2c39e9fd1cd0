"""geoswe_amd: GeoSWE's fused finite-volume time step, rebuilt for AMD MI355X (gfx950).

Public API mirrors the reference's (src/geoswe/__init__.py:38-102) for the production
scheme: ``Config``, ``Mesh2D``, ``Solver2D``, ``CompressedSolver``, ``RainfallForcing``,
``StageBoundary``. All compute runs in ``_lib/libgeoswe_hip.so`` through the C ABI of
``include/geoswe_abi.h``; there is no CPU fallback.
"""
from .config import Config, Mesh2D, G, H_MIN
from .forcing import NativeRain, RainfallForcing, StageBoundary

__version__ = "0.1.0"

__all__ = ["Config", "Mesh2D", "Solver2D", "CompressedSolver", "RainfallForcing", "StageBoundary", "G", "H_MIN"]


def __getattr__(name):
    if name == "Solver2D":
        from .solver import Solver2D
        return Solver2D
    if name == "CompressedSolver":
        from .compressed import CompressedSolver
        return CompressedSolver
    raise AttributeError(name)
