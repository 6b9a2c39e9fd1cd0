"""Config and Mesh2D: the reference's configuration surface for the production step.

Mirrors ``Config`` (src/geoswe/solver.py:629-860), its numeric guards
(``_check_numeric_inputs``, solver.py:564-626) and ``Mesh2D`` (src/geoswe/mesh.py:45-87).
Only the production scheme (first-order HLLC with SRM well balancing, forward Euler,
plain shallow water) has a kernel here; other scheme choices are accepted by name and
rejected with NotImplementedError when a solver is built, so a typo still fails early.
"""
from __future__ import annotations

import math
import warnings
from dataclasses import dataclass
from typing import Optional

import numpy as np

G = 9.81
H_MIN = 1.0e-10

_BC_OK = {"extrapolate", "periodic", "wall", "fall", "dirichlet"}
_ENUM_OK = {"time": {"euler", "ssprk3"}, "flux": {"lf", "hllc"},
            "recon": {"first", "muscl", "linear2", "linear3", "linear5", "weno5"},
            "rk_storage": {"low_storage", "high_storage"}, "sigma_stages": {"all", "first"},
            "wb_method": {"audusse", "srm"}, "pde": {"igr", "baseline"},
            "sigma_bc": {"neumann", "periodic"}, "dtype": {"float32", "float64"},
            "bc_x": _BC_OK, "bc_y": _BC_OK, "friction": {"manning_implicit"}}


def check_numeric_inputs(where="Config", **vals):
    """Bound the numeric inputs (solver.py:564-626): raises ValueError, warns on odd values."""
    for name, v in vals.items():
        if v is not None and not math.isfinite(float(v)) and name != "friction_velocity_cap_ms":
            raise ValueError(f"{where}.{name}={v!r} is not finite")
    f = lambda n: None if vals.get(n) is None else float(vals[n])
    cfl = f("cfl")
    if cfl is not None and not (0.0 < cfl <= 1.0):
        raise ValueError(f"{where}.cfl={cfl} is outside (0, 1]: the explicit step is unstable above "
                         f"the Courant limit")
    if cfl is not None and cfl > 0.5:
        warnings.warn(f"{where}.cfl={cfl} is above 0.5; with the max(|u|,|v|) norm the 2D limit is "
                      f"near 0.5", stacklevel=3)
    g = f("g")
    if g is not None and g <= 0.0:
        raise ValueError(f"{where}.g={g} must be positive")
    h_min = f("h_min")
    if h_min is not None and h_min <= 0.0:
        raise ValueError(f"{where}.h_min={h_min} must be positive (wet/dry threshold and divisor)")
    hmc = f("h_min_cfl")
    if hmc is not None and hmc < 0.0:
        raise ValueError(f"{where}.h_min_cfl={hmc} must be >= 0 (0 couples it to h_min)")
    mn = f("manning_n")
    if mn is not None and mn < 0.0:
        raise ValueError(f"{where}.manning_n={mn} must be >= 0")
    vcap = f("friction_velocity_cap_ms")
    if vcap is not None and vcap <= 0.0:
        raise ValueError(f"{where}.friction_velocity_cap_ms={vcap} must be positive (inf disables the cap)")
    sc = f("storage_courant")
    if sc is not None and sc < 0.0:
        raise ValueError(f"{where}.storage_courant={sc} must be >= 0")
    rain = f("rainfall")
    if rain is not None and rain > 1.0e-3:
        warnings.warn(f"{where}.rainfall={rain} is in m/s, not mm/h ({rain * 3.6e6:,.0f} mm/h)",
                      stacklevel=3)


@dataclass
class Config:
    """Solver configuration; ``Config()`` is the production flood scheme (solver.py:629-860)."""
    pde: str = "baseline"
    flux: str = "hllc"
    recon: str = "first"
    time: str = "euler"
    alpha: float = 0.0
    cfl: float = 0.5
    bc_x: str = "extrapolate"
    bc_y: str = "extrapolate"
    bc_x_left: Optional[np.ndarray] = None
    bc_x_right: Optional[np.ndarray] = None
    g: float = G
    h_min: float = H_MIN
    h_min_cfl: float = 0.0
    sigma_max_iter: int = 10
    sigma_tol: float = 0.0
    sigma_bc: str = "neumann"
    sigma_h_min: float = 0.0
    wb_method: str = "srm"
    friction: Optional[str] = None
    manning_n: float = 0.0
    friction_velocity_cap_ms: float = 15.0
    friction_quadratic_alpha: bool = True
    storage_courant: float = 0.0
    storage_dt_ref: float = 0.0
    rainfall: float = 0.0
    rainfall_forcing: Optional[object] = None
    manning_field: Optional[np.ndarray] = None
    stage_boundary: Optional[object] = None
    well_balanced: bool = True
    sigma_stages: str = "all"
    rk_storage: str = "low_storage"
    dtype: Optional[str] = None
    sigma_halo_every: int = 1

    def __post_init__(self):
        if self.dtype is None:
            self.dtype = "float32"          # this package computes on the GPU
        if not isinstance(self.dtype, str):
            try:
                self.dtype = np.dtype(self.dtype).name
            except TypeError:
                pass
        # float32 wet/dry floor defaults (solver.py:799-806)
        if self.dtype == "float32" and self.h_min == H_MIN:
            self.h_min = 1.0e-6
            if self.h_min_cfl == 0.0:
                self.h_min_cfl = self.h_min
        if isinstance(self.friction, str) and self.friction.lower() == "none":
            self.friction = None
        elif self.friction == "manning":
            self.friction = "manning_implicit"
        for k, ok in _ENUM_OK.items():
            v = getattr(self, k, None)
            if v is None and k == "friction":
                continue
            if not isinstance(v, str) or v not in ok:
                raise ValueError(f"Config.{k}={v!r} invalid; expected one of {sorted(ok)}")
        if self.friction is None and (self.manning_n > 0.0 or self.manning_field is not None):
            warnings.warn("Config: a Manning roughness is set but friction is off, so the run is "
                          "frictionless; pass friction='manning' to apply it", stacklevel=3)
        check_numeric_inputs(where="Config", cfl=self.cfl, g=self.g, h_min=self.h_min,
                             h_min_cfl=self.h_min_cfl, manning_n=self.manning_n,
                             friction_velocity_cap_ms=self.friction_velocity_cap_ms,
                             storage_courant=self.storage_courant, rainfall=self.rainfall)
        if self.well_balanced and self.recon != "first":
            warnings.warn(f"Config.recon={self.recon!r} has no effect with well_balanced=True",
                          stacklevel=3)

    def h_cfl(self) -> float:
        """The CFL depth floor cfl_dt uses (solver.py:2578-2585)."""
        h = self.h_min_cfl if self.h_min_cfl > 0.0 else self.h_min
        if self.dtype == "float32" and h < 1.0e-6:
            h = 1.0e-6
        return h

    def require_production_scheme(self, who: str) -> None:
        """Refuse the configurations this package has no kernel for."""
        bad = []
        if self.flux != "hllc":
            bad.append(f"flux={self.flux!r}")
        if not self.well_balanced or self.wb_method != "srm":
            bad.append(f"well_balanced={self.well_balanced}, wb_method={self.wb_method!r}")
        if self.time != "euler":
            bad.append(f"time={self.time!r}")
        if self.pde == "igr" and self.alpha > 0.0:
            bad.append("pde='igr'")
        if self.storage_courant > 0.0:
            bad.append("storage_courant > 0")
        if "dirichlet" in (self.bc_x, self.bc_y):
            bad.append("bc 'dirichlet' (1D only)")
        if bad:
            raise NotImplementedError(
                f"{who}: the MI355X step implements the production scheme only (HLLC + SRM, "
                f"first order, forward Euler, plain SWE); not {', '.join(bad)}")


@dataclass
class Mesh2D:
    """Uniform cell-centred grid of ``nx x ny`` cells with ``ngh`` ghost layers (mesh.py:45-87)."""
    nx: int
    ny: int
    dx: float
    dy: float
    x0: float = 0.0
    y0: float = 0.0
    ngh: int = 4

    def __post_init__(self):
        self.x = self.x0 + (np.arange(self.nx) + 0.5) * self.dx
        self.y = self.y0 + (np.arange(self.ny) + 0.5) * self.dy
        self.shape = (self.nx, self.ny)

    def interior(self, arr: np.ndarray) -> np.ndarray:
        if arr.ndim != 2:
            raise ValueError(f"Mesh2D.interior expects a 2-D scalar field, got ndim={arr.ndim}")
        g = self.ngh
        return arr[g:-g, g:-g] if g > 0 else arr

    def pad(self, arr: np.ndarray) -> np.ndarray:
        if arr.ndim != 2:
            raise ValueError(f"Mesh2D.pad expects a 2-D scalar field, got ndim={arr.ndim}")
        out = np.zeros((self.nx + 2 * self.ngh, self.ny + 2 * self.ngh), dtype=arr.dtype)
        out[self.ngh:self.ngh + self.nx, self.ngh:self.ngh + self.ny] = arr
        return out
