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


# Numeric bounds of the configuration (the accept/reject rules of solver.py:564-626): each
# row is a field, the interval its value must lie in, and what breaks outside it. A value of
# None is not checked; every value given must be finite except the velocity cap (inf, or a
# NaN that no speed exceeds, switches the cap off); a value is refused when it is known to lie
# outside its interval.
_BOUNDS = (
    ("cfl", (0.0, False, 1.0, True), "the explicit step is unstable above the Courant limit"),
    ("g", (0.0, False, math.inf, True), "gravity must be positive"),
    ("h_min", (0.0, False, math.inf, True), "it is the wet/dry threshold and a divisor"),
    ("h_min_cfl", (0.0, True, math.inf, True), "0 means: use h_min"),
    ("manning_n", (0.0, True, math.inf, True), "a roughness cannot be negative"),
    ("friction_velocity_cap_ms", (0.0, False, math.inf, True), "inf switches the cap off"),
    ("storage_courant", (0.0, True, math.inf, True), "a Courant number cannot be negative"),
)
# Values that are legal but usually a mistake: (field, threshold above which to warn, why).
_ODD = (
    ("cfl", 0.5, lambda v: "the 2D limit of the max(|u|,|v|) norm is about 0.5"),
    ("rainfall", 1.0e-3, lambda v: f"rainfall is in m/s; this is {v * 3.6e6:,.0f} mm/h"),
)


# spellings of the friction switch: "none" in any case means off, "manning" the point-implicit
# Manning law
_FRICTION_NAMES = {"manning": "manning_implicit"}
# Settings that a configuration accepts but that have no effect in it (warnings, as the
# reference gives them): (condition, message).
_IGNORED = (
    (lambda c: c.friction is None and (c.manning_n > 0.0 or c.manning_field is not None),
     lambda c: "Manning roughness given with friction off: the run is frictionless (friction='manning' "
               "applies it)"),
    (lambda c: c.well_balanced and c.recon != "first",
     lambda c: f"recon={c.recon!r} is ignored: the well-balanced face states are first order"),
)


def _interval(lo, lo_closed, hi, hi_closed) -> str:
    return f"{'[' if lo_closed else '('}{lo:g}, {hi:g}{']' if hi_closed else ')'}"


def check_numeric_inputs(where="Config", **vals):
    """Check numeric configuration values against ``_BOUNDS`` (ValueError) and ``_ODD``
    (a warning), after solver.py:564-626."""
    given = {k: float(v) for k, v in vals.items() if v is not None}
    for name, v in given.items():
        if not math.isfinite(v) and name != "friction_velocity_cap_ms":
            raise ValueError(f"{where}.{name}={vals[name]!r}: a finite value is required")
    for name, (lo, lo_closed, hi, hi_closed), why in _BOUNDS:
        if name not in given:
            continue
        v = given[name]
        if v < lo or (v == lo and not lo_closed) or v > hi or (v == hi and not hi_closed):
            raise ValueError(f"{where}.{name}={v!r} lies outside {_interval(lo, lo_closed, hi, hi_closed)}: {why}")
    for name, limit, why in _ODD:
        if name in given and given[name] > limit:
            warnings.warn(f"{where}.{name}={given[name]!r}: {why(given[name])}", stacklevel=3)


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
    # No reference counterpart: "float64" with dtype "float32" steps the float32 planes with
    # float64 arithmetic (each stored value rounded once to float32); dense layout only.
    compute_dtype: Optional[str] = None

    def __post_init__(self):
        # the float32 build is the default here (the package computes on the GPU); other
        # spellings of a dtype are normalised to its name
        if self.dtype is None:
            self.dtype = "float32"
        elif not isinstance(self.dtype, str):
            try:
                self.dtype = np.dtype(self.dtype).name
            except TypeError:
                pass                         # refused by the choice check below
        if self.dtype == "float32" and self.h_min == H_MIN:
            # float32 wet/dry floor, with the CFL floor coupled to it (solver.py:799-806)
            self.h_min = 1.0e-6
            if self.h_min_cfl == 0.0:
                self.h_min_cfl = self.h_min
        if isinstance(self.friction, str):
            self.friction = None if self.friction.lower() == "none" else _FRICTION_NAMES.get(self.friction,
                                                                                           self.friction)
        for name, allowed in _ENUM_OK.items():
            v = getattr(self, name, None)
            if name == "friction" and v is None:
                continue                     # no friction
            if not (isinstance(v, str) and v in allowed):
                raise ValueError(f"Config.{name}={v!r} is not one of {sorted(allowed)}")
        if self.compute_dtype not in (None, self.dtype) and not (self.dtype == "float32"
                                                                 and self.compute_dtype == "float64"):
            raise ValueError(f"Config.compute_dtype={self.compute_dtype!r} with dtype={self.dtype!r}: the arithmetic "
                             f"is the state's dtype, or float64 for a float32 state")
        check_numeric_inputs(where="Config", cfl=self.cfl, g=self.g, h_min=self.h_min,
                             h_min_cfl=self.h_min_cfl, manning_n=self.manning_n,
                             friction_velocity_cap_ms=self.friction_velocity_cap_ms,
                             storage_courant=self.storage_courant, rainfall=self.rainfall)
        for ignored, why in _IGNORED:
            if ignored(self):
                warnings.warn(f"Config: {why(self)}", stacklevel=3)

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
