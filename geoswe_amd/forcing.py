"""Rain and stage forcings with the reference's lookup semantics.

``RainfallForcing`` follows src/geoswe/forcing.py:93-216 (piecewise-constant segments,
``bisect_right - 1`` clamped, mm/h / 3.6e6 -> m/s) and ``StageBoundary`` follows
forcing.py:219-377 (linear interpolation clamped at the ends; the marked cells get
``h = max(eta - b, 0)`` and zero momentum after each step). On the device the same
lookups run inside the step scheduler, from tables built here by ``device_tables``.
"""
from __future__ import annotations

import bisect
import warnings
from dataclasses import dataclass

import numpy as np


def _require_time_order(who, t):
    t = np.asarray(t, dtype=np.float64)
    if t.ndim != 1 or t.size == 0:
        raise ValueError(f"{who}: time_s must be a non-empty 1-D array")
    if t.size > 1 and bool(np.any(np.diff(t) < 0)):
        k = int(np.argmax(np.diff(t) < 0)) + 1
        raise ValueError(f"{who}: time_s must be non-decreasing; time_s[{k}]={t[k]} < time_s[{k-1}]={t[k-1]}")


def _series(who, time_s, values, ranks, what):
    """A forcing's (times, values) pair: lists become float64 arrays; the values must have one
    of the ranks in ``ranks`` (any rank for None) and one leading row per time; the times must
    be a non-empty, non-decreasing 1-D array (forcing.py:113-137, 247-254)."""
    t = np.asarray(time_s, dtype=np.float64) if isinstance(time_s, (list, tuple)) else time_s
    v = np.asarray(values, dtype=np.float64) if isinstance(values, (list, tuple)) else values
    rank = int(getattr(v, "ndim", 1))
    if ranks is not None and rank not in ranks:
        want = " or ".join({1: "(nt,)", 3: "(nt, nx, ny)"}[r] for r in ranks)
        raise ValueError(f"{who}: {what} of shape {tuple(np.shape(v))} is not {want}")
    rows = int(np.shape(v)[0]) if rank else 1
    if rows != len(t):
        raise ValueError(f"{who}: {what} has {rows} rows for {len(t)} times")
    _require_time_order(who, t)
    return t, v


@dataclass
class RainfallForcing:
    """Time-varying rainfall: uniform ``rate_mm_h`` (nt,) or gridded (nt, nx, ny)."""
    time_s: np.ndarray
    rate_mm_h: np.ndarray

    def __post_init__(self):
        self.time_s, self.rate_mm_h = _series("RainfallForcing", self.time_s, self.rate_mm_h, (1, 3), "rate_mm_h")

    @classmethod
    def from_uniform_constant(cls, rate_mm_h: float, t_end: float = 1e9):
        return cls(time_s=np.array([0.0, t_end]), rate_mm_h=np.array([rate_mm_h, rate_mm_h], dtype=np.float64))

    @property
    def gridded(self) -> bool:
        return np.asarray(self.rate_mm_h).ndim == 3

    def index_at(self, t: float) -> int:
        tl = list(np.asarray(self.time_s, dtype=np.float64).tolist())
        idx = max(0, bisect.bisect_right(tl, float(t)) - 1)
        return min(idx, len(tl) - 1)

    def rate_at_time(self, t: float):
        """Rain rate in m/s at t: a float (uniform) or an (nx, ny) array (gridded)."""
        idx = self.index_at(t)
        r = np.asarray(self.rate_mm_h)
        if r.ndim == 1:
            return float(r[idx]) / 3.6e6
        return r[idx] / 3.6e6


@dataclass
class NativeRain:
    """Rain on a native grid, the run driver's bundle (compressed_solver.py:3157-3183):
    ``table`` (nt, npix) rates in m/s, one row per time of ``time_s`` (piecewise constant, as
    :class:`RainfallForcing`), and ``lookup`` (nx, ny) the table column each cell reads.
    A flat-solver forcing (``CompressedSolver.set_rain``)."""
    time_s: np.ndarray
    table: np.ndarray
    lookup: np.ndarray

    def __post_init__(self):
        self.time_s = np.asarray(self.time_s, dtype=np.float64)
        self.table = np.asarray(self.table)
        self.lookup = np.asarray(self.lookup)
        if self.table.ndim != 2 or self.table.shape[0] != len(self.time_s):
            raise ValueError(f"NativeRain: table has shape {self.table.shape}; it must be (nt, npix) with nt = "
                             f"{len(self.time_s)}")
        if self.lookup.size and (self.lookup.min() < 0 or self.lookup.max() >= self.table.shape[1]):
            raise ValueError(f"NativeRain: lookup columns must lie in [0, {self.table.shape[1]})")
        _require_time_order("NativeRain", self.time_s)


@dataclass
class StageBoundary:
    """Water level eta(t) imposed on ``cells`` (M, 2) padded (i, j) indices after each step."""
    cells: np.ndarray
    time_s: np.ndarray
    stage_m: np.ndarray
    bed_b: np.ndarray

    def __post_init__(self):
        self.time_s, self.stage_m = _series("StageBoundary", self.time_s, self.stage_m, None, "stage_m")
        self.cells = np.asarray(self.cells, dtype=np.int64).reshape(-1, 2)

    @classmethod
    def from_mask(cls, mask, mesh, bed, time_s, stage_m):
        mask = np.asarray(mask, dtype=bool)
        bed = np.asarray(bed)
        if mask.shape != (mesh.nx, mesh.ny) or bed.shape != mask.shape:
            raise ValueError(f"StageBoundary.from_mask: mask and bed must have shape ({mesh.nx}, {mesh.ny})")
        return cls(cells=np.argwhere(mask) + mesh.ngh, time_s=time_s, stage_m=stage_m, bed_b=bed[mask])

    def stage_at_time(self, t: float) -> float:
        tl = list(np.asarray(self.time_s, dtype=np.float64).tolist())
        sl = list(np.asarray(self.stage_m, dtype=np.float64).tolist())
        t = float(t)
        if (t < tl[0] or t > tl[-1]) and not getattr(self, "_oor_warned", False):
            warnings.warn(f"StageBoundary: t={t:.0f}s is outside [{tl[0]:.0f}, {tl[-1]:.0f}]; clamping",
                          RuntimeWarning, stacklevel=2)
            self._oor_warned = True
        if t <= tl[0]:
            return sl[0]
        if t >= tl[-1]:
            return sl[-1]
        i = bisect.bisect_left(tl, t)
        t0, t1, s0, s1 = tl[i - 1], tl[i], sl[i - 1], sl[i]
        return s0 + (t - t0) / (t1 - t0) * (s1 - s0)


def rain_tables(forcing, rainfall: float, cfl: float, g: float, dmin: float, dtype):
    """Host tables of a rain configuration: (t, rate_m_s, cap, frames or None).

    ``cap`` is the rain-film step bound ``(cfl*dmin)**(2/3)/(g*R)**(1/3)`` per segment with
    the reference's host expression (solver.py:3005-3027), inf where it is dry.
    Returns None when there is no rain (``Config.rainfall_forcing`` overrides ``rainfall``,
    solver.py:1903-1914).
    """
    def cap_of(rmax):
        if rmax <= 0.0:
            return np.inf
        return (cfl * dmin) ** (2.0 / 3.0) / (g * rmax) ** (1.0 / 3.0)

    if isinstance(forcing, NativeRain):
        # one bound from the largest rate of the whole table (compressed_solver.py:3429-3436)
        t = forcing.time_s
        frames = forcing.table.astype(dtype)
        cap = cap_of(float(frames.max())) if frames.size else np.inf
        return t, np.zeros(len(t)), np.full(len(t), cap), frames
    if forcing is not None:
        t = np.asarray(forcing.time_s, dtype=np.float64)
        r = np.asarray(forcing.rate_mm_h)
        if r.ndim == 1:
            rate = np.array([float(x) / 3.6e6 for x in r], dtype=np.float64)
            return t, rate, np.array([cap_of(x) for x in rate]), None
        frames = np.stack([(r[k] / 3.6e6).astype(dtype) for k in range(r.shape[0])])
        caps = np.array([cap_of(float(frames[k].max())) for k in range(frames.shape[0])])
        return t, np.zeros(len(t)), caps, frames
    if rainfall > 0.0:
        return np.array([0.0]), np.array([float(rainfall)]), np.array([cap_of(float(rainfall))]), None
    return None
