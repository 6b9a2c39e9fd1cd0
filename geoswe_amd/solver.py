"""Solver2D: the dense-grid solver API of GeoSWE on the MI355X fused step.

Drop-in for ``geoswe.Solver2D`` (src/geoswe/solver.py:1273-3032) on the production
scheme. State, bed, Manning classes and forcings live in device memory allocated through
libgeoswe_hip; every step is one ``gs_dense_step`` launch (residual + rain + update +
friction + wet/dry + stage + next-step CFL), and ``run`` keeps dt on the device so it
synchronises once per batch of steps instead of once per step.

Device layout: each field is an interior-only plane of ``nx`` rows with a row pitch
rounded up to 64 elements (256 B in float32); ghost cells are synthesised inside the
kernel from the boundary kinds, so ``q`` (padded, ghosts filled) is assembled on request.
"""
from __future__ import annotations

import ctypes as C
import math
import time
import warnings
from typing import Callable, Optional

import numpy as np

from . import _lib as L
from .config import Config, Mesh2D
from .forcing import NativeRain, rain_tables

_PITCH_ALIGN = 64


def _pitch(ny: int) -> int:
    return ((ny + _PITCH_ALIGN - 1) // _PITCH_ALIGN) * _PITCH_ALIGN


def host_bc_fill(q: np.ndarray, ngh: int, bc_x: str, bc_y: str) -> None:
    """Ghost fill of a padded (3, nxp, nyp) host state, x then y (bc.py:50-98)."""
    for axis, kind in ((1, bc_x), (2, bc_y)):
        n = q.shape[axis]
        sl = lambda a, b: (slice(None),) * axis + (slice(a, b),)
        if kind == "periodic":
            q[sl(0, ngh)] = q[sl(n - 2 * ngh, n - ngh)]
            q[sl(n - ngh, n)] = q[sl(ngh, 2 * ngh)]
        elif kind == "extrapolate":
            q[sl(0, ngh)] = q[sl(ngh, ngh + 1)]
            q[sl(n - ngh, n)] = q[sl(n - ngh - 1, n - ngh)]
        elif kind == "wall":
            lo = np.flip(q[sl(ngh, 2 * ngh)], axis=axis).copy()
            hi = np.flip(q[sl(n - 2 * ngh, n - ngh)], axis=axis).copy()
            lo[axis] = -lo[axis]
            hi[axis] = -hi[axis]
            q[sl(0, ngh)] = lo
            q[sl(n - ngh, n)] = hi
        elif kind == "fall":
            q[sl(0, ngh)] = 0
            q[sl(n - ngh, n)] = 0


def host_pad_bed(b_int: np.ndarray, ngh: int, bc_x: str, bc_y: str) -> np.ndarray:
    """Padded bed, ghosts edge-replicated or wrapped on periodic axes (solver.py:1466-1484)."""
    nx, ny = b_int.shape
    b = np.zeros((nx + 2 * ngh, ny + 2 * ngh), dtype=b_int.dtype)
    b[ngh:ngh + nx, ngh:ngh + ny] = b_int
    if bc_x == "periodic":
        b[:ngh, :] = b[-2 * ngh:-ngh, :]
        b[-ngh:, :] = b[ngh:2 * ngh, :]
    else:
        b[:ngh, :] = b[ngh:ngh + 1, :]
        b[-ngh:, :] = b[-ngh - 1:-ngh, :]
    if bc_y == "periodic":
        b[:, :ngh] = b[:, -2 * ngh:-ngh]
        b[:, -ngh:] = b[:, ngh:2 * ngh]
    else:
        b[:, :ngh] = b[:, ngh:ngh + 1]
        b[:, -ngh:] = b[:, -ngh - 1:-ngh]
    return b


class _Forcing:
    """Device copies of the rain and stage tables of a Config, plus the ABI struct."""

    def __init__(self, cfg: Config, dmin: float, dtype, frame_shape, frame_pitch=None):
        self.keep = []
        self.s = L.GsForcing()
        self.dtype = np.dtype(dtype)
        self.cfl, self.g, self.dmin = float(cfg.cfl), float(cfg.g), float(dmin)
        self.rain_host = rain_tables(cfg.rainfall_forcing, cfg.rainfall, cfg.cfl, cfg.g, dmin, dtype)
        if self.rain_host is not None:
            t, rate, cap, frames = self.rain_host
            dt_ = L.DeviceArray.from_host(np.ascontiguousarray(t, np.float64))
            dr = L.DeviceArray.from_host(np.ascontiguousarray(rate, np.float64))
            dc = L.DeviceArray.from_host(np.ascontiguousarray(cap, np.float64))
            self.keep += [dt_, dr, dc]
            self.s.n_rain = len(t)
            self.s.rain_t, self.s.rain_rate, self.s.rain_cap = dt_.ptr, dr.ptr, dc.ptr
            if frames is not None and frames.ndim == 2:       # a native table read through a lookup
                df = L.DeviceArray.from_host(np.ascontiguousarray(frames))
                self.keep.append(df)
                self.s.rain_gridded = 1
                self.s.rain_frames = df.ptr
                self.s.rain_frame_stride = frames.shape[1]
                self.s.rain_row_pitch = frames.shape[1]
            elif frames is not None:
                if tuple(frames.shape[1:]) != tuple(frame_shape):
                    raise ValueError(f"gridded rain must be (nt, nx, ny) = (nt, {frame_shape[0]}, "
                                     f"{frame_shape[1]}); got {tuple(frames.shape)}")
                nt = frames.shape[0]
                pitch = frame_pitch or frames.shape[2]
                buf = np.zeros((nt, frames.shape[1], pitch), dtype=frames.dtype)
                buf[:, :, :frames.shape[2]] = frames
                df = L.DeviceArray.from_host(buf)
                self.keep.append(df)
                self.s.rain_gridded = 1
                self.s.rain_frames = df.ptr
                self.s.rain_frame_stride = frames.shape[1] * pitch
                self.s.rain_row_pitch = pitch
        sb = cfg.stage_boundary
        if isinstance(sb, (list, tuple)):
            if len(sb) != 1:
                raise NotImplementedError("one StageBoundary series per solver is supported")
            sb = sb[0]
        self.stage = sb
        if sb is not None:
            st = L.DeviceArray.from_host(np.ascontiguousarray(sb.time_s, np.float64))
            sv = L.DeviceArray.from_host(np.ascontiguousarray(sb.stage_m, np.float64))
            self.keep += [st, sv]
            self.s.n_stage = len(sb.time_s)
            self.s.stage_t, self.s.stage_v = st.ptr, sv.ptr

    def use_global_rain_cap(self) -> None:
        """One rain-film bound for the whole run, from the largest rate of the table as the step
        reads it (state dtype): the flat solver's rule (compressed_solver.py:3429-3436), where
        the dense one bounds each step by its own segment's rate (solver.py:3005-3027)."""
        if self.rain_host is None:
            return
        t, rate, cap, frames = self.rain_host
        rmax = float(frames.max()) if frames is not None else float(np.max(rate.astype(self.dtype)))
        c = np.inf if rmax <= 0.0 else (self.cfl * self.dmin) ** (2.0 / 3.0) / (self.g * rmax) ** (1.0 / 3.0)
        self.rain_host = (t, rate, np.full(len(t), c), frames)
        self.keep[2].upload(np.full(self.s.n_rain, c))

    def drop_rain_cap(self) -> None:
        """No rain-film step bound (the reference's cached replay applies none by default)."""
        if self.rain_host is not None:
            self.keep[2].upload(np.full(self.s.n_rain, np.inf))

    def rain_rate_at(self, t: float):
        """(rate, rmax) at time t on the host, for the host-side conveniences."""
        if self.rain_host is None:
            return None, 0.0
        import bisect
        tl = list(self.rain_host[0].tolist())
        idx = min(max(0, bisect.bisect_right(tl, float(t)) - 1), len(tl) - 1)
        if self.rain_host[3] is None:
            r = float(self.rain_host[1][idx])
            return r, r
        fr = self.rain_host[3][idx]
        return fr, float(fr.max())


class Solver2D:
    """Two-dimensional shallow-water solver on the dense padded grid (solver.py:1273)."""

    def __init__(self, mesh: Mesh2D, cfg: Config, q0: np.ndarray, b: np.ndarray, comm=None, dims=None,
                 track_max_depth: bool = True, tile_rows: int = 0, device_init=None):
        """``q0`` (3, nx, ny) and ``b`` (nx, ny) host arrays, as the reference takes them.

        ``device_init=(h0, bed_writer)`` instead builds a state at rest with uniform depth
        ``h0`` and a bed written on the device by ``bed_writer(ptr, pitch, dtype_code)``;
        ``q0`` and ``b`` are then None (used for grids too large to stage through the host).
        """
        if comm is not None and getattr(comm, "size", 1) > 1:
            raise NotImplementedError("multi-GPU domain decomposition is not implemented; one GPU "
                                      "runs the whole domain")
        cfg.require_production_scheme("Solver2D")
        self.mesh, self.cfg = mesh, cfg
        if device_init is None:
            q0 = np.asarray(q0)
            b = np.asarray(b)
            if tuple(q0.shape) != (3, mesh.nx, mesh.ny):
                raise ValueError(f"q0 must have shape (3, nx, ny) = (3, {mesh.nx}, {mesh.ny}); got {tuple(q0.shape)}")
            if tuple(b.shape) != (mesh.nx, mesh.ny):
                raise ValueError(f"the bed must have shape (nx, ny) = ({mesh.nx}, {mesh.ny}); got {tuple(b.shape)}")
            for what, arr in (("the bed", b), ("the initial state q0", q0)):
                if not bool(np.isfinite(arr).all()):
                    raise ValueError(f"Solver2D: {what} holds non-finite values (NaN or inf)")
        if mesh.ngh < 2:
            raise ValueError(f"ngh={mesh.ngh} is too small: the SRM-HLLC stencil reads two ghost layers")
        if mesh.nx < 4 or mesh.ny < 4:
            raise ValueError("the grid needs at least 4 x 4 interior cells")
        rf = cfg.rainfall_forcing
        if isinstance(rf, NativeRain):
            raise NotImplementedError("rain on a native grid is read through a per-cell lookup by the flat solver "
                                      "only (CompressedSolver.set_rain)")
        if rf is not None and np.asarray(rf.rate_mm_h).ndim == 3 and tuple(rf.rate_mm_h.shape[1:]) != (mesh.nx, mesh.ny):
            raise ValueError(f"gridded rain must be (nt, nx, ny) = (nt, {mesh.nx}, {mesh.ny}); got "
                             f"{tuple(rf.rate_mm_h.shape)}")
        self.dtype = np.dtype(cfg.dtype)
        # float32 planes stepped with float64 arithmetic (Config.compute_dtype="float64")
        self.compute64 = self.dtype == np.float32 and cfg.compute_dtype == "float64"
        nx, ny = mesh.nx, mesh.ny
        self.pitch = _pitch(ny)
        self._plane = (nx, self.pitch)
        self.stream = None
        self._bufs = [[L.DeviceArray(self._plane, self.dtype, zero=device_init is None) for _ in range(3)]
                      for _ in range(2)]
        dcode = L.GS_F32 if self.dtype == np.float32 else L.GS_F64
        if device_init is None:
            host = np.zeros((3, nx, self.pitch), self.dtype)
            host[:, :, :ny] = q0.astype(self.dtype, copy=False)
            for k in range(3):
                self._bufs[0][k].upload(host[k])
                self._bufs[1][k].upload(host[k])
            del host
            bh = np.zeros(self._plane, self.dtype)
            bh[:, :ny] = b.astype(self.dtype, copy=False)
            self._bed = L.DeviceArray.from_host(bh)
            self._b_host = b.astype(self.dtype, copy=True)
        else:
            h0, writer = device_init
            for s_ in range(2):
                for k, v in enumerate((float(h0), 0.0, 0.0)):
                    L.call("gs_fill", C.c_void_p(self._bufs[s_][k].ptr), dcode, nx, self.pitch, self.pitch,
                           C.c_double(v), None)
            self._bed = L.DeviceArray(self._plane, self.dtype, zero=True)
            writer(self._bed.ptr, self.pitch, dcode)
            L.sync()
            self._b_host = None
        self._flags_host = None          # (nx, pitch) uint8
        self._flags = None
        self._n_field = None
        self._cls = None                 # (nx, pitch) uint8 class plane: tables of 33..256 classes
        self._ntab = L.DeviceArray((L.MAX_CLASSES_PLANE,), self.dtype, zero=True)
        self._n_classes = 1
        self._manning_uniform = float(cfg.manning_n)
        self.inside_mask = None
        self._max_h = L.DeviceArray(self._plane, self.dtype, zero=True) if track_max_depth else None
        self._any_step = False
        self._tile_rows = int(tile_rows)
        self._fo = _Forcing(cfg, min(mesh.dx, mesh.dy), self.dtype, (nx, ny), self.pitch)
        if self._fo.stage is not None:
            self._mark_stage(self._fo.stage)
        if cfg.manning_field is not None:
            mf = np.asarray(cfg.manning_field)
            exp = (nx + 2 * mesh.ngh, ny + 2 * mesh.ngh)
            if tuple(mf.shape) != exp:
                raise ValueError(f"Config.manning_field must be the padded {exp} array, got {tuple(mf.shape)}")
            g = mesh.ngh
            self._set_n_plane(mf[g:g + nx, g:g + ny])
        self._upload_ntab(np.array([self._manning_uniform]))
        self._clock = L.new_clock(0.0)
        self._lam_valid = False
        self._target = None
        self.diagnostics = {"wallclock": 0.0, "steps": 0}

    # ------------------------------------------------------------------ structs
    def _phys(self) -> L.GsPhys:
        cfg = self.cfg
        fr = cfg.friction == "manning_implicit"
        return L.phys_struct(cfg.g, cfg.h_min, cfg.h_cfl(), cfg.cfl, min(self.mesh.dx, self.mesh.dy),
                             cfg.friction_velocity_cap_ms, fr, cfg.friction_quadratic_alpha, True, False,
                             self._n_classes)

    def _desc(self) -> L.GsDense:
        d = L.GsDense()
        d.nx, d.ny, d.pitch = self.mesh.nx, self.mesh.ny, self.pitch
        d.bc_x, d.bc_y = L.GS_BC[self.cfg.bc_x], L.GS_BC[self.cfg.bc_y]
        d.dtype = L.GS_F64 if self.dtype == np.float64 else (L.GS_F32X64 if self.compute64 else L.GS_F32)
        d.tile_rows = self._tile_rows
        d.dx, d.dy = float(self.mesh.dx), float(self.mesh.dy)
        for s in range(2):
            for k in range(3):
                d.q[s][k] = self._bufs[s][k].ptr
        d.bed = self._bed.ptr
        d.flags = self._flags.ptr if self._flags is not None else None
        d.ntab = self._ntab.ptr
        d.n_field = self._n_field.ptr if self._n_field is not None else None
        d.max_h = self._max_h.ptr if self._max_h is not None else None
        d.cls = self._cls.ptr if self._cls is not None else None
        return d

    # ------------------------------------------------------------------ set-up
    def _ensure_flags(self):
        if self._flags_host is None:
            self._flags_host = np.zeros(self._plane, np.uint8)

    def _push_flags(self):
        self._flags = L.DeviceArray.from_host(self._flags_host)

    def _upload_ntab(self, table):
        t = np.zeros(L.MAX_CLASSES_PLANE, self.dtype)
        t[:len(table)] = np.asarray(table, dtype=np.float64).astype(self.dtype)
        self._ntab.upload(t)
        self._n_classes = max(1, len(table))

    def _set_n_plane(self, n_int):
        h = np.zeros(self._plane, self.dtype)
        h[:, :self.mesh.ny] = np.asarray(n_int, dtype=np.float64).astype(self.dtype)
        self._n_field = L.DeviceArray.from_host(h)
        self._cls = None

    def _set_class_plane(self, cls_int, table):
        """Classes 0..255 in a uint8 plane of their own (more than the flags byte's 32)."""
        c = np.zeros(self._plane, np.uint8)
        c[:, :self.mesh.ny] = np.asarray(cls_int, dtype=np.uint8)
        self._cls = L.DeviceArray.from_host(c)
        self._n_field = None
        self._upload_ntab(np.asarray(table)[:L.MAX_CLASSES_PLANE])   # uint8 ids reach 255 at most

    def _mark_stage(self, sb):
        g = self.mesh.ngh
        c = np.asarray(sb.cells, dtype=np.int64).reshape(-1, 2)
        if c.size and (c[:, 0].min() < g or c[:, 0].max() >= g + self.mesh.nx or c[:, 1].min() < g
                       or c[:, 1].max() >= g + self.mesh.ny):
            raise ValueError("StageBoundary.cells must index the padded grid (unpadded index + ngh) and lie "
                             "in the interior")
        self._ensure_flags()
        self._flags_host[c[:, 0] - g, c[:, 1] - g] |= L.FLAG_STAGE
        self._push_flags()

    def set_inside_mask(self, mask, cfl_robust_pct=None):
        """Active set (solver.py:1486-1522): cells outside get a zero residual and no rain."""
        if cfl_robust_pct is not None:
            raise NotImplementedError("cfl_robust_pct (percentile CFL) is not supported by the fused step")
        self._ensure_flags()
        if mask is None:
            self.inside_mask = None
            self._flags_host &= np.uint8(~L.FLAG_INACTIVE & 0xFF)
            self._push_flags()
            return
        nx, ny = self.mesh.nx, self.mesh.ny
        mask = np.asarray(mask)
        if mask.shape != (nx, ny):
            raise ValueError(f"inside_mask must have shape ({nx}, {ny}); got {mask.shape}")
        self.inside_mask = mask.astype(bool)
        fl = self._flags_host[:, :ny]
        fl &= np.uint8(~L.FLAG_INACTIVE & 0xFF)
        fl |= np.where(self.inside_mask, 0, L.FLAG_INACTIVE).astype(np.uint8)
        self._push_flags()

    def set_cfl_ghost_mask(self, mask):
        """Leave interior cells out of the CFL wave-speed maximum (solver.py:1611-1633).

        ``mask`` is an unpadded (nx, ny) bool array (True = skipped), or None to clear.
        """
        self._ensure_flags()
        ny = self.mesh.ny
        fl = self._flags_host[:, :ny]
        fl &= np.uint8(~L.FLAG_CFL_SKIP & 0xFF)
        if mask is not None:
            mask = np.asarray(mask)
            if mask.shape != (self.mesh.nx, ny):
                raise ValueError(f"cfl_ghost_mask must have shape ({self.mesh.nx}, {ny}); got {mask.shape}")
            fl |= np.where(mask.astype(bool), L.FLAG_CFL_SKIP, 0).astype(np.uint8)
        self._push_flags()
        self._lam_valid = False

    def set_manning_table(self, cls_padded, table):
        """Manning n as a uint8 class per padded cell plus a table (solver.py:1524-1556)."""
        g, nx, ny = self.mesh.ngh, self.mesh.nx, self.mesh.ny
        cls_padded = np.asarray(cls_padded)
        table = np.asarray(table, dtype=np.float32)
        exp = (nx + 2 * g, ny + 2 * g)
        if tuple(cls_padded.shape) != exp:
            raise ValueError(f"set_manning_table expects a PADDED (nxp, nyp)={exp} class array; got "
                             f"{tuple(cls_padded.shape)}")
        if cls_padded.dtype != np.uint8:
            raise ValueError(f"set_manning_table: cls_padded must be uint8; got {cls_padded.dtype}")
        if int(cls_padded.max()) >= int(table.size):
            raise ValueError(f"set_manning_table: class id {int(cls_padded.max())} out of range for a "
                             f"{int(table.size)}-entry table (OOB device read)")
        if self.cfg.friction is None:
            warnings.warn("set_manning_table: friction is off (Config.friction is None), so this roughness "
                          "table will not be applied", stacklevel=2)
        if table.size > L.MAX_CLASSES:
            # more classes than the flags byte holds: a class plane of their own (up to 256, the
            # reference's uint8 class array)
            self._set_class_plane(cls_padded[g:g + nx, g:g + ny], table.astype(np.float64))
            self._manning_table = table
            return
        self._ensure_flags()
        fl = self._flags_host[:, :ny]
        fl &= np.uint8(~L.FLAG_CLASS_MASK & 0xFF)
        fl |= cls_padded[g:g + nx, g:g + ny].astype(np.uint8)
        self._push_flags()
        self._n_field = None
        self._cls = None
        self._upload_ntab(table.astype(np.float64))
        self._manning_table = table

    def set_manning(self, n):
        """Manning n from a scalar or an unpadded (nx, ny) array; switches friction on (solver.py:1558-1591)."""
        nx, ny = self.mesh.nx, self.mesh.ny
        nh = np.asarray(n, dtype=np.float64)
        if nh.ndim == 0:
            nh = np.full((nx, ny), float(nh))
        if nh.shape != (nx, ny):
            raise ValueError(f"set_manning expects a scalar or an unpadded ({nx}, {ny}) array; got {nh.shape}")
        if not np.all(np.isfinite(nh)) or float(nh.min()) < 0.0:
            raise ValueError("set_manning: Manning's n must be finite and non-negative")
        if self.cfg.friction is None:
            self.cfg.friction = "manning_implicit"
        vals, inv = np.unique(nh, return_inverse=True)
        if vals.size <= L.MAX_CLASSES and self.dtype == np.float32:
            self._ensure_flags()
            fl = self._flags_host[:, :ny]
            fl &= np.uint8(~L.FLAG_CLASS_MASK & 0xFF)
            fl |= inv.reshape(nx, ny).astype(np.uint8)
            self._push_flags()
            self._n_field = None
            self._cls = None
            self._upload_ntab(vals.astype(np.float32).astype(np.float64))
        elif vals.size <= L.MAX_CLASSES_PLANE and self.dtype == np.float32:
            self._set_class_plane(inv.reshape(nx, ny), vals.astype(np.float32).astype(np.float64))
        else:
            self._set_n_plane(nh)

    # ------------------------------------------------------------------ state access
    def _clockv(self) -> L.GsClock:
        return L.read_clock(self._clock)

    @property
    def t(self) -> float:
        return float(self._clockv().t)

    @property
    def nsteps(self) -> int:
        return int(self._clockv().nsteps)

    def _cur(self, c: Optional[L.GsClock] = None) -> int:
        c = c or self._clockv()
        return int(c.nsteps) & 1

    def _read_planes(self, which: int) -> np.ndarray:
        ny = self.mesh.ny
        return np.stack([self._bufs[which][k].download()[:, :ny] for k in range(3)])

    @property
    def q_interior(self) -> np.ndarray:
        """Conserved state without ghosts, (3, nx, ny), as a host array."""
        return self._read_planes(self._cur())

    @property
    def q(self) -> np.ndarray:
        """Padded (3, nxp, nyp) host copy of the state with the ghost cells filled."""
        g = self.mesh.ngh
        nx, ny = self.mesh.nx, self.mesh.ny
        out = np.zeros((3, nx + 2 * g, ny + 2 * g), self.dtype)
        out[:, g:g + nx, g:g + ny] = self.q_interior
        host_bc_fill(out, g, self.cfg.bc_x, self.cfg.bc_y)
        return out

    @property
    def b(self) -> np.ndarray:
        bh = self._b_host if self._b_host is not None else self._bed.download()[:, :self.mesh.ny]
        return host_pad_bed(bh, self.mesh.ngh, self.cfg.bc_x, self.cfg.bc_y)

    def depth(self) -> np.ndarray:
        """Water depth (nx, ny) as a host array."""
        return self._bufs[self._cur()][0].download()[:, :self.mesh.ny]

    def max_depth(self) -> np.ndarray:
        if self._max_h is None:
            raise RuntimeError("max_depth: this solver was built with track_max_depth=False")
        if not self._any_step:
            return self.depth()
        return self._max_h.download()[:, :self.mesh.ny]

    def set_state(self, q_interior: np.ndarray) -> None:
        """Replace the current state (3, nx, ny) in both buffers."""
        nx, ny = self.mesh.nx, self.mesh.ny
        q_interior = np.asarray(q_interior)
        if q_interior.shape != (3, nx, ny):
            raise ValueError(f"set_state expects (3, {nx}, {ny})")
        host = np.zeros((3, nx, self.pitch), self.dtype)
        host[:, :, :ny] = q_interior.astype(self.dtype, copy=False)
        for s in range(2):
            for k in range(3):
                self._bufs[s][k].upload(host[k])
        self._lam_valid = False

    # ------------------------------------------------------------------ stepping
    def _ensure_lam(self):
        if not self._lam_valid:
            L.call("gs_dense_cfl", C.byref(self._desc()), C.byref(self._phys()), self._fo.s,
                   C.c_void_p(self._clock.ptr), self.stream)
            self._lam_valid = True

    def _raise_if_bad(self, c: L.GsClock):
        if c.status & L.ST_NONFINITE:
            lam = np.uint64(c.lam_bits).view(np.float64)
            raise FloatingPointError(f"cfl_dt: non-finite max wave speed {lam} at t={c.t:.3f}s -- NaN/Inf in "
                                     f"the state; check forcing/inputs")

    def cfl_dt(self) -> float:
        """CFL step of the current state (solver.py:2569-2693), from the device wave speed."""
        self._ensure_lam()
        c = self._clockv()
        lam = float(np.uint64(c.lam_bits).view(np.float64))
        if not math.isfinite(lam):
            self._raise_if_bad(c)
            raise FloatingPointError(f"cfl_dt: non-finite max wave speed {lam}")
        floor = (self.cfg.g * self.cfg.h_cfl()) ** 0.5
        if lam < floor:
            lam = floor
        return self.cfg.cfl * min(self.mesh.dx, self.mesh.dy) / (lam + 1.0e-12)

    def _open_mode(self):
        if self._target != "open":
            L.call("gs_clock_set_target", C.c_void_p(self._clock.ptr), C.c_double(math.inf), C.c_double(0.0),
                   C.c_int64(-1), C.byref(self._phys()), self._fo.s, self.stream)
            self._target = "open"

    def step(self, dt: Optional[float] = None) -> float:
        """Advance one step of size dt (the CFL step if None), solver.py:2695-2973.

        With an explicit dt nothing is read back from the device.
        """
        self._ensure_lam()
        self._open_mode()
        if dt is not None:
            dt = float(dt)
            L.call("gs_clock_set_dt", C.c_void_p(self._clock.ptr), C.c_double(dt), C.byref(self._phys()),
                   self._fo.s, self.stream)
        else:
            c = self._clockv()
            self._raise_if_bad(c)
            dt = float(c.dt)
        L.call("gs_dense_step", C.byref(self._desc()), C.byref(self._phys()), self._fo.s,
               C.c_void_p(self._clock.ptr), self.stream)
        self._any_step = True
        self.last_dt = dt
        return dt

    def advance(self, nsteps: int, sync: bool = True) -> None:
        """Enqueue ``nsteps`` CFL steps back to back with no host synchronisation in between."""
        self._ensure_lam()
        self._open_mode()
        d, p = C.byref(self._desc()), C.byref(self._phys())
        clk = C.c_void_p(self._clock.ptr)
        for _ in range(int(nsteps)):
            L.call("gs_dense_step", d, p, self._fo.s, clk, self.stream)
        self._any_step = True
        if sync:
            L.sync(self.stream)
            self._raise_if_bad(self._clockv())

    def capture_steps(self, nsteps: int) -> L.StepGraph:
        """Record ``nsteps`` CFL steps (gs_dense_step launches) into a HIP graph whose
        ``replay()`` enqueues them on this solver's stream with no host work per step: the
        reference's captured fixed-dt loop (tests/test_gpu_dense_no_host_sync.py). dt, t and
        the wave speed stay in the device clock, so a replay is the same launches run
        directly, bit for bit. Synchronise (``sync``) before reading the state."""
        if int(nsteps) < 1:
            raise ValueError("capture_steps: nsteps must be >= 1")
        self._ensure_lam()
        self._open_mode()
        L.sync(self.stream)
        d, p = self._desc(), self._phys()
        clk = C.c_void_p(self._clock.ptr)
        fo = self._fo.s

        def record(cap):
            for _ in range(int(nsteps)):
                L.call("gs_dense_step", C.byref(d), C.byref(p), fo, clk, cap)

        g = L.StepGraph(record, self.stream)
        g._keep = (d, p, self)      # the recorded arguments point at this solver's buffers
        self._any_step = True
        return g

    def sync(self) -> None:
        L.sync(self.stream)
        self._raise_if_bad(self._clockv())

    def run(self, t_end: float, max_steps: int = 10 ** 7, callback: Optional[Callable] = None,
            dt_max: Optional[float] = None, batch: int = 64) -> int:
        """Step to t_end with CFL steps (solver.py:2975-3003); dt is planned on the device."""
        t0 = time.perf_counter()
        self._ensure_lam()
        n0 = self.nsteps
        if callback is not None:
            steps = 0
            while steps < max_steps:
                L.call("gs_clock_set_target", C.c_void_p(self._clock.ptr), C.c_double(float(t_end)),
                       C.c_double(float(dt_max) if dt_max else 0.0), C.c_int64(1), C.byref(self._phys()),
                       self._fo.s, self.stream)
                c = self._clockv()
                self._raise_if_bad(c)
                if c.status & L.ST_DONE:
                    break
                L.call("gs_dense_step", C.byref(self._desc()), C.byref(self._phys()), self._fo.s,
                       C.c_void_p(self._clock.ptr), self.stream)
                self._any_step = True
                steps += 1
                callback(self, steps)
        else:
            L.call("gs_clock_set_target", C.c_void_p(self._clock.ptr), C.c_double(float(t_end)),
                   C.c_double(float(dt_max) if dt_max else 0.0), C.c_int64(int(max_steps)),
                   C.byref(self._phys()), self._fo.s, self.stream)
            d, p = C.byref(self._desc()), C.byref(self._phys())
            clk = C.c_void_p(self._clock.ptr)
            while True:
                c = self._clockv()
                self._raise_if_bad(c)
                if c.status & L.ST_DONE:
                    break
                for _ in range(batch):
                    L.call("gs_dense_step", d, p, self._fo.s, clk, self.stream)
                self._any_step = True
        self._target = None
        steps = self.nsteps - n0
        self.diagnostics["wallclock"] = time.perf_counter() - t0
        self.diagnostics["steps"] = steps
        return steps

    # ------------------------------------------------------------------ parity helpers
    def residual(self) -> np.ndarray:
        """SRM-HLLC residual (3, nx, ny) of the current state (inactive cells 0)."""
        nx, ny = self.mesh.nx, self.mesh.ny
        r = [L.DeviceArray(self._plane, np.float64 if self.compute64 else self.dtype, zero=True) for _ in range(3)]
        L.call("gs_dense_residual", C.byref(self._desc()), C.byref(self._phys()), self._cur(),
               C.c_void_p(r[0].ptr), C.c_void_p(r[1].ptr), C.c_void_p(r[2].ptr), self.stream)
        return np.stack([x.download()[:, :ny] for x in r])
