"""Synthetic benchmark workloads, built on the device.

* ``dam_break_c1``: 256x256 circular dam break, flat bed, walls, no friction (host arrays).
* ``rain_on_slope_c2``: everywhere-wet rain on a 1e-3 slope with seeded fBm roughness
  (Hurst 0.8, 0.5 m amplitude), Manning n = 0.03, 50 mm/h, free-outflow edges. The bed is
  written by gs_gen_bed on the device, so the 640 M-cell grid never goes through the host.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from .config import Config, Mesh2D

C2_SEED = 20261202


@dataclass
class FbmBed:
    """Parameters of a gs_gen_bed field; ``mu``/``scale`` normalise the raw fBm."""
    seed: int
    octaves: int = 12
    L0: float = 4096.0
    hurst: float = 0.8
    amplitude: float = 0.5
    slope_x: float = 1.0e-3
    dx: float = 3.0
    offset: float = 0.0
    mu: float = 0.0
    scale: float = 1.0

    def normalise(self, nx: int, ny: int, samples: int = 256) -> "FbmBed":
        """Zero mean and max |.| = amplitude over a strided sample of the raw fBm, on the device."""
        stride = max(1, min(nx, ny) // samples)
        sx, sy = (nx + stride - 1) // stride, (ny + stride - 1) // stride
        buf = L.DeviceArray((sx, sy), np.float64)
        L.call("gs_gen_bed", C.c_void_p(buf.ptr), L.GS_F64, sx, sy, sy, 0, 0, stride, C.c_uint32(self.seed),
               self.octaves, C.c_double(self.L0), C.c_double(self.hurst), C.c_double(0.0), C.c_double(1.0),
               C.c_double(0.0), C.c_double(self.dx), C.c_double(0.0), None)
        raw = buf.download()
        self.mu = float(raw.mean())
        self.scale = float(self.amplitude / np.abs(raw - self.mu).max())
        return self

    def write(self, ptr: int, pitch: int, dcode: int, nx: int, ny: int, i0: int = 0, j0: int = 0) -> None:
        L.call("gs_gen_bed", C.c_void_p(ptr), dcode, nx, ny, pitch, i0, j0, 1, C.c_uint32(self.seed), self.octaves,
               C.c_double(self.L0), C.c_double(self.hurst), C.c_double(self.mu), C.c_double(self.scale),
               C.c_double(self.slope_x), C.c_double(self.dx), C.c_double(self.offset), None)

    def host_crop(self, i0, j0, nx, ny, dtype=np.float32) -> np.ndarray:
        """The same field on the host, through the device generator (small crops)."""
        buf = L.DeviceArray((nx, ny), np.dtype(dtype))
        self.write(buf.ptr, ny, L.GS_F32 if np.dtype(dtype) == np.float32 else L.GS_F64, nx, ny, i0, j0)
        return buf.download()


def dam_break_c1(n: int = 256, dtype: str = "float32"):
    """(mesh, cfg, q0, b) of the circular dam break: h = 10 inside r < 40 cells, 1 outside."""
    i = np.arange(n)[:, None]
    j = np.arange(n)[None, :]
    h = np.where(np.hypot(i - (n - 1) / 2.0, j - (n - 1) / 2.0) < 40, 10.0, 1.0)
    q0 = np.stack([h, 0 * h, 0 * h])
    mesh = Mesh2D(n, n, 1.0, 1.0, ngh=4)
    cfg = Config(bc_x="wall", bc_y="wall", cfl=0.5, dtype=dtype)
    return mesh, cfg, q0, np.zeros((n, n))


def rain_on_slope_c2(nx: int = 25600, ny: int = 25000, dtype: str = "float32", track_max_depth: bool = False,
                     tile_rows: int = 0, bed: FbmBed = None):
    """Solver2D of the rain-on-slope case, state and bed built on the device."""
    from .solver import Solver2D
    bed = bed or FbmBed(seed=C2_SEED).normalise(nx, ny)
    mesh = Mesh2D(nx, ny, 3.0, 3.0, ngh=4)
    cfg = Config(bc_x="fall", bc_y="fall", cfl=0.5, dtype=dtype, friction="manning", manning_n=0.03,
                 rainfall=50.0 / 3.6e6)
    s = Solver2D(mesh, cfg, None, None, track_max_depth=track_max_depth, tile_rows=tile_rows,
                 device_init=(0.1, lambda ptr, pitch, dcode: bed.write(ptr, pitch, dcode, nx, ny)))
    s.bed_spec = bed
    return s
