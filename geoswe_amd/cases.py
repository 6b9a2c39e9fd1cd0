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
               C.c_double(0.0), C.c_double(self.dx), C.c_double(0.0), 0, None)
        raw = buf.download()
        self.mu = float(raw.mean())
        self.scale = float(self.amplitude / np.abs(raw - self.mu).max())
        return self

    def write(self, ptr: int, pitch: int, dcode: int, nx: int, ny: int, i0: int = 0, j0: int = 0,
              round_f32: bool = False) -> None:
        L.call("gs_gen_bed", C.c_void_p(ptr), dcode, nx, ny, pitch, i0, j0, 1, C.c_uint32(self.seed), self.octaves,
               C.c_double(self.L0), C.c_double(self.hurst), C.c_double(self.mu), C.c_double(self.scale),
               C.c_double(self.slope_x), C.c_double(self.dx), C.c_double(self.offset), int(round_f32), None)

    def host_crop(self, i0, j0, nx, ny, dtype=np.float32) -> np.ndarray:
        """The same field on the host, through the device generator (small crops)."""
        buf = L.DeviceArray((nx, ny), np.dtype(dtype))
        self.write(buf.ptr, ny, L.GS_F32 if np.dtype(dtype) == np.float32 else L.GS_F64, nx, ny, i0, j0)
        return buf.download()


def dam_break_c1(n: int = 256, dtype: str = "float32"):
    """(mesh, cfg, q0, b) of the circular dam break: h = 10 inside r < 40 cells, 1 outside;
    dry floor 1e-6 m in either precision."""
    i = np.arange(n)[:, None]
    j = np.arange(n)[None, :]
    h = np.where(np.hypot(i - (n - 1) / 2.0, j - (n - 1) / 2.0) < 40, 10.0, 1.0)
    q0 = np.stack([h, 0 * h, 0 * h])
    mesh = Mesh2D(n, n, 1.0, 1.0, ngh=4)
    cfg = Config(bc_x="wall", bc_y="wall", cfl=0.5, dtype=dtype, h_min=1.0e-6, h_min_cfl=1.0e-6)
    return mesh, cfg, q0, np.zeros((n, n))


def rain_on_slope_c2(nx: int = 25600, ny: int = 25000, dtype: str = "float32", track_max_depth: bool = False,
                     tile_rows: int = 0, bed: FbmBed = None, f32_inputs: bool = False, compute_dtype: str = None):
    """Solver2D of the rain-on-slope case, state and bed built on the device.

    ``f32_inputs`` rounds every input (bed, h0, n, rain rate) to float32 first and takes the
    float32 run's wet/dry thresholds (h_min = h_min_cfl = 1e-6 instead of the float64
    default 1e-10): a float64 run on exactly the inputs and parameters of the float32 run,
    which isolates the arithmetic error. ``compute_dtype="float64"`` with dtype float32 steps
    the float32 planes with float64 arithmetic."""
    from .solver import Solver2D
    bed = bed or FbmBed(seed=C2_SEED).normalise(nx, ny)
    r = (lambda x: float(np.float32(x))) if f32_inputs else (lambda x: x)
    thr = dict(h_min=1.0e-6, h_min_cfl=1.0e-6) if f32_inputs else {}
    mesh = Mesh2D(nx, ny, 3.0, 3.0, ngh=4)
    cfg = Config(bc_x="fall", bc_y="fall", cfl=0.5, dtype=dtype, friction="manning", manning_n=r(0.03),
                 rainfall=r(50.0 / 3.6e6), compute_dtype=compute_dtype, **thr)
    s = Solver2D(mesh, cfg, None, None, track_max_depth=track_max_depth, tile_rows=tile_rows,
                 device_init=(r(0.1), lambda ptr, pitch, dcode: bed.write(ptr, pitch, dcode, nx, ny,
                                                                          round_f32=f32_inputs)))
    s.bed_spec = bed
    return s


C3_SEED = 20261203
TIDE_PERIOD_S = 44712.0     # M2 tide, 12.42 h


@dataclass
class CoastalCase:
    """Host arrays of the compressed coastal workload (unpadded (n, n) fields)."""
    mesh: Mesh2D
    cfg: Config
    q0: np.ndarray            # (3, n, n) float32
    bed: np.ndarray           # (n, n) float32
    active: np.ndarray        # (n, n) bool: land above -2 m plus the nearshore band
    cls: np.ndarray           # (n, n) uint8 Manning class
    ntab: np.ndarray          # (8,) Manning n per class
    sea_edge: np.ndarray      # (n, n) bool: stage cells (active box-edge cells below -2 m)

    def dense(self, track_max_depth: bool = False):
        """Solver2D with the mask, classes and a CFL that skips inactive cells (so that the
        dense and flat layouts take the same steps)."""
        from .solver import Solver2D
        s = Solver2D(self.mesh, self.cfg, self.q0, self.bed, track_max_depth=track_max_depth)
        g = self.mesh.ngh
        cls_pad = np.zeros((self.mesh.nx + 2 * g, self.mesh.ny + 2 * g), np.uint8)
        cls_pad[g:-g, g:-g] = self.cls
        s.set_manning_table(cls_pad, self.ntab)
        s.set_inside_mask(self.active)
        s.set_cfl_ghost_mask(~self.active)
        return s

    def flat(self, track_max_depth: bool = False, sweep_rows: int = None):
        from .compressed import CompressedSolver
        return CompressedSolver.from_dense(self.dense(), cfl_linf=True, track_max_depth=track_max_depth,
                                           sweep_rows=sweep_rows)


def attach_coastal_forcings(case: CoastalCase, f, band: int = 32, n_clamp: int = 10000, n_drain: int = 5000,
                            seed: int = 20261207, fields: bool = False) -> dict:
    """A coastal driver's forcing tail on the flat solver `f` of `case` (before its first
    step). Sponge: a `band`-cell strip along the box edges relaxing toward still water at
    eta = 0 (keep = 1 - 0.05 r, amb = 0.05 r max(-b, 0), r rising from 0 to 1 toward the
    edge). IDW gauge ring: every stored inactive cell beside an active one gets still water
    at the inverse-square-distance mean of four corner tide gauges (M2 tide, phase lags of
    0-45 min, sampled every 360 s). Clamp: `n_clamp` seeded active cells capped at 1.5 m.
    Infiltration: 0-10 mm/h by Manning class. Karst drain: `n_drain` seeded active cells
    capped at 0.5 m. Green-Ampt by Manning class (Ks 0-20 mm/h, suction 5-30 cm, deficit
    0.2-0.4, storage capped at 0.3 m) with a 2-hour drain-tau on every third active cell.
    Returns the number of cells each one touches (``"cells"``) and, with
    `fields`, the same forcings on the unpadded grid (``"fields"``: keep, amb, ring cells,
    bed, weights, times and stages, clamp and drain depth fields, +inf where none)."""
    fm = f.fm
    g, n = case.mesh.ngh, case.mesh.nx
    rng = np.random.default_rng(seed)
    ij = fm.ij.astype(np.int64)
    ii, jj = ij[:, 0] - g, ij[:, 1] - g
    inner = (ii >= 0) & (ii < n) & (jj >= 0) & (jj < n)
    dist = np.minimum(np.minimum(ii, jj), np.minimum(n - 1 - ii, n - 1 - jj))
    in_band = inner & (dist < band)
    r = np.where(in_band, (band - dist) / band, 0.0)
    bed_f = np.zeros(fm.n_stored, np.float32)
    bed_f[inner] = case.bed[ii[inner], jj[inner]]
    keep = np.ones((fm.nxp, fm.nyp), np.float32)
    amb = np.zeros((fm.nxp, fm.nyp), np.float32)
    keep[ij[in_band, 0], ij[in_band, 1]] = 1.0 - 0.05 * r[in_band]
    amb[ij[in_band, 0], ij[in_band, 1]] = 0.05 * r[in_band] * np.maximum(-bed_f[in_band], 0.0)
    f.set_sponge({"keep": keep, "amb": amb})
    del keep, amb
    # ring: stored inactive cells with an active 4-neighbour, inside the box
    act_pad = np.zeros((fm.nxp, fm.nyp), bool)
    act_pad[ij[fm.active, 0], ij[fm.active, 1]] = True
    cand = np.nonzero(~fm.active & inner)[0]
    ci, cj = ij[cand, 0], ij[cand, 1]
    near = act_pad[ci + 1, cj] | act_pad[ci - 1, cj] | act_pad[ci, cj + 1] | act_pad[ci, cj - 1]
    ring = cand[near]
    del act_pad
    gauges = np.array([[0.0, 0.0], [0.0, n - 1.0], [n - 1.0, 0.0], [n - 1.0, n - 1.0]])
    d2 = (ii[ring, None] - gauges[None, :, 0]) ** 2 + (jj[ring, None] - gauges[None, :, 1]) ** 2 + 1.0
    w = 1.0 / d2
    w /= w.sum(axis=1, keepdims=True)
    tc = np.arange(-3600.0, 2 * TIDE_PERIOD_S + 3600.0, 360.0)
    lag = np.array([0.0, 900.0, 1800.0, 2700.0])
    S = np.sin(2.0 * np.pi * (tc[None, :] - lag[:, None]) / TIDE_PERIOD_S).astype(np.float32)
    f.set_ring({"i": ij[ring, 0], "j": ij[ring, 1], "bed": bed_f[ring], "wg": w.astype(np.float32), "NG": 4,
                "n": len(ring), "t_common": tc, "stage_all": S})
    act_ids = np.nonzero(fm.active)[0]
    cl = rng.choice(act_ids, size=min(n_clamp, len(act_ids)), replace=False)
    f.set_clamp({"rows": ij[cl, 0], "cols": ij[cl, 1], "hmax": np.full(len(cl), 1.5)})
    f.set_infil({"tab": np.linspace(0.0, 10.0, 8) / 3.6e6})
    dr = rng.choice(act_ids, size=min(n_drain, len(act_ids)), replace=False)
    f.set_drain({"idx": dr, "h_tgt": 0.5})
    ks = np.linspace(0.0, 20.0, 8) / 3.6e6
    psi = np.linspace(0.05, 0.3, 8)
    dth = np.linspace(0.2, 0.4, 8)
    cls_pad = np.zeros((fm.nxp, fm.nyp), np.uint8)
    cls_pad[g:g + n, g:g + n] = case.cls
    tau = np.zeros((fm.nxp, fm.nyp), np.float32)
    tau[ij[act_ids[::3], 0], ij[act_ids[::3], 1]] = 1.0 / 7200.0
    f.set_ga_drain({"cls_pad": cls_pad, "Ks_t": ks, "psi_t": psi, "dth_t": dth, "F_pad": None,
                    "Fmax_pad": np.full((fm.nxp, fm.nyp), 0.3, np.float32), "inv_tau_pad": tau, "mode": "fused"})
    del cls_pad
    out = {"cells": {"sponge_band": int(in_band.sum()), "gauge_ring": int(len(ring)), "clamp": int(len(cl)),
                     "drain": int(len(dr)), "infiltration": int(fm.n_active), "green_ampt": int(fm.n_active),
                     "drain_tau": int(len(act_ids[::3]))}}
    if fields:
        kg = np.ones((n, n))
        ag = np.zeros((n, n))
        kg[ii[in_band], jj[in_band]] = 1.0 - 0.05 * r[in_band]
        ag[ii[in_band], jj[in_band]] = 0.05 * r[in_band] * np.maximum(-bed_f[in_band], 0.0)
        cg = np.full((n, n), np.inf)
        cg[ii[cl], jj[cl]] = 1.5
        dg = np.full((n, n), np.inf)
        dg[ii[dr], jj[dr]] = 0.5
        taug = np.zeros((n, n))
        taug[ii[act_ids[::3]], jj[act_ids[::3]]] = np.float32(1.0 / 7200.0)
        out["fields"] = {"keep": kg.astype(np.float32), "amb": ag.astype(np.float32), "clamp": cg, "drain": dg,
                         "ga": {"cls": case.cls, "Ks": ks, "psi": psi, "dth": dth, "F": None,
                                "Fmax": np.full((n, n), 0.3), "inv_tau": taug},
                         "infil": np.linspace(0.0, 10.0, 8) / 3.6e6,
                         "ring": {"cells": np.stack([ii[ring], jj[ring]], axis=1), "bed": bed_f[ring],
                                  "w": w.astype(np.float32), "t_common": tc, "stage_all": S}}
    return out


def _taxicab_dilate(mask: np.ndarray, r: int) -> np.ndarray:
    """Cross (4-neighbour) dilation by r iterations = taxicab distance <= r."""
    from scipy.ndimage import distance_transform_cdt
    if not mask.any():
        return mask.copy()
    return distance_transform_cdt(~mask, metric="taxicab") <= r


def coastal_c3(n: int = 16384, seed: int = C3_SEED, band: int = 64, active_share: float = 0.45,
               n_sites: int = 256, rain_frames: int = 3, dx: float = 3.0) -> CoastalCase:
    """The seeded compressed coastal case.

    Bed: value-noise fBm (Hurst 0.8) generated on the device, mapped piecewise-linearly so
    that the -2 m contour sits at the quantile that makes land (b > -2 m) plus its
    `band`-cell nearshore band cover ~`active_share` of the box, with elevations spanning
    -10 .. 60 m. Manning: 8 classes n = linspace(0.02, 0.12, 8) by Voronoi over `n_sites`
    seeded sites (device). Rain: a smooth correlated field scaled to 0..80 mm/h, a new frame
    every 300 s. Tide: eta(t) = sin(2 pi t / 12.42 h) on the active box-edge cells below
    -2 m. Initial state: still water at eta = 0 where b < 0, dry elsewhere. Free-outflow
    ('fall') box edges, h_min 1e-6, float32.
    """
    from .forcing import RainfallForcing, StageBoundary
    rng = np.random.default_rng(seed)
    spec = FbmBed(seed=seed, octaves=max(4, int(np.log2(n)) - 1), L0=max(64.0, n / 4.0), slope_x=0.0, dx=dx)
    raw_dev = L.DeviceArray((n, n), np.float32)
    spec.write(raw_dev.ptr, n, L.GS_F32, n, n)
    raw = raw_dev.download()
    del raw_dev
    # choose the land quantile on a coarse copy so that land + band ~ active_share
    f = max(1, n // 1024)
    coarse = raw[::f, ::f]
    lo, hi = 0.0, 1.0
    for _ in range(18):
        mid = 0.5 * (lo + hi)
        thr = np.quantile(coarse, 1.0 - mid)
        share = _taxicab_dilate(coarse > thr, max(1, band // f)).mean()
        lo, hi = (mid, hi) if share < active_share else (lo, mid)
    land_frac = 0.5 * (lo + hi)
    r_star = float(np.quantile(raw[::max(1, n // 2048), ::max(1, n // 2048)], 1.0 - land_frac))
    rmin, rmax = float(raw.min()), float(raw.max())
    below = raw < r_star
    bed = np.where(below, -2.0 + (raw - r_star) * (8.0 / max(r_star - rmin, 1e-12)),
                   -2.0 + (raw - r_star) * (62.0 / max(rmax - r_star, 1e-12))).astype(np.float32)
    del raw, below
    land = bed > -2.0
    active = _taxicab_dilate(land, band)
    # Manning classes by Voronoi (device)
    sx = rng.uniform(0, n, n_sites).astype(np.float32)
    sy = rng.uniform(0, n, n_sites).astype(np.float32)
    sc = rng.integers(0, 8, n_sites).astype(np.uint8)
    dsx, dsy, dsc = (L.DeviceArray.from_host(a) for a in (sx, sy, sc))
    dcls = L.DeviceArray((n, n), np.uint8)
    L.call("gs_gen_voronoi", C.c_void_p(dcls.ptr), n, n, n, 0, 0, C.c_void_p(dsx.ptr), C.c_void_p(dsy.ptr),
           C.c_void_p(dsc.ptr), n_sites, None)
    cls = dcls.download()
    ntab = np.linspace(0.02, 0.12, 8)
    # rain frames (mm/h), smooth correlated fields in [0, 80]
    frames = []
    for k in range(rain_frames):
        rs = FbmBed(seed=seed + 1000 + k, octaves=4, L0=max(32.0, n / 2.0), hurst=0.8, slope_x=0.0)
        d = L.DeviceArray((n, n), np.float32)
        rs.write(d.ptr, n, L.GS_F32, n, n)
        fr = d.download()
        a, b = float(fr.min()), float(fr.max())
        frames.append(np.clip((fr - a) * (80.0 / max(b - a, 1e-12)), 0.0, 80.0).astype(np.float32))
    rain = RainfallForcing(time_s=np.arange(rain_frames) * 300.0, rate_mm_h=np.stack(frames))
    # tide on the active box-edge cells below -2 m
    edge = np.zeros((n, n), bool)
    edge[0, :] = edge[-1, :] = edge[:, 0] = edge[:, -1] = True
    sea_edge = edge & active & (bed < -2.0)
    mesh = Mesh2D(n, n, dx, dx, ngh=4)
    ts = np.linspace(0.0, 2.0 * TIDE_PERIOD_S, 4001)
    stage = StageBoundary.from_mask(sea_edge, mesh, bed, ts, np.sin(2.0 * np.pi * ts / TIDE_PERIOD_S))
    cfg = Config(bc_x="fall", bc_y="fall", cfl=0.5, dtype="float32", friction="manning",
                 rainfall_forcing=rain, stage_boundary=stage)
    h0 = np.maximum(-bed, 0.0).astype(np.float32)
    q0 = np.stack([h0, np.zeros_like(h0), np.zeros_like(h0)])
    return CoastalCase(mesh=mesh, cfg=cfg, q0=q0, bed=bed, active=active, cls=cls, ntab=ntab, sea_edge=sea_edge)


C4_SEED = 20261204


def ensemble_sides(n_members: int = 64, seed: int = C4_SEED, lo_exp: float = 8.0, span_exp: float = 4.0,
                   quantum: int = 64) -> list:
    """Member box sides L_k = 2^(lo_exp + span_exp*u_k) (u_k seeded uniform), rounded to a
    multiple of `quantum`: log-uniform between 2^lo_exp and 2^(lo_exp+span_exp)."""
    u = np.random.default_rng(seed).uniform(0.0, 1.0, n_members)
    return [int(max(quantum, round(2.0 ** (lo_exp + span_exp * x) / quantum) * quantum)) for x in u]


def ensemble_c4(n_members: int = 64, seed: int = C4_SEED, sides: list = None, rain_frames: int = 3,
                log=None) -> list:
    """CoastalCase members of the batched-ensemble workload: member k is the coastal
    generator at side L_k with seed C3_SEED + k and a nearshore band scaled with the box
    (64 cells at 16384), its rain frames scaled by U[0.5, 2] and each class's Manning n by
    1 + U[-0.3, 0.3] (seeded)."""
    from .forcing import RainfallForcing
    sides = sides or ensemble_sides(n_members, seed)
    rng = np.random.default_rng(seed + 1)
    out = []
    for k, n in enumerate(sides):
        band = max(4, int(round(64 * n / 16384)))
        case = coastal_c3(n=n, seed=C3_SEED + k, band=band, n_sites=max(16, min(256, n // 16)),
                          rain_frames=rain_frames)
        rs = float(rng.uniform(0.5, 2.0))
        nf = 1.0 + rng.uniform(-0.3, 0.3, len(case.ntab))
        rf = case.cfg.rainfall_forcing
        case.cfg.rainfall_forcing = RainfallForcing(time_s=rf.time_s,
                                                    rate_mm_h=(np.asarray(rf.rate_mm_h) * rs).astype(np.float32))
        case.ntab = case.ntab * nf
        out.append(case)
        if log:
            log(f"member {k}: {n}x{n}, active {case.active.mean():.1%}, rain x{rs:.2f}")
    return out
