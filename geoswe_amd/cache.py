"""The flat cache in the reference's on-disk layout, and its cached replay.

``save_cache`` writes a :class:`CompressedSolver`'s mesh, state, bed, Manning classes and
forcing bundles as ``.npy`` files plus ``meta.json`` under the names and dtypes of
``geoswe.compressed_solver.save_cache`` (compressed_solver.py:2504-2569); ``from_cache``
and ``run_cached`` (2616-2859) load such a directory into a solver and run it, as the
reference's cached replay does: no rain step cap unless ``dt_max="rain"``, the ring mode the
reference's environment variables select, the ring, sponge, rain and karst-drain bundles
applied, Green-Ampt, the stage clamp and constant-rate
infiltration not stored (the reference warns about the same three).

Rain bundles hold a native table ``(nt, npix)`` in m/s and a per-cell index into it; a
uniform rate is written as a one-pixel table, (nx, ny) frames as a table of nx*ny columns.
On loading, the step reads the table through the per-cell index (a :class:`NativeRain`).
"""
from __future__ import annotations

import json
import os
import warnings
from typing import Optional

import numpy as np

from . import _lib as L
from .config import Config, Mesh2D
from .flatmesh import build_flat_mesh
from .forcing import NativeRain, rain_tables


def _nbr_int16(fm) -> np.ndarray:
    """(N, 4) int16 neighbour deltas of the stored cells (i+1, j), (i-1, j), (i, j+1),
    (i, j-1), -32768 where the neighbour is not stored (compressed_mesh.py:160-177 with
    nbr_to_int16_delta, compressed_rhs.py:194-235)."""
    from .adjoint import flat_neighbours
    nb = flat_neighbours(fm).astype(np.int64)
    k = np.arange(fm.n_stored, dtype=np.int64)[:, None]
    d = np.where(nb >= 0, nb - k, -32768)
    live = nb >= 0
    if live.any() and (np.abs(d[live]) >= 32768).any():
        raise ValueError("save_cache: a neighbour lies 32768 or more flat ids away (padded rows too wide for "
                         "the int16 table)")
    return d.astype(np.int16)


def save_cache(solver, cache_dir: str, x0: float = 0.0, y0: float = 0.0, crs_wkt: str = "") -> None:
    """Write `solver` (its current state) to `cache_dir` in the reference's single-rank
    cache layout: nbr, is_active, ij_active, q0-q2, bed_f, mcls_f, m_tab, meta.json
    (no_sigma), and the ring, sponge, rain and drain bundles it has."""
    s = solver
    if s.dtype != np.float32:
        raise ValueError("save_cache: the cache holds float32 state; this solver is float64")
    if s._n_field is not None:
        raise ValueError("save_cache: the solver has a per-cell Manning n plane, which the cache's class table "
                         "cannot hold")
    if s.cfg.stage_boundary is not None:
        raise ValueError("save_cache: a StageBoundary is not part of the cache format (the cached replay takes "
                         "its open boundary from the gauge ring)")
    omitted = [n for n, v in (("Green-Ampt infiltration (set_ga_drain)", s._ga), ("the stage clamp (set_clamp)",
                                                                               s._clamp),
                              ("constant-rate infiltration (set_infil)", s._infil)) if v is not None]
    if omitted:
        warnings.warn(f"save_cache: {' and '.join(omitted)} cannot be written to a cache, and run_cached cannot "
                      f"apply them; the replay will take a different trajectory from this run", stacklevel=2)
    fm = s.fm
    N = fm.n_stored
    g = fm.ngh
    nx, ny = s.mesh.nx, s.mesh.ny
    os.makedirs(cache_dir, exist_ok=True)

    def sv(name, a):
        np.save(os.path.join(cache_dir, name + ".npy"), np.asarray(a))

    sv("nbr", _nbr_int16(fm))
    sv("is_active", fm.active.astype(np.uint8))
    sv("ij_active", fm.ij.astype(np.int32))
    q = s.q_flat()
    for k in range(3):
        sv(f"q{k}", q[k].astype(np.float32))
    sv("bed_f", s._bed.download().astype(np.float32))
    sv("mcls_f", s._cls.download().astype(np.uint8) if s._cls is not None
       else (s._flags_base & np.uint8(L.FLAG_CLASS_MASK)).astype(np.uint8))
    sv("m_tab", s._ntab.download()[:s._n_classes].astype(np.float32))
    rt = rain_tables(s.cfg.rainfall_forcing, s.cfg.rainfall, s.cfg.cfl, s.cfg.g, min(s.mesh.dx, s.mesh.dy),
                     np.float32)
    meta = dict(nxp=int(fm.nxp), nyp=int(fm.nyp), ngh=int(g), dx=float(s.mesh.dx), x0=float(x0), y0=float(y0),
                crs_wkt=str(crs_wkt), nx_glob=int(nx), ny_glob=int(ny), has_ring=s._ring is not None,
                has_sponge=s._sponge is not None, has_rain=rt is not None, no_sigma=True, N=int(N), nranks=1,
                mpi=False)
    if s._drain is not None:
        hd = s._drain[0].download()
        ids = np.nonzero(s._drain[1])[0]
        sv("drain_idx", ids.astype(np.int32))
        sv("drain_htgt", hd[ids].astype(np.float32))
        meta["has_drain"] = True
        meta["drain_h_tgt"] = 0.5
    if s._ring is not None:
        r = s._ring
        sv("ring_rflat", r["ids"].astype(np.int32))
        sv("ring_rbed", r["bed"].download()[r["rows"]].astype(np.float32))
        sv("ring_rwg", r["w"].download().reshape(-1, r["ng"])[r["rows"]].reshape(-1).astype(np.float32))
        sv("ring_t_common", r["t"].download())
        sv("ring_stage_all", r["stage"].download().reshape(r["ng"], r["nt"]).astype(np.float32))
        meta["ring_NG"] = int(r["ng"])
        meta["ring_n"] = int(len(r["ids"]))
    if s._sponge is not None:
        ka = s._sponge[0].download().reshape(N, 2)
        sv("sponge_keep_f", ka[:, 0].astype(np.float32))
        sv("sponge_amb_f", ka[:, 1].astype(np.float32))
    if rt is not None:
        t, rate, _, frames = rt
        if isinstance(s.cfg.rainfall_forcing, NativeRain):
            native = frames.astype(np.float32)
            lookup = s._rain_lookup.download().astype(np.int32)
        elif frames is None:
            native = rate.astype(np.float32)[:, None]
            lookup = np.zeros(N, np.int32)
        else:
            native = frames.reshape(frames.shape[0], nx * ny).astype(np.float32)
            ii, jj = fm.ij[:, 0].astype(np.int64) - g, fm.ij[:, 1].astype(np.int64) - g
            inb = fm.active & (ii >= 0) & (ii < nx) & (jj >= 0) & (jj < ny)
            lookup = np.where(inb, ii * ny + jj, 0).astype(np.int32)
        sv("rain_native_rate_dev", native)
        sv("rain_lookup_flat", lookup)
        sv("rain_t_s", np.asarray(t, np.float64))
    with open(os.path.join(cache_dir, "meta.json"), "w") as f:
        json.dump(meta, f)


def from_cache(cache_dir: str, cfl: float = 0.5, h_min: float = 1e-6, g: float = 9.81,
               rain_cap: bool = False, track_max_depth: bool = True, cfl_linf: bool = False, sweep_rows=None):
    """A :class:`CompressedSolver` over a single-rank cache directory (the loading part of
    run_cached, compressed_solver.py:2616-2780). The cache's stored set must be its active
    set dilated twice (the reference builds it so when the mask stays two cells clear of
    the box edge). `rain_cap`: apply the rain-film step bound (run_cached's
    ``dt_max="rain"``); the cached replay has none by default."""
    from .compressed import CompressedSolver
    with open(os.path.join(cache_dir, "meta.json")) as f:
        meta = json.load(f)
    if int(meta.get("nranks", 1)) != 1 or meta.get("mpi"):
        raise RuntimeError(f"cache {cache_dir} was saved for nranks={meta.get('nranks')}; one GPU runs one rank")
    if not meta.get("no_sigma", False) and os.path.exists(os.path.join(cache_dir, "sig_f.npy")):
        if np.load(os.path.join(cache_dir, "sig_f.npy")).any():
            raise NotImplementedError("from_cache: sub-grid storage (sig_f) is not part of this solver")
    if os.environ.get("SWE_GHOST_ETA_RAMP_MMHR"):
        raise NotImplementedError("SWE_GHOST_ETA_RAMP_MMHR is set but the ring stage is constant; unset it")

    def ld(name):
        return np.load(os.path.join(cache_dir, name + ".npy"))

    ij = ld("ij_active").astype(np.int64)
    act = ld("is_active") != 0
    N = len(act)
    if int(meta.get("N", N)) != N or len(ij) != N:
        raise RuntimeError(f"cache {cache_dir}: meta N={meta.get('N')}, is_active {N}, ij_active {len(ij)} rows "
                           f"(torn or mixed-vintage cache)")
    nxp, nyp, ngh = int(meta["nxp"]), int(meta["nyp"]), int(meta["ngh"])
    inside = np.zeros((nxp, nyp), bool)
    inside[ij[act, 0], ij[act, 1]] = True
    fm = build_flat_mesh(inside, ngh, ring=2, sweep_rows=sweep_rows)
    if fm.n_stored != N or not np.array_equal(fm.ij, ij) or not np.array_equal(fm.active, act):
        raise ValueError(f"cache {cache_dir}: its stored cells are not its active cells dilated twice (a mask that "
                         f"reaches the box edge is stored without ghost layers there)")
    nx, ny = nxp - 2 * ngh, nyp - 2 * ngh
    dx = float(meta["dx"])
    mesh = Mesh2D(nx, ny, dx, dx, ngh=ngh)
    rain = None
    if meta.get("has_rain"):
        native = ld("rain_native_rate_dev").astype(np.float32)
        lookup_flat = ld("rain_lookup_flat").astype(np.int64)
        # the (nx, ny) column map the solver takes, from the per-cell one
        lk = np.zeros((nx, ny), np.int64)
        ii, jj = ij[act, 0] - ngh, ij[act, 1] - ngh
        inb = (ii >= 0) & (ii < nx) & (jj >= 0) & (jj < ny)
        lk[ii[inb], jj[inb]] = lookup_flat[act][inb]
        rain = NativeRain(time_s=ld("rain_t_s").astype(np.float64), table=native, lookup=lk)
    cfg = Config(bc_x="fall", bc_y="fall", cfl=cfl, g=g, h_min=h_min, dtype="float32", friction="manning")
    q = np.stack([ld(f"q{k}").astype(np.float32) for k in range(3)])
    m_tab = ld("m_tab").astype(np.float64)
    mcls = fm.unpack(ld("mcls_f").astype(np.uint8))
    if len(m_tab) > L.MAX_CLASSES:     # more classes than the flags byte holds: a class plane
        s = CompressedSolver(mesh, cfg, fm, fm.unpack(q), fm.unpack(ld("bed_f").astype(np.float32)),
                             np.zeros_like(mcls), m_tab, None, cfl_linf, track_max_depth, cls_pad=mcls)
    else:
        s = CompressedSolver(mesh, cfg, fm, fm.unpack(q), fm.unpack(ld("bed_f").astype(np.float32)),
                             mcls, m_tab, None, cfl_linf, track_max_depth)
    if rain is not None:
        s.set_rain(rain)
        if not rain_cap:
            s._fo.drop_rain_cap()
    # the ring mode of the replay, from the reference's environment variables (2730-2735)
    eta = os.environ.get("GEOSWE_RING_ETA", os.environ.get("SWE_GHOST_ETA"))
    mode = os.environ.get("GEOSWE_RING_BC", "auto").lower()
    if os.environ.get("SWE_GHOST_UV") in ("1", "true", "True") and mode in ("auto", "stage"):
        mode = "stage_uv"
    cls_env = os.environ.get("GEOSWE_RING_CLASSIFY")
    s.set_ring_mode(mode, eta=None if eta is None else float(eta),
                    classify=None if cls_env is None else float(cls_env))
    if meta.get("has_sponge"):
        s.set_sponge({"keep": fm.unpack(ld("sponge_keep_f"), fill=1.0), "amb": fm.unpack(ld("sponge_amb_f"))})
    if meta.get("has_ring"):
        rflat = ld("ring_rflat").astype(np.int64)
        ng = int(meta["ring_NG"])
        keep = rflat >= 0
        s.set_ring({"i": ij[rflat[keep], 0], "j": ij[rflat[keep], 1], "bed": ld("ring_rbed")[keep],
                    "wg": ld("ring_rwg").reshape(-1, ng)[keep], "NG": ng, "n": int(keep.sum()),
                    "t_common": ld("ring_t_common"), "stage_all": ld("ring_stage_all")})
    if meta.get("has_drain") and os.path.exists(os.path.join(cache_dir, "drain_idx.npy")):
        idx = ld("drain_idx")
        if os.path.exists(os.path.join(cache_dir, "drain_htgt.npy")):
            s.set_drain({"idx": idx, "h_tgt_arr": ld("drain_htgt")})
        else:
            s.set_drain({"idx": idx, "h_tgt": float(meta.get("drain_h_tgt", 0.5))})
    return s


def run_cached(cache_dir: str, *, t_end: float, frame_every_s: float = 0.0, out_dir: Optional[str] = None,
               cfl: float = 0.5, h_min: float = 1e-6, g: float = 9.81, checkpoint_every_s: float = 0.0,
               ckpt_dir: Optional[str] = None, resume: bool = False, dt_max=None, say=print):
    """Load a cache and run it to `t_end` (run_cached, compressed_solver.py:2616-2859):
    frames into `out_dir` every `frame_every_s`, checkpoints into `ckpt_dir`, ``resume``
    from the last one; ``dt_max`` a number, or ``"rain"`` for the rain-film step bound.
    Returns the solver."""
    s = from_cache(cache_dir, cfl=cfl, h_min=h_min, g=g, rain_cap=(dt_max == "rain"))
    if say is not None:
        say(f"  [cache] {cache_dir}: {s.n_active} active of {s.n_stored} stored cells")
    s.run(t_end, dt_max=None if isinstance(dt_max, str) else dt_max, checkpoint_every_s=checkpoint_every_s,
          ckpt_dir=ckpt_dir, resume=resume, out_dir=out_dir, frame_every_s=frame_every_s)
    return s
