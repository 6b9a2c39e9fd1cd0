"""CompressedSolver: the compressed active-cell API of GeoSWE on the MI355X fused step.

Drop-in for ``geoswe.CompressedSolver`` (src/geoswe/compressed_solver.py:2882-3450) on
the production scheme: ``from_dense`` packs a dense solver's active set (its inside
mask, or every cell) into the flat layout of ``flatmesh``; ``run``/``step`` call
``gs_flat_step``, whose arithmetic is the dense kernel's, so the residual is bitwise the
dense one at a fixed state.
"""
from __future__ import annotations

import ctypes as C
import json
import math
import os
import time
from typing import Optional

import numpy as np

from . import _lib as L
from .config import Config, check_numeric_inputs
from .flatmesh import build_flat_mesh
from .forcing import NativeRain, RainfallForcing
from .solver import Solver2D, _Forcing, host_bc_fill


class CompressedSolver:
    """Flat-storage solver over the active cells plus a 2-cell ring (compressed_solver.py:2882)."""

    def __init__(self, mesh, cfg: Config, fm, q_pad: np.ndarray, b_pad: np.ndarray, flags_pad: np.ndarray,
                 ntab: np.ndarray, n_field_pad: Optional[np.ndarray], cfl_linf: bool, track_max_depth: bool = True,
                 h_min_cfl: Optional[float] = None, cls_pad: Optional[np.ndarray] = None):
        cfg.require_production_scheme("CompressedSolver")
        self.mesh, self.cfg, self.fm = mesh, cfg, fm
        self.dtype = np.dtype(cfg.dtype)
        self.cfl_linf = bool(cfl_linf)
        # The CFL velocity-divisor floor of the flat layout (CompressedStepper,
        # compressed_solver.py:599-604): the argument, else SWE_HMIN_CFL, else h_min, floored at
        # 1e-6 in float32. As there, the dense Config's h_min_cfl does not carry over.
        if h_min_cfl is None:
            env = os.environ.get("SWE_HMIN_CFL", "")
            h_min_cfl = float(env) if env else float(cfg.h_min)
        h_min_cfl = float(h_min_cfl)
        if not (h_min_cfl > 0.0 and math.isfinite(h_min_cfl)):
            raise ValueError(f"CompressedSolver: h_min_cfl={h_min_cfl} must be positive and finite")
        self.h_min_cfl = max(h_min_cfl, 1.0e-6) if self.dtype == np.float32 else h_min_cfl
        N = fm.n_stored
        qf = fm.pack(q_pad).astype(self.dtype)
        self._bufs = [[L.DeviceArray.from_host(np.ascontiguousarray(qf[k])) for k in range(3)] for _ in range(2)]
        self._bed = L.DeviceArray.from_host(np.ascontiguousarray(fm.pack(b_pad).astype(self.dtype)))
        # per-cell flags (Manning class, stage, CFL skip); none at all when every byte is zero
        # (uniform n, no stage cells, no CFL mask): the kernels then read no flags plane
        fl = np.ascontiguousarray(fm.pack(flags_pad).astype(np.uint8))
        self._flags = L.DeviceArray.from_host(fl) if fl.any() else None
        self._flags_base = fl              # without the coastal-forcing bit (set_sponge / set_clamp)
        self._sponge = self._clamp = self._force_ring = self._force_ring_row = self._ring = None
        self._infil = self._drain = self._ga = None
        self._rain_lookup = None
        self._rmode = None
        self._tables_taken = False         # an Ensemble or GaugeMisfit holds this solver's device pointers
        t = np.zeros(L.MAX_CLASSES_PLANE, self.dtype)
        t[:len(ntab)] = ntab
        self._ntab = L.DeviceArray.from_host(t)
        self._n_classes = max(1, len(ntab))
        self._n_field = (L.DeviceArray.from_host(np.ascontiguousarray(fm.pack(n_field_pad).astype(self.dtype)))
                         if n_field_pad is not None else None)
        # Manning classes 0..255 in a plane of their own (tables of more than 32 classes)
        self._cls = (L.DeviceArray.from_host(np.ascontiguousarray(fm.pack(cls_pad).astype(np.uint8)))
                     if cls_pad is not None else None)
        self._max_h = L.DeviceArray((N,), self.dtype, zero=True) if track_max_depth else None
        self._runs = [L.DeviceArray.from_host(a) for a in (fm.run_k0, fm.run_len, fm.run_i, fm.run_j0,
                                                            fm.run_d.reshape(-1), fm.chunk_run)]
        self._tiles = [L.DeviceArray.from_host(a) for a in (fm.tile_e0, fm.tile_irc.reshape(-1), fm.ent.reshape(-1))]
        # "tiles": the step sweeps 64-column windows down the rows, each face once (default);
        # "chunks": it walks 60-id chunks of the run table (a second traversal, same bits)
        self.traversal = "tiles"
        self._fo = self._forcing()
        self.stream = None
        self._clock = L.new_clock(0.0)
        self._lam_valid = False
        self._target = None
        self._any_step = False
        self.diagnostics = {"wallclock": 0.0, "steps": 0}

    # ------------------------------------------------------------------ construction
    @classmethod
    def from_dense(cls, s: Solver2D, cfl_linf: bool = False, track_max_depth: bool = True, sweep_rows: int = None,
                   h_min_cfl: Optional[float] = None, **_ignored):
        """Pack a dense solver (its current state, bed, Manning, mask, forcings). `sweep_rows`:
        rows per row-sweep tile of the step (None: chosen from the domain's size). `h_min_cfl`:
        the CFL velocity-divisor floor of the flat wave speed (the reference's
        CompressedStepper argument; None: SWE_HMIN_CFL, else h_min)."""
        mesh, cfg = s.mesh, s.cfg
        if getattr(s, "compute64", False):
            raise NotImplementedError("CompressedSolver.from_dense: float64 arithmetic on float32 planes "
                                      "(Config.compute_dtype='float64') is a dense-layout mode")
        g, nx, ny = mesh.ngh, mesh.nx, mesh.ny
        nxp, nyp = nx + 2 * g, ny + 2 * g
        inside = np.zeros((nxp, nyp), bool)
        inside[g:g + nx, g:g + ny] = s.inside_mask if s.inside_mask is not None else True
        fm = build_flat_mesh(inside, g, ring=2, sweep_rows=sweep_rows)
        q_pad = s.q                                  # ghosts filled from the boundary kinds
        b_pad = s.b
        flags = np.zeros((nxp, nyp), np.uint8)
        if s._flags_host is not None:
            flags[g:g + nx, g:g + ny] = s._flags_host[:, :ny] & np.uint8(~L.FLAG_INACTIVE & 0xFF)
        ntab = s._ntab.download()[:s._n_classes]
        nf = None
        if s._n_field is not None:
            nf = np.zeros((nxp, nyp), s.dtype)
            nf[g:g + nx, g:g + ny] = s._n_field.download()[:, :ny]
        cp = None
        if s._cls is not None:
            cp = np.zeros((nxp, nyp), np.uint8)
            cp[g:g + nx, g:g + ny] = s._cls.download()[:, :ny]
        out = cls(mesh, cfg, fm, q_pad, b_pad, flags, ntab, nf, cfl_linf, track_max_depth, h_min_cfl, cls_pad=cp)
        c = s._clockv()
        if c.t != 0.0:
            L.call("gs_clock_reset", C.c_void_p(out._clock.ptr), C.c_double(float(c.t)), None)
        return out

    def _forcing(self) -> _Forcing:
        """Device rain and stage tables; the rain-film step bound from the largest rate of the
        whole table, as the reference's flat run sets it (compressed_solver.py:3429-3436)."""
        fo = _Forcing(self.cfg, min(self.mesh.dx, self.mesh.dy), self.dtype, (self.mesh.nx, self.mesh.ny))
        fo.use_global_rain_cap()
        return fo

    @property
    def n_active(self) -> int:
        return self.fm.n_active

    @property
    def n_stored(self) -> int:
        return self.fm.n_stored

    def set_rain(self, forcing):
        """Replace the rain forcing (compressed_solver.py:3157-3183): a :class:`RainfallForcing`
        (uniform, or one (nx, ny) frame per time), a :class:`NativeRain`, or the run driver's
        bundle ``{"native_rate_dev", "lookup_dev", "t_s"}`` (a rate table on its native grid in
        m/s, one row per time, and the (nx, ny) table column of each cell), which the step
        reads through a per-cell lookup. Before the first step, as there."""
        self._require_unbuilt("set_rain")
        if isinstance(forcing, dict):
            _h = lambda a: a.get() if hasattr(a, "get") else np.asarray(a)
            forcing = NativeRain(time_s=_h(forcing["t_s"]), table=_h(forcing["native_rate_dev"]),
                                 lookup=_h(forcing["lookup_dev"]))
        lookup = None
        if isinstance(forcing, NativeRain):
            fm, nx, ny = self.fm, self.mesh.nx, self.mesh.ny
            if forcing.lookup.shape != (nx, ny):
                raise ValueError(f"set_rain: lookup has shape {forcing.lookup.shape}, not ({nx}, {ny})")
            ii = fm.ij[:, 0].astype(np.int64) - fm.ngh
            jj = fm.ij[:, 1].astype(np.int64) - fm.ngh
            inb = fm.active & (ii >= 0) & (ii < nx) & (jj >= 0) & (jj < ny)
            lk = np.where(inb, forcing.lookup[np.clip(ii, 0, nx - 1), np.clip(jj, 0, ny - 1)], 0)
            lookup = L.DeviceArray.from_host(np.ascontiguousarray(lk.astype(np.int32)))
        self.cfg.rainfall_forcing = forcing
        self._fo = self._forcing()
        self._rain_lookup = lookup
        self._lam_valid = False
        self._target = None
        return self

    # ------------------------------------------------------------------ coastal forcings
    def _require_unbuilt(self, who: str) -> None:
        if self._any_step or self._tables_taken:
            raise RuntimeError(f"CompressedSolver.{who}: the solver has already stepped or been taken into an "
                               f"Ensemble/GaugeMisfit, which would not see the new forcing. Attach the forcings "
                               f"before the first run, or build a fresh solver.")

    def set_sponge(self, sponge):
        """Open-boundary sponge (compressed_solver.py:3148-3156): ``{"keep": k, "amb": a}``, two
        fields on the padded grid (the shape of the dense solver's ``q[0]``). After each update,
        ``h <- keep*h + amb`` (two roundings) and ``hu, hv <- keep*hu, keep*hv`` on every stored
        cell with ``keep != 1`` or ``amb != 0``, ring cells included, as the reference's band
        kernel (compressed_solver.py:1320-1326, 1641-1645). ``None`` removes it."""
        self._require_unbuilt("set_sponge")
        new = None
        if sponge is not None:
            kf = self.fm.pack(np.asarray(sponge["keep"])).astype(self.dtype)
            af = self.fm.pack(np.asarray(sponge["amb"])).astype(self.dtype)
            band = (kf != 1) | (af != 0)
            if band.any():
                new = (L.DeviceArray.from_host(np.ascontiguousarray(np.stack([kf, af], axis=1))), band)
        self._sponge = new
        self._rebuild_force()
        return self

    def set_clamp(self, b):
        """Stage clamp (compressed_solver.py:3216-3217, 3293-3298): ``{"rows", "cols", "hmax"}``,
        cells in padded coordinates with their depth cap. After the update and the sponge, a
        listed cell with ``h > hmax`` gets ``h = hmax`` and its momentum scaled by ``hmax/h``
        (stage_clamp_flat, compressed_solver.py:424-440). Cells that are not stored are
        skipped, as there. ``None`` removes it."""
        self._require_unbuilt("set_clamp")
        new = None
        if b is not None:
            fm = self.fm
            rows = np.asarray(b["rows"], np.int64).ravel()
            cols = np.asarray(b["cols"], np.int64).ravel()
            hmax = np.asarray(b["hmax"], np.float64).ravel()
            if not (len(rows) == len(cols) == len(hmax)):
                raise ValueError(f"set_clamp: rows, cols, hmax have lengths {len(rows)}, {len(cols)}, {len(hmax)}")
            if len(rows) and ((rows < 0).any() or (rows >= fm.nxp).any() or (cols < 0).any()
                              or (cols >= fm.nyp).any()):
                raise ValueError(f"set_clamp: a cell lies outside the padded grid ({fm.nxp}, {fm.nyp})")
            ids = self._flat_ids(rows, cols)
            hit = ids >= 0
            hx = np.full(fm.n_stored, np.inf, self.dtype)
            hx[ids[hit]] = hmax[hit].astype(self.dtype)
            cells = ~(np.isnan(hx) | (hx == np.inf))        # h > +inf or NaN never clamps
            if cells.any():
                new = (L.DeviceArray.from_host(hx), cells)
        self._clamp = new
        self._rebuild_force()
        return self

    def set_ring(self, ring):
        """IDW gauge ring (compressed_solver.py:3146-3147, 3265-3268): ``{"i", "j", "bed", "wg",
        "NG", "n", "t_common", "stage_all"}`` as the run driver builds it: ring cells in padded
        coordinates with their bed, inverse-distance weights ``wg`` (``n`` rows of ``NG``, flat
        or 2-D), the common time grid and the ``(NG, len(t_common))`` gauge stages. Each step,
        after the sponge and before the clamp, a ring cell gets still water at the weighted
        stage of the step's start time, ``h = max(0, sum_g w_g s_g - bed)``, ``hu = hv = 0``
        (ring_bc_flat, 287-307; stage interpolation, 2338-2348). Ring cells must lie outside
        the active set (the reference's ring is the stored ring around it); cells that are not
        stored are skipped, as there. ``None`` (or ``n == 0``) removes it."""
        self._require_unbuilt("set_ring")
        new = None
        if ring and int(ring.get("n", 0)) > 0:
            fm = self.fm
            n, ng = int(ring["n"]), int(ring["NG"])
            ri = np.asarray(ring["i"], np.int64).ravel()
            rj = np.asarray(ring["j"], np.int64).ravel()
            bed = np.asarray(ring["bed"], np.float64).ravel()
            w = np.asarray(ring["wg"], np.float64).reshape(-1)
            tc = np.ascontiguousarray(ring["t_common"], np.float64).ravel()
            st = np.asarray(ring["stage_all"], np.float64)
            if ng < 1 or len(ri) != n or len(rj) != n or len(bed) != n or w.size != n * ng:
                raise ValueError(f"set_ring: n={n}, NG={ng} but i, j, bed, wg have {len(ri)}, {len(rj)}, "
                                 f"{len(bed)}, {w.size} entries")
            if len(tc) < 2 or st.shape != (ng, len(tc)):
                raise ValueError(f"set_ring: stage_all has shape {st.shape}, expected ({ng}, {len(tc)}) with "
                                 f"at least two times")
            if n and ((ri < 0).any() or (ri >= fm.nxp).any() or (rj < 0).any() or (rj >= fm.nyp).any()):
                raise ValueError(f"set_ring: a ring cell lies outside the padded grid ({fm.nxp}, {fm.nyp})")
            ids = self._flat_ids(ri, rj)
            rows = np.nonzero(ids >= 0)[0]
            ids = ids[rows]
            if fm.active[ids].any():
                raise ValueError("set_ring: ring cells must lie outside the active set")
            if len(np.unique(ids)) != len(ids):
                raise ValueError("set_ring: a cell is listed twice")
            dt = self.dtype
            new = None if not len(ids) else dict(ids=ids, rows=rows.astype(np.int32), ng=ng, nt=len(tc),
                       bed=L.DeviceArray.from_host(np.ascontiguousarray(bed.astype(dt))),
                       w=L.DeviceArray.from_host(np.ascontiguousarray(w.astype(dt))),
                       t=L.DeviceArray.from_host(tc),
                       stage=L.DeviceArray.from_host(np.ascontiguousarray(st.astype(dt))))
        self._ring = new
        self._rebuild_force()
        return self

    def set_infil(self, infil):
        """Constant-rate infiltration per land-cover class (compressed_solver.py:3212, the
        ``infiltrate_flat`` kernel 143-158): ``{"tab": rate}``, one rate in m/s per Manning
        class. Each step, after the update and the running max and before the sponge, an
        active cell with ``h >= h_min`` loses ``rate[class]*dt`` of depth (floored at 0) with
        its velocity kept. ``None`` removes it."""
        self._require_unbuilt("set_infil")
        new = None
        if infil is not None and infil.get("tab") is not None:
            tab = np.asarray(infil["tab"], np.float64).ravel()
            if self._n_field is not None:
                raise ValueError("set_infil: the solver has a per-cell Manning n plane and no class per cell")
            cap = L.MAX_CLASSES_PLANE if self._cls is not None else L.MAX_CLASSES
            if len(tab) > cap:
                raise ValueError(f"set_infil: {len(tab)} rates for at most {cap} classes")
            t = np.zeros(L.MAX_CLASSES_PLANE, self.dtype)
            t[:len(tab)] = tab
            new = L.DeviceArray.from_host(t)
        self._infil = new
        return self

    def set_drain(self, drain):
        """Karst drain cap (compressed_solver.py:3211, 1688-1694; the ``drain_cap`` kernel
        269-286): ``{"idx": flat ids, "h_tgt": depth}`` or ``{"idx", "h_tgt_arr": per-cell
        depths}`` (default 0.5 m). Each step, after the clamp, a listed cell deeper than its
        target is set to it with its velocity kept (momentum dropped below 1e-12 m).
        ``None`` or an empty ``idx`` removes it."""
        self._require_unbuilt("set_drain")
        new = None
        if drain is not None and np.asarray(drain["idx"]).size > 0:
            idx = np.asarray(drain["idx"], np.int64).ravel()
            if (idx < 0).any() or (idx >= self.fm.n_stored).any():
                raise ValueError(f"set_drain: flat ids must lie in [0, {self.fm.n_stored})")
            if drain.get("h_tgt_arr") is not None:
                ht = np.asarray(drain["h_tgt_arr"], np.float64).ravel()
                if len(ht) != len(idx):
                    raise ValueError(f"set_drain: {len(ht)} target depths for {len(idx)} cells")
            else:
                ht = np.full(len(idx), float(drain.get("h_tgt", 0.5)))
            hd = np.full(self.fm.n_stored, np.inf, self.dtype)
            hd[idx] = ht.astype(self.dtype)
            cells = ~(np.isnan(hd) | (hd == np.inf))
            if cells.any():
                new = (L.DeviceArray.from_host(hd), cells)
        self._drain = new
        self._rebuild_force()
        return self

    def set_ga_drain(self, b):
        """Green-Ampt infiltration and linear-reservoir drain (compressed_solver.py:3214-3215,
        3282-3291; kernels 320-421): ``{"cls_pad", "Ks_t", "psi_t", "dth_t", "F_pad",
        "Fmax_pad"?, "inv_tau_pad"?, "mode"}`` with padded-grid fields (soil class, initial
        cumulative infiltration F, its cap, the drain rate 1/tau) and per-class tables.
        ``mode``: ``"fused"`` (default: both), ``"ga"`` (infiltration only) or ``"drain"``
        (drain only). Each step, after the sponge and the gauge ring and before the clamp,
        on every active cell with h > 0. F is state: checkpoints carry it. The drain's decay
        uses an exponential built from IEEE operations (bitwise the oracle's) where the
        reference uses the fast ``__expf``. ``None`` removes it."""
        self._require_unbuilt("set_ga_drain")
        new = None
        if b is not None:
            fm, dt = self.fm, self.dtype
            mode = b.get("mode", "fused")
            if mode not in ("fused", "ga", "drain"):
                raise ValueError(f"set_ga_drain: mode must be 'fused', 'ga' or 'drain', got {mode!r}")
            new = {"mode": mode}
            if mode in ("fused", "ga"):
                cls = fm.pack(np.asarray(b["cls_pad"]))
                nc = int(cls.max()) + 1 if cls.size else 1
                tabs = [np.asarray(b[k], np.float64).ravel() for k in ("Ks_t", "psi_t", "dth_t")]
                if any(len(t) < nc for t in tabs):
                    raise ValueError(f"set_ga_drain: the soil classes reach {nc - 1} but Ks_t, psi_t, dth_t have "
                                     f"{[len(t) for t in tabs]} entries")
                new["cls"] = L.DeviceArray.from_host(np.ascontiguousarray(cls.astype(np.uint8)))
                new["ks"], new["psi"], new["dth"] = (L.DeviceArray.from_host(np.ascontiguousarray(t.astype(dt)))
                                                     for t in tabs)
                F0 = b.get("F_pad")
                F = fm.pack(np.asarray(F0)).astype(dt) if F0 is not None else np.zeros(fm.n_stored, dt)
                new["F"] = L.DeviceArray.from_host(np.ascontiguousarray(F))
                if b.get("Fmax_pad") is not None:
                    new["fmax"] = L.DeviceArray.from_host(np.ascontiguousarray(fm.pack(np.asarray(b["Fmax_pad"]))
                                                                               .astype(dt)))
            if mode in ("fused", "drain") and b.get("inv_tau_pad") is not None:
                new["inv_tau"] = L.DeviceArray.from_host(
                    np.ascontiguousarray(fm.pack(np.asarray(b["inv_tau_pad"])).astype(dt)))
        self._ga = new
        return self

    def ga_F(self) -> Optional[np.ndarray]:
        """The Green-Ampt cumulative infiltration per stored cell (None without Green-Ampt)."""
        if self._ga is None or "F" not in self._ga:
            return None
        L.sync(self.stream)
        return self._ga["F"].download()

    def set_ring_mode(self, mode: str = "auto", eta: Optional[float] = None, classify: Optional[float] = None):
        """The outer boundary of the ring in the reference's cached replay (_build_ring_bc,
        compressed_solver.py:486-575, applied at the top of every step, 2134-2144; chosen
        there by GEOSWE_RING_BC, GEOSWE_RING_ETA and GEOSWE_RING_CLASSIFY). ``"auto"`` or
        ``"stage"``: still water at `eta` with zero momentum on every ring cell;
        ``"extrapolate"``: each ring cell beside the active set takes the state of its first
        active neighbour (i+1, i-1, j+1, j-1) and, once, its bed; ``"hybrid"``: those ring
        cells whose bed lies below `classify` (default `eta`) get the stage, the others
        extrapolate; ``"stage_uv"``: the stage, then the velocity of the neighbour
        (ghost_uv, 450-465); ``"off"``, or any stage mode without `eta`, leaves the ring
        dry. The ring cells' stored values between steps are the previous step's."""
        from .adjoint import flat_neighbours
        self._require_unbuilt("set_ring_mode")
        mode = (mode or "auto").lower()
        if mode not in ("auto", "stage", "extrapolate", "hybrid", "stage_uv", "off", "none"):
            raise ValueError(f"set_ring_mode: mode {mode!r}: expected auto|stage|extrapolate|hybrid|stage_uv|off")
        self._rmode = None
        if getattr(self, "_bed0", None) is not None:    # a second call starts from the packed bed
            self._bed.upload(self._bed0)
        if mode in ("off", "none") or (mode != "extrapolate" and eta is None):
            return self
        fm = self.fm
        act = fm.active
        nb = flat_neighbours(fm)
        best = np.full(fm.n_stored, -1, np.int64)
        for d in range(4):                    # the first active neighbour of each ring cell
            cand = nb[:, d].astype(np.int64)
            ok = (cand >= 0) & ~act & (best < 0)
            ok[ok] = act[cand[ok]]
            best = np.where(ok, cand, best)
        gi = np.nonzero(best >= 0)[0]
        gn = best[gi]
        if getattr(self, "_bed0", None) is None:      # the packed bed, before any replication
            self._bed0 = self._bed.download()
        bed = self._bed0.copy()
        if mode == "extrapolate":
            bed[gi] = bed[gn]
            ids, nbs, kind = gi, gn, np.zeros(len(gi), np.uint8)
        elif mode == "hybrid":
            thr = self.dtype.type(classify if classify is not None else eta)
            water = bed[gi] < thr
            mi, mn = gi[~water], gn[~water]
            bed[mi] = bed[mn]
            ids = np.concatenate([mi, gi[water]])
            nbs = np.concatenate([mn, np.full(int(water.sum()), -1)])
            kind = np.concatenate([np.zeros(len(mi), np.uint8), np.ones(int(water.sum()), np.uint8)])
        else:
            ids = np.nonzero(~act)[0]
            nbs = best[ids]
            kind = np.where((nbs >= 0) & (mode == "stage_uv"), 2, 1).astype(np.uint8)
        self._bed.upload(bed)
        if len(ids):
            self._rmode = dict(ids=L.DeviceArray.from_host(np.ascontiguousarray(ids, np.int64)),
                               nb=L.DeviceArray.from_host(np.ascontiguousarray(nbs, np.int64)),
                               kind=L.DeviceArray.from_host(np.ascontiguousarray(kind)), n=len(ids),
                               eta=float(eta) if eta is not None else 0.0, mode=mode,
                               host=(ids, nbs, kind))
        return self

    def _flat_ids(self, rows: np.ndarray, cols: np.ndarray) -> np.ndarray:
        """Flat id of each padded (row, col), -1 where the cell is not stored."""
        fm = self.fm
        # stored cells are numbered row by row, columns ascending: their keys are sorted
        key = fm.ij[:, 0].astype(np.int64) * fm.nyp + fm.ij[:, 1]
        want = rows * fm.nyp + cols
        pos = np.minimum(np.searchsorted(key, want), fm.n_stored - 1)
        return np.where(key[pos] == want, pos, -1)

    def _rebuild_force(self) -> None:
        force = np.zeros(self.fm.n_stored, bool)
        for f in (self._sponge, self._clamp, self._drain):
            if f is not None:
                force |= f[1]
        fl = self._flags_base
        if force.any():
            fl = fl | (force.astype(np.uint8) * np.uint8(L.FLAG_FORCE))
        self._flags = L.DeviceArray.from_host(np.ascontiguousarray(fl)) if fl.any() else None
        listed = force & ~self.fm.active
        if self._ring is not None:
            listed[self._ring["ids"]] = True
        ids = np.nonzero(listed)[0].astype(np.int64)
        self._force_ring = L.DeviceArray.from_host(ids) if len(ids) else None
        self._force_ring_row = None
        if self._ring is not None and len(ids):
            row = np.full(self.fm.n_stored, -1, np.int32)
            row[self._ring["ids"]] = self._ring["rows"]
            self._force_ring_row = L.DeviceArray.from_host(np.ascontiguousarray(row[ids]))

    @property
    def has_coastal_forcing(self) -> bool:
        return any(f is not None for f in (self._sponge, self._clamp, self._ring, self._infil, self._drain, self._ga,
                                           self._rmode))

    # ------------------------------------------------------------------ structs
    def _phys(self) -> L.GsPhys:
        cfg = self.cfg
        # the flat wave speed skips h < h_min and divides by max(h, h_min_cfl); its floor is
        # sqrt(g*h_min) (compressed_solver.py:200-240, 713-719); dmin is the reference's dx on its
        # square cells
        return L.phys_struct(cfg.g, cfg.h_min, self.h_min_cfl, cfg.cfl, min(self.mesh.dx, self.mesh.dy),
                             cfg.friction_velocity_cap_ms, cfg.friction == "manning_implicit",
                             cfg.friction_quadratic_alpha, True, not self.cfl_linf, self._n_classes,
                             floor_depth=cfg.h_min)

    def _desc(self) -> L.GsFlat:
        f = L.GsFlat()
        fm = self.fm
        f.n_stored, f.n_runs = fm.n_stored, len(fm.run_k0)
        f.dtype = L.GS_F32 if self.dtype == np.float32 else L.GS_F64
        f.nyp, f.ngh = fm.nyp, fm.ngh
        f.dx, f.dy = float(self.mesh.dx), float(self.mesh.dy)
        for s in range(2):
            for k in range(3):
                f.q[s][k] = self._bufs[s][k].ptr
        f.bed, f.ntab = self._bed.ptr, self._ntab.ptr
        f.flags = self._flags.ptr if self._flags is not None else None
        f.n_field = self._n_field.ptr if self._n_field is not None else None
        f.max_h = self._max_h.ptr if self._max_h is not None else None
        f.cls = self._cls.ptr if self._cls is not None else None
        f.run_k0, f.run_len, f.run_i, f.run_j0, f.run_d, f.chunk_run = (r.ptr for r in self._runs)
        if self.traversal not in ("tiles", "chunks"):
            raise ValueError(f"traversal must be 'tiles' or 'chunks', got {self.traversal!r}")
        f.n_tiles = len(fm.tile_e0) if self.traversal == "tiles" else 0
        f.tile_e0, f.tile_irc, f.ent = (t.ptr for t in self._tiles)
        f.sponge = self._sponge[0].ptr if self._sponge is not None else None
        f.clamp = self._clamp[0].ptr if self._clamp is not None else None
        f.infil = self._infil.ptr if self._infil is not None else None
        f.drain = self._drain[0].ptr if self._drain is not None else None
        ga = self._ga or {}
        f.rain_lookup = self._rain_lookup.ptr if self._rain_lookup is not None else None
        rm = self._rmode
        if rm is not None:
            f.n_ring_mode, f.ring_mode_eta = rm["n"], rm["eta"]
            f.ring_mode_ids, f.ring_mode_nb, f.ring_mode_kind = rm["ids"].ptr, rm["nb"].ptr, rm["kind"].ptr
        f.ga_cls, f.ga_ks, f.ga_psi, f.ga_dth, f.ga_F, f.ga_fmax, f.ga_inv_tau = (
            ga[k].ptr if k in ga else None for k in ("cls", "ks", "psi", "dth", "F", "fmax", "inv_tau"))
        f.n_force_ring = self._force_ring.shape[0] if self._force_ring is not None else 0
        f.force_ring = self._force_ring.ptr if self._force_ring is not None else None
        r = self._ring
        if r is not None and self._force_ring_row is not None:
            f.force_ring_row = self._force_ring_row.ptr
            f.ring_ng, f.ring_nt = r["ng"], r["nt"]
            f.ring_bed, f.ring_w, f.ring_t, f.ring_stage = r["bed"].ptr, r["w"].ptr, r["t"].ptr, r["stage"].ptr
        return f

    # ------------------------------------------------------------------ state
    def _clockv(self):
        return L.read_clock(self._clock)

    @property
    def t(self) -> float:
        return float(self._clockv().t)

    @property
    def nsteps(self) -> int:
        return int(self._clockv().nsteps)

    def q_flat(self) -> np.ndarray:
        cur = int(self._clockv().nsteps) & 1
        return np.stack([self._bufs[cur][k].download() for k in range(3)])

    def depth(self) -> np.ndarray:
        """Depth on the unpadded (nx, ny) grid; cells outside the stored set are 0."""
        g, nx, ny = self.mesh.ngh, self.mesh.nx, self.mesh.ny
        return self.fm.unpack(self.q_flat()[0])[g:g + nx, g:g + ny]

    def max_depth(self) -> np.ndarray:
        g, nx, ny = self.mesh.ngh, self.mesh.nx, self.mesh.ny
        if self._max_h is None or not self._any_step:
            return self.depth()
        return self.fm.unpack(self._max_h.download())[g:g + nx, g:g + ny]

    # ------------------------------------------------------------------ stepping
    def _ensure_lam(self):
        if not self._lam_valid:
            L.call("gs_flat_cfl", C.byref(self._desc()), C.byref(self._phys()), self._fo.s,
                   C.c_void_p(self._clock.ptr), self.stream)
            self._lam_valid = True

    def _raise_if_bad(self, c):
        if c.status & L.ST_NONFINITE:
            raise FloatingPointError(f"CompressedSolver: non-finite wave speed at t={c.t:.3f}s "
                                     f"(NaN/Inf in an active cell)")

    def _open_mode(self):
        if self._target != "open":
            L.call("gs_clock_set_target", C.c_void_p(self._clock.ptr), C.c_double(math.inf), C.c_double(0.0),
                   C.c_int64(-1), C.byref(self._phys()), self._fo.s, self.stream)
            self._target = "open"

    def step(self, dt: Optional[float] = None) -> float:
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
        L.call("gs_flat_step", C.byref(self._desc()), C.byref(self._phys()), self._fo.s,
               C.c_void_p(self._clock.ptr), self.stream)
        self._any_step = True
        return dt

    def advance(self, nsteps: int, sync: bool = True) -> None:
        self._ensure_lam()
        self._open_mode()
        d, p, clk = C.byref(self._desc()), C.byref(self._phys()), C.c_void_p(self._clock.ptr)
        for _ in range(int(nsteps)):
            L.call("gs_flat_step", d, p, self._fo.s, clk, self.stream)
        self._any_step = True
        if sync:
            L.sync(self.stream)
            self._raise_if_bad(self._clockv())

    def launch_info(self) -> dict:
        """The step kernel's launch for this solver: ``kernel`` ("chunks", "sweep" or
        "sweep_ext"), ``blocks`` (its grid) and ``resident`` (that kernel's resident grid on
        this device; ``blocks > resident`` means one wavefront per tile)."""
        k, b, r = C.c_int(0), C.c_int64(0), C.c_int64(0)
        L.call("gs_flat_launch_info", C.byref(self._desc()), C.byref(self._phys()), self._fo.s, C.byref(k),
               C.byref(b), C.byref(r))
        return {"kernel": ("chunks", "sweep", "sweep_ext")[k.value], "blocks": int(b.value),
                "resident": int(r.value)}

    def capture_steps(self, nsteps: int) -> L.StepGraph:
        """Record ``nsteps`` CFL steps (gs_flat_step, with its ring passes) into a HIP graph;
        ``replay()`` enqueues them on this solver's stream, bit for bit the direct launches."""
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
                L.call("gs_flat_step", C.byref(d), C.byref(p), fo, clk, cap)

        g = L.StepGraph(record, self.stream)
        g._keep = (d, p, self)
        self._any_step = True
        return g

    def sync(self) -> None:
        L.sync(self.stream)
        self._raise_if_bad(self._clockv())

    # ------------------------------------------------------------------ checkpoints
    def save_cache(self, cache_dir: str, x0: float = 0.0, y0: float = 0.0, crs_wkt: str = "") -> None:
        """Write this solver in the reference's cache layout (compressed_solver.py:2504-2569);
        :func:`geoswe_amd.cache.run_cached` replays it."""
        from .cache import save_cache
        save_cache(self, cache_dir, x0=x0, y0=y0, crs_wkt=crs_wkt)

    @classmethod
    def from_cache(cls, cache_dir: str, **kw) -> "CompressedSolver":
        """A solver over a cache directory (see :func:`geoswe_amd.cache.from_cache`)."""
        from .cache import from_cache
        return from_cache(cache_dir, **kw)

    def save_checkpoint(self, ckpt_dir: str, fidx: int = 0, next_frame: Optional[float] = None) -> None:
        """Write the current state to ``ckpt_dir`` in the reference's single-rank layout
        (compressed_solver.py:1789-1880): ``ckpt_r00.npz`` holding ``t``, ``steps``, ``q0``,
        ``q1``, ``q2`` (flat arrays in the mesh's stored order), ``max_h`` when the running
        maximum is tracked and ``F`` with Green-Ampt, published by an atomic rename, then
        ``ckpt_meta.json``
        with ``t``, ``steps``, ``fidx``, ``next_frame`` and ``nranks``."""
        os.makedirs(ckpt_dir, exist_ok=True)
        c = self._clockv()
        q = self.q_flat()
        arrays = {"t": np.float64(c.t), "steps": np.int64(c.nsteps), "q0": q[0], "q1": q[1], "q2": q[2]}
        if self._max_h is not None:
            arrays["max_h"] = self._max_h.download()
        if self._ga is not None and "F" in self._ga:
            arrays["F"] = self._ga["F"].download()
        tmp = os.path.join(ckpt_dir, "ckpt_r00_tmp")
        np.savez(tmp, **arrays)
        os.replace(tmp + ".npz", os.path.join(ckpt_dir, "ckpt_r00.npz"))
        meta = dict(t=float(c.t), steps=int(c.nsteps), fidx=int(fidx),
                    next_frame=float(c.t if next_frame is None else next_frame), nranks=1)
        mt = os.path.join(ckpt_dir, "ckpt_meta.json.tmp")
        with open(mt, "w") as f:
            json.dump(meta, f)
        os.replace(mt, os.path.join(ckpt_dir, "ckpt_meta.json"))

    def load_checkpoint(self, ckpt_dir: str) -> None:
        """Restore the state written by :meth:`save_checkpoint` (or by the reference's run on
        one rank for the same mesh; compressed_solver.py:1881-1932): the same rank count, the
        array's (t, steps) stamp equal to the meta's (else a torn checkpoint), the state in
        both step buffers. Stepping on continues bit for bit as
        the run that wrote it."""
        with open(os.path.join(ckpt_dir, "ckpt_meta.json")) as f:
            meta = json.load(f)
        if int(meta.get("nranks", 1)) != 1:
            raise RuntimeError(f"checkpoint nranks={meta.get('nranks')} != 1; resume with the same rank count")
        z = np.load(os.path.join(ckpt_dir, "ckpt_r00.npz"))
        t, steps = float(meta["t"]), int(meta["steps"])
        if "t" in z.files and (abs(float(z["t"]) - t) > 1e-9 or int(z["steps"]) != steps):
            raise RuntimeError(f"checkpoint state is at t={float(z['t']):.3f}s/step {int(z['steps'])}, meta says "
                               f"t={t:.3f}s/step {steps}: torn checkpoint")
        N = self.fm.n_stored
        for k in range(3):
            a = z[f"q{k}"]
            if a.shape != (N,):
                raise ValueError(f"checkpoint q{k} has shape {a.shape}, the mesh stores {N} cells")
        if self._ga is not None and "F" in self._ga and "F" not in z.files:
            # as the reference (compressed_solver.py:1911-1918): resuming would reset the
            # cumulative infiltration silently
            raise RuntimeError("load_checkpoint: Green-Ampt is on but the checkpoint holds no cumulative "
                               "infiltration F (written without Green-Ampt); resuming would reset it to zero")
        L.call("gs_clock_reset", C.c_void_p(self._clock.ptr), C.c_double(t), self.stream)
        c = self._clockv()
        c.nsteps = steps
        L.call("gs_memcpy_h2d", C.c_void_p(self._clock.ptr), C.byref(c), C.c_size_t(C.sizeof(c)), self.stream)
        # both buffers: the ring cells a step does not write then hold the checkpoint's values
        # whichever buffer a later step reads (the active cells of the other one are rewritten)
        for k in range(3):
            a = np.ascontiguousarray(z[f"q{k}"], dtype=self.dtype)
            for par in (0, 1):
                self._bufs[par][k].upload(a, self.stream)
        if self._ga is not None and "F" in self._ga:
            self._ga["F"].upload(np.ascontiguousarray(z["F"], dtype=self.dtype), self.stream)
        if self._max_h is not None:
            if "max_h" in z.files:
                self._max_h.upload(np.ascontiguousarray(z["max_h"], dtype=self.dtype), self.stream)
            else:
                import warnings
                warnings.warn("load_checkpoint: the checkpoint has no max_h; the running maximum covers only "
                              "the steps after the resume", RuntimeWarning)
        L.sync(self.stream)
        self._lam_valid = False
        self._target = None
        self._any_step = self._any_step or steps > 0
        return meta

    def _save_frame(self, fdir: str, idx: int, t: float, fp32: bool) -> None:
        """One depth frame as the reference writes it (compressed_solver.py:1762-1768): the
        interior (nx, ny) depth, float16 unless `fp32`, in a compressed .npz holding `h`."""
        h = self.depth().astype(np.float32 if fp32 else np.float16)
        np.savez_compressed(os.path.join(fdir, f"depth_{idx:05d}_t{int(round(t)):07d}_r00.npz"), h=h)

    def run(self, t_end: float, dt_max: Optional[float] = None, max_steps: int = 10 ** 9, batch: int = 64,
            checkpoint_every_s: float = 0.0, ckpt_dir: Optional[str] = None, resume: bool = False,
            out_dir: Optional[str] = None, frame_every_s: float = 0.0, frame_fp32: bool = False) -> int:
        """Step to t_end (compressed_solver.py:3359-3450); dt planned on the device.
        ``checkpoint_every_s`` > 0 with ``ckpt_dir`` writes a checkpoint whenever simulated
        time has advanced that far since the last one (checked every ``batch`` steps) and
        at the end; ``resume`` first restores ``ckpt_dir`` when it holds one.
        ``frame_every_s`` > 0 with ``out_dir`` writes depth frames and their manifest in the
        reference's layout (``<out_dir>/frames_parallel/depth_<index:05d>_t<seconds:07d>_r00.npz``,
        compressed_solver.py:1734-1768, 2395-2406, 2450-2451): a frame after the first step
        that reaches each frame time, and one at the end; the clock is then checked after
        every step, as the reference's loop does. Frames do not change the trajectory."""
        if dt_max is not None:
            check_numeric_inputs(where="CompressedSolver.run", cfl=None)
        if (checkpoint_every_s or resume) and not ckpt_dir:
            raise ValueError("checkpoint_every_s and resume need ckpt_dir")
        frames_on = frame_every_s > 0
        if frames_on and not out_dir:
            raise ValueError("frame_every_s needs out_dir")
        fidx, next_frame, last_ft = 0, (0.0 if frames_on else math.inf), -1.0
        if resume and os.path.exists(os.path.join(ckpt_dir, "ckpt_meta.json")):
            meta = self.load_checkpoint(ckpt_dir)
            if frames_on:
                fidx, next_frame, last_ft = int(meta["fidx"]), float(meta["next_frame"]), float(meta["t"])
        fdir = None
        if frames_on:
            fdir = os.path.join(out_dir, "frames_parallel")
            os.makedirs(fdir, exist_ok=True)
            nx, ny = self.mesh.nx, self.mesh.ny
            with open(os.path.join(fdir, "manifest.json"), "w") as mf:
                json.dump({"nx_orig": nx, "ny_orig": ny, "nx_glob": nx, "ny_glob": ny, "dx": float(self.mesh.dx),
                           "x0": 0.0, "y0": 0.0, "crs_wkt": "", "nranks": 1, "frame_every_s": float(frame_every_s),
                           "ranks": [{"rank": 0, "cx": 0, "cy": 0, "i0": 0, "j0": 0, "nx": nx, "ny": ny}]}, mf)
            batch = 1
        t0 = time.perf_counter()
        self._ensure_lam()
        n0 = self.nsteps
        next_ckpt = self.t + checkpoint_every_s if checkpoint_every_s else math.inf
        L.call("gs_clock_set_target", C.c_void_p(self._clock.ptr), C.c_double(float(t_end)),
               C.c_double(float(dt_max) if dt_max else 0.0), C.c_int64(int(max_steps)), C.byref(self._phys()),
               self._fo.s, self.stream)
        d, p, clk = C.byref(self._desc()), C.byref(self._phys()), C.c_void_p(self._clock.ptr)
        while True:
            c = self._clockv()
            self._raise_if_bad(c)
            if c.status & L.ST_DONE:
                break
            if frames_on and c.nsteps > n0 and c.t >= next_frame - 1e-9:
                self._save_frame(fdir, fidx, c.t, frame_fp32)
                fidx, next_frame, last_ft = fidx + 1, next_frame + frame_every_s, c.t
            if c.t >= next_ckpt:
                self.save_checkpoint(ckpt_dir, fidx, next_frame if frames_on else None)
                next_ckpt = c.t + checkpoint_every_s
            for _ in range(batch):
                L.call("gs_flat_step", d, p, self._fo.s, clk, self.stream)
            self._any_step = True
        c = self._clockv()
        if frames_on and c.nsteps > n0 and c.t >= next_frame - 1e-9:     # the last step reached a frame time
            self._save_frame(fdir, fidx, c.t, frame_fp32)
            fidx, next_frame, last_ft = fidx + 1, next_frame + frame_every_s, c.t
        if frames_on and c.t > last_ft + 1e-6:
            self._save_frame(fdir, fidx, c.t, frame_fp32)
        if checkpoint_every_s:
            self.save_checkpoint(ckpt_dir, fidx, next_frame if frames_on else None)   # final: a planned stop loses nothing
        self._target = None
        steps = self.nsteps - n0
        self.diagnostics["wallclock"] = time.perf_counter() - t0
        self.diagnostics["steps"] = steps
        return steps

    def residual(self) -> np.ndarray:
        """Residual (3, N) of the current flat state (ring cells 0)."""
        N = self.fm.n_stored
        r = [L.DeviceArray((N,), self.dtype, zero=True) for _ in range(3)]
        cur = int(self._clockv().nsteps) & 1
        L.call("gs_flat_residual", C.byref(self._desc()), C.byref(self._phys()), cur, C.c_void_p(r[0].ptr),
               C.c_void_p(r[1].ptr), C.c_void_p(r[2].ptr), self.stream)
        return np.stack([x.download() for x in r])
