"""Batched ensemble of compressed (flat) domains on one MI355X.

The reference advances independent scenarios one after another, each through
``CompressedSolver.run`` (src/geoswe/compressed_solver.py:3359-3450) with one host-driven
launch sequence per step. ``Ensemble`` advances all members together: one launch of
``k_ens_step`` per step walks every member's sweep tiles on a resident grid, each member keeps
its own device clock (its own CFL dt, rain frame and tide level), the last workgroup plans
every member's next dt, and the steps are replayed from a captured hipGraph. The cell
arithmetic is the flat step's own device code, so member k's state is bit for bit the one
its plain ``CompressedSolver`` run produces in the same number of steps.
"""
from __future__ import annotations

import ctypes as C
import math
from typing import Optional, Sequence

from . import _lib as L
from .compressed import CompressedSolver
from .flatmesh import mesh_sweep_tiles, sweep_rows_for


class Ensemble:
    """Members are ``CompressedSolver`` objects; they stay usable (``depth()``, ``t``,
    ``nsteps``) and must outlive the ensemble, which steps their buffers in place."""

    def __init__(self, members: Sequence[CompressedSolver], graph_batch: int = 50, sweep_rows: Optional[int] = None):
        members = list(members)
        if not members:
            raise ValueError("Ensemble needs at least one member")
        dt0 = members[0].dtype
        for k, m in enumerate(members):
            if not isinstance(m, CompressedSolver):
                raise TypeError(f"member {k} is {type(m).__name__}, not CompressedSolver")
            if m.dtype != dt0:
                raise ValueError(f"member {k} has dtype {m.dtype}, member 0 {dt0}")
            if m._rmode is not None:
                raise NotImplementedError(f"member {k} has ring modes (set_ring_mode); the ensemble step does not "
                                          f"apply them")
            if m._force_ring is not None:
                raise NotImplementedError(f"member {k} forces ring cells (gauge ring, or sponge or clamp outside its active set); "
                                          f"the ensemble step updates active cells only")
        self.members = members
        for m in members:
            m._tables_taken = True
        self.graph_batch = int(graph_batch)
        for m in members:                      # first dt exactly as a plain open run plans it
            m._ensure_lam()
            m._open_mode()
        n = len(members)
        self._phys = (L.GsPhys * n)(*[m._phys() for m in members])
        self._forc = (L.GsForcing * n)(*[m._fo.s for m in members])
        self._clocks = (C.c_void_p * n)(*[m._clock.ptr for m in members])
        self._h = C.c_void_p()
        self._tile_cache = {}
        # Tile height for the whole ensemble (or the caller's): members run side by side, so a
        # small member needs no short tiles to fill the GPU (its own tables are sized for it
        # alone); autotune(heights=...) may pick another.
        if sweep_rows is not None and not (1 <= int(sweep_rows) <= 1024):
            raise ValueError(f"sweep_rows must be in 1..1024, got {sweep_rows}")
        self._build(int(sweep_rows) if sweep_rows is not None
                    else sweep_rows_for(sum(int(m.fm.tile_irc[:, 1].sum()) for m in members)),
                    exact=sweep_rows is not None)

    def _build(self, rows: int, exact: bool = True) -> None:
        """The native ensemble over the members' tiles of `rows` rows (`exact`: every member
        re-tiled to that height; else members already at least as tall keep their tiles)."""
        self.close()
        members = self.members
        self.sweep_rows = int(rows)
        descs = (L.GsFlat * len(members))(*[m._desc() for m in members])
        tiles = []
        for k, m in enumerate(members):
            tall = int(m.fm.tile_irc[:, 1].max(initial=0))
            if m.n_active == 0 or (tall == rows if exact else tall >= rows):
                continue                       # keep the member's own tiles
            key = (k, rows)
            if key not in self._tile_cache:
                e0, irc, ent = mesh_sweep_tiles(m.fm, rows)
                self._tile_cache[key] = (len(e0), [L.DeviceArray.from_host(a) for a in (e0, irc.reshape(-1),
                                                                                        ent.reshape(-1))])
            nt, arrs = self._tile_cache[key]
            tiles.append(arrs)
            d = descs[k]
            d.n_tiles = nt
            d.tile_e0, d.tile_irc, d.ent = (a.ptr for a in arrs)
        self._descs, self._tiles = descs, tiles
        h = C.c_void_p()
        L.call("gs_ens_create", self._descs, self._phys, self._forc, self._clocks, len(members), C.byref(h))
        self._h = h

    # ------------------------------------------------------------------ properties
    @property
    def n_members(self) -> int:
        return len(self.members)

    @property
    def n_active(self) -> int:
        return sum(m.n_active for m in self.members)

    @property
    def n_stored(self) -> int:
        return sum(m.n_stored for m in self.members)

    @property
    def launch(self) -> dict:
        b, c, t = C.c_int(), C.c_int(), C.c_int64()
        L.call("gs_ens_get_launch", self._h, C.byref(b), C.byref(c), C.byref(t))
        return {"blocks": b.value, "tiles_per_pass": c.value, "total_tiles": t.value}

    def set_launch(self, blocks: int = 0, tiles_per_pass: int = 1) -> None:
        L.call("gs_ens_set_launch", self._h, int(blocks), int(tiles_per_pass))

    # ------------------------------------------------------------------ stepping
    def _check(self) -> None:
        bad = [k for k, m in enumerate(self.members) if m._clockv().status & L.ST_NONFINITE]
        if bad:
            raise FloatingPointError(f"Ensemble: non-finite wave speed in member(s) {bad}")

    def step(self, sync: bool = True) -> None:
        L.call("gs_ens_step", self._h, None)
        for m in self.members:
            m._any_step = True
        if sync:
            L.sync(None)
            self._check()

    def advance(self, nsteps: int, graph_batch: Optional[int] = None, sync: bool = True) -> None:
        """nsteps steps of every member that is not done, replayed from a hipGraph of
        ``graph_batch`` steps (0: plain launches)."""
        gb = self.graph_batch if graph_batch is None else int(graph_batch)
        L.call("gs_ens_run", self._h, C.c_int64(int(nsteps)), gb, None)
        for m in self.members:
            m._any_step = True
        if sync:
            L.sync(None)
            self._check()

    def run(self, t_end: float, dt_max: Optional[float] = None, max_steps: int = -1, batch: int = 100) -> list:
        """Every member to t_end (its own dt sequence, clipped to t_end like
        CompressedSolver.run); one host check per `batch` steps. Returns the steps taken
        by each member."""
        n0 = [m.nsteps for m in self.members]
        for m in self.members:
            L.call("gs_clock_set_target", C.c_void_p(m._clock.ptr), C.c_double(float(t_end)),
                   C.c_double(float(dt_max) if dt_max else 0.0), C.c_int64(int(max_steps)), C.byref(m._phys()),
                   m._fo.s, None)
            m._target = "run"
        while True:
            self.advance(batch, sync=True)
            if all(m._clockv().status & L.ST_DONE for m in self.members):
                break
        return [m.nsteps - a for m, a in zip(self.members, n0)]

    def open(self) -> None:
        """Back to open-ended stepping (no t_end) after run()."""
        for m in self.members:
            m._target = None
            m._open_mode()

    def autotune(self, trial_steps: int = 5, heights: Optional[Sequence[int]] = None) -> dict:
        """Time the candidate launch shapes on a snapshot of the state (restored
        afterwards) and keep the fastest. With `heights`, also the tile height: every member
        re-tiled to each height in turn, the launch shapes timed for each, and the fastest
        height and shape kept (the results are the same bits at every height and shape)."""
        def native():
            best = C.c_double()
            L.call("gs_ens_autotune", self._h, int(trial_steps), None, C.byref(best))
            return best.value if math.isfinite(best.value) else math.inf
        if not heights:
            ms = native()
            out = self.launch
            out["ms_per_step"] = ms if math.isfinite(ms) else None
            return out
        trials = {}
        for rows in heights:
            self._build(int(rows))
            trials[int(rows)] = (native(), self.launch)
        rows = min(trials, key=lambda r: trials[r][0])
        self._build(rows)
        ms, launch = trials[rows]
        self.set_launch(launch["blocks"], launch["tiles_per_pass"])
        out = self.launch
        out["ms_per_step"] = ms if math.isfinite(ms) else None
        out["sweep_rows"] = rows
        out["trials"] = {r: round(v[0], 4) for r, v in trials.items()}
        return out

    def close(self) -> None:
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            L.call("gs_ens_destroy", h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_noalloc(members: Sequence[CompressedSolver], nsteps: int, sync: bool = True) -> None:
    """nsteps steps of every member through ``gs_ens_run_noalloc``: the ensemble step with
    no device allocation (its member table lives in a static device array of the library),
    for callers that cannot allocate. Members use their own tile tables; bitwise the steps
    an ``Ensemble`` of the same members takes."""
    members = list(members)
    if not members:
        raise ValueError("run_noalloc needs at least one member")
    for k, m in enumerate(members):
        if not isinstance(m, CompressedSolver):
            raise TypeError(f"member {k} is {type(m).__name__}, not CompressedSolver")
        if m._rmode is not None or m._force_ring is not None:
            raise NotImplementedError(f"member {k} has ring modes or forced ring cells; the ensemble step updates "
                                      f"active cells only")
        m._tables_taken = True
        m._ensure_lam()
        m._open_mode()
    n = len(members)
    descs = (L.GsFlat * n)(*[m._desc() for m in members])
    phys = (L.GsPhys * n)(*[m._phys() for m in members])
    forc = (L.GsForcing * n)(*[m._fo.s for m in members])
    clocks = (C.c_void_p * n)(*[m._clock.ptr for m in members])
    L.call("gs_ens_run_noalloc", descs, phys, forc, clocks, n, C.c_int64(int(nsteps)), None)
    for m in members:
        m._any_step = True
    if sync:
        L.sync(None)
        bad = [k for k, m in enumerate(members) if m._clockv().status & L.ST_NONFINITE]
        if bad:
            raise FloatingPointError(f"run_noalloc: non-finite wave speed in member(s) {bad}")
