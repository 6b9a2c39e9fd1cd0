"""Host builder of the compressed active-cell ("flat") mesh and its run table.

Stored set and order follow the reference's CompressedSWE / CompressedMesh2D
(src/geoswe/compressed_rhs.py:241-369, compressed_mesh.py:108-187): the active mask is
dilated twice with the 4-neighbour cross, and stored cells are numbered row-major over
the padded grid with j fastest. The dilation here runs on the padded grid, so an active
cell on the domain edge keeps its two ghost layers in the ring (the reference dilates
the interior-shaped mask and falls back to a self-neighbour there).

Instead of the reference's (N, 4) int16 neighbour deltas, consecutive active cells of
one row whose rows i+1, i-1, i+2, i-2 sit at the same flat-id offsets form a run; the
kernel reads (k0, len, i, j0, offsets) per run. Along j the neighbours of an active cell
are k+-1 and k+-2 by construction of the dilation.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

CHUNK = 60   # flat ids written per wavefront (the kernel's chunk stride)


def cross_dilate(mask: np.ndarray, iterations: int) -> np.ndarray:
    """Binary dilation with the 4-neighbour cross (scipy.ndimage default structure)."""
    m = np.asarray(mask, dtype=bool).copy()
    for _ in range(iterations):
        d = m.copy()
        d[1:, :] |= m[:-1, :]
        d[:-1, :] |= m[1:, :]
        d[:, 1:] |= m[:, :-1]
        d[:, :-1] |= m[:, 1:]
        m = d
    return m


@dataclass
class FlatMesh:
    nxp: int
    nyp: int
    ngh: int
    n_stored: int
    n_active: int
    ij: np.ndarray            # (N, 2) int32 padded coordinates of stored cells
    active: np.ndarray        # (N,) bool
    run_k0: np.ndarray        # int64
    run_len: np.ndarray       # int32
    run_i: np.ndarray         # int32
    run_j0: np.ndarray        # int32
    run_d: np.ndarray         # (R, 4) int32 offsets of rows i+1, i-1, i+2, i-2
    chunk_run: np.ndarray     # int32, first run ending past each chunk start
    tile_e0: np.ndarray = None    # (T,) int64 first window-row entry (row i0 - 2) of each sweep tile
    tile_irc: np.ndarray = None   # (T, 3) int32 padded row i0, rows swept, padded column of lane 0
    ent: np.ndarray = None        # (E, 3) uint64 first stored flat id, stored-lane mask, active-lane mask

    def pack(self, padded: np.ndarray) -> np.ndarray:
        """Gather a padded (..., nxp, nyp) field into flat order."""
        if tuple(padded.shape[-2:]) != (self.nxp, self.nyp):
            raise ValueError(f"pack: field has last axes {tuple(padded.shape[-2:])}, not ({self.nxp}, {self.nyp})")
        return padded[..., self.ij[:, 0], self.ij[:, 1]]

    def unpack(self, flat: np.ndarray, fill=0.0) -> np.ndarray:
        """Scatter a flat (..., N) field back to the padded grid."""
        if flat.shape[-1] != self.n_stored:
            raise ValueError(f"unpack: last axis {flat.shape[-1]} != n_stored {self.n_stored}")
        out = np.full(flat.shape[:-1] + (self.nxp, self.nyp), fill, dtype=flat.dtype)
        out[..., self.ij[:, 0], self.ij[:, 1]] = flat
        return out


def build_flat_mesh(inside_padded: np.ndarray, ngh: int, ring: int = 2, sweep_rows: int = None,
                    native: bool = True) -> FlatMesh:
    """Stored set, flat order, run table and row-sweep tiles of an active mask given on the
    padded grid. `native`: the library's host builder (gs_flat_mesh_build, one pass with
    rolling rows); otherwise the NumPy restatement below, which produces the same arrays."""
    if native:
        return _build_native(inside_padded, ngh, ring, sweep_rows)
    inside = np.asarray(inside_padded, dtype=bool)
    nxp, nyp = inside.shape
    if ring < 2:
        raise ValueError("ring must be >= 2 (the SRM stencil reaches two cells)")
    if inside[:2].any() or inside[-2:].any() or inside[:, :2].any() or inside[:, -2:].any():
        raise ValueError("active cells must lie at least two cells inside the padded grid (ngh >= 2)")
    if nyp >= 2 ** 31 // 4:
        raise ValueError("padded rows too long")
    stored = cross_dilate(inside, ring)
    n = int(stored.sum())
    idmap = np.full((nxp, nyp), -1, dtype=np.int64)
    idmap[stored] = np.arange(n, dtype=np.int64)
    ij = np.argwhere(stored).astype(np.int32)
    act = inside[ij[:, 0], ij[:, 1]]
    ka = np.nonzero(act)[0].astype(np.int64)
    ia = ij[ka, 0].astype(np.int64)
    ja = ij[ka, 1].astype(np.int64)
    d = np.stack([idmap[ia + 1, ja] - ka, idmap[ia - 1, ja] - ka,
                  idmap[ia + 2, ja] - ka, idmap[ia - 2, ja] - ka], axis=1)
    if (np.stack([idmap[ia + 1, ja], idmap[ia - 1, ja], idmap[ia + 2, ja], idmap[ia - 2, ja],
                  idmap[ia, ja + 1], idmap[ia, ja - 1], idmap[ia, ja + 2], idmap[ia, ja - 2]]) < 0).any():
        raise AssertionError("ring does not cover the +-2 stencil of every active cell")
    # along j the neighbours are consecutive ids (dilation invariant)
    assert np.all(idmap[ia, ja + 1] == ka + 1) and np.all(idmap[ia, ja - 2] == ka - 2)
    if np.abs(d).max(initial=0) >= 2 ** 31:
        raise ValueError("row offsets exceed int32")
    if ka.size:
        brk = np.ones(ka.size, dtype=bool)
        brk[1:] = ((ka[1:] != ka[:-1] + 1) | (ia[1:] != ia[:-1]) | (ja[1:] != ja[:-1] + 1)
                   | np.any(d[1:] != d[:-1], axis=1))
        starts = np.nonzero(brk)[0]
        ends = np.append(starts[1:], ka.size)
        run_k0 = ka[starts]
        run_len = (ends - starts).astype(np.int32)
        run_i = ia[starts].astype(np.int32)
        run_j0 = ja[starts].astype(np.int32)
        run_d = d[starts].astype(np.int32)
    else:
        run_k0 = np.zeros(0, np.int64)
        run_len = run_i = run_j0 = np.zeros(0, np.int32)
        run_d = np.zeros((0, 4), np.int32)
    nchunks = (n + CHUNK - 1) // CHUNK
    run_end = run_k0 + run_len
    chunk_run = np.searchsorted(run_end, np.arange(nchunks, dtype=np.int64) * CHUNK, side="right").astype(np.int32)
    if run_k0.size:
        lo = run_k0 + run_d.min(axis=1)
        hi = run_end - 1 + run_d.max(axis=1)
        if lo.min() < 0 or hi.max() >= n:
            raise AssertionError("run offsets leave the stored range")
    tile_e0, tile_irc, ent = build_sweep_tiles(stored, inside, sweep_rows)
    return FlatMesh(nxp=nxp, nyp=nyp, ngh=ngh, n_stored=n, n_active=int(act.sum()), ij=ij, active=act,
                    run_k0=run_k0.astype(np.int64), run_len=run_len, run_i=run_i, run_j0=run_j0,
                    run_d=np.ascontiguousarray(run_d), chunk_run=chunk_run,
                    tile_e0=tile_e0, tile_irc=tile_irc, ent=ent)


def _build_native(inside_padded: np.ndarray, ngh: int, ring: int, sweep_rows) -> FlatMesh:
    import ctypes as C
    from . import _lib as L
    inside = np.ascontiguousarray(np.asarray(inside_padded, dtype=bool), dtype=np.uint8)
    if inside.ndim != 2:
        raise ValueError("inside mask must be 2-D (padded grid)")
    nxp, nyp = inside.shape
    if sweep_rows is not None and sweep_rows < 1:
        raise ValueError("rows must be >= 1")
    out = L.GsFlatMeshHost()
    lib = L.lib()
    rc = lib.gs_flat_mesh_build(C.c_void_p(inside.ctypes.data), nxp, nyp, int(ring),
                                int(sweep_rows) if sweep_rows else -1, TILE_TARGET, MAX_SWEEP_ROWS, C.byref(out))
    if rc:
        msg = lib.gs_last_error().decode()
        raise (AssertionError if rc == 4 else ValueError)(msg)

    def arr(ptr, ctype, shape, dtype):
        count = int(np.prod(shape))
        if count == 0:
            return np.zeros(shape, dtype)
        a = np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ctype)), shape=(count,))
        return a.copy().view(dtype).reshape(shape)

    try:
        n, nr, nc, nt, ne = out.n_stored, out.n_runs, out.n_chunks, out.n_tiles, out.n_entries
        fm = FlatMesh(nxp=nxp, nyp=nyp, ngh=ngh, n_stored=n, n_active=out.n_active,
                      ij=arr(out.ij, C.c_int32, (n, 2), np.int32),
                      active=arr(out.active, C.c_uint8, (n,), np.uint8).astype(bool),
                      run_k0=arr(out.run_k0, C.c_int64, (nr,), np.int64),
                      run_len=arr(out.run_len, C.c_int32, (nr,), np.int32),
                      run_i=arr(out.run_i, C.c_int32, (nr,), np.int32),
                      run_j0=arr(out.run_j0, C.c_int32, (nr,), np.int32),
                      run_d=arr(out.run_d, C.c_int32, (nr, 4), np.int32),
                      chunk_run=arr(out.chunk_run, C.c_int32, (nc,), np.int32),
                      tile_e0=arr(out.tile_e0, C.c_int64, (nt,), np.int64),
                      tile_irc=arr(out.tile_irc, C.c_int32, (nt, 3), np.int32),
                      ent=arr(out.ent, C.c_uint64, (ne, 3), np.uint64))
    finally:
        lib.gs_flat_mesh_free(C.byref(out))
    return fm


WRITE_LANES = np.uint64(((1 << 62) - 1) & ~3)   # lanes 2..61 of a 64-lane sweep window
TILE_TARGET = 8192   # tiles a domain should give at least (~1.5 resident wavefronts of an MI355X each)
MAX_SWEEP_ROWS = 32  # tallest tile: each tile re-reads 4 stencil rows and re-evaluates one x-face row;
                     # on MI355X 32-row tiles run the 16384^2 coastal step 3-4 % faster than 64-row
                     # ones (twice the tiles, a shorter tail of the last round of wavefronts)


def _window_masks(m: np.ndarray, nwin: int) -> np.ndarray:
    """(rows, nwin) uint64: bit l of window k = m[:, CHUNK*k + l], l = 0..63."""
    rows = m.shape[0]
    v = np.lib.stride_tricks.as_strided(m, shape=(rows, nwin, 64),
                                        strides=(m.strides[0], CHUNK * m.strides[1], m.strides[1]))
    return np.packbits(v, axis=2, bitorder="little").view("<u8")[..., 0]


def sweep_rows_for(busy: int) -> int:
    """Tile height for `busy` window rows with an active lane: the largest power of two
    <= MAX_SWEEP_ROWS that still gives about TILE_TARGET tiles."""
    rows = 1
    while rows < MAX_SWEEP_ROWS and int(busy) // (2 * rows) >= TILE_TARGET:
        rows *= 2
    return rows


def mesh_sweep_tiles(fm: "FlatMesh", rows: int):
    """Sweep tiles of an existing flat mesh at another tile height (same stored set)."""
    stored = np.zeros((fm.nxp, fm.nyp), bool)
    stored[fm.ij[:, 0], fm.ij[:, 1]] = True
    inside = np.zeros((fm.nxp, fm.nyp), bool)
    a = fm.ij[fm.active]
    inside[a[:, 0], a[:, 1]] = True
    return build_sweep_tiles(stored, inside, rows)


def build_sweep_tiles(stored: np.ndarray, active: np.ndarray, rows: int = None):
    """Row-sweep tiles of the flat step.

    The padded rows are cut into bands of `rows` rows. In each band, windows of 64 lanes are
    placed from the left: a window starts two columns before the band's first active column
    not yet covered (lanes 2..61 are written, so it covers that column and the next 59), and
    the next window starts from the first active column after those 60. Every active column
    of the band thus belongs to exactly one window, with as few windows as 60-column windows
    allow (placing each at the first uncovered column is optimal for covering points with
    intervals of one length). For every row of a window one entry holds the flat id of the
    row's first stored cell at or after the window's column c0 and the stored- and
    active-lane masks: the stored cells of a row are consecutive flat ids, so lane l's id is
    base + popcount(stored & lanes below l). A tile is a window trimmed to the band's rows
    with an active lane; its entries are rows i0-2 .. i0+nr+1 (the stencil's two rows each
    side), stored contiguously. Tiles are ordered band-major, so the tiles that run side by
    side read neighbouring columns of the same rows.

    `rows=None` picks the tile height from the domain: the largest power of two <= 32 that
    still gives about TILE_TARGET tiles of the column grid CHUNK*k, so that a small domain run
    alone keeps every CU busy (a tile of r rows evaluates 1 + 1/r x-faces per row instead
    of 2).
    """
    stored = np.asarray(stored, dtype=bool)
    active = np.asarray(active, dtype=bool)
    nxp, nyp = stored.shape
    if rows is not None and rows < 1:
        raise ValueError("rows must be >= 1")
    if rows is None:
        nwin = max(1, -(-(nyp - 4) // CHUNK))
        A = np.zeros((nxp, CHUNK * nwin + 4), bool)
        A[:, :nyp] = active
        rows = sweep_rows_for(np.count_nonzero(_window_masks(A, nwin) & WRITE_LANES))   # grid windows
        del A
    if active[:, :2].any():
        raise AssertionError("a sweep window would start left of the padded grid")
    width = nyp + 64                      # room for a window at any active column
    S = np.zeros((nxp, width), bool)
    S[:, :nyp] = stored
    Aw = np.zeros((nxp, width), bool)
    Aw[:, :nyp] = active
    row_start = np.zeros(nxp, np.int64)
    row_start[1:] = np.cumsum(stored.sum(axis=1, dtype=np.int64))[:-1]
    lane = np.arange(64)
    i0s, nrs, c0s, ents = [], [], [], []
    for r0 in range(0, nxp, rows):
        r1 = min(r0 + rows, nxp)
        cols = np.flatnonzero(active[r0:r1].any(axis=0))
        if cols.size == 0:
            continue
        lo, hi = max(r0 - 2, 0), min(r1 + 2, nxp)
        before = np.zeros((hi - lo, width + 1), np.int64)      # stored cells left of a column
        np.cumsum(S[lo:hi], axis=1, out=before[:, 1:])
        k = 0
        while k < cols.size:
            c0 = int(cols[k]) - 2
            k = int(np.searchsorted(cols, c0 + 2 + CHUNK))
            live = Aw[r0:r1, c0 + 2:c0 + 62].any(axis=1)
            first = int(np.argmax(live))
            last = r1 - r0 - 1 - int(np.argmax(live[::-1]))
            i0, nr = r0 + first, last - first + 1
            if i0 < 2 or i0 + nr + 1 > nxp - 1:
                raise AssertionError("a sweep tile's stencil rows leave the padded grid")
            rr = np.arange(i0 - 2, i0 + nr + 2)
            e = np.empty((rr.size, 3), np.uint64)
            e[:, 0] = (row_start[rr] + before[rr - lo, c0]).astype(np.uint64)
            cols64 = c0 + lane
            e[:, 1] = np.packbits(S[rr[:, None], cols64], axis=1, bitorder="little").view("<u8")[:, 0]
            e[:, 2] = np.packbits(Aw[rr[:, None], cols64], axis=1, bitorder="little").view("<u8")[:, 0] & WRITE_LANES
            i0s.append(i0)
            nrs.append(nr)
            c0s.append(c0)
            ents.append(e)
    nt = len(i0s)
    cnt = np.asarray(nrs, np.int64) + 4
    tile_e0 = np.zeros(nt, np.int64)
    if nt:
        tile_e0[1:] = np.cumsum(cnt)[:-1]
    ent = np.concatenate(ents) if ents else np.zeros((0, 3), np.uint64)
    tile_irc = np.stack([np.asarray(i0s, np.int64), np.asarray(nrs, np.int64), np.asarray(c0s, np.int64)],
                        axis=1).astype(np.int32).reshape(nt, 3)
    return tile_e0, np.ascontiguousarray(tile_irc), np.ascontiguousarray(ent)
