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


def build_flat_mesh(inside_padded: np.ndarray, ngh: int, ring: int = 2) -> FlatMesh:
    """Stored set, flat order and run table of an active mask given on the padded grid."""
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
    return FlatMesh(nxp=nxp, nyp=nyp, ngh=ngh, n_stored=n, n_active=int(act.sum()), ij=ij, active=act,
                    run_k0=run_k0.astype(np.int64), run_len=run_len, run_i=run_i, run_j0=run_j0,
                    run_d=np.ascontiguousarray(run_d), chunk_run=chunk_run)
