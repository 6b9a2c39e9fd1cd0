"""CPU mirror of the device bed generator (test infrastructure only).

Restates gs_gen_bed (geoswe_amd/csrc/gs_util.hip) in NumPy with the same IEEE double
operations in the same order, so a CPU crop of a synthetic case is bit-identical to the
device field: Philox2x32-10 lattice values (Salmon et al., SC'11), value-noise fBm with a
smoothstep-weighted bilinear interpolation, float64 throughout, rounded once to the
state dtype. Pinned by the published Philox2x32-10 known-answer vectors
(tests/test_oracle_invariants.py) and by the device comparison in tests/test_gpu_cases.py.
"""
from __future__ import annotations

import numpy as np

M32 = np.uint64(0xFFFFFFFF)


def philox2x32(c0, c1, key):
    """Philox2x32 with 10 rounds; returns both output words (uint32 arrays)."""
    c0 = np.asarray(c0, dtype=np.uint64) & M32
    c1 = np.asarray(c1, dtype=np.uint64) & M32
    k = np.asarray(key, dtype=np.uint64) & M32
    for _ in range(10):
        p = np.uint64(0xD256D193) * c0
        hi = p >> np.uint64(32)
        lo = p & M32
        c0, c1 = (hi ^ k ^ c1) & M32, lo
        k = (k + np.uint64(0x9E3779B9)) & M32
    return c0.astype(np.uint32), c1.astype(np.uint32)


def lattice(seed: int, octave: int, ix, iy):
    c0 = np.asarray(ix, dtype=np.int64).astype(np.uint64) & M32
    c1 = (np.asarray(iy, dtype=np.int64).astype(np.uint64) & M32) ^ np.uint64((octave << 24) & 0xFFFFFFFF)
    key = (seed + 0x1000193 * octave) & 0xFFFFFFFF
    r, _ = philox2x32(c0, c1, key)
    return r.astype(np.float64) * (2.0 / 4294967296.0) - 1.0


def fbm(gi, gj, seed: int, octaves: int, L0: float, hurst: float):
    """Raw value-noise fBm at global cells (gi, gj) (broadcastable int arrays)."""
    gain = 2.0 ** (-hurst)
    gi = np.asarray(gi, dtype=np.int64)
    gj = np.asarray(gj, dtype=np.int64)
    acc = np.zeros(np.broadcast(gi, gj).shape)
    amp, L = 1.0, float(L0)
    for k in range(octaves):
        fx = (gi.astype(np.float64) + 0.5) / L
        fy = (gj.astype(np.float64) + 0.5) / L
        x0, y0 = np.floor(fx), np.floor(fy)
        tx, ty = fx - x0, fy - y0
        sx = (tx * tx) * (3.0 - 2.0 * tx)
        sy = (ty * ty) * (3.0 - 2.0 * ty)
        ix, iy = x0.astype(np.int64), y0.astype(np.int64)
        v00, v10 = lattice(seed, k, ix, iy), lattice(seed, k, ix + 1, iy)
        v01, v11 = lattice(seed, k, ix, iy + 1), lattice(seed, k, ix + 1, iy + 1)
        a = v00 + sx * (v10 - v00)
        b = v01 + sx * (v11 - v01)
        acc = acc + amp * (a + sy * (b - a))
        amp = amp * gain
        L = L * 0.5
    return acc


def bed_crop(i0, j0, nx, ny, seed, octaves, L0, hurst, mu, scale, slope_x, dx, offset, dtype=np.float32, stride=1,
             round_f32=False):
    """b = slope_x*x + scale*(fbm - mu) + offset on the crop [i0, i0+nx) x [j0, j0+ny)."""
    gi = i0 + np.arange(nx, dtype=np.int64)[:, None] * stride
    gj = j0 + np.arange(ny, dtype=np.int64)[None, :] * stride
    v = (fbm(gi, gj, seed, octaves, L0, hurst) - mu) * scale
    b = slope_x * ((gi.astype(np.float64) + 0.5) * dx) + v + offset
    if round_f32:
        b = b.astype(np.float32)
    return b.astype(dtype)


def voronoi_crop(i0, j0, nx, ny, sx, sy, scls):
    """Class of the nearest site per cell (float32 squared distance, ties to the lowest
    index), the CPU mirror of gs_gen_voronoi."""
    x = (np.arange(nx, dtype=np.int64) + i0).astype(np.float32) + np.float32(0.5)
    y = (np.arange(ny, dtype=np.int64) + j0).astype(np.float32) + np.float32(0.5)
    best = np.full((nx, ny), np.inf, np.float32)
    cls = np.zeros((nx, ny), np.uint8)
    for t in range(len(sx)):
        dx = x[:, None] - np.float32(sx[t])
        dy = y[None, :] - np.float32(sy[t])
        d = dx * dx + dy * dy
        m = d < best
        best = np.where(m, d, best)
        cls = np.where(m, np.uint8(scls[t]), cls)
    return cls
