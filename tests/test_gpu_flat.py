"""Flat (compressed active-cell) layout against the dense layout and the oracle.

The flat kernel shares the dense kernel's face arithmetic, so at a fixed state its
residual equals the dense residual bit for bit on every active cell, and with the L-inf
CFL norm on both layouts a run takes the same steps and ends in the same state.
"""
import numpy as np
import pytest

from oracle import swe_oracle as O

pytestmark = pytest.mark.gpu


def rough_state(nx, ny, seed=7, depth=0.8, rough=1.2, vel=0.04):
    rng = np.random.default_rng(seed)
    bed = rng.random((nx, ny)) * rough
    h = np.maximum(depth - bed, 0.0)
    wet = h > 1e-6
    return np.stack([h, (rng.random((nx, ny)) - 0.5) * vel * wet, (rng.random((nx, ny)) - 0.5) * vel * wet]), bed


def irregular_mask(nx, ny, seed=3):
    rng = np.random.default_rng(seed)
    i = np.arange(nx)[:, None]
    j = np.arange(ny)[None, :]
    m = (i - nx / 2) ** 2 / (0.4 * nx) ** 2 + (j - ny / 2) ** 2 / (0.45 * ny) ** 2 < 1.0
    m &= rng.random((nx, ny)) > 0.04
    m[nx // 3, :] = False
    return m


def dense(q0, bed, bc=("fall", "fall"), **kw):
    from geoswe_amd import Config, Mesh2D, Solver2D
    cfg = Config(bc_x=bc[0], bc_y=bc[1], **kw)
    return Solver2D(Mesh2D(bed.shape[0], bed.shape[1], 2.0, 2.0), cfg, q0, bed)


@pytest.mark.parametrize("masked", [False, True])
def test_flat_residual_bitwise_dense(gpu, masked):
    from geoswe_amd import CompressedSolver
    q0, bed = rough_state(70, 90, seed=13)
    s = dense(q0, bed)
    if masked:
        s.set_inside_mask(irregular_mask(70, 90))
    r_dense = s.residual()
    c = CompressedSolver.from_dense(s, cfl_linf=True)
    r_flat = c.residual()
    act = c.fm.active
    g = s.mesh.ngh
    ij = c.fm.ij[act]
    want = r_dense[:, ij[:, 0] - g, ij[:, 1] - g]
    assert np.array_equal(r_flat[:, act], want)
    assert np.all(r_flat[:, ~act] == 0)
    assert c.n_active == (int(s.inside_mask.sum()) if masked else 70 * 90)


@pytest.mark.parametrize("masked", [False, True])
def test_flat_run_matches_dense(gpu, masked):
    """Same step count and bit-identical state after a run (L-inf CFL norm on both)."""
    from geoswe_amd import CompressedSolver
    q0, bed = rough_state(64, 80, seed=17, depth=0.7, rough=0.9, vel=0.3)
    mk = lambda: dense(q0, bed, friction="manning", manning_n=0.03, rainfall=60.0 / 3.6e6)
    if masked:
        # inactive cells at rest: the dense carry leaves them unchanged, as the flat ring
        m = irregular_mask(64, 80)
        q0 = q0 * m
    s = mk()
    c_src = mk()
    if masked:
        for x in (s, c_src):
            x.set_inside_mask(m)
            x.set_cfl_ghost_mask(~m)      # the dense CFL otherwise includes inactive cells
    c = CompressedSolver.from_dense(c_src, cfl_linf=True)
    nd = s.run(4.0)
    nc = c.run(4.0)
    assert nd == nc > 5
    assert s.t == c.t
    hd = s.depth()
    hc = c.depth()
    act = c.fm.active
    g = s.mesh.ngh
    ij = c.fm.ij[act]
    assert np.array_equal(hc[ij[:, 0] - g, ij[:, 1] - g], hd[ij[:, 0] - g, ij[:, 1] - g])


def test_flat_classes_gridded_rain_stage(gpu):
    """C3-shaped features on the flat layout: Manning classes, gridded rain, tidal stage."""
    from geoswe_amd import CompressedSolver, RainfallForcing, StageBoundary
    nx, ny = 60, 60
    q0, bed = rough_state(nx, ny, seed=23, depth=0.5, rough=1.0, vel=0.1)
    rng = np.random.default_rng(2)
    frames = rng.random((2, nx, ny)) * 80.0
    sea = np.zeros((nx, ny), bool)
    sea[0, :] = True
    from geoswe_amd import Mesh2D
    mesh = Mesh2D(nx, ny, 2.0, 2.0)
    sb = StageBoundary.from_mask(sea, mesh, bed, np.array([0.0, 5.0]), np.array([0.6, 1.1]))
    mk = lambda: dense(q0, bed, friction="manning", rainfall_forcing=RainfallForcing(np.array([0.0, 1.0]), frames),
                       stage_boundary=sb)   # reads q0 when called
    n = np.choose(rng.integers(0, 8, (nx, ny)), np.linspace(0.02, 0.12, 8))
    m = irregular_mask(nx, ny, seed=8)
    m[0, :] = True
    q0 = q0 * m                      # inactive cells at rest
    s = mk()
    s.set_manning(n)
    s.set_inside_mask(m)
    s.set_cfl_ghost_mask(~m)
    c = CompressedSolver.from_dense(s, cfl_linf=True)
    s.run(2.5)
    c.run(2.5)
    assert s.nsteps == c.nsteps and s.t == c.t
    g = s.mesh.ngh
    act = c.fm.active
    ij = c.fm.ij[act]
    assert np.array_equal(c.depth()[ij[:, 0] - g, ij[:, 1] - g], s.depth()[ij[:, 0] - g, ij[:, 1] - g])
    # and the dense run against the oracle
    case = O.DenseCase(h=q0[0], hu=q0[1], hv=q0[2], bed=bed, dx=2.0, dy=2.0, bc_x="fall", bc_y="fall",
                       manning=np.asarray(n, np.float32), rain=(np.array([0.0, 1.0]), frames), active=m,
                       stage={"cells": np.argwhere(sea), "time_s": np.array([0.0, 5.0]),
                              "stage_m": np.array([0.6, 1.1])})
    o = O.DenseOracle(case, O.Physics(friction=True), np.float32)
    o.cfl_skip = ~m
    o.run(2.5)
    assert np.array_equal(s.q_interior, o.interior)


def _traversal_masks(nx, ny):
    rng = np.random.default_rng(41)
    out = [("irregular", irregular_mask(nx, ny, seed=5))]
    m = np.zeros((nx, ny), bool)          # islands closer together than one 64-column window
    m[3:9, 5:15] = m[2:12, 30:36] = m[20:40, 50:58] = True
    m[30, 60:ny - 1] = True
    m[nx - 6:, 100:110] = True
    out.append(("islands", m))
    out.append(("full", np.ones((nx, ny), bool)))
    m = rng.random((nx, ny)) > 0.35         # speckle: most window rows are partly stored
    out.append(("speckle", m))
    return out


@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_sweep_tiles_equal_chunk_traversal(gpu, dtype):
    """The row-sweep traversal and the chunk traversal of the run table give the same
    state, running maximum, time and step count, for ragged, islanded, full and speckled
    masks (three windows wide), tiles of 1, 5 and 64 rows, with Manning classes, gridded
    rain and a tidal stage row."""
    from geoswe_amd import CompressedSolver, RainfallForcing, StageBoundary, Mesh2D
    nx, ny = 48, 150
    q0, bed = rough_state(nx, ny, seed=29, depth=0.6, rough=1.0, vel=0.2)
    rng = np.random.default_rng(4)
    frames = rng.random((2, nx, ny)) * 80.0
    n = np.choose(rng.integers(0, 8, (nx, ny)), np.linspace(0.02, 0.12, 8))
    sea = np.zeros((nx, ny), bool)
    sea[0, :] = True
    sb = StageBoundary.from_mask(sea, Mesh2D(nx, ny, 2.0, 2.0), bed, np.array([0.0, 5.0]), np.array([0.5, 1.0]))
    for name, m in _traversal_masks(nx, ny):
        m = m.copy()
        m[0, :] = True
        for rows in (1, 5, 64):
            runs = []
            for trav in ("tiles", "chunks"):
                s = dense(q0 * m, bed, friction="manning", dtype=dtype, stage_boundary=sb,
                          rainfall_forcing=RainfallForcing(np.array([0.0, 1.0]), frames))
                s.set_manning(n)
                s.set_inside_mask(m)
                s.set_cfl_ghost_mask(~m)
                c = CompressedSolver.from_dense(s, cfl_linf=True, track_max_depth=True, sweep_rows=rows)
                c.traversal = trav
                c.run(1.5)
                runs.append((c.q_flat(), c.max_depth(), c.t, c.nsteps))
            (qa, ma, ta, na), (qb, mb, tb, nb) = runs
            print(name, rows, dtype, na, flush=True)
            assert na == nb > 3 and ta == tb, (name, rows)
            assert np.array_equal(qa, qb), (name, rows)
            assert np.array_equal(ma, mb), (name, rows)


def test_checkpoint_resume_is_bitwise(gpu, tmp_path):
    """A run stopped by max_steps after writing checkpoints, resumed from the last one in a
    fresh solver, ends bit for bit where the uninterrupted run ends (state, running maximum,
    time, step count); gridded rain, Manning classes, a tidal stage row and an irregular
    mask; a torn checkpoint (array stamp and meta disagree) is refused."""
    import json
    from geoswe_amd import CompressedSolver, RainfallForcing, StageBoundary, Mesh2D
    nx, ny = 48, 70
    q0, bed = rough_state(nx, ny, seed=31, depth=0.6, rough=1.0, vel=0.2)
    rng = np.random.default_rng(6)
    frames = rng.random((3, nx, ny)) * 80.0
    n = np.choose(rng.integers(0, 8, (nx, ny)), np.linspace(0.02, 0.12, 8))
    m = irregular_mask(nx, ny, seed=12)
    m[0, :] = True
    sea = np.zeros((nx, ny), bool)
    sea[0, :] = True
    sb = StageBoundary.from_mask(sea, Mesh2D(nx, ny, 2.0, 2.0), bed, np.array([0.0, 40.0]), np.array([0.5, 1.0]))

    def mk():
        s = dense(q0 * m, bed, friction="manning", stage_boundary=sb,
                  rainfall_forcing=RainfallForcing(np.array([0.0, 6.0, 14.0]), frames))
        s.set_manning(n)
        s.set_inside_mask(m)
        s.set_cfl_ghost_mask(~m)
        return CompressedSolver.from_dense(s, cfl_linf=True, track_max_depth=True)

    T = 30.0
    a = mk()
    a.run(T, batch=8)
    d = str(tmp_path / "ckpt")
    b = mk()
    b.run(T, max_steps=37, batch=8, ckpt_dir=d, checkpoint_every_s=3.0)
    assert b.nsteps == 37 and b.t < T and a.nsteps > 37
    meta = json.load(open(tmp_path / "ckpt" / "ckpt_meta.json"))
    assert meta["steps"] == 37 and meta["t"] == b.t and meta["nranks"] == 1
    c = mk()
    c.run(T, batch=8, ckpt_dir=d, resume=True)
    assert c.nsteps == a.nsteps and c.t == a.t
    assert np.array_equal(c.q_flat(), a.q_flat())
    assert np.array_equal(c.max_depth(), a.max_depth())
    meta["steps"] = 36
    json.dump(meta, open(tmp_path / "ckpt" / "ckpt_meta.json", "w"))
    with pytest.raises(RuntimeError, match="torn"):
        mk().load_checkpoint(d)


def test_frames_in_reference_layout(gpu, tmp_path):
    """run(out_dir=, frame_every_s=) writes the reference's frames: a float16 interior depth
    after the first step reaching each frame time and one at the end, named by index and
    rounded time, with a manifest; the framed run ends bit for bit where an unframed run
    ends."""
    import json
    import os
    from geoswe_amd import CompressedSolver
    q0, bed = rough_state(40, 56, seed=37, depth=0.5, rough=0.8, vel=0.2)
    m = irregular_mask(40, 56, seed=14)

    def mk():
        s = dense(q0 * m, bed, friction="manning", manning_n=0.03, rainfall=30.0 / 3.6e6)
        s.set_inside_mask(m)
        s.set_cfl_ghost_mask(~m)
        return CompressedSolver.from_dense(s, cfl_linf=True)

    T, every = 20.0, 3.0
    a = mk()
    a.run(T)
    b = mk()
    b.run(T, out_dir=str(tmp_path), frame_every_s=every)
    assert b.nsteps == a.nsteps and b.t == a.t and np.array_equal(b.q_flat(), a.q_flat())
    # the expected frames, one step at a time (each call plans dt exactly as the run did)
    c = mk()
    want, nxt, last = [], 0.0, -1.0
    while c.t < T - 1e-12:
        c.run(T, max_steps=1)
        if c.t >= nxt - 1e-9:
            want.append((len(want), c.t, c.depth().astype(np.float16)))
            nxt, last = nxt + every, c.t
    if c.t > last + 1e-6:
        want.append((len(want), c.t, c.depth().astype(np.float16)))
    assert c.nsteps == a.nsteps and np.array_equal(c.q_flat(), a.q_flat())
    fdir = tmp_path / "frames_parallel"
    names = sorted(f for f in os.listdir(fdir) if f.startswith("depth_"))
    assert names == [f"depth_{i:05d}_t{int(round(t)):07d}_r00.npz" for i, t, _ in want]
    for (i, t, h), name in zip(want, names):
        z = np.load(fdir / name)
        assert z["h"].dtype == np.float16 and np.array_equal(z["h"], h), name
    man = json.load(open(fdir / "manifest.json"))
    assert man["nranks"] == 1 and man["frame_every_s"] == every and man["ranks"][0]["nx"] == 40


def test_per_tile_launch_equals_chunk_traversal(gpu):
    """A domain with more than four resident rounds of tiles launches one wavefront per tile
    (2048 x 1040 cells, one-row tiles: about 36 900 tiles); its steps equal the chunk
    traversal's bit for bit, for the plain and the extended sweep (Green-Ampt attached)."""
    from geoswe_amd import CompressedSolver
    nx, ny = 2048, 1040
    q0, bed = rough_state(nx, ny, seed=81, depth=0.6, rough=0.8, vel=0.2)
    rng = np.random.default_rng(8)
    cls = rng.integers(0, 3, (nx, ny)).astype(np.uint8)
    for ext in (False, True):
        runs = []
        for trav in ("tiles", "chunks"):
            s = dense(q0, bed, friction="manning", manning_n=0.03, rainfall=20.0 / 3.6e6)
            c = CompressedSolver.from_dense(s, cfl_linf=True, sweep_rows=1)
            if ext:
                g = s.mesh.ngh
                c.set_ga_drain({"cls_pad": np.pad(cls, g), "Ks_t": [0.0, 1e-5, 4e-5], "psi_t": [0.1, 0.1, 0.2],
                                "dth_t": [0.3, 0.3, 0.3], "F_pad": None,
                                "inv_tau_pad": np.full((nx + 2 * g, ny + 2 * g), 1.0 / 600.0), "mode": "fused"})
            c.traversal = trav
            info = c.launch_info()
            if trav == "tiles":
                assert info["kernel"] == ("sweep_ext" if ext else "sweep")
                assert info["blocks"] > info["resident"], info     # one wavefront per tile
            else:
                assert info["kernel"] == "chunks"
            c.advance(6)
            runs.append((c.q_flat(), c.nsteps, c.t))
        (qa, na, ta), (qb, nb, tb) = runs
        print("ext", ext, na, flush=True)
        assert na == nb == 6 and ta == tb and np.array_equal(qa, qb)
