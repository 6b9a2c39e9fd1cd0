"""The flat cache in the reference's layout: written by save_cache, replayed by run_cached.

A replay of a cache written at t = 0 takes the same steps and ends in the same state, bit
for bit, as the solver that wrote it (with a rain rate whose film bound never binds, since
the cached replay applies none); the files carry the reference's names, dtypes and shapes.
"""
import json
import os

import numpy as np
import pytest

from test_gpu_flat import dense, irregular_mask, rough_state

pytestmark = pytest.mark.gpu

NX, NY = 56, 70


def _solver(gridded=True, forcings=True):
    from geoswe_amd import CompressedSolver, RainfallForcing
    q0, bed = rough_state(NX, NY, seed=51, depth=0.7, rough=1.0, vel=0.3)
    m = irregular_mask(NX, NY, seed=9)
    q0 = q0 * m
    rng = np.random.default_rng(5)
    if gridded:
        rain = RainfallForcing(np.array([0.0, 0.7, 1.4]), rng.random((3, NX, NY)) * 40.0)
    else:
        rain = RainfallForcing(np.array([0.0, 1.0]), np.array([20.0, 35.0]))
    s = dense(q0, bed, friction="manning", rainfall_forcing=rain)
    g = s.mesh.ngh
    cls = rng.integers(0, 5, (NX, NY)).astype(np.uint8)
    s.set_manning_table(np.pad(cls, g), np.linspace(0.02, 0.06, 5).astype(np.float32))
    s.set_inside_mask(m)
    s.set_cfl_ghost_mask(~m)
    c = CompressedSolver.from_dense(s, cfl_linf=False, track_max_depth=True)
    if forcings:
        keep = np.ones((NX + 2 * g, NY + 2 * g))
        amb = np.zeros_like(keep)
        keep[g:g + 10] = 0.9
        amb[g:g + 10] = 0.04
        c.set_sponge({"keep": keep, "amb": amb})
        near = np.zeros_like(m)
        for ax in (0, 1):
            for sh in (1, -1):
                near |= np.roll(m, sh, ax)
        cells = np.argwhere(~m & near)
        w = rng.random((len(cells), 2))
        w /= w.sum(axis=1, keepdims=True)
        c.set_ring({"i": cells[:, 0] + g, "j": cells[:, 1] + g, "bed": bed[cells[:, 0], cells[:, 1]], "wg": w,
                    "NG": 2, "n": len(cells), "t_common": np.array([-1.0, 1.0, 3.0]),
                    "stage_all": 0.5 + 0.3 * rng.random((2, 3))})
        ids = np.nonzero(c.fm.active)[0][::17]
        c.set_drain({"idx": ids, "h_tgt_arr": 0.3 + 0.2 * rng.random(len(ids))})
    return c


@pytest.mark.parametrize("gridded", [True, False])
def test_cache_replay_is_bitwise_the_run(gpu, tmp_path, gridded):
    from geoswe_amd.cache import run_cached
    a = _solver(gridded)
    d = str(tmp_path / "cache")
    a.save_cache(d)
    a.run(2.5)
    b = run_cached(d, t_end=2.5, say=None)
    assert b.nsteps == a.nsteps > 5 and b.t == a.t
    assert np.array_equal(b.q_flat(), a.q_flat())
    assert np.array_equal(b.max_depth(), a.max_depth())


def test_cache_files_in_the_reference_layout(gpu, tmp_path):
    a = _solver()
    d = tmp_path / "cache"
    a.save_cache(str(d))
    N = a.n_stored
    meta = json.load(open(d / "meta.json"))
    for key in ("nxp", "nyp", "ngh", "dx", "x0", "y0", "crs_wkt", "nx_glob", "ny_glob", "has_ring", "has_sponge",
                "has_rain", "no_sigma", "N", "nranks", "ring_NG", "ring_n", "has_drain"):
        assert key in meta, key
    assert meta["N"] == N and meta["nranks"] == 1 and meta["has_ring"] and meta["has_sponge"] and meta["has_rain"]
    want = {"nbr": (np.int16, (N, 4)), "is_active": (np.uint8, (N,)), "ij_active": (np.int32, (N, 2)),
            "q0": (np.float32, (N,)), "q1": (np.float32, (N,)), "q2": (np.float32, (N,)),
            "bed_f": (np.float32, (N,)), "mcls_f": (np.uint8, (N,)), "m_tab": (np.float32, (5,)),
            "ring_rflat": (np.int32, (meta["ring_n"],)), "ring_rwg": (np.float32, (meta["ring_n"] * 2,)),
            "ring_stage_all": (np.float32, (2, 3)), "sponge_keep_f": (np.float32, (N,)),
            "sponge_amb_f": (np.float32, (N,)), "rain_native_rate_dev": (np.float32, (3, NX * NY)),
            "rain_lookup_flat": (np.int32, (N,)), "rain_t_s": (np.float64, (3,)), "drain_idx": (np.int32, None)}
    for name, (dt, shape) in want.items():
        a_ = np.load(d / f"{name}.npy")
        assert a_.dtype == dt, name
        assert shape is None or a_.shape == shape, (name, a_.shape)
    nbr = np.load(d / "nbr.npy").astype(np.int64)
    live = nbr != -32768
    k = np.arange(N)[:, None]
    ij = np.load(d / "ij_active.npy").astype(np.int64)
    tgt = (k + nbr)[live]
    off = np.array([[1, 0], [-1, 0], [0, 1], [0, -1]])
    src = np.broadcast_to(k, nbr.shape)[live]
    dirn = np.broadcast_to(np.arange(4), nbr.shape)[live]
    assert np.array_equal(ij[tgt], ij[src] + off[dirn])          # every delta lands on its face neighbour


def test_cache_refusals_and_warnings(gpu, tmp_path):
    from geoswe_amd import CompressedSolver
    from geoswe_amd.cache import from_cache
    c = _solver(forcings=False)
    c.set_infil({"tab": [1e-6] * 5})
    with pytest.warns(UserWarning, match="infiltration"):
        c.save_cache(str(tmp_path / "a"))
    q0, bed = rough_state(30, 30, seed=1)
    s = dense(q0, bed, dtype="float64")
    with pytest.raises(ValueError, match="float32"):
        CompressedSolver.from_dense(s).save_cache(str(tmp_path / "b"))
    os.environ["SWE_GHOST_ETA_RAMP_MMHR"] = "5"
    try:
        with pytest.raises(NotImplementedError, match="constant"):
            from_cache(str(tmp_path / "a"))
    finally:
        del os.environ["SWE_GHOST_ETA_RAMP_MMHR"]


@pytest.mark.parametrize("mode", ["hybrid", "extrapolate"])
def test_cache_replay_takes_the_ring_mode_from_the_environment(gpu, tmp_path, monkeypatch, mode):
    """GEOSWE_RING_BC / GEOSWE_RING_ETA select the replay's ring mode as in the reference; the
    replay equals a solver given the same mode by set_ring_mode."""
    from geoswe_amd.cache import run_cached
    d = str(tmp_path / "cache")
    _solver(forcings=False).save_cache(d)
    monkeypatch.setenv("GEOSWE_RING_BC", mode)
    monkeypatch.setenv("GEOSWE_RING_ETA", "0.55")
    b = run_cached(d, t_end=2.0, say=None)
    assert b._rmode is not None and b._rmode["mode"] == mode
    a = _solver(forcings=False)
    a._fo.drop_rain_cap()
    a.set_ring_mode(mode, eta=0.55)
    a.run(2.0)
    assert a.nsteps == b.nsteps and a.t == b.t
    act = a.fm.active
    assert np.array_equal(a.q_flat()[:, act], b.q_flat()[:, act])
