"""K plain steps recorded into a HIP graph and replayed equal K direct launches byte for byte:
the state, the running maximum and the whole device clock (t, dt, step count, wave speed,
forcing values). The reference pins the same property for its captured fixed-dt dense step
(tests/test_gpu_dense_no_host_sync.py:1-96). Replays alternate with eager work on the same
stream (direct steps, residual launches, read-backs), and every replay is checked."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rough_state(nx, ny, seed=5):
    rng = np.random.default_rng(seed)
    bed = rng.random((nx, ny)) * 0.8
    h = np.maximum(0.7 - bed, 0.0)
    wet = h > 1e-6
    return np.stack([h, (rng.random((nx, ny)) - 0.5) * 0.3 * wet, (rng.random((nx, ny)) - 0.5) * 0.3 * wet]), bed


def dense_pair(dtype="float32"):
    from geoswe_amd import Config, Mesh2D, RainfallForcing, Solver2D
    q0, bed = rough_state(70, 90)
    out = []
    for _ in range(2):
        rf = RainfallForcing(time_s=np.array([0.0, 0.5]), rate_mm_h=np.array([40.0, 90.0]))
        cfg = Config(bc_x="wall", bc_y="fall", dtype=dtype, friction="manning", manning_n=0.03,
                     rainfall_forcing=rf)
        s = Solver2D(Mesh2D(70, 90, 2.0, 2.0), cfg, q0, bed, track_max_depth=True)
        s.set_manning(np.where(np.arange(90)[None, :] < 45, 0.02, 0.05) * np.ones((70, 1)))
        out.append(s)
    return out


def same(a, b):
    assert a.nsteps == b.nsteps
    assert a._clock.download().tobytes() == b._clock.download().tobytes()
    assert np.array_equal(a.q_interior, b.q_interior)
    assert np.array_equal(a.max_depth(), b.max_depth())


@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_dense_graph_replay_equals_direct_launches(gpu, dtype):
    direct, graphed = dense_pair(dtype)
    g = graphed.capture_steps(7)
    assert g.nodes == 7
    for rep in range(3):
        g.replay()
        graphed.sync()
        direct.advance(7)
        same(direct, graphed)
        graphed.residual()                 # eager kernels between replays
        direct.advance(1)                  # and a direct step, so the next replay starts at odd parity
        graphed.advance(1)
        same(direct, graphed)
    g.replay(2)
    graphed.sync()
    direct.advance(14)
    same(direct, graphed)
    assert direct.nsteps == 3 * 8 + 14
    g.close()


def test_flat_graph_replay_equals_direct_launches(gpu):
    from geoswe_amd import CompressedSolver
    pair = []
    for s in dense_pair():
        m = np.ones((70, 90), bool)
        m[20:30, 10:70] = False
        s.set_inside_mask(m)
        pair.append(CompressedSolver.from_dense(s, cfl_linf=True))
    direct, graphed = pair
    g = graphed.capture_steps(5)
    for rep in range(3):
        g.replay()
        graphed.sync()
        direct.advance(5)
        assert direct._clock.download().tobytes() == graphed._clock.download().tobytes()
        assert np.array_equal(direct.q_flat(), graphed.q_flat())
        assert np.array_equal(direct.max_depth(), graphed.max_depth())
        graphed.residual()
    g.close()
