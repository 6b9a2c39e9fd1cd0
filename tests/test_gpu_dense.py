"""Dense-layout HIP step against the CPU oracle (bit-exact in float32).

Every operation of the step is one IEEE-754 rounding (no FMA contraction, correctly
rounded division and square root, the reproducible h^(-4/3)), so the float32 GPU state
must equal the NumPy oracle's bit for bit; float64 is compared to 1e-12 relative.
"""
from pathlib import Path

import numpy as np
import pytest

from oracle import swe_oracle as O

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"


def rough_state(nx, ny, seed=7, depth=0.8, rough=1.2, vel=0.04):
    """Rough, partially dry state (the reference's test_srm_cpu_matches_gpu.py:44-57 recipe)."""
    rng = np.random.default_rng(seed)
    bed = rng.random((nx, ny)) * rough
    h = np.maximum(depth - bed, 0.0)
    wet = h > 1e-6
    hu = (rng.random((nx, ny)) - 0.5) * vel * wet
    hv = (rng.random((nx, ny)) - 0.5) * vel * wet
    return np.stack([h, hu, hv]), bed


def make(q0, bed, dx=2.0, dtype="float32", bc=("wall", "wall"), ngh=4, **cfgkw):
    from geoswe_amd import Config, Mesh2D, Solver2D
    nx, ny = bed.shape
    cfg = Config(bc_x=bc[0], bc_y=bc[1], dtype=dtype, **cfgkw)
    return Solver2D(Mesh2D(nx, ny, dx, dx, ngh=ngh), cfg, q0, bed)


def oracle_residual(q0, bed, dx, dtype, bc, h_min):
    ngh = 4
    nx, ny = bed.shape
    qp = np.zeros((3, nx + 2 * ngh, ny + 2 * ngh), dtype)
    qp[:, ngh:-ngh, ngh:-ngh] = q0.astype(dtype)
    O.bc_fill(qp, ngh, *bc)
    bp = O.pad_bed(bed.astype(dtype), ngh, *bc, dtype=dtype)
    return O.residual(qp, bp, dx, dx, 9.81, h_min)[:, ngh:-ngh, ngh:-ngh]


@pytest.mark.parametrize("bc", [("wall", "wall"), ("fall", "extrapolate"), ("periodic", "periodic"),
                                ("extrapolate", "fall"), ("wall", "periodic")])
@pytest.mark.parametrize("shape", [(40, 40), (67, 131), (5, 200)])
def test_residual_f32_bitexact(gpu, bc, shape):
    q0, bed = rough_state(*shape)
    s = make(q0, bed, bc=bc)
    got = s.residual()
    want = oracle_residual(q0, bed, 2.0, np.float32, bc, 1e-6)
    assert got.dtype == np.float32
    assert np.array_equal(got, want), np.abs(got - want).max()
    assert np.abs(want).max() > 1e-3


@pytest.mark.parametrize("bc", [("wall", "wall"), ("periodic", "fall")])
def test_residual_f64(gpu, bc):
    q0, bed = rough_state(53, 61, seed=11)
    s = make(q0, bed, dtype="float64", bc=bc, h_min=1e-6)
    got = s.residual()
    want = oracle_residual(q0, bed, 2.0, np.float64, bc, 1e-6)
    assert np.abs(got - want).max() <= 1e-13 * np.abs(want).max()


def test_residual_f64_golden_reference(gpu):
    """The reference NumPy residual at its own rough 40x40 state (fixture), through the GPU."""
    z = np.load(GOLD / "rhs_srm40_f64.npz")
    s = make(z["q0"], z["bed"], dtype="float64", bc=("wall", "wall"), h_min=1e-6)
    got = s.residual()
    ref = z["rhs"][:, 2:-2, 2:-2]
    assert np.abs(got - ref).max() < 1e-12


def run_both(q0, bed, nsteps, dtype=np.float32, bc=("wall", "wall"), dx=2.0, manning=None, rain=None,
             active=None, stage=None, **cfgkw):
    """Run the GPU solver and the oracle side by side; returns (solver, oracle)."""
    from geoswe_amd import RainfallForcing, StageBoundary
    kw = dict(cfgkw)
    if manning is not None:
        kw["friction"] = "manning"
    if rain is not None:
        if np.isscalar(rain):
            kw["rainfall"] = rain
        else:
            kw["rainfall_forcing"] = RainfallForcing(time_s=rain[0], rate_mm_h=rain[1])
    if stage is not None:
        cells = np.asarray(stage["cells"]) + 4
        kw["stage_boundary"] = StageBoundary(cells=cells, time_s=stage["time_s"], stage_m=stage["stage_m"],
                                             bed_b=bed[tuple(np.asarray(stage["cells"]).T)])
    s = make(q0, bed, dx=dx, dtype=np.dtype(dtype).name, bc=bc, **kw)
    if manning is not None:
        s.set_manning(manning)
    if active is not None:
        s.set_inside_mask(active)
    for _ in range(nsteps):
        s.step()
    ph = O.Physics(h_min=s.cfg.h_min, h_cfl=s.cfg.h_cfl(), cfl=s.cfg.cfl, friction=manning is not None,
                   vcap=s.cfg.friction_velocity_cap_ms)
    case = O.DenseCase(h=q0[0], hu=q0[1], hv=q0[2], bed=bed, dx=dx, dy=dx, bc_x=bc[0], bc_y=bc[1],
                       manning=(np.dtype(dtype).type(manning) if np.isscalar(manning) else
                                (np.asarray(manning, np.float32) if manning is not None else 0.0)),
                       rain=rain, active=active, stage=stage)
    o = O.DenseOracle(case, ph, dtype)
    o.run_steps(nsteps)
    return s, o


def test_dam_break_trajectory_bitexact(gpu):
    N = 96
    i = np.arange(N)[:, None]
    j = np.arange(80)[None, :]
    h = np.where(np.hypot(i - 47.5, j - 39.5) < 15, 10.0, 1.0)
    q0 = np.stack([h, 0 * h, 0 * h])
    s, o = run_both(q0, 0 * h, 60, dx=1.0)
    assert np.array_equal(s.q_interior, o.interior)
    assert s.t == o.t and s.nsteps == 60


def test_friction_rain_mask_stage_bitexact(gpu):
    nx, ny = 64, 72
    q0, bed = rough_state(nx, ny, seed=5, depth=0.7, rough=1.0, vel=0.3)
    rng = np.random.default_rng(1)
    n = np.choose(rng.integers(0, 5, (nx, ny)), [0.02, 0.035, 0.05, 0.08, 0.12])
    active = np.ones((nx, ny), bool)
    active[40:50, 10:30] = False
    frames = (rng.random((3, nx, ny)) * 80.0).astype(np.float64)
    rain = (np.array([0.0, 0.4, 0.9]), frames)
    cells = np.stack(np.nonzero(np.zeros((nx, ny), bool) | (np.arange(nx)[:, None] == 0))).T
    stage = {"cells": cells, "time_s": np.array([0.0, 10.0]), "stage_m": np.array([0.9, 1.4])}
    s, o = run_both(q0, bed, 40, bc=("wall", "fall"), manning=n, rain=rain, active=active, stage=stage)
    assert np.array_equal(s.q_interior, o.interior)
    assert s.t == o.t


def test_velocity_cap_and_linear_friction_bitexact(gpu):
    nx, ny = 48, 48
    q0, bed = rough_state(nx, ny, seed=9, depth=0.5, rough=0.4, vel=40.0)   # |U| far above the 15 m/s cap
    s, o = run_both(q0, bed, 20, manning=0.03, rain=50.0 / 3.6e6, friction_quadratic_alpha=True)
    assert np.array_equal(s.q_interior, o.interior)
    from oracle.swe_oracle import Physics
    s2, o2 = None, None
    from geoswe_amd import Config
    q0b, bedb = rough_state(nx, ny, seed=10, depth=0.5, rough=0.4, vel=2.0)
    s2 = make(q0b, bedb, friction="manning", manning_n=0.05, friction_quadratic_alpha=False)
    for _ in range(15):
        s2.step()
    case = O.DenseCase(h=q0b[0], hu=q0b[1], hv=q0b[2], bed=bedb, dx=2.0, dy=2.0, manning=np.float32(0.05))
    o2 = O.DenseOracle(case, Physics(friction=True, quadratic=False), np.float32)
    o2.run_steps(15)
    assert np.array_equal(s2.q_interior, o2.interior)


def test_f64_trajectory(gpu):
    q0, bed = rough_state(40, 56, seed=3, depth=0.9, rough=0.6, vel=0.2)
    s, o = run_both(q0, bed, 25, dtype=np.float64, manning=0.04, rain=30.0 / 3.6e6, bc=("extrapolate", "wall"),
                    h_min=1e-6)
    # float64 division and square root are correctly rounded on the GPU too: bit-exact
    assert np.array_equal(s.q_interior, o.interior), np.abs(s.q_interior - o.interior).max()
    assert s.t == o.t


def test_lake_at_rest_f32(gpu):
    nx = ny = 64
    i = np.arange(nx)[:, None]
    j = np.arange(ny)[None, :]
    bed = 0.8 * np.exp(-((i - 30) ** 2 + (j - 34) ** 2) / 60.0) + 0.3 * np.sin(i / 4.0) * np.cos(j / 6.0)
    bed = bed - bed.min()
    h = np.maximum(1.0 - bed, 0.0)
    s = make(np.stack([h, 0 * h, 0 * h]), bed, dx=1.0)
    s.advance(200)
    q = s.q_interior
    eta = q[0] + bed.astype(np.float32)
    wet = q[0] > 1e-3
    assert np.abs(q[1]).max() < 1e-5 and np.abs(q[2]).max() < 1e-5
    assert np.abs(eta[wet] - 1.0).max() < 1e-5


def test_run_to_t_end_matches_oracle(gpu):
    q0, bed = rough_state(50, 50, seed=21, depth=0.6, rough=0.5, vel=0.5)
    from geoswe_amd import Config
    s = make(q0, bed, friction="manning", manning_n=0.03, rainfall=120.0 / 3.6e6)
    n = s.run(3.0, dt_max=0.05, batch=7)
    case = O.DenseCase(h=q0[0], hu=q0[1], hv=q0[2], bed=bed, dx=2.0, dy=2.0, manning=np.float32(0.03),
                       rain=120.0 / 3.6e6)
    o = O.DenseOracle(case, O.Physics(friction=True), np.float32)
    no = o.run(3.0, dt_max=0.05)
    assert n == no and s.t == o.t and abs(s.t - 3.0) < 1e-9
    assert np.array_equal(s.q_interior, o.interior)


def test_explicit_dt_and_advance_equal(gpu):
    q0, bed = rough_state(33, 70, seed=2)
    a = make(q0, bed)
    b = make(q0, bed)
    a.advance(12)
    for _ in range(12):
        b.step()
    assert np.array_equal(a.q_interior, b.q_interior) and a.t == b.t
    c = make(q0, bed)
    for dt in (0.01, 0.02, 0.005):
        assert c.step(dt) == dt
    case = O.DenseCase(h=q0[0], hu=q0[1], hv=q0[2], bed=bed, dx=2.0, dy=2.0)
    o = O.DenseOracle(case, O.Physics(), np.float32)
    for dt in (0.01, 0.02, 0.005):
        o.step(dt)
    assert np.array_equal(c.q_interior, o.interior) and c.t == o.t


def test_nonfinite_state_raises(gpu):
    q0, bed = rough_state(20, 20)
    q0[1, 5, 5] = np.inf
    with pytest.raises(ValueError):
        make(q0, bed)
    q0, bed = rough_state(20, 20)
    s = make(q0, bed)
    q = s.q_interior
    q[0, 7, 9] = np.nan
    s.set_state(q)
    with pytest.raises(FloatingPointError):
        s.run(10.0)


def test_max_depth_tracks_running_maximum(gpu):
    """Running maximum over the states after each step, as Solver2D._update_max_depth."""
    q0, bed = rough_state(30, 30, seed=4, depth=0.6)
    s = make(q0, bed)
    assert np.array_equal(s.max_depth(), s.depth())          # before any step: the depth
    mx = np.zeros(bed.shape, np.float32)
    for _ in range(10):
        s.step()
        mx = np.maximum(mx, s.q_interior[0])
    assert np.array_equal(s.max_depth(), mx)
