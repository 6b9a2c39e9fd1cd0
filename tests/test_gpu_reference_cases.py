"""The HIP step on the reference's own golden cases and on the configurations' exact set-ups.

* A full reference ``Solver2D.step`` with a Manning class table and the 15 m/s velocity cap
  engaged (both friction roots), and a ``Solver2D.run`` on a dry bed under rain where the
  rain-film step bound sets the first steps: the float64 GPU path against the reference's
  vectors, float32 and float64 bit for bit against the oracle.
* The 256^2 circular dam break at its configured size (r < 40 cells, walls, 2000 steps):
  float32 and float64 bit for bit against the oracle, float64 against the reference's own
  trajectory at steps 1, 10 and 100.
* The flat layout's CFL rule (skip h < h_min, velocity over max(h, h_min_cfl), floor
  sqrt(g h_min)) with h_min_cfl != h_min in both velocity norms, and its rain-film bound from
  the whole table's largest rate, against the oracle's flat rules.
"""
from pathlib import Path

import numpy as np
import pytest

from oracle import swe_oracle as O
from test_oracle_golden import _face_velocity_divergence, _vcap_oracle

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"


def _vcap_solver(z, quadratic, dtype):
    from geoswe_amd import Config, Mesh2D, Solver2D
    q0, bed = z["q0"], z["bed"]
    cfg = Config(bc_x="wall", bc_y="extrapolate", dtype=dtype, h_min=1e-6, cfl=0.45, friction="manning",
                 friction_quadratic_alpha=quadratic)
    s = Solver2D(Mesh2D(q0.shape[1], q0.shape[2], 2.0, 2.0, ngh=4), cfg, q0, bed)
    s.set_manning_table(np.pad(z["cls"], 4, mode="edge"), z["tab"])
    return s


@pytest.mark.parametrize("quadratic", [True, False])
def test_vcap_class_table_step_against_reference(gpu, quadratic):
    z = np.load(GOLD / "step_vcap_f64.npz")
    tag = "quad" if quadratic else "lin"
    s = _vcap_solver(z, quadratic, "float64")
    dt = s.step()
    assert abs(dt - float(z[f"dt_{tag}"])) <= 1e-15 * dt
    want = z[f"q1_{tag}"]
    scale = np.abs(want).max(axis=(1, 2), keepdims=True)
    err = (np.abs(s.q_interior - want) / scale).max(axis=0)
    skip = _face_velocity_divergence(z["q0"], z["bed"])
    assert err[~skip].max() < 1e-13
    for dtype in (np.float32, np.float64):
        g = _vcap_solver(z, quadratic, np.dtype(dtype).name)
        g.step()
        o = _vcap_oracle(z, quadratic, dtype=dtype)
        o.step(o.cfl_dt())
        assert np.array_equal(g.q_interior, o.interior), dtype


def _rain_case(z, dtype, bc="wall"):
    from geoswe_amd import Config, Mesh2D, RainfallForcing, Solver2D
    bed = z["bed"]
    rf = RainfallForcing(time_s=z["rain_t"], rate_mm_h=z["rain_mm_h"])
    cfg = Config(bc_x=bc, bc_y=bc, dtype=dtype, h_min=1e-6, cfl=0.5, friction="manning", manning_n=0.03,
                 rainfall_forcing=rf)
    s = Solver2D(Mesh2D(bed.shape[0], bed.shape[1], 2.0, 2.0, ngh=4), cfg, np.zeros((3,) + bed.shape), bed)
    zero = np.zeros_like(bed)
    oc = O.DenseCase(h=zero, hu=zero, hv=zero, bed=bed, dx=2.0, dy=2.0, bc_x=bc, bc_y=bc,
                     manning=np.dtype(dtype).type(0.03), rain=(z["rain_t"], z["rain_mm_h"]))
    return s, oc


def test_rain_on_dry_bed_run_against_reference(gpu):
    z = np.load(GOLD / "rain_dry_run_f64.npz")
    s, _ = _rain_case(z, "float64")
    times = []
    n = s.run(float(z["t_end"]), callback=lambda sv, k: times.append(sv.t))
    assert n == int(z["steps"])
    assert np.abs(np.array(times) - z["ts"]).max() < 1e-11
    scale = np.abs(z["q"]).max(axis=(1, 2), keepdims=True)
    assert (np.abs(s.q_interior - z["q"]) / scale).max() < 1e-7


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_rain_on_dry_bed_run_bitwise_oracle(gpu, dtype):
    z = np.load(GOLD / "rain_dry_run_f64.npz")
    s, oc = _rain_case(z, np.dtype(dtype).name)
    n = s.run(float(z["t_end"]))
    o = O.DenseOracle(oc, O.Physics(h_min=1e-6, h_cfl=1e-6, cfl=0.5, friction=True), dtype)
    assert n == o.run(float(z["t_end"]))
    assert s.t == o.t
    assert np.array_equal(s.q_interior, o.interior)
    assert o.dts[0] == O.rain_step_cap(50.0 / 3.6e6, o.ph, 2.0)     # the bound set the first step


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_c1_dam_break_exact_config(gpu, dtype):
    """256^2, h = 10 m inside r < 40 cells and 1 m outside, walls, no friction, CFL 0.5, dry floor
    1e-6 m, 2000 steps: bit for bit the oracle at steps 1, 10 and 100 (run alongside) and at
    step 2000 (the oracle's state digest, tests/golden/make_c1_oracle_digest.py); float64
    within 1e-12 of the reference's own trajectory at steps 1, 10 and 100."""
    import hashlib
    import json
    from geoswe_amd import Solver2D
    from geoswe_amd.cases import dam_break_c1
    name = np.dtype(dtype).name
    mesh, cfg, q0, b = dam_break_c1(256, name)
    s = Solver2D(mesh, cfg, q0, b)
    case = O.DenseCase(h=q0[0], hu=q0[1], hv=q0[2], bed=b, dx=1.0, dy=1.0, bc_x="wall", bc_y="wall")
    o = O.DenseOracle(case, O.Physics(h_min=1e-6, h_cfl=1e-6, cfl=0.5), dtype)
    gold = np.load(GOLD / "c1_traj_f64.npz")
    done = 0
    for k in (1, 10, 100):
        s.advance(k - done)
        o.run_steps(k - done)
        done = k
        assert s.nsteps == k and s.t == o.t, k
        assert np.array_equal(s.q_interior, o.interior), k
        if dtype == np.float64:
            for c, nm in enumerate(("h", "hu", "hv")):
                assert np.abs(s.q_interior[c] - gold[f"{nm}{k}"]).max() < 1e-12, (k, nm)
    if dtype == np.float64:
        assert np.abs(np.array(o.dts) - gold["dts"]).max() < 1e-15
    s.advance(2000 - done)
    want = json.loads((GOLD / "c1_oracle_2000.json").read_text())[name]
    got = np.ascontiguousarray(s.q_interior)
    assert s.nsteps == want["steps"] == 2000
    assert s.t == want["t"]
    assert hashlib.sha256(got.tobytes()).hexdigest() == want["state_sha256"]


def _film_state(nx=64, ny=56, seed=21):
    """A rough wet state plus thin fast films (3e-5 m carrying 3e-3 m^2/s) whose speed the flat rule's
    velocity divisor max(h, h_min_cfl) = 1e-3 m turns into 3 m/s, above the deep cells' wave
    speed: the flat rule's maximum is set by them, the dense rule skips them, and at
    h_min_cfl = h_min they would move at 100 m/s."""
    rng = np.random.default_rng(seed)
    bed = rng.random((nx, ny)) * 0.6
    h = np.maximum(0.5 - bed, 0.0)
    hu = (rng.random((nx, ny)) - 0.5) * 0.2 * h
    hv = (rng.random((nx, ny)) - 0.5) * 0.2 * h
    films = rng.choice(nx * ny, 25, replace=False)
    h.flat[films] = 3e-5
    hu.flat[films] = 3e-3
    hv.flat[films] = -1e-3
    return np.stack([h, hu, hv]), bed


@pytest.mark.parametrize("l2", [False, True])
def test_flat_cfl_rule_with_h_min_cfl(gpu, l2):
    from geoswe_amd import CompressedSolver, Config, Mesh2D, Solver2D
    q0, bed = _film_state()
    cfg = Config(bc_x="fall", bc_y="fall", dtype="float32", friction="manning", manning_n=0.03)
    s = Solver2D(Mesh2D(bed.shape[0], bed.shape[1], 2.0, 2.0), cfg, q0, bed)
    c = CompressedSolver.from_dense(s, cfl_linf=not l2, h_min_cfl=1e-3)
    c.advance(30)
    case = O.DenseCase(h=q0[0], hu=q0[1], hv=q0[2], bed=bed, dx=2.0, dy=2.0, bc_x="fall", bc_y="fall",
                       manning=np.float32(0.03))

    def oracle(**kw):
        o = O.DenseOracle(case, O.Physics(friction=True, cfl_l2=l2, **kw), np.float32)
        o.run_steps(30)
        return o

    o = oracle(h_min=1e-6, h_cfl=1e-3, flat_cfl=True)
    g = c.mesh.ngh
    got = c.fm.unpack(c.q_flat())[:, g:-g, g:-g]
    assert c.t == o.t and c.nsteps == 30
    assert np.array_equal(got, o.interior)
    # the rule matters here: the dense rule, or the flat one at h_min_cfl = h_min, step differently
    assert oracle(h_min=1e-6, h_cfl=1e-3).t != o.t
    assert oracle(h_min=1e-6, h_cfl=1e-6, flat_cfl=True).t != o.t


def test_flat_rain_bound_from_whole_table(gpu):
    """The flat run bounds dt by the rain film of the table's largest rate (150 mm/h of the
    second segment from the first step on), the dense run by the current segment's (50 mm/h)."""
    from geoswe_amd import CompressedSolver
    z = np.load(GOLD / "rain_dry_run_f64.npz")
    s, oc = _rain_case(z, "float32", bc="fall")
    c = CompressedSolver.from_dense(s, cfl_linf=True)
    n = c.run(float(z["t_end"]))
    o = O.DenseOracle(oc, O.Physics(h_min=1e-6, h_cfl=1e-6, cfl=0.5, friction=True, flat_cfl=True), np.float32)
    assert n == o.run(float(z["t_end"]))
    assert c.t == o.t
    g = c.mesh.ngh
    assert np.array_equal(c.fm.unpack(c.q_flat())[:, g:-g, g:-g], o.interior)
    assert o.dts[0] == O.rain_step_cap(float(np.float32(150.0 / 3.6e6)), o.ph, 2.0)
    assert s.run(float(z["t_end"])) > 0 and s.t == float(z["t_end"])
