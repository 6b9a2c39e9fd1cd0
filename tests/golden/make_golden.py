"""Write golden vectors from the reference's own NumPy path (run once, in the build container).

Imports GeoSWE from /root/reference/src with GEOSWE_BACKEND=numpy and records, as small
.npz fixtures next to this script:

* rhs_srm40_{f64,f32}.npz  -- Solver2D._rhs at the rough, partially dry 40x40 state of the
  reference's tests/test_srm_cpu_matches_gpu.py:44-57 (seed 7, dx=2, wall BCs, ngh=2).
* rhs_bowl64_f64.npz        -- the same at the 64x64 bowl + dam state of
  tests/test_gpu_compressed_equiv.py:37-44 (extrapolate BCs, ngh=4), plus one full step
  with Manning friction and rain.
* ritter_srm_f64.npz        -- dry-bed dam break (800 x 4 strip, dx 0.025, h_min 1e-14) run to
  t = 1 s by Solver2D.run with the production scheme.
* c1_traj_f64.npz           -- the 256x256 circular dam break (walls, no friction), h at
  steps 1, 10 and 100 and the dt sequence, from Solver2D.step(cfl_dt()).
* flatmesh_small.npz        -- CompressedMesh2D flat order and E/W/N/S neighbour table on a
  small irregular mask after the ring-2 dilation of CompressedSWE.
* forcing_samples.npz       -- RainfallForcing.rate_at_time / StageBoundary.stage_at_time.
* step_vcap_f64.npz         -- one full Solver2D.step on a rough 48x40 state with a Manning
  class table and a fifth of the cells above the 15 m/s velocity cap, quadratic and linear
  friction roots, with the residual it starts from.
* rain_dry_run_f64.npz      -- Solver2D.run on a dry sloping bed under a two-segment rain
  series: the rain-film step bound sets the first steps (step times and final state).

c1_oracle_2000.json (the oracle's own C1 state after 2000 steps) is written by
make_c1_oracle_digest.py, which runs the oracle, not the reference.

The reference is not available where the GPU tests run; only these data files are.
Usage: python tests/golden/make_golden.py
"""
import os
import sys
import time
from pathlib import Path

import numpy as np

REF_SRC = Path("/root/reference/src")
OUT = Path(__file__).resolve().parent


def main():
    os.environ["GEOSWE_BACKEND"] = "numpy"
    sys.path.insert(0, str(REF_SRC))
    from geoswe import Config, Mesh2D, Solver2D
    from geoswe.forcing import RainfallForcing, StageBoundary
    from geoswe.compressed_mesh import CompressedMesh2D

    # 1. rough partially-dry 40x40 state (test_srm_cpu_matches_gpu.py:44-57)
    nx = ny = 40
    rng = np.random.default_rng(7)
    bed = rng.random((nx, ny)) * 1.2
    q0 = np.zeros((3, nx, ny))
    q0[0] = np.maximum(0.8 - bed, 0.0)
    wet = q0[0] > 1e-6
    q0[1] = (rng.random((nx, ny)) - 0.5) * 0.04 * wet
    q0[2] = (rng.random((nx, ny)) - 0.5) * 0.04 * wet
    for dt, tag in (("float64", "f64"), ("float32", "f32")):
        mesh = Mesh2D(nx=nx, ny=ny, dx=2.0, dy=2.0, ngh=2)
        cfg = Config(bc_x="wall", bc_y="wall", dtype=dt, h_min=1e-6)
        s = Solver2D(mesh, cfg, q0.astype(dt), bed.astype(dt))
        s._apply_bc()
        t0 = time.perf_counter()
        rhs = s._rhs(s.q)
        el = time.perf_counter() - t0
        np.savez_compressed(OUT / f"rhs_srm40_{tag}.npz", q0=q0.astype(dt), bed=bed.astype(dt),
                            q_pad=s.q, b_pad=s.b, rhs=rhs, ngh=2, dx=2.0, h_min=1e-6, seconds=el)

    # 2. bowl + dam 64x64 (test_gpu_compressed_equiv.py:37-44), residual and one full step
    NX = 64
    ii, jj = np.meshgrid(np.arange(NX), np.arange(NX), indexing="ij")
    coef = 5.0 / ((NX / 2.0) ** 2)
    bedb = coef * ((ii - NX / 2.0) ** 2 + (jj - NX / 2.0) ** 2)
    eta = np.where(ii < NX // 2, 1.2, 0.4)
    hb = np.maximum(eta - bedb, 0.0)
    qb = np.zeros((3, NX, NX))
    qb[0] = hb
    mesh = Mesh2D(nx=NX, ny=NX, dx=1.0, dy=1.0, ngh=4)
    rain = 40.0 / 3.6e6
    cfg = Config(bc_x="extrapolate", bc_y="extrapolate", dtype="float64", h_min=1e-6, cfl=0.4,
                 friction="manning_implicit", manning_n=0.03, rainfall=rain)
    s = Solver2D(mesh, cfg, qb, bedb)
    s._apply_bc()
    rhs = s._rhs(s.q)
    dt = float(s.cfl_dt())
    s.step(dt)
    np.savez_compressed(OUT / "rhs_bowl64_f64.npz", q0=qb, bed=bedb, rhs=rhs, dt=dt, q1=s.q_interior.copy(),
                        ngh=4, dx=1.0, h_min=1e-6, manning=0.03, rain=rain, cfl=0.4)

    # 3. C1 circular dam break, float64 reference NumPy path, first 100 steps
    N = 256
    i = np.arange(N)[:, None]
    j = np.arange(N)[None, :]
    r = np.hypot(i - 127.5, j - 127.5)
    h1 = np.where(r < 40, 10.0, 1.0)
    q1 = np.zeros((3, N, N))
    q1[0] = h1
    mesh = Mesh2D(nx=N, ny=N, dx=1.0, dy=1.0, ngh=4)
    cfg = Config(bc_x="wall", bc_y="wall", dtype="float64", h_min=1e-6, cfl=0.5)
    s = Solver2D(mesh, cfg, q1, np.zeros((N, N)))
    keep = {}
    dts = []
    t0 = time.perf_counter()
    for k in range(1, 101):
        dt = float(s.cfl_dt())
        dts.append(dt)
        s.step(dt)
        if k in (1, 10, 100):
            keep[f"h{k}"] = s.q_interior[0].copy()
            keep[f"hu{k}"] = s.q_interior[1].copy()
            keep[f"hv{k}"] = s.q_interior[2].copy()
    el = time.perf_counter() - t0
    np.savez_compressed(OUT / "c1_traj_f64.npz", dts=np.array(dts), seconds_100_steps=el, **keep)

    # 3b. dry-bed dam break (Ritter) through the production SRM-HLLC scheme, float64
    nx, dx = 800, 0.025
    hR = np.zeros((nx, 4))
    hR[: nx // 2] = 1.0
    qR = np.zeros((3, nx, 4))
    qR[0] = hR
    mesh = Mesh2D(nx=nx, ny=4, dx=dx, dy=dx, ngh=4)
    cfg = Config(bc_x="extrapolate", bc_y="wall", dtype="float64", h_min=1e-14, cfl=0.4)
    s = Solver2D(mesh, cfg, qR, np.zeros((nx, 4)))
    n = s.run(t_end=1.0)
    np.savez_compressed(OUT / "ritter_srm_f64.npz", h=s.q_interior[0].copy(), hu=s.q_interior[1].copy(),
                        steps=n, t=s.t, dx=dx)

    # 4. flat order and neighbours on a small irregular mask (CompressedSWE ring=2)
    from scipy.ndimage import binary_dilation
    rng = np.random.default_rng(3)
    m = np.zeros((20, 18), bool)
    m[4:15, 3:12] = True
    m[8:12, 12:15] = True
    m[rng.random(m.shape) < 0.1] = False
    m[6, 5] = False
    stored = binary_dilation(m, iterations=2)
    mesh = Mesh2D(nx=20, ny=18, dx=1.0, dy=1.0, ngh=4)
    cm = CompressedMesh2D(mesh, stored)
    np.savez_compressed(OUT / "flatmesh_small.npz", inside=m, stored=stored, ij_active=cm._ij_host,
                        neighbors=np.asarray(cm.neighbors), ngh=4)

    # 5. forcing lookups
    rf = RainfallForcing(time_s=np.array([0.0, 300.0, 300.0, 900.0]), rate_mm_h=np.array([10.0, 50.0, 80.0, 0.0]))
    ts = np.array([-5.0, 0.0, 1.0, 299.999, 300.0, 600.0, 900.0, 1e6])
    rates = np.array([rf.rate_at_time(t) for t in ts])
    sb = StageBoundary(cells=np.zeros((1, 2), int), time_s=np.array([0.0, 3600.0, 7200.0]),
                       stage_m=np.array([0.0, 1.0, -0.5]), bed_b=np.zeros(1))
    import warnings
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        ts2 = np.array([-10.0, 0.0, 1800.0, 3600.0, 5000.5, 7200.0, 9000.0])
        stages = np.array([sb.stage_at_time(t) for t in ts2])
    np.savez_compressed(OUT / "forcing_samples.npz", rain_t=rf.time_s, rain_mm_h=rf.rate_mm_h, t=ts, rates=rates,
                        stage_t=sb.time_s, stage_m=sb.stage_m, t2=ts2, stages=stages)

    step_vcap_classes(Config, Mesh2D, Solver2D)
    rain_dry_bed_run(Config, Mesh2D, Solver2D, RainfallForcing)
    print("golden vectors written to", OUT)


def vcap_state():
    """48 x 40 rough state: beds of a few decimetres, depths 2 cm - 82 cm with 8 % dry cells,
    velocities up to 4 m/s and a fifth of the cells at 16-24 m/s (above the 15 m/s cap)."""
    nx, ny = 48, 40
    rng = np.random.default_rng(11)
    i = np.arange(nx)[:, None]
    j = np.arange(ny)[None, :]
    bed = 0.2 * np.sin(i / 6.0) * np.cos(j / 5.0) + 0.05 * rng.random((nx, ny))
    h = 0.02 + 0.8 * rng.random((nx, ny))
    h[rng.random((nx, ny)) < 0.08] = 0.0
    u = rng.uniform(-4.0, 4.0, (nx, ny))
    v = rng.uniform(-4.0, 4.0, (nx, ny))
    fast = rng.random((nx, ny)) < 0.2
    u[fast] = rng.uniform(16.0, 24.0, fast.sum()) * np.sign(rng.standard_normal(fast.sum()))
    cls = rng.integers(0, 5, (nx, ny)).astype(np.uint8)
    tab = np.array([0.01, 0.025, 0.04, 0.07, 0.12])
    return np.stack([h, h * u, h * v]), bed, cls, tab


def step_vcap_classes(Config, Mesh2D, Solver2D):
    """One full Solver2D.step (BC, residual, Euler, point-implicit Manning friction from a
    class table, velocity cap, wet/dry; solver.py:2695-2973, friction 2917-2950) in float64,
    with the quadratic and with the linearised friction root, and the residual it starts from."""
    q0, bed, cls, tab = vcap_state()
    ngh = 4
    out = dict(q0=q0, bed=bed, cls=cls, tab=tab, ngh=ngh, dx=2.0, cfl=0.45, h_min=1e-6, vcap=15.0)
    for quad in (True, False):
        mesh = Mesh2D(nx=q0.shape[1], ny=q0.shape[2], dx=2.0, dy=2.0, ngh=ngh)
        cfg = Config(bc_x="wall", bc_y="extrapolate", dtype="float64", h_min=1e-6, cfl=0.45,
                     friction="manning_implicit", friction_quadratic_alpha=quad)
        s = Solver2D(mesh, cfg, q0.copy(), bed)
        s.set_manning_table(np.pad(cls, ngh, mode="edge"), tab)
        dt = float(s.cfl_dt())
        s._apply_bc()
        rhs = s._rhs(s.q)[:, ngh:-ngh, ngh:-ngh].copy()
        s.step(dt)
        tag = "quad" if quad else "lin"
        out[f"rhs_{tag}"] = rhs
        out[f"dt_{tag}"] = dt
        out[f"q1_{tag}"] = s.q_interior.copy()
        out[f"tab32_{tag}"] = np.asarray(s._manning_tab)
    np.savez_compressed(OUT / "step_vcap_f64.npz", **out)


def rain_dry_bed_run(Config, Mesh2D, Solver2D, RainfallForcing):
    """Solver2D.run (solver.py:2975-3027) on a dry sloping bed under a two-segment rain series,
    float64: the rain-film step bound (cfl*dmin)^(2/3)/(g*R)^(1/3) of the current segment sets
    the first steps, the CFL step the later ones; the time of every step and the final state."""
    nx, ny = 32, 24
    rng = np.random.default_rng(5)
    i = np.arange(nx)[:, None]
    bed = 0.02 * (nx - i) + 0.01 * rng.random((nx, ny))
    q0 = np.zeros((3, nx, ny))
    rf = RainfallForcing(time_s=np.array([0.0, 60.0]), rate_mm_h=np.array([50.0, 150.0]))
    mesh = Mesh2D(nx=nx, ny=ny, dx=2.0, dy=2.0, ngh=4)
    cfg = Config(bc_x="wall", bc_y="wall", dtype="float64", h_min=1e-6, cfl=0.5, friction="manning_implicit",
                 manning_n=0.03, rainfall_forcing=rf)
    s = Solver2D(mesh, cfg, q0, bed)
    ts = []
    n = s.run(t_end=150.0, callback=lambda sv, k: ts.append(float(sv.t)))
    np.savez_compressed(OUT / "rain_dry_run_f64.npz", bed=bed, rain_t=rf.time_s, rain_mm_h=rf.rate_mm_h,
                        t_end=150.0, steps=n, ts=np.array(ts), q=s.q_interior.copy(), dx=2.0, cfl=0.5,
                        h_min=1e-6, manning=0.03)


if __name__ == "__main__":
    main()
