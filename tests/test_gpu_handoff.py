"""The grid hand-off of the CFL maximum under many workgroups.

Every workgroup folds its wave-speed maximum into the clock with an atomic max and then
counts itself done; the last one plans the next dt from the total. Only atomics cross
workgroups, so the protocol relies on each workgroup's maximum being performed before its
count. These tests put one fast cell (u = 40 m/s against sqrt(g h) = 3.1 m/s elsewhere) in
a different workgroup each round, on grids of thousands of workgroups, and check the dt
the device planned against the oracle's CFL rule on the downloaded state: the dt of the
CFL pass that follows set_state and the dt planned after every step (dense layout), and
the dt planned after every step of the flat layout.
"""
import numpy as np
import pytest

from oracle import swe_oracle as O

pytestmark = pytest.mark.gpu


def spike_state(nx, ny, rng, i, j):
    h = 1.0 + 0.01 * rng.random((nx, ny))
    hu = 0.01 * (rng.random((nx, ny)) - 0.5)
    hv = 0.01 * (rng.random((nx, ny)) - 0.5)
    hu[i, j] = 40.0 * h[i, j]
    return np.stack([h, hu, hv]).astype(np.float32)


def make(q0, dx=2.0):
    from geoswe_amd import Config, Mesh2D, Solver2D
    nx, ny = q0.shape[1:]
    return Solver2D(Mesh2D(nx, ny, dx, dx), Config(bc_x="wall", bc_y="wall", dtype="float32"), q0,
                    np.zeros((nx, ny)))


def test_dense_planned_dt_many_workgroups(gpu):
    # 6000 x 6000: 100 strips across x 94 of 64 rows = 9400 wavefronts, 2350 workgroups
    nx, ny = 6000, 6000
    rng = np.random.default_rng(11)
    s = make(spike_state(nx, ny, rng, 0, 0))
    ph = O.Physics()
    for _ in range(4):
        i, j = int(rng.integers(nx)), int(rng.integers(ny))
        q = spike_state(nx, ny, rng, i, j)
        s.set_state(q)   # the CFL pass of the new state runs with the next step
        assert s.step() == O.dt_from_lam(O.lam_max(q, ph), ph, 2.0), (i, j)
        q1 = s.q_interior
        assert s._clockv().dt == O.dt_from_lam(O.lam_max(q1, ph), ph, 2.0), (i, j)


def test_flat_planned_dt_many_workgroups(gpu):
    # 3000 x 3000 fully active: about 9 400 row-sweep tiles of 16 rows, one wavefront each
    from geoswe_amd import CompressedSolver
    nx, ny = 3000, 3000
    rng = np.random.default_rng(12)
    ph = O.Physics(flat_cfl=True)
    for _ in range(2):
        i, j = int(rng.integers(nx)), int(rng.integers(ny))
        c = CompressedSolver.from_dense(make(spike_state(nx, ny, rng, i, j)), cfl_linf=True)
        g = c.mesh.ngh
        for _ in range(3):
            c.step()
            q1 = c.fm.unpack(c.q_flat())[:, g:g + nx, g:g + ny]
            assert c._clockv().dt == O.dt_from_lam(O.lam_max(q1, ph), ph, 2.0), (i, j)
