"""More Manning classes than the flags byte holds (the reference's uint8 class plane takes
up to 256, solver.py:1524-1556): set_manning_table with a 40-entry table stores n per cell
(a float plane read instead of the 32-entry class table), and the dense and flat steps are
bitwise the oracle run with that per-cell n."""
import numpy as np
import pytest

from oracle import swe_oracle as O

pytestmark = pytest.mark.gpu


def case(nx=60, ny=70, seed=5):
    rng = np.random.default_rng(seed)
    bed = rng.random((nx, ny)) * 0.4
    h = np.maximum(0.6 - bed, 0.0)
    wet = h > 1e-6
    hu = (rng.random((nx, ny)) - 0.5) * 0.3 * wet
    hv = (rng.random((nx, ny)) - 0.5) * 0.3 * wet
    cls = rng.integers(0, 40, size=(nx + 8, ny + 8)).astype(np.uint8)
    table = np.linspace(0.02, 0.10, 40).astype(np.float32)
    return np.stack([h, hu, hv]), bed, cls, table


def solver(q0, bed, cls, table, bc_y="wall"):
    from geoswe_amd import Config, Mesh2D, Solver2D
    nx, ny = bed.shape
    cfg = Config(bc_x="fall", bc_y=bc_y, dtype="float32", friction="manning_implicit")
    s = Solver2D(Mesh2D(nx, ny, 2.0, 2.0, ngh=4), cfg, q0, bed)
    s.set_manning_table(cls, table)
    return s


def oracle(q0, bed, n_cell, flat=False, bc_y="wall"):
    c = O.DenseCase(h=q0[0], hu=q0[1], hv=q0[2], bed=bed, dx=2.0, dy=2.0, bc_x="fall", bc_y=bc_y,
                    manning=n_cell)
    return O.DenseOracle(c, O.Physics(friction=True, flat_cfl=flat), np.float32)


def test_dense_forty_classes_bitwise_oracle(gpu):
    q0, bed, cls, table = case()
    s = solver(q0, bed, cls, table)
    assert s._n_field is not None          # the per-cell plane, not the class table
    n_cell = table[cls[4:-4, 4:-4]]
    s.advance(40)
    o = oracle(q0, bed, n_cell)
    o.run_steps(40)
    assert s.t == o.t and np.array_equal(s.q_interior, o.interior)


def test_flat_forty_classes_bitwise_oracle(gpu):
    # fall edges: the flat ring is frozen at its packed value, which equals the dense
    # ghosts only for free-outflow edges
    from geoswe_amd import CompressedSolver
    q0, bed, cls, table = case(seed=6)
    c = CompressedSolver.from_dense(solver(q0, bed, cls, table, bc_y="fall"), cfl_linf=True)
    c.advance(40)
    o = oracle(q0, bed, table[cls[4:-4, 4:-4]], flat=True, bc_y="fall")
    o.run_steps(40)
    g = c.mesh.ngh
    assert c.t == o.t
    assert np.array_equal(c.fm.unpack(c.q_flat())[:, g:-g, g:-g], o.interior)
