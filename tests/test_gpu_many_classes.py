"""More Manning classes than the flags byte holds (32). The reference keeps up to 256
classes as a uint8 plane (solver.py:1524-1556); so does this build: set_manning_table
with a 40-entry table stores the class ids in a plane of their own (GsDense.cls,
GsFlat.cls) and the dense and flat steps are bitwise the oracle run with n = table[cls]
per cell. Also: an ensemble member with a class plane equals its plain run, the cache
round trip keeps the plane, infiltration rates reach past class 31, and set_manning with
more than 256 distinct values falls back to a per-cell n plane."""
import numpy as np
import pytest

from oracle import swe_oracle as O

pytestmark = pytest.mark.gpu


def case(nx=60, ny=70, seed=5, n_classes=40):
    rng = np.random.default_rng(seed)
    bed = rng.random((nx, ny)) * 0.4
    h = np.maximum(0.6 - bed, 0.0)
    wet = h > 1e-6
    hu = (rng.random((nx, ny)) - 0.5) * 0.3 * wet
    hv = (rng.random((nx, ny)) - 0.5) * 0.3 * wet
    cls = rng.integers(0, n_classes, size=(nx + 8, ny + 8)).astype(np.uint8)
    table = np.linspace(0.02, 0.10, n_classes).astype(np.float32)
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
    assert s._cls is not None and s._n_field is None      # the class plane, not n per cell
    s.advance(40)
    o = oracle(q0, bed, table[cls[4:-4, 4:-4]])
    o.run_steps(40)
    assert s.t == o.t and np.array_equal(s.q_interior, o.interior)


def test_flat_forty_classes_bitwise_oracle(gpu):
    # fall edges: the flat ring is frozen at its packed value, which equals the dense
    # ghosts only for free-outflow edges
    from geoswe_amd import CompressedSolver
    q0, bed, cls, table = case(seed=6)
    c = CompressedSolver.from_dense(solver(q0, bed, cls, table, bc_y="fall"), cfl_linf=True)
    assert c._cls is not None and c._n_field is None
    c.advance(40)
    o = oracle(q0, bed, table[cls[4:-4, 4:-4]], flat=True, bc_y="fall")
    o.run_steps(40)
    g = c.mesh.ngh
    assert c.t == o.t
    assert np.array_equal(c.fm.unpack(c.q_flat())[:, g:-g, g:-g], o.interior)


def test_flat_infiltration_by_class_past_31(gpu):
    """Infiltration rates per class read through the class plane (ids up to 39)."""
    from geoswe_amd import CompressedSolver
    q0, bed, cls, table = case(seed=7)
    rates = np.linspace(0.0, 40.0, 40) / 3.6e6
    c = CompressedSolver.from_dense(solver(q0, bed, cls, table, bc_y="fall"), cfl_linf=True)
    c.set_infil({"tab": rates})
    c.advance(20)
    inner = cls[4:-4, 4:-4].astype(np.int64)
    oc = O.DenseCase(h=q0[0], hu=q0[1], hv=q0[2], bed=bed, dx=2.0, dy=2.0, bc_x="fall", bc_y="fall",
                     manning=table[inner], infil=(inner, rates))
    o = O.DenseOracle(oc, O.Physics(friction=True, flat_cfl=True), np.float32)
    o.run_steps(20)
    g = c.mesh.ngh
    assert c.t == o.t
    assert np.array_equal(c.fm.unpack(c.q_flat())[:, g:-g, g:-g], o.interior)


def test_ensemble_member_with_class_plane(gpu):
    from geoswe_amd import CompressedSolver
    from geoswe_amd.ensemble import Ensemble
    q0, bed, cls, table = case(seed=8)
    q1, bed1, cls1, table1 = case(nx=50, ny=64, seed=9, n_classes=20)
    mk = [lambda: CompressedSolver.from_dense(solver(q0, bed, cls, table, bc_y="fall"), cfl_linf=True),
          lambda: CompressedSolver.from_dense(solver(q1, bed1, cls1, table1, bc_y="fall"), cfl_linf=True)]
    plain = [f() for f in mk]
    for m in plain:
        m.advance(25)
    ens = Ensemble([f() for f in mk], graph_batch=0)
    ens.advance(25)
    for a, b in zip(plain, ens.members):
        assert a.t == b.t and np.array_equal(a.q_flat(), b.q_flat())


def test_cache_round_trip_keeps_class_plane(gpu, tmp_path):
    from geoswe_amd import CompressedSolver
    from geoswe_amd.cache import from_cache, save_cache
    q0, bed, cls, table = case(seed=10)
    a = CompressedSolver.from_dense(solver(q0, bed, cls, table, bc_y="fall"), cfl_linf=True)
    save_cache(a, str(tmp_path))
    b = from_cache(str(tmp_path), cfl_linf=True)
    assert b._cls is not None
    a.advance(15)
    b.advance(15)
    assert a.t == b.t and np.array_equal(a.q_flat(), b.q_flat())


def test_more_than_256_values_use_an_n_plane(gpu):
    q0, bed, _, _ = case(seed=11)
    from geoswe_amd import Config, Mesh2D, Solver2D
    nx, ny = bed.shape
    s = Solver2D(Mesh2D(nx, ny, 2.0, 2.0), Config(bc_x="fall", bc_y="wall", dtype="float32"), q0, bed)
    n = (0.02 + 0.08 * np.random.default_rng(3).random((nx, ny))).astype(np.float32)
    s.set_manning(n)
    assert s._n_field is not None and s._cls is None
    s.advance(20)
    o = oracle(q0, bed, n)
    o.run_steps(20)
    assert s.t == o.t and np.array_equal(s.q_interior, o.interior)
