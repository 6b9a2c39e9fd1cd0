"""The compressed coastal workload (fBm DEM, active land + nearshore band, Manning classes
by Voronoi, gridded rain, tidal stage) at test size: dense and flat layouts bit-identical,
and the dense run bit-exact against the oracle."""
import numpy as np
import pytest

from oracle import generators as GN
from oracle import swe_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def case(gpu):
    from geoswe_amd.cases import coastal_c3
    return coastal_c3(n=192, band=8, n_sites=24, rain_frames=2)


def test_voronoi_generator_matches_cpu_mirror(gpu):
    import ctypes as C
    from geoswe_amd import _lib as L
    rng = np.random.default_rng(0)
    sx = rng.uniform(0, 300, 40).astype(np.float32)
    sy = rng.uniform(0, 300, 40).astype(np.float32)
    sc = rng.integers(0, 8, 40).astype(np.uint8)
    d = [L.DeviceArray.from_host(a) for a in (sx, sy, sc)]
    out = L.DeviceArray((57, 91), np.uint8)
    L.call("gs_gen_voronoi", C.c_void_p(out.ptr), 57, 91, 91, 120, 33, C.c_void_p(d[0].ptr), C.c_void_p(d[1].ptr),
           C.c_void_p(d[2].ptr), 40, None)
    assert np.array_equal(out.download(), GN.voronoi_crop(120, 33, 57, 91, sx, sy, sc))


def test_case_shape(case):
    share = case.active.mean()
    assert 0.3 < share < 0.6
    assert case.bed.min() >= -10.0 - 1e-3 and case.bed.max() <= 60.0 + 1e-3
    assert case.sea_edge.any() and np.all(case.active[case.sea_edge])
    assert set(np.unique(case.cls)) <= set(range(8))


def test_dense_flat_bitwise(case):
    d = case.dense()
    f = case.flat()
    act = f.fm.active
    g = case.mesh.ngh
    ij = f.fm.ij[act] - g

    def same_residual():
        rd = d.residual()
        rf = f.residual()
        return np.array_equal(rf[:, act], rd[:, ij[:, 0], ij[:, 1]])

    assert same_residual()                       # t = 0
    for n in (1, 29):
        d.advance(n)
        f.advance(n)
        assert d.nsteps == f.nsteps and d.t == f.t
        assert np.array_equal(f.depth()[ij[:, 0], ij[:, 1]], d.depth()[ij[:, 0], ij[:, 1]])
        assert same_residual()


def test_dense_matches_oracle(case):
    d = case.dense()
    d.advance(20)
    sb = case.cfg.stage_boundary
    rf = case.cfg.rainfall_forcing
    oc = O.DenseCase(h=case.q0[0], hu=case.q0[1], hv=case.q0[2], bed=case.bed, dx=3.0, dy=3.0, bc_x="fall",
                     bc_y="fall", manning=case.ntab.astype(np.float32)[case.cls], rain=(rf.time_s, rf.rate_mm_h),
                     active=case.active,
                     stage={"cells": np.argwhere(case.sea_edge), "time_s": sb.time_s, "stage_m": sb.stage_m})
    o = O.DenseOracle(oc, O.Physics(friction=True), np.float32)
    o.cfl_skip = ~case.active
    o.run_steps(20)
    assert np.array_equal(d.q_interior, o.interior)
    assert d.t == o.t
