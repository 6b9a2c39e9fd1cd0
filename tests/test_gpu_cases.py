"""Synthetic workload generators on the device against their CPU mirror, and a crop of the
rain-on-slope benchmark case stepped on the GPU against the oracle."""
import numpy as np
import pytest

from oracle import generators as GN
from oracle import swe_oracle as O

pytestmark = pytest.mark.gpu


def spec():
    from geoswe_amd.cases import FbmBed
    return FbmBed(seed=20261202, mu=0.0123, scale=0.37)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("origin", [(0, 0), (12345, 20000), (25000, 24900)])
def test_bed_generator_matches_cpu_mirror(gpu, dtype, origin):
    b = spec()
    got = b.host_crop(origin[0], origin[1], 97, 131, dtype)
    want = GN.bed_crop(origin[0], origin[1], 97, 131, b.seed, b.octaves, b.L0, b.hurst, b.mu, b.scale, b.slope_x,
                       b.dx, b.offset, dtype)
    assert np.array_equal(got, want)


def test_bed_generator_float32_rounded_inputs(gpu):
    """round_f32: the float64 field holds exactly the float32 field's values."""
    from geoswe_amd import _lib as L
    b = spec()
    d64 = L.DeviceArray((40, 50), np.float64)
    b.write(d64.ptr, 50, L.GS_F64, 40, 50, 100, 200, round_f32=True)
    assert np.array_equal(d64.download(), b.host_crop(100, 200, 40, 50, np.float32).astype(np.float64))


def test_normalisation_bounds_amplitude(gpu):
    from geoswe_amd.cases import FbmBed
    b = FbmBed(seed=5).normalise(2048, 2000)
    crop = b.host_crop(0, 0, 512, 500, np.float64)
    noise = crop - b.slope_x * ((np.arange(512)[:, None] + 0.5) * b.dx)
    assert np.abs(noise).max() < 0.6 and noise.std() > 0.02


def test_rain_on_slope_crop_bitexact(gpu):
    """A 300x280 crop of the benchmark case on the GPU equals the float32 oracle bit for bit."""
    from geoswe_amd.cases import FbmBed, rain_on_slope_c2
    b = FbmBed(seed=20261202).normalise(25600, 25000)
    nx, ny = 300, 280
    s = rain_on_slope_c2(nx, ny, "float32", bed=b, track_max_depth=True)
    s.advance(25)
    bed = GN.bed_crop(0, 0, nx, ny, b.seed, b.octaves, b.L0, b.hurst, b.mu, b.scale, b.slope_x, b.dx, b.offset,
                      np.float32)
    h = np.full((nx, ny), 0.1)
    case = O.DenseCase(h=h, hu=0 * h, hv=0 * h, bed=bed, dx=3.0, dy=3.0, bc_x="fall", bc_y="fall",
                       manning=np.float32(0.03), rain=50.0 / 3.6e6)
    o = O.DenseOracle(case, O.Physics(friction=True), np.float32)
    o.run_steps(25)
    assert np.array_equal(s.q_interior, o.interior)
    assert s.t == o.t


def test_rain_on_slope_crop_f64_thresholds(gpu):
    """float64 at its own wet/dry threshold (h_min 1e-10): the 'fall' edge faces whose
    reconstructed ghost depth lies between 1e-10 and 1e-6 take the wet HLLC branch, as the
    oracle does at the same threshold; the float32-parameter float64 run (f32_inputs) stays
    within 1e-5 m of the float32 run."""
    from geoswe_amd.cases import FbmBed, rain_on_slope_c2
    b = FbmBed(seed=20261202).normalise(25600, 25000)
    nx, ny = 300, 280
    bed = GN.bed_crop(0, 0, nx, ny, b.seed, b.octaves, b.L0, b.hurst, b.mu, b.scale, b.slope_x, b.dx, b.offset,
                      np.float64, round_f32=False)
    # the crop holds edge faces in the band between the two thresholds
    T = np.float64
    iy = T(1.0 / 3.0)
    gz = (T(0.5) * (bed[:, 1] - bed[:, 0])) * iy
    hlf = O.srm_face(bed[:, 0], bed[:, 0], 0 * gz, gz, 0 * gz, np.full(nx, 0.1), T(0.5) * (T(1) / iy), T(9.81),
                     T(1e-10))[0]
    assert ((hlf > 1e-10) & (hlf < 1e-6)).sum() >= 2
    s = rain_on_slope_c2(nx, ny, "float64", bed=b)
    s.advance(12)
    h = np.full((nx, ny), 0.1)
    case = O.DenseCase(h=h, hu=0 * h, hv=0 * h, bed=bed, dx=3.0, dy=3.0, bc_x="fall", bc_y="fall",
                       manning=0.03, rain=50.0 / 3.6e6)
    o = O.DenseOracle(case, O.Physics(friction=True, h_min=1e-10, h_cfl=1e-10), np.float64)
    o.run_steps(12)
    assert np.array_equal(s.q_interior, o.interior), np.abs(s.q_interior - o.interior).max()
    s32 = rain_on_slope_c2(nx, ny, "float32", bed=b)
    s64 = rain_on_slope_c2(nx, ny, "float64", bed=b, f32_inputs=True)
    s32.advance(12)
    s64.advance(12)
    assert np.abs(s32.depth() - s64.depth()).max() < 1e-5
