"""float32 planes stepped with float64 arithmetic (Config(dtype="float32",
compute_dtype="float64"), GS_F32X64): bit for bit the float64 oracle whose state is rounded
to float32 after every step, on a crop of the 640 M-cell benchmark case and on a walled
dam break with Manning classes; closer to the float64 run than the float32 arithmetic."""
import numpy as np
import pytest

from oracle import generators as GN
from oracle import swe_oracle as O

pytestmark = pytest.mark.gpu


def _crop_case(nx, ny, b):
    bed = GN.bed_crop(0, 0, nx, ny, b.seed, b.octaves, b.L0, b.hurst, b.mu, b.scale, b.slope_x, b.dx, b.offset,
                      np.float32)
    h = np.full((nx, ny), 0.1, np.float32)
    return O.DenseCase(h=h, hu=0 * h, hv=0 * h, bed=bed, dx=3.0, dy=3.0, bc_x="fall", bc_y="fall",
                       manning=np.float32(0.03), rain=50.0 / 3.6e6)


def test_mixed_rain_on_slope_crop_bitexact(gpu):
    from geoswe_amd.cases import FbmBed, rain_on_slope_c2
    b = FbmBed(seed=20261202).normalise(25600, 25000)
    nx, ny = 300, 280
    s = rain_on_slope_c2(nx, ny, "float32", bed=b, track_max_depth=True, compute_dtype="float64")
    assert s.compute64 and s.dtype == np.float32
    s.advance(25)
    o = O.DenseOracle(_crop_case(nx, ny, b), O.Physics(friction=True), np.float64, store_dtype=np.float32)
    o.run_steps(25)
    got = s.q_interior
    assert got.dtype == np.float32
    assert np.array_equal(got.astype(np.float64), o.interior)
    assert s.t == o.t
    assert np.array_equal(s.max_depth() >= got[0], np.ones_like(got[0], bool))


def test_mixed_dam_break_with_classes_and_walls(gpu):
    from geoswe_amd import Config, Mesh2D, Solver2D
    nx, ny = 96, 80
    i = np.arange(nx)[:, None]
    j = np.arange(ny)[None, :]
    bed = (0.02 * np.sin(i / 7.0) * np.cos(j / 5.0)).astype(np.float32)
    h = np.where((i - 48) ** 2 + (j - 40) ** 2 < 15 ** 2, 2.0, 0.4).astype(np.float32) - bed
    q0 = np.stack([h, 0 * h, 0 * h]).astype(np.float32)
    cfg = Config(bc_x="wall", bc_y="wall", dtype="float32", compute_dtype="float64", friction="manning")
    s = Solver2D(Mesh2D(nx, ny, 1.0, 1.0), cfg, q0, bed)
    tab = np.array([0.02, 0.05, 0.11], np.float32)
    cls = ((i + 2 * j) % 3).astype(np.uint8) * np.ones((nx, ny), np.uint8)
    s.set_manning_table(np.pad(cls, 4, mode="edge"), tab)
    s.advance(60)
    case = O.DenseCase(h=q0[0], hu=q0[1], hv=q0[2], bed=bed, dx=1.0, dy=1.0, bc_x="wall", bc_y="wall",
                       manning=tab[cls])
    o = O.DenseOracle(case, O.Physics(friction=True), np.float64, store_dtype=np.float32)
    o.run_steps(60)
    assert np.array_equal(s.q_interior.astype(np.float64), o.interior)
    assert s.t == o.t and s.nsteps == 60


def test_mixed_closer_to_float64_than_float32(gpu):
    from geoswe_amd.cases import FbmBed, rain_on_slope_c2
    b = FbmBed(seed=20261202).normalise(25600, 25000)
    nx, ny = 512, 500
    out = {}
    for name, kw in (("f32", dict(dtype="float32")), ("mixed", dict(dtype="float32", compute_dtype="float64")),
                     ("f64", dict(dtype="float64", f32_inputs=True))):
        s = rain_on_slope_c2(nx, ny, bed=b, **kw)
        s.advance(200)
        out[name] = s.depth().astype(np.float64)
    e32 = np.abs(out["f32"] - out["f64"])
    emx = np.abs(out["mixed"] - out["f64"])
    assert np.percentile(emx, 99.9) < np.percentile(e32, 99.9)
    assert np.median(emx) < np.median(e32)


def test_mixed_refused_by_the_flat_layout(gpu):
    from geoswe_amd import CompressedSolver
    from geoswe_amd.cases import FbmBed, rain_on_slope_c2
    s = rain_on_slope_c2(64, 64, "float32", bed=FbmBed(seed=20261202).normalise(25600, 25000),
                         compute_dtype="float64")
    with pytest.raises(NotImplementedError, match="dense-layout"):
        CompressedSolver.from_dense(s)
