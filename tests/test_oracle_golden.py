"""The CPU oracle against golden vectors written by the reference's own NumPy path.

Fixtures come from tests/golden/make_golden.py (reference GeoSWE, GEOSWE_BACKEND=numpy).
Where the oracle follows the production kernel rather than the NumPy path (SM clamp,
face velocity, both-dry faces, float64 source accumulator; see the oracle's header),
differences are bounded by round-off and the tolerance is stated per comparison.
"""
from pathlib import Path

import numpy as np
import pytest

from oracle import swe_oracle as O

G = Path(__file__).resolve().parent / "golden"
W = (slice(None), slice(2, -2), slice(2, -2))     # the kernel window


def test_residual_rough_partially_dry_f64():
    """test_srm_cpu_matches_gpu.py:44-59 state: the reference bounds CPU vs kernel by 1e-12."""
    z = np.load(G / "rhs_srm40_f64.npz")
    r = O.residual(z["q_pad"], z["b_pad"], 2.0, 2.0, 9.81, 1e-6)
    d = np.abs(r[W] - z["rhs"][W]).max()
    assert d < 1e-12, d
    assert np.abs(z["rhs"][W]).max() > 1e-2        # the state is not trivially at rest


def test_residual_rough_partially_dry_f32():
    """float32: at least as accurate as the reference's own float32 NumPy residual, both
    measured against the reference's float64 residual of the same state."""
    z = np.load(G / "rhs_srm40_f32.npz")
    z64 = np.load(G / "rhs_srm40_f64.npz")
    r = O.residual(z["q_pad"], z["b_pad"], 2.0, 2.0, 9.81, 1e-6)
    assert r.dtype == np.float32
    err_mine = np.abs(r[W].astype(np.float64) - z64["rhs"][W]).max(axis=(1, 2))
    err_ref = np.abs(z["rhs"][W] - z64["rhs"][W]).max(axis=(1, 2))
    assert np.all(err_mine <= 1.05 * err_ref), (err_mine, err_ref)
    scale = np.abs(z64["rhs"][W]).max()
    assert np.all(err_mine < 1e-5 * scale)
    # and within float32 round-off of the reference's float32 path itself
    assert np.abs(r[W].astype(np.float64) - z["rhs"][W]).max() < 2e-5 * scale


def test_bowl_residual_and_step_f64():
    """Bowl + dam (test_gpu_compressed_equiv.py:37-44) with Manning and rain, one full step."""
    z = np.load(G / "rhs_bowl64_f64.npz")
    q0, bed, rain = z["q0"], z["bed"], float(z["rain"])
    qp = np.zeros((3, 72, 72))
    qp[:, 4:-4, 4:-4] = q0
    O.bc_fill(qp, 4, "extrapolate", "extrapolate")
    bp = O.pad_bed(bed, 4, "extrapolate", "extrapolate", np.float64)
    r = O.residual(qp, bp, 1.0, 1.0, 9.81, 1e-6)
    ref = z["rhs"].copy()
    ref[0, 4:-4, 4:-4] -= rain                      # the reference's _rhs includes the rain
    assert np.abs(r[W] - ref[W]).max() < 1e-12
    case = O.DenseCase(h=q0[0], hu=q0[1], hv=q0[2], bed=bed, dx=1.0, dy=1.0, bc_x="extrapolate",
                       bc_y="extrapolate", manning=0.03, rain=rain)
    ph = O.Physics(h_min=1e-6, h_cfl=1e-6, cfl=0.4, friction=True)
    o = O.DenseOracle(case, ph, np.float64)
    assert o.cfl_dt() == float(z["dt"])
    o.step(float(z["dt"]))
    # friction: the NumPy path's alpha = 2/(1+sqrt(1+4x)) vs the kernel's 2/(sqrt(1+2(2x))+1)
    assert np.abs(o.interior - z["q1"]).max() < 1e-13


def test_c1_dam_break_trajectory_f64():
    """256^2 circular dam break (walls): h after 1, 10, 100 steps and every dt."""
    z = np.load(G / "c1_traj_f64.npz")
    N = 256
    i = np.arange(N)[:, None]
    j = np.arange(N)[None, :]
    h = np.where(np.hypot(i - 127.5, j - 127.5) < 40, 10.0, 1.0)
    case = O.DenseCase(h=h, hu=0 * h, hv=0 * h, bed=0 * h, dx=1.0, dy=1.0, bc_x="wall", bc_y="wall")
    o = O.DenseOracle(case, O.Physics(), np.float64)
    dts = []
    for k in range(1, 101):
        dt = o.next_dt()
        dts.append(dt)
        o.step(dt)
        if k in (1, 10, 100):
            for c, name in enumerate(("h", "hu", "hv")):
                assert np.abs(o.interior[c] - z[f"{name}{k}"]).max() < 1e-12, (k, name)
    assert np.abs(np.array(dts) - z["dts"]).max() < 1e-15


def test_forcing_lookups():
    from geoswe_amd.forcing import RainfallForcing, StageBoundary
    z = np.load(G / "forcing_samples.npz")
    rf = RainfallForcing(time_s=z["rain_t"], rate_mm_h=z["rain_mm_h"])
    assert np.array_equal(np.array([rf.rate_at_time(t) for t in z["t"]]), z["rates"])
    got = np.array([float(z["rain_mm_h"][O.rain_index(z["rain_t"], t)]) / 3.6e6 for t in z["t"]])
    assert np.array_equal(got, z["rates"])
    with pytest.warns(RuntimeWarning):
        sb = StageBoundary(cells=np.zeros((1, 2), int), time_s=z["stage_t"], stage_m=z["stage_m"], bed_b=np.zeros(1))
        s1 = np.array([sb.stage_at_time(t) for t in z["t2"]])
    assert np.array_equal(s1, z["stages"])
    s2 = np.array([O.stage_at_time(z["stage_t"], z["stage_m"], t) for t in z["t2"]])
    assert np.array_equal(s2, z["stages"])


def test_flat_mesh_order_and_neighbours():
    """Flat order and E/W/N/S neighbours equal the reference CompressedMesh2D's."""
    from geoswe_amd.flatmesh import build_flat_mesh
    z = np.load(G / "flatmesh_small.npz")
    g = int(z["ngh"])
    inside = np.zeros((z["inside"].shape[0] + 2 * g, z["inside"].shape[1] + 2 * g), bool)
    inside[g:-g, g:-g] = z["inside"]
    fm = build_flat_mesh(inside, g, ring=2)
    assert np.array_equal(fm.ij, z["ij_active"])
    nb = z["neighbors"]
    act = np.nonzero(fm.active)[0]
    # every active cell's E/W neighbours from the run table, N/S are k+-1
    for r in range(len(fm.run_k0)):
        for t in range(fm.run_len[r]):
            k = fm.run_k0[r] + t
            assert nb[k, 0] == k + fm.run_d[r, 0] and nb[k, 1] == k + fm.run_d[r, 1]
            assert nb[k, 2] == k + 1 and nb[k, 3] == k - 1
            assert nb[nb[k, 0], 0] == k + fm.run_d[r, 2] and nb[nb[k, 1], 1] == k + fm.run_d[r, 3]
    assert sum(int(x) for x in fm.run_len) == act.size == int(z["inside"].sum())


def test_cache_neighbour_table_is_the_reference_one():
    """The cache's nbr.npy (int16 deltas, -32768 for none) over every stored cell equals the
    reference's CompressedMesh2D neighbour table of the golden mask, delta-encoded as
    nbr_to_int16_delta does (compressed_rhs.py:194-235)."""
    from geoswe_amd.cache import _nbr_int16
    from geoswe_amd.flatmesh import build_flat_mesh
    z = np.load(G / "flatmesh_small.npz")
    g = int(z["ngh"])
    inside = np.zeros((z["inside"].shape[0] + 2 * g, z["inside"].shape[1] + 2 * g), bool)
    inside[g:-g, g:-g] = z["inside"]
    fm = build_flat_mesh(inside, g, ring=2)
    nb = z["neighbors"].astype(np.int64)
    k = np.arange(fm.n_stored, dtype=np.int64)[:, None]
    want = np.where(nb >= 0, nb - k, -32768).astype(np.int16)
    got = _nbr_int16(fm)
    assert got.dtype == np.int16 and np.array_equal(got, want)


def _vcap_oracle(z, quadratic, vcap=15.0, dtype=np.float64):
    q0, bed, cls = z["q0"], z["bed"], z["cls"]
    n = z["tab32_quad"][cls]                        # the reference casts the class table to float32
    case = O.DenseCase(h=q0[0], hu=q0[1], hv=q0[2], bed=bed, dx=2.0, dy=2.0, bc_x="wall", bc_y="extrapolate",
                       manning=n)
    ph = O.Physics(h_min=1e-6, h_cfl=1e-6, cfl=0.45, friction=True, vcap=vcap, quadratic=quadratic)
    return O.DenseOracle(case, ph, dtype)


def _face_velocity_divergence(q0, bed):
    """Cells touching a face where a side's reconstructed depth is <= h_min while its cell is
    wet: there the reference's NumPy flux gates the velocity on the face depth and its kernel
    (which the oracle follows) on the cell depth (flux.py:98-101 vs rhs_cuda.py:931-938)."""
    g, hm = 4, 1e-6
    qp = np.zeros((3, q0.shape[1] + 2 * g, q0.shape[2] + 2 * g))
    qp[:, g:-g, g:-g] = q0
    O.bc_fill(qp, g, "wall", "extrapolate")
    b = O.pad_bed(bed, g, "wall", "extrapolate", np.float64)
    h = qp[0]
    gzx = np.zeros_like(b)
    gzy = np.zeros_like(b)
    gzx[1:-1] = (0.5 * (b[2:] - b[:-2])) * 0.5
    gzy[:, 1:-1] = (0.5 * (b[:, 2:] - b[:, :-2])) * 0.5
    bad = np.zeros(h.shape, bool)
    xl, xr, *_ = O.srm_face(b[:-1], b[1:], gzx[:-1], gzx[1:], h[:-1], h[1:], 1.0, 9.81, hm)
    fx = ((xl <= hm) & (h[:-1] > hm)) | ((xr <= hm) & (h[1:] > hm))
    bad[:-1] |= fx
    bad[1:] |= fx
    yl, yr, *_ = O.srm_face(b[:, :-1], b[:, 1:], gzy[:, :-1], gzy[:, 1:], h[:, :-1], h[:, 1:], 1.0, 9.81, hm)
    fy = ((yl <= hm) & (h[:, :-1] > hm)) | ((yr <= hm) & (h[:, 1:] > hm))
    bad[:, :-1] |= fy
    bad[:, 1:] |= fy
    return bad[g:-g, g:-g]


@pytest.mark.parametrize("quadratic", [True, False])
def test_update_with_class_table_and_velocity_cap_f64(quadratic):
    """The update tail of one reference Solver2D.step (solver.py:2917-2962: Euler, point-implicit
    Manning friction from a class table, the 15 m/s velocity cap, wet/dry) applied to the
    reference's own residual: the oracle's ``update`` equals it to float64 round-off, with the
    quadratic and with the linear root. A fifth of the cells are above the cap. The fixture
    discriminates: without the cap, or with the other root, the update misses it by orders of
    magnitude more than the tolerance."""
    z = np.load(G / "step_vcap_f64.npz")
    tag = "quad" if quadratic else "lin"
    n = z["tab32_quad"][z["cls"]]
    dt = float(z[f"dt_{tag}"])
    want = z[f"q1_{tag}"]
    scale = np.abs(want).max(axis=(1, 2), keepdims=True)

    def upd(vcap, quad):
        ph = O.Physics(h_min=1e-6, h_cfl=1e-6, friction=True, vcap=vcap, quadratic=quad)
        return O.update(z["q0"], z[f"rhs_{tag}"], dt, ph, n_cell=n)

    err = np.abs(upd(15.0, quadratic) - want) / scale
    assert err.max() < 1e-14, err.max()
    h, hu, hv = z["q0"]
    assert (np.hypot(hu, hv) > 15.0 * np.maximum(h, 1e-6)).sum() > 200
    for other in (upd(np.inf, quadratic), upd(15.0, not quadratic)):
        assert (np.abs(other - want) / scale).max() > 1e-6


@pytest.mark.parametrize("quadratic", [True, False])
def test_step_with_class_table_and_velocity_cap_f64(quadratic):
    """The whole step (BC fill, residual, update) against the reference's: the dt exactly, the
    state to float64 round-off on every cell except those next to a face where the NumPy flux
    and the kernel gate the velocity differently (a documented divergence of the reference's
    own two paths; 4 of 1920 cells here)."""
    z = np.load(G / "step_vcap_f64.npz")
    tag = "quad" if quadratic else "lin"
    o = _vcap_oracle(z, quadratic)
    dt = o.cfl_dt()
    assert abs(dt - float(z[f"dt_{tag}"])) <= 1e-15 * dt
    o.step(dt)
    want = z[f"q1_{tag}"]
    scale = np.abs(want).max(axis=(1, 2), keepdims=True)
    err = (np.abs(o.interior - want) / scale).max(axis=0)
    skip = _face_velocity_divergence(z["q0"], z["bed"])
    assert skip.sum() <= 8
    assert err[~skip].max() < 1e-13, err[~skip].max()


def test_step_with_class_table_and_velocity_cap_f32_close():
    """The float32 oracle step stays within float32 round-off of the reference's float64 step
    (away from the face-velocity divergence)."""
    z = np.load(G / "step_vcap_f64.npz")
    o = _vcap_oracle(z, True, dtype=np.float32)
    o.step(float(z["dt_quad"]))
    want = z["q1_quad"]
    scale = np.abs(want).max(axis=(1, 2), keepdims=True)
    err = (np.abs(o.interior.astype(np.float64) - want) / scale).max(axis=0)
    assert err[~_face_velocity_divergence(z["q0"], z["bed"])].max() < 2e-6


def test_rain_on_dry_bed_run_f64():
    """Solver2D.run on a dry bed under a two-segment rain (solver.py:2975-3027): the reference's
    step times, step count and final state. The first two steps sit at the rain-film bound
    (cfl*dmin)^(2/3)/(g*R)^(1/3) of the first segment, the later ones at the CFL step."""
    z = np.load(G / "rain_dry_run_f64.npz")
    bed = z["bed"]
    zero = np.zeros_like(bed)
    case = O.DenseCase(h=zero, hu=zero, hv=zero, bed=bed, dx=2.0, dy=2.0, bc_x="wall", bc_y="wall",
                       manning=0.03, rain=(z["rain_t"], z["rain_mm_h"]))
    ph = O.Physics(h_min=1e-6, h_cfl=1e-6, cfl=0.5, friction=True)
    o = O.DenseOracle(case, ph, np.float64)
    n = o.run(float(z["t_end"]))
    assert n == int(z["steps"])
    ts = np.cumsum(o.dts)
    assert np.abs(ts - z["ts"]).max() < 1e-11
    cap = O.rain_step_cap(50.0 / 3.6e6, ph, 2.0)
    assert o.dts[0] == cap and o.dts[1] <= cap and o.dts[2] < 0.6 * cap
    # thin films (1e-4 - 2e-2 m) crossing the wet/dry threshold: the reference's NumPy flux and its
    # kernel (which the oracle follows) differ below round-off at such faces (velocity gate,
    # both-dry faces), which the run amplifies to ~2e-8 of each field's scale in 36 steps
    scale = np.abs(z["q"]).max(axis=(1, 2), keepdims=True)
    assert (np.abs(o.interior - z["q"]) / scale).max() < 1e-7
