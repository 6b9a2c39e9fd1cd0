"""Invariants that pin the CPU oracle independently of the golden vectors.

Each check is chosen so that a plausible slip (a dropped term, a wrong sign or index, a
swapped operand) fails it: physical-flux consistency and mirror symmetry of HLLC, the
Ritter dry dam-break front speed, lake at rest over a bump (well balancing), the friction
root of its quadratic, exact rain volume in a closed basin, the CFL closed form, ghost
fills built by hand, the h^(-4/3) accuracy, Philox2x32-10 known-answer vectors.
"""
import math

import numpy as np
import pytest

from oracle import swe_oracle as O
from oracle import generators as GN

G = 9.81


def phys_flux(h, u, v, g=G):
    return h * u, h * u * u + 0.5 * g * h * h, h * u * v


@pytest.mark.parametrize("dtype", [np.float64])
def test_hllc_consistency_on_equal_states(dtype):
    rng = np.random.default_rng(0)
    h = rng.uniform(0.1, 5.0, 200).astype(dtype)
    u = rng.uniform(-3, 3, 200).astype(dtype)
    v = rng.uniform(-3, 3, 200).astype(dtype)
    Fh, FN, FT = O.hllc(h, u, v, h, u, v, dtype(G), dtype(1e-10))
    for got, want in zip((Fh, FN, FT), phys_flux(h, u, v)):
        assert np.allclose(got, want, rtol=1e-12, atol=1e-12)


def test_hllc_mirror_antisymmetry():
    """F(qR*, qL*) with mirrored velocities = -F(qL, qR) for mass and transverse momentum."""
    rng = np.random.default_rng(1)
    n = 500
    hL, hR = rng.uniform(0.01, 3, n), rng.uniform(0.01, 3, n)
    uL, uR = rng.uniform(-2, 2, n), rng.uniform(-2, 2, n)
    vL, vR = rng.uniform(-2, 2, n), rng.uniform(-2, 2, n)
    g, hm = np.float64(G), np.float64(1e-10)
    Fh, FN, FT = O.hllc(hL, uL, vL, hR, uR, vR, g, hm)
    Mh, MN, MT = O.hllc(hR, -uR, vR, hL, -uL, vL, g, hm)
    assert np.allclose(Fh, -Mh, atol=1e-12)
    assert np.allclose(FN, MN, atol=1e-12)          # normal momentum flux is even under mirroring
    assert np.allclose(FT, -MT, atol=1e-12)


def test_hllc_both_dry_gives_zero_flux_and_rest_gives_pressure():
    g, hm = np.float64(G), np.float64(1e-6)
    z = np.zeros(1)
    assert np.allclose(O.hllc(z, z, z, z, z, z, g, hm), 0.0)
    h = np.array([2.0])
    Fh, FN, FT = O.hllc(h, z, z, h, z, z, g, hm)
    assert Fh[0] == 0.0 and FT[0] == 0.0 and FN[0] == pytest.approx(0.5 * G * 4.0, rel=1e-15)


def test_ritter_dry_dam_break_front():
    """Dry-bed dam break: the run equals the reference's production-scheme run (golden, the
    same first-order SRM-HLLC), the front is within 8 % of Ritter's x0 + 2 sqrt(g h0) t
    (the first-order scheme lags ~6 % at every resolution; the reference's own Ritter test,
    test_hllc_dry_front.py:10-58, uses MUSCL + SSPRK3 instead), mass is conserved and
    depths stay in [0, h0]."""
    from pathlib import Path
    z = np.load(Path(__file__).resolve().parent / "golden" / "ritter_srm_f64.npz")
    nx, ny, dx = 800, 4, 0.025
    h = np.zeros((nx, ny))
    h[: nx // 2] = 1.0
    case = O.DenseCase(h=h, hu=0 * h, hv=0 * h, bed=0 * h, dx=dx, dy=dx, bc_x="extrapolate", bc_y="wall")
    o = O.DenseOracle(case, O.Physics(h_min=1e-14, h_cfl=1e-14, cfl=0.4), np.float64)
    mass0 = o.interior[0].sum()
    n = o.run(1.0)
    assert n == int(z["steps"]) and o.t == float(z["t"])
    # the NumPy path's SM clamp and face-velocity rebuild differ from the kernel form at the
    # front tip only (3e-11 measured)
    assert np.abs(o.interior[0] - z["h"]).max() < 1e-9 and np.abs(o.interior[1] - z["hu"]).max() < 1e-9
    hf = o.interior[0][:, 1]
    x = (np.arange(nx) + 0.5) * dx
    d_num = x[hf > 1e-8][-1] - (nx // 2) * dx
    d_exact = 2.0 * math.sqrt(G * 1.0) * 1.0
    assert abs(d_num - d_exact) <= 0.08 * d_exact, (d_num, d_exact)
    assert abs(o.interior[0].sum() - mass0) < 1e-12 * mass0
    assert o.interior[0].min() >= 0.0 and o.interior[0].max() <= 1.0 + 1e-6


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-10), (np.float32, 2e-5)])
def test_lake_at_rest_over_bump(dtype, tol):
    """test_well_balanced.py:6-29: still water over a Gaussian bump stays still (30 steps)."""
    nx = ny = 40
    x = np.arange(nx)[:, None]
    y = np.arange(ny)[None, :]
    bed = 0.8 * np.exp(-((x - 20) ** 2 + (y - 20) ** 2) / 30.0) + 20.0
    h = np.maximum(21.0 - bed, 0.0)
    case = O.DenseCase(h=h, hu=0 * h, hv=0 * h, bed=bed, dx=1.0, dy=1.0, bc_x="wall", bc_y="wall")
    hm = 1e-10 if dtype == np.float64 else 1e-6
    o = O.DenseOracle(case, O.Physics(h_min=hm, h_cfl=hm), dtype)
    o.run_steps(30)
    q = o.interior
    assert np.abs(q[1]).max() < tol and np.abs(q[2]).max() < tol
    assert np.abs((q[0] + bed.astype(dtype)) - 21.0)[q[0] > 0].max() < tol


def test_wet_dry_shoreline_at_rest():
    """Lake at rest with a dry island: nothing moves (SRM with dry faces)."""
    nx = ny = 30
    x = np.arange(nx)[:, None]
    bed = np.broadcast_to(np.where(x < 15, 0.0, 2.0), (nx, ny)).astype(float)
    h = np.maximum(1.0 - bed, 0.0)
    case = O.DenseCase(h=h, hu=0 * h, hv=0 * h, bed=bed, dx=1.0, dy=1.0, bc_x="wall", bc_y="wall")
    o = O.DenseOracle(case, O.Physics(h_min=1e-10, h_cfl=1e-10), np.float64)
    o.run_steps(20)
    assert np.abs(o.interior[1]).max() < 1e-12 and np.abs(o.interior[0] - h).max() < 1e-12


def test_friction_alpha_is_root_of_quadratic():
    """alpha solves x*alpha^2 + alpha - 1 = 0 with x = dt*Cf*|U| (|U_new| = |U| - dt*Cf*|U_new|^2,
    solver.py:295-300), for a uniform state."""
    h, hu, n, dt = 0.3, 0.6, 0.05, 0.5
    q = np.array([[[h]], [[hu]], [[0.0]]])
    r = np.zeros_like(q)
    out = O.update(q, r, dt, O.Physics(friction=True, vcap=math.inf), n_cell=n)
    alpha = out[1, 0, 0] / hu
    cf = G * n * n * h ** (-4.0 / 3.0)
    U = hu / h
    assert abs(dt * cf * U * alpha * alpha + alpha - 1.0) < 1e-13
    lin = O.update(q, r, dt, O.Physics(friction=True, quadratic=False, vcap=math.inf), n_cell=n)
    assert lin[1, 0, 0] / hu == pytest.approx(1.0 / (1.0 + dt * cf * U), rel=1e-14)


def test_dry_cells_get_zero_momentum():
    q = np.array([[[1e-8, 0.5]], [[0.3, 0.2]], [[-0.1, 0.1]]])
    out = O.update(q, np.zeros_like(q), 0.1, O.Physics(h_min=1e-6, friction=True), n_cell=0.03)
    assert out[1, 0, 0] == 0.0 and out[2, 0, 0] == 0.0 and out[0, 0, 0] == 1e-8
    assert out[1, 0, 1] > 0.0                    # wet: decays, sign kept


def test_rain_volume_closed_basin():
    """Walls all round: the volume gain over k steps is R * sum(dt) * area exactly (float64)."""
    nx, ny = 24, 20
    rng = np.random.default_rng(4)
    bed = rng.uniform(0, 0.3, (nx, ny))
    h = np.maximum(0.5 - bed, 0.0)
    R = 100.0 / 3.6e6
    case = O.DenseCase(h=h, hu=0 * h, hv=0 * h, bed=bed, dx=2.0, dy=2.0, rain=R)
    o = O.DenseOracle(case, O.Physics(h_min=1e-10, h_cfl=1e-10, friction=True), np.float64)
    v0 = o.interior[0].sum() * 4.0
    o.run_steps(15)
    gain = o.interior[0].sum() * 4.0 - v0
    assert gain == pytest.approx(R * sum(o.dts) * nx * ny * 4.0, rel=1e-10)


def test_cfl_closed_form():
    h = np.array([[1.0, 4.0]])
    q = np.stack([h, np.array([[2.0, -1.0]]), np.array([[0.5, 3.0]])])
    ph = O.Physics(cfl=0.5)
    lam = O.lam_max(q, ph)
    assert lam == max(2.0 + math.sqrt(G * 1.0), 0.75 + math.sqrt(G * 4.0))
    assert O.dt_from_lam(lam, ph, 2.0) == 0.5 * 2.0 / (lam + 1e-12)
    assert O.dt_from_lam(0.0, ph, 2.0) == 0.5 * 2.0 / (math.sqrt(G * 1e-6) + 1e-12)
    with pytest.raises(FloatingPointError):
        O.dt_from_lam(math.nan, ph, 2.0)
    qn = q.copy()
    qn[0, 0, 0] = np.nan
    assert math.isnan(O.lam_max(qn, ph))


def test_cfl_l2_and_flat_rules_closed_form():
    """The L2 velocity norm and the flat layout's skip/divisor rule on hand-built cells
    (3-4-5 velocities, a thin film between h_min and h_min_cfl, a sub-h_min film)."""
    c = math.sqrt(G * 4.0)
    #            3-4-5 cell      thin film          below h_min
    h = np.array([[4.0,          2.0e-6,            5.0e-7]])
    hu = np.array([[12.0,        2.0e-6,            1.0]])
    hv = np.array([[-16.0,       0.0,               1.0]])
    q = np.stack([h, hu, hv])
    dense = O.Physics(h_min=1e-6, h_cfl=1e-3)
    assert O.lam_max(q, dense) == 4.0 + c                                   # L-inf; films skipped
    assert O.lam_max(q, O.Physics(h_min=1e-6, h_cfl=1e-3, cfl_l2=True)) == 5.0 + c
    # flat: only h < h_min is skipped; the film's velocity divides by h_min_cfl = 1e-3, not by h
    flat = O.Physics(h_min=1e-6, h_cfl=1e-3, flat_cfl=True)
    film = 2.0e-6 / 1.0e-3 + math.sqrt(G * 2.0e-6)
    assert O.lam_max(q[:, :, 1:], flat) == np.float64(film)
    assert O.lam_max(q, flat) == 4.0 + c
    assert O.lam_max(q, O.Physics(h_min=1e-6, h_cfl=1e-3, flat_cfl=True, cfl_l2=True)) == 5.0 + c
    # with h_min_cfl = h_min the flat rule is the dense one
    same = q[:, :, 1:].copy()
    assert O.lam_max(same, O.Physics(h_min=1e-6, h_cfl=1e-6, flat_cfl=True)) == \
        O.lam_max(same, O.Physics(h_min=1e-6, h_cfl=1e-6))
    # floors: sqrt(g*h_cfl) dense, sqrt(g*h_min) flat
    assert O.dt_from_lam(0.0, dense, 3.0) == 0.5 * 3.0 / (math.sqrt(G * 1e-3) + 1e-12)
    assert O.dt_from_lam(0.0, flat, 3.0) == 0.5 * 3.0 / (math.sqrt(G * 1e-6) + 1e-12)


@pytest.mark.parametrize("R,dmin,cfl", [(50 / 3.6e6, 3.0, 0.5), (1e-3, 1.0, 0.9), (2e-7, 30.0, 0.25)])
def test_rain_step_cap_is_the_film_cfl_step(R, dmin, cfl):
    """The rain-film bound (solver.py:3005-3027) is the fixed point dt = cfl*dmin/sqrt(g*R*dt):
    the CFL step of the film R*dt that the step itself lays down."""
    ph = O.Physics(cfl=cfl)
    dt = O.rain_step_cap(R, ph, dmin)
    assert dt == pytest.approx(cfl * dmin / math.sqrt(G * R * dt), rel=1e-14)
    assert O.rain_step_cap(0.0, ph, dmin) == math.inf
    assert O.rain_step_cap(-1.0, ph, dmin) == math.inf


def test_ghost_fill_by_hand():
    q = np.arange(3 * 6 * 5, dtype=float).reshape(3, 6, 5) + 1.0
    for kind in ("wall", "fall", "extrapolate", "periodic"):
        p = q.copy()
        O.bc_fill(p, 2, kind, "extrapolate")
        J = slice(2, 3)
        if kind == "wall":
            assert np.array_equal(p[0, 1, J], q[0, 2, J]) and np.array_equal(p[0, 0, J], q[0, 3, J])
            assert np.array_equal(p[1, 1, J], -q[1, 2, J]) and np.array_equal(p[2, 1, J], q[2, 2, J])
        elif kind == "fall":
            assert np.all(p[:, :2, J] == 0) and np.all(p[:, -2:, J] == 0)
        elif kind == "extrapolate":
            assert np.array_equal(p[:, 0, J], q[:, 2, J]) and np.array_equal(p[:, -1, J], q[:, 3, J])
        else:
            assert np.array_equal(p[:, 0, J], q[:, 2, J]) and np.array_equal(p[:, 1, J], q[:, 3, J])
            assert np.array_equal(p[:, 4, J], q[:, 2, J])
    b = O.pad_bed(np.array([[1.0, 2.0], [3.0, 4.0]]), 2, "wall", "periodic", np.float64)
    assert b[0, 2] == 1.0 and b[5, 3] == 4.0 and b[2, 0] == 1.0 and b[2, 1] == 2.0


@pytest.mark.parametrize("dtype,ulps", [(np.float32, 13), (np.float64, 17)])
def test_pow_m43_accuracy(dtype, ulps):
    rng = np.random.default_rng(0)
    h = (10 ** rng.uniform(-6, 4, 200000)).astype(dtype)
    ex = h.astype(np.float64) ** (-4.0 / 3.0)
    got = O.pow_m43(h)
    assert got.dtype == dtype
    err = np.abs(got.astype(np.float64) - ex) / np.spacing(ex.astype(dtype)).astype(np.float64)
    assert err.max() < ulps


def test_philox2x32_known_answers():
    """Random123 kat_vectors for philox2x32_10."""
    kat = [((0, 0, 0), (0xFF1DAE59, 0x6CD10DF2)),
           ((0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF), (0x2C3F628B, 0xAB4FD7AD)),
           ((0x243F6A88, 0x85A308D3, 0x13198A2E), (0xDD7CE038, 0xF62A4C12))]
    for (c0, c1, k), want in kat:
        a, b = GN.philox2x32(c0, c1, k)
        assert (int(a), int(b)) == want


def test_fbm_is_continuous_and_seeded():
    f1 = GN.fbm(np.arange(512)[:, None], np.arange(7)[None, :], 11, 6, 256.0, 0.8)
    f2 = GN.fbm(np.arange(512)[:, None], np.arange(7)[None, :], 11, 6, 256.0, 0.8)
    f3 = GN.fbm(np.arange(512)[:, None], np.arange(7)[None, :], 12, 6, 256.0, 0.8)
    assert np.array_equal(f1, f2) and not np.array_equal(f1, f3)
    # smoothstep-bilinear value noise: neighbouring cells differ by far less than the range
    assert np.abs(np.diff(f1, axis=0)).max() < 0.25 * (f1.max() - f1.min())


def test_flat_mesh_builder_invariants():
    from geoswe_amd.flatmesh import build_flat_mesh, cross_dilate
    rng = np.random.default_rng(9)
    inside = np.zeros((40, 50), bool)
    inside[4:36, 4:46] = rng.random((32, 42)) > 0.2
    fm = build_flat_mesh(inside, 4)
    stored = cross_dilate(inside, 2)
    assert fm.n_stored == stored.sum() and fm.n_active == inside.sum()
    assert np.array_equal(fm.unpack(fm.pack(stored.astype(float))), stored.astype(float))
    k = np.arange(fm.n_stored)
    ij = fm.ij
    for r in range(len(fm.run_k0)):
        for t in range(fm.run_len[r]):
            kk = fm.run_k0[r] + t
            i, j = ij[kk]
            assert inside[i, j]
            for off, di in zip(fm.run_d[r], (1, -1, 2, -2)):
                assert tuple(ij[kk + off]) == (i + di, j)
            for dj in (1, 2):
                assert tuple(ij[kk + dj]) == (i, j + dj) and tuple(ij[kk - dj]) == (i, j - dj)
    end = fm.run_k0 + fm.run_len
    for c in range(len(fm.chunk_run)):
        r = fm.chunk_run[c]
        assert r == len(end) or end[r] > c * 60
        assert r == 0 or end[r - 1] <= c * 60


def test_planted_mistakes_are_caught(monkeypatch):
    """A sign slip in the bed source or a swapped velocity fails the lake-at-rest / symmetry pins."""
    orig = O.srm_face

    def bad_face(*a, **k):
        hLf, hRf, gate, srcL, srcR = orig(*a, **k)
        return hLf, hRf, gate, -srcL, srcR
    monkeypatch.setattr(O, "srm_face", bad_face)
    with pytest.raises((AssertionError, FloatingPointError)):
        test_lake_at_rest_over_bump(np.float64, 1e-10)
    monkeypatch.setattr(O, "srm_face", orig)
    orig_h = O.hllc
    monkeypatch.setattr(O, "hllc", lambda hL, uNL, uTL, hR, uNR, uTR, g, hm: orig_h(hL, uTL, uNL, hR, uTR, uNR, g, hm))
    with pytest.raises((AssertionError, FloatingPointError)):
        test_ritter_dry_dam_break_front()


def _tile_walk(fm, inside):
    """Visit every (tile, row, lane) of the sweep tiles by the kernel's own rule and check
    ids against the row-major numbering; returns how often each active cell was written."""
    stored = np.zeros((fm.nxp, fm.nyp), bool)
    stored[fm.ij[:, 0], fm.ij[:, 1]] = True
    idmap = np.full((fm.nxp, fm.nyp), -1, np.int64)
    idmap[fm.ij[:, 0], fm.ij[:, 1]] = np.arange(fm.n_stored)
    written = np.zeros((fm.nxp, fm.nyp), np.int64)
    for t in range(len(fm.tile_e0)):
        i0, nr, c0 = (int(v) for v in fm.tile_irc[t])
        assert 1 <= nr and c0 >= 0
        for r in range(-2, nr + 2):
            base, sm, am = (int(v) for v in fm.ent[fm.tile_e0[t] + r + 2])
            i = i0 + r
            for lane in range(64):
                j = c0 + lane
                st = (sm >> lane) & 1
                assert st == (j < fm.nyp and stored[i, j])
                if st:
                    assert base + bin(sm & ((1 << lane) - 1)).count("1") == idmap[i, j]
                if 0 <= r < nr and (am >> lane) & 1:   # halo rows' active masks are never read
                    assert 2 <= lane < 62 and inside[i, j]
                    written[i, j] += 1
                    for di, dj in ((1, 0), (-1, 0), (2, 0), (-2, 0), (0, 1), (0, -1), (0, 2), (0, -2)):
                        assert stored[i + di, j + dj]
    return written


@pytest.mark.parametrize("rows", [1, 3, 8, 64])
def test_sweep_tiles_cover_every_active_cell_once(rows):
    """Row-sweep tiles: lane ids from base + popcount(stored lanes below) equal the
    row-major numbering, stored-lane masks equal the stored set, and every active cell
    is written by exactly one (tile, row, lane) whose +-2 stencil is stored."""
    from geoswe_amd.flatmesh import build_flat_mesh
    rng = np.random.default_rng(31)
    masks = []
    m = np.zeros((70, 150), bool)                   # ragged blob wider than two windows
    m[5:64, 4:140] = rng.random((59, 136)) > 0.25
    masks.append(m)
    m = np.zeros((30, 200), bool)                   # islands closer together than a window
    m[4:8, 10:20] = m[4:9, 40:45] = m[12:26, 58:66] = m[20, 120:196] = True
    masks.append(m)
    m = np.zeros((12, 12), bool)                    # single cell
    m[6, 6] = True
    masks.append(m)
    m = np.zeros((40, 67), bool)                    # active up to the last allowed column/row
    m[2:38, 2:65] = True
    masks.append(m)
    for inside in masks:
        fm = build_flat_mesh(inside, 2 if inside[:, 2].any() else 4, sweep_rows=rows)
        written = _tile_walk(fm, inside)
        assert np.array_equal(written, inside.astype(np.int64))
        assert np.all(fm.tile_irc[:, 1] <= rows)
        # the same tables rebuilt from the mesh alone (an ensemble's taller tiles)
        from geoswe_amd.flatmesh import mesh_sweep_tiles
        e0, irc, ent = mesh_sweep_tiles(build_flat_mesh(inside, fm.ngh, sweep_rows=1), rows)
        assert np.array_equal(e0, fm.tile_e0) and np.array_equal(irc, fm.tile_irc) and np.array_equal(ent, fm.ent)


def test_sweep_rows_for_domain_size():
    from geoswe_amd.flatmesh import MAX_SWEEP_ROWS, TILE_TARGET, sweep_rows_for
    assert sweep_rows_for(0) == 1 and sweep_rows_for(2 * TILE_TARGET - 1) == 1
    assert sweep_rows_for(2 * TILE_TARGET) == 2 and sweep_rows_for(10 ** 9) == MAX_SWEEP_ROWS == 32
    for busy in (1, 5000, 40000, 300000, 2000000):
        r = sweep_rows_for(busy)
        assert r == MAX_SWEEP_ROWS or busy // r >= TILE_TARGET > busy // (2 * r) or r == 1


def test_sweep_tiles_random_masks():
    """Property check of the sweep-tile builder over random masks, sizes and tile heights
    (hypothesis): every active cell written exactly once, ids and stored masks exact."""
    hyp = pytest.importorskip("hypothesis")
    from hypothesis import given, settings, strategies as st
    from geoswe_amd.flatmesh import build_flat_mesh

    @settings(max_examples=25, deadline=None, derandomize=True)
    @given(nx=st.integers(1, 30), ny=st.integers(1, 140), density=st.floats(0.05, 1.0),
           rows=st.sampled_from([1, 2, 7, 64]), seed=st.integers(0, 2 ** 16))
    def check(nx, ny, density, rows, seed):
        rng = np.random.default_rng(seed)
        inside = np.zeros((nx + 8, ny + 8), bool)
        inside[4:4 + nx, 4:4 + ny] = rng.random((nx, ny)) < density
        if not inside.any():
            inside[4, 4] = True
        fm = build_flat_mesh(inside, 4, sweep_rows=rows)
        assert np.array_equal(_tile_walk(fm, inside), inside.astype(np.int64))
        # windows placed at the first uncovered active column of each band: never more
        # tiles than windows on the fixed column grid 60k, and per band the fewest that
        # 60-column windows allow (a window writes 60 consecutive columns, so active
        # columns pairwise 60 or more apart need one window each: a set of that many
        # proves the count minimal)
        nxp, nyp = inside.shape
        grid = best = 0
        for r0 in range(0, nxp, rows):
            cols = np.flatnonzero(inside[r0:r0 + rows].any(axis=0))
            grid += len({(c - 2) // 60 for c in cols})
            k = band = 0
            while k < cols.size:                   # greedy cover, restated
                band += 1
                k = int(np.searchsorted(cols, cols[k] + 60))
            need = 0                              # columns pairwise 60 or more apart
            last = -10 ** 9
            for c in cols:
                if c - last >= 60:
                    need += 1
                    last = c
            assert band == need
            best += band
        assert len(fm.tile_e0) == best <= grid

    check()


def _same_mesh(a, b):
    for f in ("nxp", "nyp", "ngh", "n_stored", "n_active"):
        assert getattr(a, f) == getattr(b, f), f
    for f in ("ij", "active", "run_k0", "run_len", "run_i", "run_j0", "run_d", "chunk_run", "tile_e0", "tile_irc",
              "ent"):
        x, y = getattr(a, f), getattr(b, f)
        assert x.dtype == y.dtype and x.shape == y.shape and np.array_equal(x, y), f


def test_native_mesh_builder_equals_numpy_builder():
    """gs_flat_mesh_build (host C++) gives every array of the NumPy builder exactly, for
    ragged, islanded, single-cell, edge-touching and random masks, automatic and fixed tile
    heights, and rejects the same inputs."""
    from geoswe_amd.flatmesh import build_flat_mesh
    rng = np.random.default_rng(77)
    masks = []
    m = np.zeros((70, 150), bool)
    m[5:64, 4:140] = rng.random((59, 136)) > 0.25
    masks.append((m, 4))
    m = np.zeros((30, 200), bool)
    m[4:8, 10:20] = m[4:9, 40:45] = m[12:26, 58:66] = m[20, 120:196] = True
    masks.append((m, 4))
    m = np.zeros((12, 12), bool)
    m[6, 6] = True
    masks.append((m, 4))
    m = np.zeros((40, 67), bool)
    m[2:38, 2:65] = True
    masks.append((m, 2))
    for k in range(6):
        nx, ny = rng.integers(1, 90), rng.integers(1, 260)
        m = np.zeros((nx + 8, ny + 8), bool)
        m[4:4 + nx, 4:4 + ny] = rng.random((nx, ny)) < rng.uniform(0.05, 1.0)
        m[4, 4] = True
        masks.append((m, 4))
    for inside, g in masks:
        for rows in (None, 1, 3, 64):
            fm = build_flat_mesh(inside, g, sweep_rows=rows, native=True)
            _same_mesh(fm, build_flat_mesh(inside, g, sweep_rows=rows, native=False))
            # stored cells in row-major order (the (row, col) -> id lookup of set_clamp)
            key = fm.ij[:, 0].astype(np.int64) * fm.nyp + fm.ij[:, 1]
            assert np.all(np.diff(key) > 0)
    bad = np.zeros((10, 10), bool)
    bad[1, 5] = True                                  # closer than two cells to the edge
    for native in (True, False):
        with pytest.raises(ValueError):
            build_flat_mesh(bad, 4, native=native)
        with pytest.raises(ValueError):
            build_flat_mesh(masks[2][0], 4, ring=1, native=native)


def test_config_compute_dtype_rule():
    """float64 arithmetic is offered for a float32 state only."""
    import pytest
    from geoswe_amd.config import Config
    assert Config(dtype="float32", compute_dtype="float64").compute_dtype == "float64"
    assert Config(dtype="float64", compute_dtype="float64").compute_dtype == "float64"
    for kw in (dict(dtype="float64", compute_dtype="float32"), dict(dtype="float32", compute_dtype="float16")):
        with pytest.raises(ValueError, match="compute_dtype"):
            Config(**kw)

