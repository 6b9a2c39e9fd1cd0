"""Open-boundary sponge and stage clamp fused into the flat step, against the oracle.

The flat solver's state on every stored interior cell (active and ring), its running
maximum and its step sequence must equal the dense oracle's with the same forcings bit for
bit, for both traversals and both precisions; the ensemble carries active-cell forcings
and refuses ring forcings, and the adjoint refuses both.
"""
import numpy as np
import pytest

from oracle import swe_oracle as O

from test_gpu_flat import dense, irregular_mask, rough_state

pytestmark = pytest.mark.gpu

NX, NY = 56, 70
T_END = 2.0


def _forcings(m, band_rows=14, h_amb=0.55, alpha=0.3):
    """Sponge over the first `band_rows` rows (active, ring and outer cells alike) ramping to
    `alpha` at row 0; clamp on a patch of active cells, a few ring cells and cells that are not
    stored (skipped by the flat layout, inert in the oracle)."""
    nx, ny = m.shape
    ramp = np.clip((band_rows - np.arange(nx)) / band_rows, 0.0, 1.0)[:, None] * np.ones((1, ny))
    keep = 1.0 - alpha * ramp
    amb = alpha * ramp * h_amb
    rows, cols = np.meshgrid(np.arange(20, 32), np.arange(25, 45, 2), indexing="ij")
    rows, cols = list(rows.ravel()), list(cols.ravel())
    ring = np.argwhere(~m & (np.roll(m, 1, 0) | np.roll(m, -1, 0)))[:6]
    outer = np.argwhere(~m)[:4]
    for i, j in list(ring) + list(outer):
        rows.append(int(i))
        cols.append(int(j))
    rows, cols = np.array(rows), np.array(cols)
    hmax = np.full(len(rows), 0.42)
    grid = np.full((nx, ny), np.inf)
    grid[rows, cols] = hmax
    return keep, amb, rows, cols, hmax, grid


def _padded(a, g, fill):
    out = np.full((a.shape[0] + 2 * g, a.shape[1] + 2 * g), fill)
    out[g:-g, g:-g] = a
    return out


def _physics(cfg):
    """The oracle's scalars from the solver's Config (h_min is 1e-6 only for float32)."""
    return O.Physics(h_min=cfg.h_min, h_cfl=cfg.h_cfl(), cfl=cfg.cfl, friction=True,
                     vcap=cfg.friction_velocity_cap_ms)


def _flat(q0, bed, m, dtype, trav, keep, amb, rows, cols, hmax, sweep_rows=None):
    from geoswe_amd import CompressedSolver
    s = dense(q0, bed, friction="manning", manning_n=0.03, rainfall=30.0 / 3.6e6, dtype=dtype)
    s.set_inside_mask(m)
    s.set_cfl_ghost_mask(~m)
    c = CompressedSolver.from_dense(s, cfl_linf=True, track_max_depth=True, sweep_rows=sweep_rows)
    g = s.mesh.ngh
    c.set_sponge({"keep": _padded(keep, g, 1.0), "amb": _padded(amb, g, 0.0)})
    c.set_clamp({"rows": rows + g, "cols": cols + g, "hmax": hmax})
    c.traversal = trav
    return c, g


@pytest.mark.parametrize("dtype", ["float32", "float64"])
@pytest.mark.parametrize("trav", ["tiles", "chunks"])
def test_sponge_and_clamp_match_oracle(gpu, dtype, trav, monkeypatch):
    q0, bed = rough_state(NX, NY, seed=31, depth=0.7, rough=1.0, vel=0.3)
    m = irregular_mask(NX, NY, seed=9)
    q0 = q0 * m                                  # inactive cells at rest
    keep, amb, rows, cols, hmax, grid = _forcings(m)
    c, g = _flat(q0, bed, m, dtype, trav, keep, amb, rows, cols, hmax, sweep_rows=5)
    assert c._force_ring is not None and c._force_ring.shape[0] > 0       # ring cells are forced
    n = c.run(T_END)

    # oracle, recording the running maximum where the step records it (before the forcings)
    mx = np.zeros((NX, NY), np.dtype(dtype))
    orig = O.coastal_force

    def recording(q, *args):
        np.maximum(mx, q[0], out=mx)
        orig(q, *args)

    monkeypatch.setattr(O, "coastal_force", recording)
    case = O.DenseCase(h=q0[0], hu=q0[1], hv=q0[2], bed=bed, dx=2.0, dy=2.0, bc_x="fall", bc_y="fall",
                       manning=0.03, rain=30.0 / 3.6e6, active=m, sponge=(keep, amb), clamp=grid)
    o = O.DenseOracle(case, _physics(c.cfg), np.dtype(dtype))
    o.cfl_skip = ~m
    o.run(T_END)
    print(dtype, trav, n, flush=True)
    assert o.nsteps == c.nsteps == n > 5 and o.t == c.t

    ij = c.fm.ij - g
    inner = (ij[:, 0] >= 0) & (ij[:, 0] < NX) & (ij[:, 1] >= 0) & (ij[:, 1] < NY)
    qf = c.q_flat()
    want = o.interior[:, ij[inner, 0], ij[inner, 1]]
    assert np.array_equal(qf[:, inner], want)
    act = c.fm.active
    ia = ij[act]
    assert np.array_equal(c.max_depth()[ia[:, 0], ia[:, 1]], mx[ia[:, 0], ia[:, 1]])
    # the forcings acted: band rows relaxed, clamp patch capped, ring cells of the band moved
    band_ring = inner & ~act & (ij[:, 0] < 5) & (keep[np.clip(ij[:, 0], 0, NX - 1), 0] < 1)
    if band_ring.any():
        assert not np.array_equal(qf[0, band_ring], c.fm.pack(np.pad(q0[0], g))[band_ring].astype(dtype))
    patch = act & (ij[:, 0] >= 20) & (ij[:, 0] < 32)
    assert np.all(qf[0, patch & np.isin(ij[:, 1], np.arange(25, 45, 2))] <= 0.42)


def test_forcing_with_identity_fields_is_the_plain_step(gpu):
    """keep = 1, amb = 0 and no finite cap flag no cell: the run is bitwise the plain run."""
    from geoswe_amd import CompressedSolver
    q0, bed = rough_state(NX, NY, seed=32, depth=0.7, rough=1.0, vel=0.3)
    runs = []
    for forced in (False, True):
        s = dense(q0, bed, friction="manning", manning_n=0.03)
        c = CompressedSolver.from_dense(s, cfl_linf=True)
        if forced:
            g = s.mesh.ngh
            c.set_sponge({"keep": np.ones((NX + 2 * g, NY + 2 * g)), "amb": np.zeros((NX + 2 * g, NY + 2 * g))})
            c.set_clamp({"rows": np.array([g + 3]), "cols": np.array([g + 3]), "hmax": np.array([np.inf])})
            assert not c.has_coastal_forcing and c._flags is None
        c.run(1.5)
        runs.append((c.q_flat(), c.nsteps))
    assert runs[0][1] == runs[1][1] and np.array_equal(runs[0][0], runs[1][0])


def test_ensemble_carries_active_forcings_and_refuses_ring_ones(gpu):
    from geoswe_amd.ensemble import Ensemble
    q0, bed = rough_state(NX, NY, seed=33, depth=0.7, rough=1.0, vel=0.3)
    m = np.zeros((NX, NY), bool)
    m[2:NX - 2, 2:NY - 2] = True
    q0 = q0 * m
    keep, amb, rows, cols, hmax, _ = _forcings(m)
    keep[~m] = 1.0                               # band on active cells only
    amb[~m] = 0.0
    inside = m[rows, cols]
    mk = lambda: _flat(q0, bed, m, "float32", "tiles", keep, amb, rows[inside], cols[inside], hmax[inside])[0]
    plain = [mk(), mk()]
    assert plain[0]._force_ring is None and plain[0].has_coastal_forcing
    for p in plain:
        p.advance(20)
    mem = [mk(), mk()]
    ens = Ensemble(mem, graph_batch=8)
    ens.advance(20)
    for p, e in zip(plain, mem):
        assert p.nsteps == e.nsteps == 20 and p.t == e.t
        assert np.array_equal(p.q_flat(), e.q_flat())
    ens.close()
    ring = _flat(q0, bed, m, "float32", "tiles", *_forcings(m)[:5])[0]
    assert ring._force_ring is not None
    with pytest.raises(NotImplementedError, match="ring cells"):
        Ensemble([ring])


def test_forcing_setters_refuse_after_stepping_and_adjoint_refuses(gpu):
    from geoswe_amd import CompressedSolver
    from geoswe_amd.adjoint import GaugeMisfit
    q0, bed = rough_state(40, 40, seed=34)
    s = dense(q0, bed, friction="manning", manning_n=0.03)
    c = CompressedSolver.from_dense(s, cfl_linf=True)
    g = s.mesh.ngh
    c.set_clamp({"rows": np.array([g + 5]), "cols": np.array([g + 5]), "hmax": np.array([0.1])})
    with pytest.raises(NotImplementedError, match="no reverse"):
        GaugeMisfit(c, np.array([int(np.nonzero(c.fm.active)[0][0])]), np.full(3, 0.01))
    with pytest.raises(ValueError, match="outside the padded grid"):
        c.set_clamp({"rows": np.array([-1]), "cols": np.array([0]), "hmax": np.array([0.1])})
    c.step()
    with pytest.raises(RuntimeError, match="already stepped"):
        c.set_sponge(None)


def test_resume_with_forced_ring_cells_is_bitwise(gpu, tmp_path):
    """A run with forced ring cells stopped at an odd step count and resumed in a fresh
    solver ends bitwise where the uninterrupted run ends; an unforced ring cell whose value in
    the checkpoint differs from the solver's own (a checkpoint written by a code whose ring
    moved) keeps that value on every later step, whichever buffer the step reads."""
    q0, bed = rough_state(NX, NY, seed=35, depth=0.7, rough=1.0, vel=0.3)
    m = irregular_mask(NX, NY, seed=9)
    q0 = q0 * m
    f = _forcings(m)[:5]
    mk = lambda: _flat(q0, bed, m, "float32", "tiles", *f)[0]
    T = 6.0
    a = mk()
    a.run(T, batch=8)
    d = tmp_path / "ckpt"
    b = mk()
    b.run(T, max_steps=13, batch=4, ckpt_dir=str(d), checkpoint_every_s=1.0)
    assert b.nsteps == 13 and a.nsteps > 13
    c = mk()
    c.run(T, batch=8, ckpt_dir=str(d), resume=True)
    assert c.nsteps == a.nsteps and c.t == a.t
    assert np.array_equal(c.q_flat(), a.q_flat()) and np.array_equal(c.max_depth(), a.max_depth())

    forced = np.zeros(c.n_stored, bool)
    forced[c._force_ring.download()] = True
    k = int(np.nonzero(~c.fm.active & ~forced)[0][0])        # an unforced ring cell
    z = dict(np.load(d / "ckpt_r00.npz"))
    z["q0"] = z["q0"].copy()
    z["q0"][k] = np.float32(0.3125)
    np.savez(d / "ckpt_r00.npz", **z)
    e = mk()
    e.load_checkpoint(str(d))
    for _ in range(2):
        e.step()
        assert e.q_flat()[0, k] == np.float32(0.3125)


@pytest.mark.parametrize("dtype", ["float32", "float64"])
@pytest.mark.parametrize("trav", ["tiles", "chunks"])
def test_gauge_ring_matches_oracle(gpu, dtype, trav):
    """The IDW gauge ring on every first-layer ring cell, with the sponge and the clamp: the
    flat state of every stored interior cell, the step count and the time equal the oracle's
    bit for bit; the common time grid is fine enough that the run crosses several samples."""
    q0, bed = rough_state(NX, NY, seed=36, depth=0.7, rough=1.0, vel=0.3)
    m = irregular_mask(NX, NY, seed=9)
    q0 = q0 * m
    keep, amb, rows, cols, hmax, grid = _forcings(m)
    c, g = _flat(q0, bed, m, dtype, trav, keep, amb, rows, cols, hmax)
    near = np.zeros_like(m)
    for ax in (0, 1):
        for sh in (1, -1):
            near |= np.roll(m, sh, ax)
    cells = np.argwhere(~m & near)                      # the ring's first layer, inside the box
    rng = np.random.default_rng(3)
    ng = 3
    w = rng.random((len(cells), ng))
    w /= w.sum(axis=1, keepdims=True)
    tc = np.array([-1.0, 0.0, 0.35, 0.8, 1.3, 1.9, 4.0])
    S = 0.45 + 0.5 * rng.random((ng, len(tc)))
    c.set_ring({"i": cells[:, 0] + g, "j": cells[:, 1] + g, "bed": bed[cells[:, 0], cells[:, 1]], "wg": w.ravel(),
                "NG": ng, "n": len(cells), "t_common": tc, "stage_all": S})
    n = c.run(T_END)
    case = O.DenseCase(h=q0[0], hu=q0[1], hv=q0[2], bed=bed, dx=2.0, dy=2.0, bc_x="fall", bc_y="fall",
                       manning=0.03, rain=30.0 / 3.6e6, active=m, sponge=(keep, amb), clamp=grid,
                       ring={"cells": cells, "bed": bed[cells[:, 0], cells[:, 1]], "w": w, "t_common": tc,
                             "stage_all": S})
    o = O.DenseOracle(case, _physics(c.cfg), np.dtype(dtype))
    o.cfl_skip = ~m
    o.run(T_END)
    print(dtype, trav, n, flush=True)
    assert o.nsteps == c.nsteps == n > 5 and o.t == c.t
    ij = c.fm.ij - g
    inner = (ij[:, 0] >= 0) & (ij[:, 0] < NX) & (ij[:, 1] >= 0) & (ij[:, 1] < NY)
    qf = c.q_flat()
    assert np.array_equal(qf[:, inner], o.interior[:, ij[inner, 0], ij[inner, 1]])
    k = c._flat_ids(cells[:, 0] + g, cells[:, 1] + g)
    assert (k >= 0).all() and (qf[0, k] > 0).any() and np.all(qf[1:, k] == 0)   # wet still-water ring


@pytest.mark.parametrize("dtype", ["float32", "float64"])
@pytest.mark.parametrize("trav", ["tiles", "chunks"])
def test_full_coastal_tail_matches_oracle(gpu, dtype, trav, monkeypatch):
    """Every coastal forcing at once (Manning classes with per-class infiltration, sponge,
    gauge ring, clamp, karst drain on active and ring cells): state of every stored interior
    cell, running maximum, steps and time equal the oracle's bit for bit."""
    from geoswe_amd import CompressedSolver
    q0, bed = rough_state(NX, NY, seed=37, depth=0.7, rough=1.0, vel=0.3)
    m = irregular_mask(NX, NY, seed=9)
    q0 = q0 * m
    rng = np.random.default_rng(8)
    cls = rng.integers(0, 6, (NX, NY)).astype(np.uint8)
    table = np.linspace(0.02, 0.07, 6).astype(np.float32)
    rates = np.array([0.0, 2e-3, 5e-3, 0.0, 1e-2, 3e-3])
    keep, amb, rows, cols, hmax, grid = _forcings(m)
    s = dense(q0, bed, friction="manning", rainfall=30.0 / 3.6e6, dtype=dtype)
    g = s.mesh.ngh
    s.set_manning_table(np.pad(cls, g), table)
    s.set_inside_mask(m)
    s.set_cfl_ghost_mask(~m)
    c = CompressedSolver.from_dense(s, cfl_linf=True, track_max_depth=True)
    c.set_sponge({"keep": _padded(keep, g, 1.0), "amb": _padded(amb, g, 0.0)})
    c.set_clamp({"rows": rows + g, "cols": cols + g, "hmax": hmax})
    near = np.zeros_like(m)
    for ax in (0, 1):
        for sh in (1, -1):
            near |= np.roll(m, sh, ax)
    cells = np.argwhere(~m & near)
    w = rng.random((len(cells), 2))
    w /= w.sum(axis=1, keepdims=True)
    tc = np.array([-1.0, 0.5, 1.2, 3.0])
    S = 0.45 + 0.5 * rng.random((2, len(tc)))
    ring = {"cells": cells, "bed": bed[cells[:, 0], cells[:, 1]], "w": w, "t_common": tc, "stage_all": S}
    c.set_ring({"i": cells[:, 0] + g, "j": cells[:, 1] + g, "bed": ring["bed"], "wg": w, "NG": 2, "n": len(cells),
                "t_common": tc, "stage_all": S})
    c.set_infil({"tab": rates})
    ij = c.fm.ij - g
    act = c.fm.active
    cand = np.nonzero(act & (ij[:, 0] > 30) & (ij[:, 0] < 40))[0][::3]
    ring_ids = np.nonzero(~act & (ij[:, 0] >= 0) & (ij[:, 0] < NX) & (ij[:, 1] >= 0) & (ij[:, 1] < NY))[0][::7]
    didx = np.concatenate([cand, ring_ids])
    ht = 0.3 + 0.2 * rng.random(len(didx))
    c.set_drain({"idx": didx, "h_tgt_arr": ht})
    c.traversal = trav
    n = c.run(T_END)

    dgrid = np.full((NX, NY), np.inf)
    dgrid[ij[didx, 0], ij[didx, 1]] = ht
    mx = np.zeros((NX, NY), np.dtype(dtype))
    orig = O.coastal_force

    def recording(q, *args):
        np.maximum(mx, q[0], out=mx)
        orig(q, *args)

    monkeypatch.setattr(O, "coastal_force", recording)
    case = O.DenseCase(h=q0[0], hu=q0[1], hv=q0[2], bed=bed, dx=2.0, dy=2.0, bc_x="fall", bc_y="fall",
                       manning=table[cls], rain=30.0 / 3.6e6, active=m, sponge=(keep, amb), clamp=grid,
                       ring=ring, infil=(cls, rates), drain=dgrid)
    o = O.DenseOracle(case, _physics(c.cfg), np.dtype(dtype))
    o.cfl_skip = ~m
    o.run(T_END)
    print(dtype, trav, n, flush=True)
    assert o.nsteps == c.nsteps == n > 5 and o.t == c.t
    inner = (ij[:, 0] >= 0) & (ij[:, 0] < NX) & (ij[:, 1] >= 0) & (ij[:, 1] < NY)
    qf = c.q_flat()
    assert np.array_equal(qf[:, inner], o.interior[:, ij[inner, 0], ij[inner, 1]])
    ia = ij[act]
    assert np.array_equal(c.max_depth()[ia[:, 0], ia[:, 1]], mx[ia[:, 0], ia[:, 1]])
    assert np.all(qf[0, didx] <= ht.astype(dtype))


def test_ensemble_carries_infiltration_and_active_drain(gpu):
    from geoswe_amd import CompressedSolver
    from geoswe_amd.ensemble import Ensemble
    q0, bed = rough_state(NX, NY, seed=38, depth=0.7, rough=1.0, vel=0.3)
    m = np.zeros((NX, NY), bool)
    m[2:NX - 2, 2:NY - 2] = True
    q0 = q0 * m
    cls = np.random.default_rng(9).integers(0, 4, (NX, NY)).astype(np.uint8)

    def mk():
        s = dense(q0, bed, friction="manning")
        g = s.mesh.ngh
        s.set_manning_table(np.pad(cls, g), np.array([0.02, 0.03, 0.05, 0.04], np.float32))
        s.set_inside_mask(m)
        s.set_cfl_ghost_mask(~m)
        c = CompressedSolver.from_dense(s, cfl_linf=True)
        c.set_infil({"tab": [0.0, 4e-3, 1e-2, 2e-3]})
        ids = np.nonzero(c.fm.active)[0][::11]
        c.set_drain({"idx": ids, "h_tgt": 0.35})
        assert c._force_ring is None
        return c

    plain = [mk(), mk()]
    for p in plain:
        p.advance(20)
    mem = [mk(), mk()]
    ens = Ensemble(mem, graph_batch=8)
    ens.advance(20)
    for p, e in zip(plain, mem):
        assert p.nsteps == e.nsteps == 20 and p.t == e.t
        assert np.array_equal(p.q_flat(), e.q_flat())
    ens.close()


def test_coastal_case_with_its_forcing_tail_matches_oracle(gpu):
    """The forcing tail the coastal benchmark attaches (infiltration by class, edge sponge,
    four-gauge IDW tide ring, Green-Ampt with drain-tau, clamp, drain) on the coastal case at
    test size, with its gridded rain and tidal stage: 25 flat steps bitwise the oracle's,
    the cumulative infiltration F included."""
    from geoswe_amd.cases import attach_coastal_forcings, coastal_c3
    case = coastal_c3(n=192, band=8, n_sites=24, rain_frames=2)
    f = case.flat()
    spec = attach_coastal_forcings(case, f, band=12, n_clamp=300, n_drain=150, fields=True)
    cells = spec["cells"]
    assert all(v > 0 for v in cells.values()), cells
    F = spec["fields"]
    f.advance(25)
    sb, rf = case.cfg.stage_boundary, case.cfg.rainfall_forcing
    oc = O.DenseCase(h=case.q0[0], hu=case.q0[1], hv=case.q0[2], bed=case.bed, dx=3.0, dy=3.0, bc_x="fall",
                     bc_y="fall", manning=case.ntab.astype(np.float32)[case.cls], rain=(rf.time_s, rf.rate_mm_h),
                     active=case.active,
                     stage={"cells": np.argwhere(case.sea_edge), "time_s": sb.time_s, "stage_m": sb.stage_m},
                     sponge=(F["keep"], F["amb"]), ring=F["ring"], clamp=F["clamp"], drain=F["drain"],
                     infil=(case.cls, F["infil"]), ga=F["ga"])
    o = O.DenseOracle(oc, _physics(f.cfg), np.float32)
    o.cfl_skip = ~case.active
    o.run_steps(25)
    assert f.nsteps == o.nsteps == 25 and f.t == o.t
    g = case.mesh.ngh
    ij = f.fm.ij - g
    n = case.mesh.nx
    inner = (ij[:, 0] >= 0) & (ij[:, 0] < n) & (ij[:, 1] >= 0) & (ij[:, 1] < n)
    assert np.array_equal(f.q_flat()[:, inner], o.interior[:, ij[inner, 0], ij[inner, 1]])
    act = f.fm.active
    assert np.array_equal(f.ga_F()[act], o.ga_F[ij[act, 0], ij[act, 1]]) and (f.ga_F()[act] > 0).any()


def _ga_setup(dtype, mode, seed=41):
    from geoswe_amd import CompressedSolver
    q0, bed = rough_state(NX, NY, seed=seed, depth=0.7, rough=1.0, vel=0.3)
    m = irregular_mask(NX, NY, seed=9)
    q0 = q0 * m
    rng = np.random.default_rng(seed)
    scls = rng.integers(0, 4, (NX, NY)).astype(np.uint8)
    Ks, psi, dth = np.array([0.0, 2e-5, 8e-5, 3e-4]), np.array([0.05, 0.11, 0.2, 0.3]), np.array([0.2, 0.3, 0.35, 0.4])
    F0 = (rng.random((NX, NY)) * 0.01) * m
    Fmax = F0 + np.where(rng.random((NX, NY)) < 0.3, 2e-4, 1.0)
    inv_tau = np.where(rng.random((NX, NY)) < 0.5, 1.0 / 40.0, 0.0)
    s = dense(q0, bed, friction="manning", manning_n=0.03, rainfall=30.0 / 3.6e6, dtype=dtype)
    s.set_inside_mask(m)
    s.set_cfl_ghost_mask(~m)
    c = CompressedSolver.from_dense(s, cfl_linf=True, track_max_depth=True)
    g = s.mesh.ngh
    c.set_ga_drain({"cls_pad": np.pad(scls, g), "Ks_t": Ks, "psi_t": psi, "dth_t": dth, "F_pad": np.pad(F0, g),
                    "Fmax_pad": np.pad(Fmax, g), "inv_tau_pad": np.pad(inv_tau, g), "mode": mode})
    ga = {"cls": scls, "Ks": None if mode == "drain" else Ks, "psi": psi, "dth": dth, "F": F0,
          "Fmax": None if mode == "drain" else Fmax, "inv_tau": None if mode == "ga" else inv_tau}
    case = O.DenseCase(h=q0[0], hu=q0[1], hv=q0[2], bed=bed, dx=2.0, dy=2.0, bc_x="fall", bc_y="fall",
                       manning=0.03, rain=30.0 / 3.6e6, active=m, ga=ga)
    return c, g, m, case


@pytest.mark.parametrize("dtype,trav,mode", [("float32", "tiles", "fused"), ("float64", "tiles", "fused"),
                                             ("float32", "chunks", "fused"), ("float64", "chunks", "fused"),
                                             ("float32", "tiles", "ga"), ("float32", "tiles", "drain")])
def test_green_ampt_drain_matches_oracle(gpu, dtype, trav, mode, monkeypatch):
    """Green-Ampt infiltration with a storage cap and drain-tau on an irregular domain: flat
    state of every stored interior cell, the cumulative infiltration F, the running maximum,
    steps and time equal the oracle's bit for bit (the drain's exponential is the IEEE-only
    one on both sides)."""
    c, g, m, case = _ga_setup(dtype, mode)
    c.traversal = trav
    n = c.run(T_END)
    mx = np.zeros((NX, NY), np.dtype(dtype))
    orig = O.coastal_force

    def recording(q, *args):
        np.maximum(mx, q[0], out=mx)
        orig(q, *args)

    monkeypatch.setattr(O, "coastal_force", recording)
    o = O.DenseOracle(case, _physics(c.cfg), np.dtype(dtype))
    o.cfl_skip = ~m
    o.run(T_END)
    print(dtype, trav, mode, n, flush=True)
    assert o.nsteps == c.nsteps == n > 5 and o.t == c.t
    ij = c.fm.ij - g
    inner = (ij[:, 0] >= 0) & (ij[:, 0] < NX) & (ij[:, 1] >= 0) & (ij[:, 1] < NY)
    qf = c.q_flat()
    assert np.array_equal(qf[:, inner], o.interior[:, ij[inner, 0], ij[inner, 1]])
    act = c.fm.active
    ia = ij[act]
    assert np.array_equal(c.max_depth()[ia[:, 0], ia[:, 1]], mx[ia[:, 0], ia[:, 1]])
    if mode != "drain":
        F = c.ga_F()
        assert np.array_equal(F[act], o.ga_F[ia[:, 0], ia[:, 1]])
        assert (F[act] > np.asarray(case.ga["F"])[ia[:, 0], ia[:, 1]]).any()       # water went into the soil


def test_green_ampt_resume_and_ensemble(gpu, tmp_path):
    """F is state: a run stopped at an odd step and resumed from its checkpoint ends bitwise
    where the uninterrupted run ends (state and F); ensemble members with Green-Ampt equal
    their plain runs."""
    from geoswe_amd.ensemble import Ensemble
    mk = lambda: _ga_setup("float32", "fused", seed=42)[0]
    a = mk()
    a.run(6.0, batch=8)
    d = str(tmp_path / "ckpt")
    b = mk()
    b.run(6.0, max_steps=13, batch=4, ckpt_dir=d, checkpoint_every_s=1.0)
    assert b.nsteps == 13 and a.nsteps > 13
    c = mk()
    c.run(6.0, batch=8, ckpt_dir=d, resume=True)
    assert c.nsteps == a.nsteps and np.array_equal(c.q_flat(), a.q_flat()) and np.array_equal(c.ga_F(), a.ga_F())
    plain = [mk(), mk()]
    for p in plain:
        p.advance(15)
    mem = [mk(), mk()]
    ens = Ensemble(mem, graph_batch=8)
    ens.advance(15)
    for p, e in zip(plain, mem):
        assert p.nsteps == e.nsteps == 15
        assert np.array_equal(p.q_flat(), e.q_flat()) and np.array_equal(p.ga_F(), e.ga_F())
    ens.close()


def test_green_ampt_ensemble_after_autotune(gpu):
    """The launch-shape autotuner restores everything its trial steps write, the Green-Ampt
    cumulative infiltration F included: members with Green-Ampt stepped after an autotune
    equal their plain runs bit for bit (state and F)."""
    from geoswe_amd.ensemble import Ensemble
    mk = lambda: _ga_setup("float32", "fused", seed=43)[0]
    plain = [mk(), mk()]
    for p in plain:
        p.advance(9)
    mem = [mk(), mk()]
    F0 = [m.ga_F().copy() for m in mem]
    ens = Ensemble(mem, graph_batch=4)
    ens.autotune(trial_steps=3)
    for m, f0 in zip(mem, F0):
        assert m.nsteps == 0 and np.array_equal(m.ga_F(), f0)
    ens.advance(9)
    for p, e in zip(plain, mem):
        assert p.nsteps == e.nsteps == 9
        assert np.array_equal(p.q_flat(), e.q_flat()) and np.array_equal(p.ga_F(), e.ga_F())
    ens.close()


def test_checkpoint_without_F_is_refused_with_green_ampt(gpu, tmp_path):
    """A checkpoint written without Green-Ampt (no F array) does not load into a solver with
    Green-Ampt: resuming would reset the cumulative infiltration (the reference raises too)."""
    c, g, m, case = _ga_setup("float32", "fused", seed=44)
    from geoswe_amd import CompressedSolver
    q0, bed = rough_state(NX, NY, seed=44, depth=0.7, rough=1.0, vel=0.3)
    s = dense(q0 * m, bed, friction="manning", manning_n=0.03, rainfall=30.0 / 3.6e6)
    s.set_inside_mask(m)
    plain = CompressedSolver.from_dense(s, cfl_linf=True, track_max_depth=True)
    plain.advance(3)
    d = str(tmp_path / "noF")
    plain.save_checkpoint(d)
    before = c.q_flat()
    with pytest.raises(RuntimeError, match="cumulative infiltration"):
        c.load_checkpoint(d)
    assert c.nsteps == 0 and np.array_equal(c.q_flat(), before)


def _native_rain(seed=61):
    rng = np.random.default_rng(seed)
    bi, bj = 7, 9
    pny = -(-NY // bj)
    npix = -(-NX // bi) * pny
    lk = (np.arange(NX)[:, None] // bi) * pny + (np.arange(NY)[None, :] // bj)
    t_s = np.array([0.0, 0.6, 1.3])
    mmh = rng.random((3, npix)) * 40.0
    return t_s, mmh, lk


@pytest.mark.parametrize("dtype,trav", [("float32", "tiles"), ("float32", "chunks"), ("float64", "tiles")])
def test_native_grid_rain_matches_oracle(gpu, dtype, trav):
    """The run driver's rain bundle (a rate table on a coarser native grid and each cell's
    column) read through the per-cell lookup: bitwise the oracle given the same rates spread
    onto the grid."""
    from geoswe_amd import CompressedSolver
    q0, bed = rough_state(NX, NY, seed=62, depth=0.7, rough=1.0, vel=0.3)
    m = irregular_mask(NX, NY, seed=9)
    q0 = q0 * m
    s = dense(q0, bed, friction="manning", manning_n=0.03, dtype=dtype)
    s.set_inside_mask(m)
    s.set_cfl_ghost_mask(~m)
    c = CompressedSolver.from_dense(s, cfl_linf=True, track_max_depth=True)
    t_s, mmh, lk = _native_rain()
    c.set_rain({"native_rate_dev": mmh / 3.6e6, "lookup_dev": lk, "t_s": t_s})
    c.traversal = trav
    n = c.run(T_END)
    case = O.DenseCase(h=q0[0], hu=q0[1], hv=q0[2], bed=bed, dx=2.0, dy=2.0, bc_x="fall", bc_y="fall",
                       manning=0.03, rain=(t_s, mmh[:, lk]), active=m)
    o = O.DenseOracle(case, _physics(c.cfg), np.dtype(dtype))
    o.cfl_skip = ~m
    o.run(T_END)
    print(dtype, trav, n, flush=True)
    assert o.nsteps == c.nsteps == n > 5 and o.t == c.t
    g = c.fm.ngh
    ij = c.fm.ij - g
    inner = (ij[:, 0] >= 0) & (ij[:, 0] < NX) & (ij[:, 1] >= 0) & (ij[:, 1] < NY)
    assert np.array_equal(c.q_flat()[:, inner], o.interior[:, ij[inner, 0], ij[inner, 1]])


def test_native_grid_rain_in_the_ensemble_and_the_cache(gpu, tmp_path):
    from geoswe_amd import CompressedSolver
    from geoswe_amd.cache import run_cached
    from geoswe_amd.ensemble import Ensemble
    q0, bed = rough_state(NX, NY, seed=63, depth=0.7, rough=1.0, vel=0.3)
    m = irregular_mask(NX, NY, seed=9)
    q0 = q0 * m
    t_s, mmh, lk = _native_rain(64)

    def mk(linf=True):
        s = dense(q0, bed, friction="manning", manning_n=0.03)
        s.set_inside_mask(m)
        s.set_cfl_ghost_mask(~m)
        c = CompressedSolver.from_dense(s, cfl_linf=linf, track_max_depth=True)
        c.set_rain({"native_rate_dev": mmh / 3.6e6, "lookup_dev": lk, "t_s": t_s})
        return c

    plain = [mk(), mk()]
    for p in plain:
        p.advance(15)
    mem = [mk(), mk()]
    ens = Ensemble(mem, graph_batch=8)
    ens.advance(15)
    for p, e in zip(plain, mem):
        assert p.nsteps == e.nsteps == 15 and np.array_equal(p.q_flat(), e.q_flat())
    ens.close()
    a = mk(linf=False)
    d = str(tmp_path / "cache")
    a.save_cache(d)
    a.run(2.5)
    b = run_cached(d, t_end=2.5, say=None)
    assert b.nsteps == a.nsteps and b.t == a.t and np.array_equal(b.q_flat(), a.q_flat())
    assert np.load(tmp_path / "cache" / "rain_native_rate_dev.npy").shape == mmh.shape


@pytest.mark.parametrize("mode,trav", [("auto", "tiles"), ("extrapolate", "tiles"), ("hybrid", "tiles"),
                                       ("stage_uv", "tiles"), ("stage_uv", "chunks")])
def test_ring_modes_match_oracle(gpu, mode, trav):
    """The cached replay's ring modes at the top of every step: active cells, steps and time
    bitwise the oracle's (which applies the same mode on its inactive ring cells, with the
    extrapolated cells' bed replicated)."""
    from geoswe_amd import CompressedSolver
    q0, bed = rough_state(NX, NY, seed=71, depth=0.7, rough=1.0, vel=0.3)
    m = irregular_mask(NX, NY, seed=9)
    q0 = q0 * m
    s = dense(q0, bed, friction="manning", manning_n=0.03, rainfall=20.0 / 3.6e6)
    s.set_inside_mask(m)
    s.set_cfl_ghost_mask(~m)
    c = CompressedSolver.from_dense(s, cfl_linf=True)
    eta = 0.45
    c.set_ring_mode(mode, eta=eta, classify=0.5 if mode == "hybrid" else None)
    ids, nbs, kind = c._rmode["host"]
    g = c.fm.ngh
    ij = c.fm.ij.astype(np.int64) - g
    bed2 = np.array(bed, np.float64)
    bf = c._bed.download()
    ring_inner = (ij[:, 0] >= 0) & (ij[:, 0] < NX) & (ij[:, 1] >= 0) & (ij[:, 1] < NY)
    bed2[ij[ring_inner, 0], ij[ring_inner, 1]] = bf[ring_inner]          # the replicated bed
    assert ring_inner[ids].all()
    rm = {"cells": ij[ids], "nb": np.where(nbs[:, None] >= 0, ij[np.maximum(nbs, 0)], -1), "kind": kind, "eta": eta}
    c.traversal = trav
    n = c.run(T_END)
    case = O.DenseCase(h=q0[0], hu=q0[1], hv=q0[2], bed=bed2.astype(np.float32), dx=2.0, dy=2.0, bc_x="fall",
                       bc_y="fall", manning=0.03, rain=20.0 / 3.6e6, active=m, ring_mode=rm)
    o = O.DenseOracle(case, _physics(c.cfg), np.float32)
    o.cfl_skip = ~m
    o.run(T_END)
    print(mode, trav, n, len(ids), flush=True)
    assert o.nsteps == c.nsteps == n > 5 and o.t == c.t
    act = c.fm.active
    ia = ij[act]
    assert np.array_equal(c.q_flat()[:, act], o.interior[:, ia[:, 0], ia[:, 1]])
    if mode != "extrapolate":
        assert (kind > 0).any()


def test_ring_mode_second_call_starts_from_the_packed_bed(gpu):
    from geoswe_amd import CompressedSolver
    q0, bed = rough_state(NX, NY, seed=72, depth=0.7, rough=1.0, vel=0.3)
    m = irregular_mask(NX, NY, seed=9)
    s = dense(q0 * m, bed, friction="manning", manning_n=0.03)
    s.set_inside_mask(m)
    c = CompressedSolver.from_dense(s, cfl_linf=True)
    b0 = c._bed.download()
    c.set_ring_mode("extrapolate")
    assert not np.array_equal(c._bed.download(), b0)
    c.set_ring_mode("auto", eta=0.5)
    assert np.array_equal(c._bed.download(), b0) and c._rmode["mode"] == "auto"
    c.set_ring_mode("off")
    assert c._rmode is None and np.array_equal(c._bed.download(), b0)
