"""Batched ensemble of uneven flat domains: every member bitwise equal to its own plain
CompressedSolver run (same steps, same per-member dt sequence), through direct launches,
hipGraph replays, a t_end run and after the launch-shape autotuner."""
import numpy as np
import pytest

from oracle import swe_oracle as O

pytestmark = pytest.mark.gpu

SIDES = [64, 192, 96, 128, 64]


@pytest.fixture(scope="module")
def cases(gpu):
    from geoswe_amd.cases import ensemble_c4
    return ensemble_c4(sides=SIDES, rain_frames=2)


def flats(cases):
    return [c.flat(track_max_depth=True) for c in cases]


def assert_same(plain, ens_members):
    for k, (p, e) in enumerate(zip(plain, ens_members)):
        assert p.nsteps == e.nsteps, k
        assert p.t == e.t, k
        assert np.array_equal(p.q_flat(), e.q_flat()), k
        assert np.array_equal(p.max_depth(), e.max_depth()), k


def test_members_differ(cases):
    """Per-member rain and roughness really differ, so per-member dt does too."""
    assert len({float(np.asarray(c.cfg.rainfall_forcing.rate_mm_h).max()) for c in cases}) == len(cases)
    assert len({tuple(np.round(c.ntab, 6)) for c in cases}) == len(cases)


@pytest.mark.parametrize("graph_batch", [0, 16])
def test_ensemble_equals_plain_runs(cases, graph_batch):
    from geoswe_amd.ensemble import Ensemble
    plain = flats(cases)
    for p in plain:
        p.advance(40)
    mem = flats(cases)
    ens = Ensemble(mem, graph_batch=graph_batch)
    ens.advance(40)
    assert_same(plain, mem)
    assert len({m.t for m in mem}) == len(mem)       # per-member dt
    ens.advance(7)                                    # graph replays resume after a partial batch
    for p in plain:
        p.advance(7)
    assert_same(plain, mem)
    ens.close()


def test_ensemble_launch_shapes_and_autotune(cases):
    from geoswe_amd.ensemble import Ensemble
    plain = flats(cases)
    for p in plain:
        p.advance(30)
    mem = flats(cases)
    ens = Ensemble(mem, graph_batch=8)
    got = ens.autotune(trial_steps=2)                 # restores the state it timed on
    assert got["blocks"] >= 1 and 1 <= got["tiles_per_pass"] <= 2
    assert all(m.nsteps == 0 for m in mem)
    ens.advance(10)
    ens.set_launch(blocks=3, tiles_per_pass=5)        # a small, odd grid: every tile still visited
    ens.advance(10)
    ens.set_launch(blocks=10 ** 6, tiles_per_pass=1)  # capped at one wavefront per tile
    ens.advance(5)
    ens.set_launch(blocks=0, tiles_per_pass=1)
    ens.advance(5, graph_batch=4)
    assert_same(plain, mem)
    ens.close()


def test_ensemble_tile_height_autotune(cases):
    """The autotuner re-tiles every member at each candidate height and keeps the fastest;
    the state it timed on is restored and the members stay bitwise their plain runs."""
    from geoswe_amd.ensemble import Ensemble
    plain = flats(cases)
    for p in plain:
        p.advance(25)
    mem = flats(cases)
    ens = Ensemble(mem, graph_batch=8)
    got = ens.autotune(trial_steps=2, heights=(8, 16, 32))
    assert got["sweep_rows"] in (8, 16, 32) and ens.sweep_rows == got["sweep_rows"]
    assert set(got["trials"]) == {8, 16, 32}
    assert all(m.nsteps == 0 for m in mem)
    ens.advance(15)
    ens.advance(10, graph_batch=0)
    assert_same(plain, mem)
    ens.close()


def test_ensemble_run_to_t_end(cases):
    from geoswe_amd.ensemble import Ensemble
    plain = flats(cases)
    steps_p = [p.run(t_end=12.5) for p in plain]
    mem = flats(cases)
    ens = Ensemble(mem, graph_batch=8)
    steps_e = ens.run(t_end=12.5, batch=16)
    assert steps_p == steps_e
    assert all(abs(m.t - 12.5) < 1e-9 for m in mem)
    assert_same(plain, mem)
    ens.close()


def test_ensemble_float64_equals_plain_runs(gpu):
    from geoswe_amd.cases import ensemble_c4
    from geoswe_amd.ensemble import Ensemble
    cases = ensemble_c4(sides=[64, 128, 96], rain_frames=2)
    for c in cases:
        c.cfg.dtype = "float64"
    plain = flats(cases)
    for p in plain:
        p.advance(25)
    mem = flats(cases)
    ens = Ensemble(mem, graph_batch=8)
    ens.advance(25)
    assert all(m.dtype == np.float64 for m in mem)
    assert_same(plain, mem)
    ens.close()


@pytest.mark.parametrize("target", [16, 1])
def test_ensemble_with_taller_tiles_than_members(cases, monkeypatch, target):
    """The ensemble sizes its tile height for all members together and builds its own tile
    tables where a member's are shorter; members' own tables stay as they are, and every
    member is still bitwise its plain run."""
    import geoswe_amd.flatmesh as FM
    import geoswe_amd.ensemble as EN
    from geoswe_amd.ensemble import Ensemble
    plain = flats(cases)
    for p in plain:
        p.advance(25)
    mem = flats(cases)
    own = [m.fm.tile_irc.copy() for m in mem]
    monkeypatch.setattr(FM, "TILE_TARGET", target)          # few tiles wanted -> tall tiles
    ens = Ensemble(mem, graph_batch=8)
    assert ens.sweep_rows == EN.sweep_rows_for(sum(int(t[:, 1].sum()) for t in own))
    if target == 1:
        assert ens.sweep_rows == FM.MAX_SWEEP_ROWS and len(ens._tiles) == len(mem)   # every member rebuilt
    assert all(np.array_equal(o, m.fm.tile_irc) for o, m in zip(own, mem))
    ens.advance(25)
    assert_same(plain, mem)
    ens.close()


def test_c4_generator_members_match_the_oracle(gpu):
    """Members of the batched-ensemble workload's generator at the small end of its size range
    (256^2 - 512^2 boxes, per-member rain scaling and Manning perturbation, gridded rain,
    tidal stage), stepped together as one ensemble: each member bit for bit the dense oracle
    of its own case on every active cell, with its own step count and time."""
    from geoswe_amd.cases import ensemble_c4
    from geoswe_amd.ensemble import Ensemble
    members = ensemble_c4(sides=[256, 384, 512], rain_frames=2)
    mem = [c.flat() for c in members]
    ens = Ensemble(mem, graph_batch=8)
    ens.advance(25)
    assert len({m.t for m in mem}) == len(mem)
    for case, m in zip(members, mem):
        sb, rf = case.cfg.stage_boundary, case.cfg.rainfall_forcing
        oc = O.DenseCase(h=case.q0[0], hu=case.q0[1], hv=case.q0[2], bed=case.bed, dx=3.0, dy=3.0, bc_x="fall",
                         bc_y="fall", manning=case.ntab.astype(np.float32)[case.cls], rain=(rf.time_s, rf.rate_mm_h),
                         active=case.active,
                         stage={"cells": np.argwhere(case.sea_edge), "time_s": sb.time_s, "stage_m": sb.stage_m})
        o = O.DenseOracle(oc, O.Physics(friction=True, flat_cfl=True), np.float32)
        o.cfl_skip = ~case.active
        o.run_steps(25)
        assert m.nsteps == 25 and m.t == o.t
        g = case.mesh.ngh
        got = m.fm.unpack(m.q_flat())[:, g:-g, g:-g]
        a = case.active
        assert np.array_equal(got[:, a], o.interior[:, a])
    ens.close()
