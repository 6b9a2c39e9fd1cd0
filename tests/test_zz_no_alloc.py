"""The ensemble run that allocates nothing (gs_ens_run_noalloc, through
geoswe_amd.ensemble.run_noalloc) against the ensemble entry points that allocate
(gs_ens_create / gs_ens_run): two calls in a row, the second with fewer and smaller members,
each bitwise equal to an Ensemble of the same members over the same steps (state, running
maximum, time and step count of every member); and the refusal beyond the static capacity."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _assert_same(ref, got):
    for k, (a, b) in enumerate(zip(ref, got)):
        assert a.nsteps == b.nsteps, k
        assert a.t == b.t, k
        assert np.array_equal(a.q_flat(), b.q_flat()), k
        assert np.array_equal(a.max_depth(), b.max_depth()), k


def test_noalloc_run_equals_ensemble(gpu):
    from geoswe_amd.cases import ensemble_c4
    from geoswe_amd.ensemble import Ensemble, run_noalloc
    for sides, steps in (([96, 160, 64, 128], 30), ([64, 80], 12)):
        print(f"noalloc: members {sides}, {steps} steps", flush=True)
        cases = ensemble_c4(sides=sides, rain_frames=2)
        ref = [c.flat(track_max_depth=True) for c in cases]
        got = [c.flat(track_max_depth=True) for c in cases]
        ens = Ensemble(ref, graph_batch=0)
        ens.advance(steps)
        ens.close()
        run_noalloc(got, steps)
        _assert_same(ref, got)
        assert len({m.t for m in got}) == len(got)      # every member kept its own dt


def test_noalloc_refuses_beyond_capacity(gpu):
    from geoswe_amd import _lib as L
    from geoswe_amd.cases import ensemble_c4
    from geoswe_amd.ensemble import run_noalloc
    m = ensemble_c4(sides=[64], rain_frames=2)[0].flat()
    with pytest.raises(L.GsError, match="exceed"):
        run_noalloc([m] * 415, 1)
    assert m.nsteps == 0
