"""The float64 torch restatement used for the autograd check of the device adjoint
(tests/torch_ref.py) equals the NumPy oracle: residual on rough partially dry states with
several boundary kinds, and the update with friction, rain and the velocity cap."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import swe_oracle as O  # noqa: E402
import torch_ref as TR  # noqa: E402


def padded_state(nx, ny, seed, bc):
    rng = np.random.default_rng(seed)
    bed = rng.random((nx, ny)) * 1.2
    h = np.maximum(0.8 - bed, 0.0)
    wet = h > 1e-6
    q = np.zeros((3, nx + 8, ny + 8))
    q[0, 4:-4, 4:-4] = h
    q[1, 4:-4, 4:-4] = (rng.random((nx, ny)) - 0.5) * 0.4 * wet
    q[2, 4:-4, 4:-4] = (rng.random((nx, ny)) - 0.5) * 0.4 * wet
    O.bc_fill(q, 4, *bc)
    return q, O.pad_bed(bed, 4, *bc, dtype=np.float64)


@pytest.mark.parametrize("bc", [("wall", "wall"), ("fall", "extrapolate"), ("periodic", "fall")])
def test_torch_residual_equals_oracle(bc):
    q, b = padded_state(37, 29, 3, bc)
    want = O.residual(q, b, 2.0, 2.0, 9.81, 1e-6)
    c = O.kernel_consts(2.0, 2.0, 9.81, 1e-6, np.float64)
    got = TR.residual(torch.from_numpy(q), torch.from_numpy(b), float(c["inv_dx"]), float(c["inv_dy"]),
                      0.5 * float(c["dxk"]), 0.5 * float(c["dyk"]), 9.81, 1e-6).numpy()
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


def test_torch_update_equals_oracle():
    rng = np.random.default_rng(9)
    n = 5000
    h = 10 ** rng.uniform(-7, 0.5, n)
    q = np.stack([h, h * rng.uniform(-30, 30, n), h * rng.uniform(-30, 30, n)])
    r = rng.normal(size=(3, n)) * 0.01
    ncell = rng.uniform(0.01, 0.12, n)
    rain = 1e-5
    dt = 0.4
    ph = O.Physics(friction=True, vcap=15.0)
    want = O.update(q, r, dt, ph, n_cell=ncell, rain=rain)
    got = TR.update(torch.from_numpy(q), torch.from_numpy(r), dt, torch.from_numpy(ncell), rain, 1e-6, 9.81,
                    15.0).numpy()
    assert np.allclose(got, want, rtol=1e-12, atol=1e-15)
