"""The device adjoint of the gauge misfit (float64 build) against autograd through the
float64 torch restatement of the step, against central finite differences of the device
forward, and through the torch.autograd.Function wrapper."""
import numpy as np
import pytest

import adjoint_cases as AC

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def problem(gpu):
    return AC.coastal_problem(n=48, S=40, K=16, dtype="float64")


def test_forward_matches_torch_and_adjoint_matches_autograd(problem):
    torch = pytest.importorskip("torch")
    case, prob = problem
    n0, b0 = prob.params()
    prob.simulate()
    J = prob.loss()
    gn, gb = prob.gradient()
    nt = torch.tensor(n0, dtype=torch.float64, requires_grad=True)
    bt = torch.tensor(b0, dtype=torch.float64, requires_grad=True)
    Jt, samples = AC.torch_misfit(case, prob, nt, bt)
    assert np.allclose(samples, prob.samples, rtol=1e-10, atol=1e-13)
    assert abs(float(Jt.detach()) - J) <= 1e-9 * J and J > 0
    Jt.backward()
    tn, tb = nt.grad.numpy(), bt.grad.numpy()
    assert np.abs(tn).max() > 0 and np.abs(tb).max() > 0
    assert np.abs(gn - tn).max() <= 1e-7 * np.abs(tn).max(), (gn, tn)
    assert np.linalg.norm(gb - tb) <= 1e-7 * np.linalg.norm(tb)
    assert np.abs(gb - tb).max() <= 1e-6 * np.abs(tb).max()


@pytest.fixture(scope="module")
def flowing(gpu):
    return AC.flowing_problem()


def test_adjoint_matches_central_differences(flowing):
    prob = flowing
    n0, b0 = prob.params()
    prob.simulate()
    gn, gb = prob.gradient()

    def J(n=None, b=None):
        prob.set_params(ntab=n if n is not None else n0, bed=b if b is not None else b0)
        prob.simulate(keep_checkpoints=False)
        return prob.loss()

    eps = 1e-6
    for k in range(len(n0)):
        e = np.zeros_like(n0)
        e[k] = eps
        fd = (J(n=n0 + e) - J(n=n0 - e)) / (2 * eps)
        assert abs(gn[k] - fd) <= 1e-5 * max(abs(fd), 1e-3 * np.abs(gn).max()), (k, gn[k], fd)
    # bed cells where the gradient is largest (near the gauges) and a few seeded ones, away
    # from exact bed ties with a neighbour (a frozen ghost ring copies the edge bed, so
    # b_R - b_L = 0 exactly there: a kink of max(b_R - b_L, 0) whose one-sided derivatives
    # differ and which a central difference averages)
    from geoswe_amd.adjoint import flat_neighbours
    nb = flat_neighbours(prob.s.fm)
    gap = np.min(np.where(nb >= 0, np.abs(b0[:, None] - b0[np.maximum(nb, 0)]), np.inf), axis=1)
    ok = gap > 1e-3
    rng = np.random.default_rng(3)
    order = [c for c in np.argsort(-np.abs(gb)) if ok[c]][:8]
    cells = np.unique(np.concatenate([order, rng.choice(np.flatnonzero(prob.s.fm.active & ok), 4,
                                                        replace=False)]))
    scale = np.abs(gb).max()
    compared = 0
    for c in cells:
        e = np.zeros_like(b0)
        e[c] = eps
        fd = (J(b=b0 + e) - J(b=b0 - e)) / (2 * eps)
        fd4 = (J(b=b0 + e / 4) - J(b=b0 - e / 4)) / (eps / 2)
        # the difference quotient carries ~1e-9 of float64 rounding of J; a state-dependent
        # kink (a wet/dry switch) within +-eps shows as fd != fd4 beyond that
        ref = max(abs(fd), 1e-2 * scale)
        if abs(fd - fd4) > 1e-4 * ref:
            continue
        compared += 1
        assert abs(gb[c] - fd) <= 1e-4 * ref, (c, gb[c], fd)
    assert compared >= 10, compared
    prob.set_params(ntab=n0, bed=b0)


def test_autograd_function_wrapper(problem):
    torch = pytest.importorskip("torch")
    from geoswe_amd.adjoint import GaugeMisfitFn
    case, prob = problem
    n0, b0 = prob.params()
    nt = torch.tensor(n0, requires_grad=True)
    bt = torch.tensor(b0, requires_grad=True)
    J = GaugeMisfitFn.apply(nt, bt, prob)
    (2.0 * J).backward()
    prob.simulate()
    gn, gb = prob.gradient()
    assert np.allclose(nt.grad.numpy(), 2.0 * gn, rtol=1e-12, atol=0)
    assert np.allclose(bt.grad.numpy(), 2.0 * gb, rtol=1e-12, atol=0)


def test_gradient_independent_of_checkpoint_interval(gpu):
    """Recomputation is deterministic: the same gradient, bit for bit, for any checkpoint
    interval (one segment, uneven segments, one step per segment)."""
    from geoswe_amd.adjoint import GaugeMisfit
    prob = AC.flowing_problem(n=32, S=24, K=24)
    prob.simulate()
    ref_n, ref_b = prob.gradient()
    for K in (7, 1, 5):
        other = GaugeMisfit(prob.s, prob.gauges, prob.dts, h_star=prob.h_star, checkpoint_every=K)
        other.simulate()
        gn, gb = other.gradient()
        assert np.array_equal(gn, ref_n) and np.array_equal(gb, ref_b), K


def test_gradient_with_native_grid_rain(gpu):
    """Rain read through a per-cell lookup into a native-grid table (the run driver's bundle)
    reaches the adjoint's recomputation as it reaches the forward run: the misfit and its
    gradient equal, bit for bit, those of the same rates given as full (nx, ny) frames."""
    from geoswe_amd import CompressedSolver, Config, Mesh2D, NativeRain, RainfallForcing, Solver2D
    from geoswe_amd.adjoint import GaugeMisfit
    n = 36
    i = np.arange(n)[:, None].astype(np.float64)
    j = np.arange(n)[None, :].astype(np.float64)
    bed = 0.3 * np.sin(i / 6.0) * np.cos(j / 5.0) + 0.004 * i
    h = 1.2 - bed + 0.05 * np.sin(i / 4.0 + j / 7.0)
    q0 = np.stack([h, h * 0.3, h * -0.2])
    t_s = np.array([0.0, 0.4, 0.9])
    lk = (np.arange(n)[:, None] // 6) * 6 + (np.arange(n)[None, :] // 6)
    mmh = np.random.default_rng(2).random((3, 36)) * 400.0
    frames = np.stack([mmh[k][lk] for k in range(3)])

    def solver(native):
        cfg = Config(bc_x="extrapolate", bc_y="extrapolate", dtype="float64", friction="manning", manning_n=0.03,
                     h_min=1e-6)
        s = Solver2D(Mesh2D(n, n, 2.0, 2.0, ngh=4), cfg, q0, bed)
        cls = np.zeros((n + 8, n + 8), np.uint8)
        cls[4:-4, 4:-4] = ((i + 2 * j) % 4).astype(np.uint8)
        s.set_manning_table(cls, np.array([0.02, 0.035, 0.05, 0.07]))
        c = CompressedSolver.from_dense(s, cfl_linf=True, track_max_depth=False)
        c.set_rain(NativeRain(time_s=t_s, table=mmh / 3.6e6, lookup=lk) if native
                   else RainfallForcing(time_s=t_s, rate_mm_h=frames))
        return c

    ref = solver(False)
    dts = np.array([ref.step() for _ in range(20)])
    assert ref.t > t_s[-1]                                    # the run crosses both rain changes
    out = []
    for native in (True, False):
        f = solver(native)
        gauges = np.sort(np.random.default_rng(4).choice(np.flatnonzero(f.fm.active), 6, replace=False))
        prob = GaugeMisfit(f, gauges, dts, h_star=np.full((20, 6), 1.0), checkpoint_every=7)
        prob.simulate()
        out.append((prob.loss(), *prob.gradient()))
    (Ja, na, ba), (Jb, nb, bb) = out
    assert Ja == Jb and np.array_equal(na, nb) and np.array_equal(ba, bb)
    assert np.abs(nb).max() > 0


def test_float32_adjoint_against_float64_and_torch(gpu):
    """The float32 build's gradient (the precision of the calibration benchmark) against the
    float64 build on the same problem (the float32 run's inputs, dt sequence, gauges and
    truth depths) and against float64 autograd through the torch restatement: the coastal
    generator at 128x128 over 200 steps, 32 gauges, 8 Manning classes."""
    torch = pytest.importorskip("torch")
    from geoswe_amd.adjoint import GaugeMisfit
    case, p32 = AC.coastal_problem(n=128, S=200, K=50, dtype="float32", n_gauges=32)
    n0, b0 = p32.params()
    p32.simulate()
    gn32, gb32 = p32.gradient()
    case.cfg.dtype = "float64"
    s64 = case.flat(track_max_depth=False)
    p64 = GaugeMisfit(s64, p32.gauges, p32.dts, h_star=p32.h_star, checkpoint_every=50)
    p64.set_params(ntab=n0, bed=b0)
    p64.simulate()
    gn64, gb64 = p64.gradient()
    assert np.abs(p32.samples - p64.samples).max() <= 1e-5 * np.abs(p64.samples).max()
    nt = torch.tensor(n0, dtype=torch.float64, requires_grad=True)
    bt = torch.tensor(b0, dtype=torch.float64, requires_grad=True)
    Jt, _ = AC.torch_misfit(case, p64, nt, bt)
    Jt.backward()
    tn, tb = nt.grad.numpy(), bt.grad.numpy()
    assert np.abs(gn64 - tn).max() <= 1e-7 * np.abs(tn).max()
    assert np.linalg.norm(gb64 - tb) <= 1e-7 * np.linalg.norm(tb)
    print(f"float32 vs float64: class n max rel {np.abs(gn32 - gn64).max() / np.abs(gn64).max():.3e}, "
          f"bed normwise {np.linalg.norm(gb32 - gb64) / np.linalg.norm(gb64):.3e}")
    assert np.abs(gn32 - gn64).max() <= 1e-2 * np.abs(gn64).max()
    assert np.linalg.norm(gb32 - gb64) <= 5e-2 * np.linalg.norm(gb64)
