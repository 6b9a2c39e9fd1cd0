"""Pins the oracle's coastal forcings (open-boundary sponge, IDW gauge ring, stage clamp:
coastal_force, ring_depth) on the CPU.

The reference's tests hold no vectors for these forcings, so each check is one the
mathematics fixes: the identity fields, the full-absorption limit, the geometric series of
a repeated relaxation, the clamp's invariants (depth cap, velocity kept, untouched cells
bitwise), the sponge-before-clamp order, the ring's stage at and outside the sample times
and midway, its inverse-distance weights, its use of the step's start time and its place
between the sponge and the clamp, and the place of all of them in the step (after the
update and the stage).
"""
import numpy as np
import pytest

from oracle import swe_oracle as O


def _state(nx=9, ny=7, seed=3, dtype=np.float32):
    rng = np.random.default_rng(seed)
    h = (rng.random((nx, ny)) * 2.0).astype(dtype)
    hu = ((rng.random((nx, ny)) - 0.5) * h).astype(dtype)
    hv = ((rng.random((nx, ny)) - 0.5) * h).astype(dtype)
    return np.stack([h, hu, hv])


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_identity_fields_change_nothing(dtype):
    q = _state(dtype=dtype)
    q0 = q.copy()
    ones, zeros = np.ones(q.shape[1:]), np.zeros(q.shape[1:])
    O.coastal_force(q, sponge=(ones, zeros), clamp=np.full(q.shape[1:], np.inf))
    assert np.array_equal(q, q0) and q.dtype == dtype


def test_full_absorption_sets_ambient_depth_at_rest():
    """keep = 0: the cell becomes the ambient state exactly (h = amb, no momentum)."""
    q = _state()
    keep = np.ones(q.shape[1:])
    amb = np.zeros(q.shape[1:])
    keep[:2] = 0.0
    amb[:2] = 0.75
    q_out = q.copy()
    O.coastal_force(q_out, sponge=(keep, amb))
    assert np.all(q_out[0, :2] == np.float32(0.75))
    assert np.all(q_out[1:, :2] == 0)
    assert np.array_equal(q_out[:, 2:], q[:, 2:])          # off the band: untouched


def test_repeated_relaxation_is_the_geometric_series():
    """n applications of h <- k h + a from h0 give k^n h0 + a (1 - k^n) / (1 - k), momentum k^n."""
    q = _state(dtype=np.float64)
    h0, hu0 = q[0].copy(), q[1].copy()
    k, a, n = 0.9, 0.02, 25
    keep = np.full(q.shape[1:], k)
    amb = np.full(q.shape[1:], a)
    for _ in range(n):
        O.coastal_force(q, sponge=(keep, amb))
    want = k ** n * h0 + a * (1 - k ** n) / (1 - k)
    assert np.allclose(q[0], want, rtol=1e-13, atol=1e-15)
    assert np.allclose(q[1], k ** n * hu0, rtol=1e-13, atol=1e-15)


def test_sponge_rounds_multiply_then_add():
    """float32: h*keep is rounded before amb is added (no fused multiply-add)."""
    h = np.float32(1.0000001)
    kp = np.float32(0.99999994)
    am = np.float32(-0.99999994)
    two = np.float32(h * kp) + am
    fused = np.float32(np.float64(h) * np.float64(kp) + np.float64(am))
    assert two != fused                                      # the case tells the two apart
    q = np.array([[[h]], [[0.0]], [[0.0]]], np.float32)
    O.coastal_force(q, sponge=(np.array([[kp]]), np.array([[am]])))
    assert q[0, 0, 0] == two


def test_clamp_caps_depth_and_keeps_velocity():
    q = _state(dtype=np.float64)
    q0 = q.copy()
    hmax = np.full(q.shape[1:], np.inf)
    hmax[:, :3] = 0.8
    O.coastal_force(q, clamp=hmax)
    over = (q0[0] > 0.8) & (hmax < np.inf)
    assert over.any() and (~over & (hmax < np.inf)).any()
    assert np.all(q[0][over] == 0.8)
    for m in (1, 2):    # velocity hu/h unchanged to rounding
        assert np.allclose(q[m][over] / q[0][over], q0[m][over] / q0[0][over], rtol=1e-14)
    assert np.array_equal(q[:, ~over], q0[:, ~over])         # at or under the cap: bitwise


def test_sponge_comes_before_clamp():
    """A sponge that lifts a cell over its cap leaves it exactly at the cap; the other order
    would leave h_max*keep + amb."""
    q = np.array([[[0.5]], [[0.1]], [[0.0]]], np.float32)
    O.coastal_force(q, sponge=(np.array([[0.5]]), np.array([[1.0]])), clamp=np.array([[1.1]]))
    assert q[0, 0, 0] == np.float32(1.1)
    assert q[1, 0, 0] == np.float32(np.float32(0.1 * np.float32(0.5)) * np.float32(np.float32(1.1) / np.float32(1.25)))


def test_clamp_only_cell_gets_the_identity_sponge():
    """A clamp cell outside the band still takes h*1 + 0, which turns a -0 depth into +0."""
    q = np.array([[[-0.0, -0.0]], [[0.0, 0.0]], [[0.0, 0.0]]], np.float32)
    O.coastal_force(q, sponge=(np.ones((1, 2)), np.zeros((1, 2))), clamp=np.array([[1.0, np.inf]]))
    assert not np.signbit(q[0, 0, 0]) and np.signbit(q[0, 0, 1])


def _case(**kw):
    nx, ny = 12, 10
    rng = np.random.default_rng(5)
    bed = rng.random((nx, ny)) * 0.3
    h = np.maximum(0.6 - bed, 0.0)
    return O.DenseCase(h=h, hu=0.05 * h, hv=np.zeros_like(h), bed=bed, dx=2.0, dy=2.0, bc_x="fall",
                       bc_y="fall", **kw)


def test_step_applies_forcing_after_update_and_stage():
    """Full absorption on rows 0-1 holds them at the ambient depth after every step, also on
    a stage cell (the stage is written first, then the sponge)."""
    nx, ny = 12, 10
    keep = np.ones((nx, ny))
    amb = np.zeros((nx, ny))
    keep[:2] = 0.0
    amb[:2] = 0.4
    stage = {"cells": np.array([[0, 3], [5, 5]]), "time_s": np.array([0.0, 10.0]), "stage_m": np.array([1.0, 1.0])}
    o = O.DenseOracle(_case(sponge=(keep, amb), stage=stage), O.Physics(friction=False), np.float32)
    plain = O.DenseOracle(_case(stage=stage), O.Physics(friction=False), np.float32)
    for _ in range(4):
        dt = o.next_dt()
        o.step(dt)
        plain.step(dt)
        q = o.interior
        assert np.all(q[0, :2] == np.float32(0.4)) and np.all(q[1:, :2] == 0)
    assert q[0, 5, 5] == plain.interior[0, 5, 5]     # a stage cell off the band keeps the stage depth
    assert not np.array_equal(q[:, 2:4], plain.interior[:, 2:4])   # the band's neighbours feel it


def test_step_clamp_caps_the_updated_depth():
    nx, ny = 12, 10
    hmax = np.full((nx, ny), np.inf)
    hmax[4:8, 4:8] = 0.45
    o = O.DenseOracle(_case(clamp=hmax), O.Physics(friction=False), np.float32)
    for _ in range(3):
        o.step(o.next_dt())
        assert np.all(o.interior[0, 4:8, 4:8] <= np.float32(0.45))


def _ring(cells, bed, w, tc, S):
    return {"cells": np.asarray(cells), "bed": np.asarray(bed, float), "w": np.asarray(w, float),
            "t_common": np.asarray(tc, float), "stage_all": np.asarray(S, float)}


def test_ring_depth_at_and_outside_the_sample_times():
    S = np.array([[0.5, 1.5, 0.9, 2.0], [1.0, 0.2, 0.4, 3.0]])
    tc = [0.0, 360.0, 720.0, 1080.0]
    r = _ring([[0, 0], [1, 1]], [0.1, 0.3], [[1.0, 0.0], [0.0, 1.0]], tc, S)
    # first sample, and before it (weight clamped to 0): the first column exactly
    for t in (0.0, -50.0):
        assert np.array_equal(O.ring_depth(r, t, np.float32), np.float32(S[[0, 1], 0]) - np.float32([0.1, 0.3]))
    # past the last sample (weight clamped to 1): the last column exactly
    assert np.array_equal(O.ring_depth(r, 5000.0, np.float32), np.float32(S[[0, 1], 3]) - np.float32([0.1, 0.3]))
    # midway between two samples: the mean (float64)
    h = O.ring_depth(r, 540.0, np.float64)
    assert np.allclose(h, [(1.5 + 0.9) / 2 - 0.1, (0.2 + 0.4) / 2 - 0.3], rtol=1e-12)


def test_ring_depth_idw_weights_and_dry_cells():
    S = np.array([[1.0, 1.0], [1.0, 1.0], [1.0, 1.0]])
    w = np.array([[0.2, 0.3, 0.5], [0.6, 0.3, 0.1]])
    r = _ring([[0, 0], [0, 1]], [0.25, 1.5], w, [0.0, 600.0], S)
    h = O.ring_depth(r, 100.0, np.float64)
    assert np.isclose(h[0], 0.75, rtol=1e-14) and h[1] == 0.0         # above the stage: dry
    r["bed"] = np.array([np.nan, 0.0])
    assert O.ring_depth(r, 100.0, np.float32)[0] == 0.0                # fmax(0, NaN) = 0


def test_ring_uses_the_step_start_time_and_sits_between_sponge_and_clamp():
    nx, ny = 12, 10
    cells = [[0, 2], [0, 3], [0, 4]]
    ring = _ring(cells, [0.0, 0.0, 0.0], [[1.0], [1.0], [1.0]], [0.0, 100.0], [[1.0, 11.0]])
    keep = np.ones((nx, ny))
    keep[0, 3] = 0.5                       # a band cell of the ring: the ring overwrites it
    hmax = np.full((nx, ny), np.inf)
    hmax[0, 4] = 0.8                       # a clamped ring cell: capped after the ring
    case = _case(active=np.pad(np.ones((nx - 2, ny), bool), ((2, 0), (0, 0))), ring=ring,
                 sponge=(keep, np.zeros((nx, ny))), clamp=hmax)
    case.h[:2] = 0.0
    case.hu[:2] = 0.0
    o = O.DenseOracle(case, O.Physics(friction=False), np.float32)
    o.cfl_skip = ~case.active
    dt1 = o.next_dt()
    o.step(dt1)
    assert np.all(o.interior[0, 0, 2:4] == np.float32(1.0)) and o.interior[0, 0, 4] == np.float32(0.8)
    assert np.all(o.interior[1:, 0, 2:5] == 0)
    o.step(o.next_dt())
    want = np.float32(np.float32(1.0) * np.float32(1 - dt1 / 100.0) + np.float32(11.0) * np.float32(dt1 / 100.0))
    assert o.interior[0, 0, 2] == want != np.float32(1.0)


def _one(h, hu, hv=0.0, dtype=np.float32):
    return np.array([[[h]], [[hu]], [[hv]]], dtype)


def test_infiltration_closed_forms():
    cls = np.zeros((1, 1), int)
    # partial loss: h = 1, rate*dt = 0.25 -> 0.75 exactly, momentum scaled by 0.75
    q = _one(1.0, 0.5)
    O.coastal_force(q, infil=(cls, [0.125], 2.0, 1e-6, None))
    assert q[0, 0, 0] == np.float32(0.75) and q[1, 0, 0] == np.float32(0.375)
    # more than the depth: dry, no momentum
    q = _one(0.1, 0.02, -0.01)
    O.coastal_force(q, infil=(cls, [1.0], 1.0, 1e-6, None))
    assert q[0, 0, 0] == 0 and q[1, 0, 0] == 0 and q[2, 0, 0] == 0
    # below h_min: untouched; zero rate: bitwise identity; inactive: untouched
    for h, rate, act in ((5e-7, 1.0, None), (0.3, 0.0, None), (0.3, 1.0, np.zeros((1, 1), bool))):
        q = _one(h, 1e-3)
        q0 = q.copy()
        O.coastal_force(q, infil=(cls, [rate], 0.1, 1e-6, act))
        assert np.array_equal(q, q0)


def test_infiltration_by_class_comes_before_the_sponge():
    cls = np.array([[0, 1]])
    q = np.array([[[1.0, 1.0]], [[0.0, 0.0]], [[0.0, 0.0]]], np.float32)
    O.coastal_force(q, sponge=(np.full((1, 2), 0.5), np.full((1, 2), 0.25)), infil=(cls, [0.0, 0.5], 1.0, 1e-6, None))
    assert q[0, 0, 0] == np.float32(0.75)                 # class 0: no loss, then sponge
    assert q[0, 0, 1] == np.float32(0.5)                  # (1 - 0.5)*0.5 + 0.25; the other order gives 0.25


def test_drain_caps_with_its_tiny_depth_guard_after_the_clamp():
    q = _one(2.0, 1.0)
    O.coastal_force(q, drain=np.array([[0.5]]))
    assert q[0, 0, 0] == np.float32(0.5) and q[1, 0, 0] == np.float32(0.25)
    q = _one(5e-13, 1e-13)                                 # above the target but under 1e-12: r = 0
    O.coastal_force(q, drain=np.array([[1e-13]]))
    assert q[0, 0, 0] == np.float32(1e-13) and q[1, 0, 0] == 0
    q = _one(0.3, 0.1)
    q0 = q.copy()
    O.coastal_force(q, drain=np.array([[0.5]]))            # under the target: untouched
    assert np.array_equal(q, q0)
    # clamp (0.5) then drain (0.8): the momentum is scaled once, by 0.5/h
    h, hu = np.float32(1.9748254), np.float32(1.905881)
    q = _one(h, hu)
    O.coastal_force(q, clamp=np.array([[0.5]]), drain=np.array([[0.8]]))
    once = np.float32(hu * np.float32(np.float32(0.5) / h))
    twice = np.float32(np.float32(hu * np.float32(np.float32(0.8) / h)) * np.float32(np.float32(0.5) / np.float32(0.8)))
    assert once != twice and q[1, 0, 0] == once and q[0, 0, 0] == np.float32(0.5)


def test_exp_ieee_within_two_ulp_of_numpy_exp():
    rng = np.random.default_rng(11)
    x = np.concatenate([np.linspace(-745.0, 709.0, 200001), rng.uniform(-50.0, 0.0, 100000),
                        rng.uniform(-1e-3, 1e-3, 10000), [0.0, -0.0, 1.0, -1.0, np.log(2.0)]])
    got, want = O.exp_ieee(x), np.exp(x)
    ok = want > 0
    ulps = np.abs(got[ok] - want[ok]) / np.spacing(want[ok])
    assert ulps.max() <= 2.0, ulps.max()
    assert O.exp_ieee(np.array([0.0]))[0] == 1.0
    sp = O.exp_ieee(np.array([-np.inf, -800.0, np.inf, 800.0, np.nan]))
    assert sp[0] == 0 and sp[1] == 0 and sp[2] == np.inf and sp[3] == np.inf and np.isnan(sp[4])


def _ga_state(dtype=np.float64):
    rng = np.random.default_rng(21)
    nx, ny = 6, 5
    h = (rng.random((nx, ny)) * 0.3 + 0.01).astype(dtype)
    hu = ((rng.random((nx, ny)) - 0.5) * h).astype(dtype)
    hv = ((rng.random((nx, ny)) - 0.5) * h).astype(dtype)
    return np.stack([h, hu, hv]), rng.integers(0, 3, (nx, ny)), (rng.random((nx, ny)) * 0.02).astype(dtype)


def test_green_ampt_step_solves_its_quadratic_and_conserves_water():
    q, cls, F = _ga_state()
    q0, F0 = q.copy(), F.copy()
    Ks, psi, dth, dt = np.array([1e-5, 3e-6, 0.0]), np.array([0.11, 0.05, 0.2]), np.array([0.3, 0.25, 0.4]), 2.0
    O.green_ampt_drain(q, F, {"cls": cls, "Ks": Ks, "psi": psi, "dth": dth}, dt)
    K, head = Ks[cls], psi[cls] + q0[0]
    a = F0 + K * dt
    F1 = F                                                      # no cap, dF < h here
    g = K > 0
    assert np.allclose((F1 * (F1 - a))[g], (K * dt * head * dth[cls])[g], rtol=1e-10)
    assert np.allclose(q[0] + F, q0[0] + F0, rtol=1e-14, atol=0)          # water moves into the soil
    assert np.array_equal(q[:, ~g], q0[:, ~g]) and np.array_equal(F[~g], F0[~g])   # K = 0: untouched
    for m in (1, 2):
        assert np.allclose(q[m] / q[0], q0[m] / q0[0], rtol=1e-14)


def test_green_ampt_caps_and_drain_tau_decay():
    q, cls, F = _ga_state()
    q0, F0 = q.copy(), F.copy()
    Fmax = F0 + 1e-7
    Fmax[0, 0] = F0[0, 0] - 1.0                                # no room left: nothing infiltrates
    inv_tau = np.full(F.shape, 1.0 / 3600.0)
    inv_tau[1, :] = 0.0
    dt = 30.0
    O.green_ampt_drain(q, F, {"cls": cls, "Ks": np.array([1e-4, 1e-4, 1e-4]), "psi": np.full(3, 0.1),
                              "dth": np.full(3, 0.3), "Fmax": Fmax, "inv_tau": inv_tau}, dt)
    assert np.all((F <= Fmax + 1e-18) | (F == F0)) and F[0, 0] == F0[0, 0]
    decay = np.exp(-dt / 3600.0)
    dF = F - F0
    assert np.allclose(q[0][2:], ((q0[0] - dF) * decay)[2:], rtol=1e-14)
    assert np.allclose(q[0][1], (q0[0] - dF)[1], rtol=1e-15)          # inv_tau = 0: no decay
    # drain alone: h * exp(-dt/tau) exactly as exp_ieee gives it, momentum by the same factor
    q2, _, F2 = _ga_state()
    q20 = q2.copy()
    O.green_ampt_drain(q2, F2, {"cls": cls, "Ks": None, "inv_tau": inv_tau}, dt)
    d = O.exp_ieee(np.array([-dt * (1.0 / 3600.0)]))[0]
    assert np.array_equal(q2[:, 2:], q20[:, 2:] * d)


def test_green_ampt_sits_between_ring_and_clamp():
    """A clamp above the Green-Ampt result leaves it alone; clamping first would cap and then
    infiltrate, ending below the cap."""
    q = np.array([[[0.5]], [[0.1]], [[0.0]]], np.float32)
    F = np.zeros((1, 1), np.float32)
    ga = {"cls": np.zeros((1, 1), int), "Ks": [1e-3], "psi": [0.1], "dth": [0.3]}
    qa = q.copy()
    O.green_ampt_drain(qa, F.copy(), ga, 10.0)
    hmax = float(qa[0, 0, 0]) + 0.01
    O.coastal_force(q, clamp=np.array([[hmax]]), ga=(ga, F, 10.0, None))
    assert q[0, 0, 0] == qa[0, 0, 0] and np.float32(hmax) < np.float32(0.5)


def test_ring_modes_closed_forms():
    """Kind 0 copies the neighbour exactly; kind 1 is still water at eta over the cell's bed;
    kind 2 is that depth with the neighbour's velocity (none from a dry neighbour)."""
    q = np.zeros((3, 4, 4), np.float32)
    q[:, 1, 1] = [0.8, 0.16, -0.08]
    q[:, 2, 2] = [1e-7, 1e-8, 0.0]                       # a dry neighbour (below h_min)
    b = np.zeros((4, 4))
    b[0, 1] = 0.2
    b[3, 3] = 0.9
    rm = {"cells": np.array([[1, 0], [0, 1], [3, 3], [1, 2], [2, 3]]),
          "nb": np.array([[1, 1], [1, 1], [-1, -1], [1, 1], [2, 2]]), "kind": np.array([0, 2, 1, 2, 2]), "eta": 0.5}
    O.ring_mode_apply(q, b, rm, 1e-6)
    assert np.array_equal(q[:, 1, 0], q[:, 1, 1])
    assert q[0, 0, 1] == np.float32(0.5) - np.float32(0.2)
    assert np.isclose(q[1, 0, 1] / q[0, 0, 1], 0.16 / 0.8, rtol=1e-6)
    assert np.isclose(q[2, 0, 1] / q[0, 0, 1], -0.08 / 0.8, rtol=1e-6)
    assert q[0, 3, 3] == 0 and q[1, 3, 3] == 0                          # bed above eta: dry
    assert q[0, 1, 2] == np.float32(0.5) and q[1, 2, 3] == 0 and q[0, 2, 3] == np.float32(0.5)


def test_exp_ieee_f32_within_two_ulp_of_exp():
    rng = np.random.default_rng(12)
    x = np.concatenate([np.linspace(-87.0, 88.0, 300001), rng.uniform(-30.0, 0.0, 100000),
                        rng.uniform(-1e-3, 1e-3, 10000), [0.0, -1.0, 1.0]]).astype(np.float32)
    got = O.exp_ieee_f32(x).astype(np.float64)
    want = np.exp(x.astype(np.float64))
    ulps = np.abs(got - want) / np.spacing(want.astype(np.float32)).astype(np.float64)
    assert ulps.max() <= 2.0, ulps.max()
    assert O.exp_ieee_f32(np.float32([0.0]))[0] == 1.0
    sp = O.exp_ieee_f32(np.float32([-np.inf, -200.0, np.inf, 100.0, np.nan]))
    assert sp[0] == 0 and sp[1] == 0 and sp[2] == np.inf and sp[3] == np.inf and np.isnan(sp[4])
    d = O.exp_ieee_f32(np.float32([-95.0, -103.0]))                    # subnormal results, rounded once
    assert np.all(np.abs(d.astype(np.float64) - np.exp([-95.0, -103.0])) <= 1.5 * 2.0 ** -149)
