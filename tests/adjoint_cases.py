"""Shared set-up of the gauge-misfit adjoint checks (test infrastructure): a coastal flat
case with a water mound on land so that flow develops within a few dozen steps, a frozen
dt sequence from a CFL run, seeded gauges among cells wet at t=0, and a truth run with
perturbed Manning classes."""
import bisect

import numpy as np


def coastal_problem(n=48, S=40, K=16, dtype="float64", n_gauges=8, seed=11, mound=True):
    from geoswe_amd.adjoint import GaugeMisfit
    from geoswe_amd.cases import coastal_c3
    case = coastal_c3(n=n, band=max(4, n // 24), n_sites=16, rain_frames=2)
    case.cfg.dtype = dtype
    if mound:
        i = np.arange(n)[:, None]
        j = np.arange(n)[None, :]
        disc = np.hypot(i - 0.6 * n, j - 0.45 * n) < 0.2 * n
        land = case.active & (case.bed > 0.0) & disc
        h = case.q0[0].astype(np.float64)
        h[land] = np.maximum(h[land], (case.bed[land] + 1.0) - case.bed[land])
        case.q0[0] = h.astype(case.q0.dtype)
    ref = case.flat(track_max_depth=False)
    dts = np.array([ref.step() for _ in range(S)])
    f = case.flat(track_max_depth=False)
    fm = f.fm
    h0 = f.q_flat()[0]
    cand = np.flatnonzero(fm.active & (h0 > 0.05))
    rng = np.random.default_rng(seed)
    gauges = np.sort(rng.choice(cand, size=min(n_gauges, len(cand)), replace=False))
    prob = GaugeMisfit(f, gauges, dts, checkpoint_every=K)
    n0 = prob.params()[0]
    n_true = n0 * (1.0 + 0.25 * rng.uniform(-1.0, 1.0, len(n0)))
    prob.set_params(ntab=n_true)
    prob.h_star = prob.simulate(keep_checkpoints=False).copy()
    prob.set_params(ntab=n0)
    return case, prob


def torch_misfit(case, prob, ntab, bed_flat):
    """J(ntab, bed_flat) through S steps of tests/torch_ref.step in float64 on the padded box."""
    import torch
    import torch_ref as TR
    from geoswe_amd import _lib as L
    from oracle import swe_oracle as O
    d = case.dense()
    fm = prob.s.fm
    g = case.mesh.ngh
    q = torch.from_numpy(d.q.astype(np.float64))
    b_base = torch.from_numpy(d.b.astype(np.float64))
    I = torch.from_numpy(fm.ij[:, 0].astype(np.int64))
    J = torch.from_numpy(fm.ij[:, 1].astype(np.int64))
    b = b_base.index_put((I, J), bed_flat)
    flags = fm.unpack(prob.s._flags.download(), fill=0)
    cls = torch.from_numpy((flags & L.FLAG_CLASS_MASK).astype(np.int64))
    n_cell = ntab[cls]
    active = torch.from_numpy(fm.unpack(fm.active.astype(np.uint8), fill=0) > 0)
    stage = torch.from_numpy((flags & L.FLAG_STAGE) > 0)
    c = O.kernel_consts(case.mesh.dx, case.mesh.dy, 9.81, case.cfg.h_min, np.float64)
    consts = dict(ix=float(c["inv_dx"]), iy=float(c["inv_dy"]), hdx=0.5 * float(c["dxk"]),
                  hdy=0.5 * float(c["dyk"]), g=9.81, hm=case.cfg.h_min, vcap=case.cfg.friction_velocity_cap_ms)
    rf = case.cfg.rainfall_forcing
    sb = case.cfg.stage_boundary
    r = np.asarray(rf.rate_mm_h)
    frames = np.stack([(r[k] / 3.6e6).astype(np.float64) for k in range(r.shape[0])])   # as rain_tables
    gi = torch.from_numpy(fm.ij[prob.gauges, 0].astype(np.int64))
    gj = torch.from_numpy(fm.ij[prob.gauges, 1].astype(np.int64))
    hstar = torch.from_numpy(prob.h_star)
    t = prob.t0
    loss = torch.zeros((), dtype=torch.float64)
    samples = []
    for s in range(prob.S):
        dt = float(prob.dts[s])
        k = min(max(bisect.bisect_right(list(rf.time_s), t) - 1, 0), len(rf.time_s) - 1)
        rain = torch.zeros_like(b_base)
        rain[g:-g, g:-g] = torch.from_numpy(frames[k])
        eta = float(sb.stage_at_time(t + dt))
        q = TR.step(q, b, n_cell, rain, dt, active, stage, eta, consts)
        t = t + dt
        hg = q[0][gi, gj]
        samples.append(hg.detach().numpy().copy())
        loss = loss + torch.sum((hg - hstar[s]) ** 2)
    return loss, np.array(samples)


def flowing_problem(n=40, S=30, K=8, dtype="float64", n_gauges=8, seed=5):
    """A fully wet, flowing flat domain (smooth bed, non-zero velocity everywhere, 8
    Manning classes): no still water and no dry cells, so the misfit is smooth in n and in
    the bed away from measure-zero branch switches, the setting where central differences
    pin the adjoint."""
    from geoswe_amd import Config, Mesh2D, Solver2D
    from geoswe_amd.adjoint import GaugeMisfit
    from geoswe_amd.compressed import CompressedSolver
    i = np.arange(n)[:, None].astype(np.float64)
    j = np.arange(n)[None, :].astype(np.float64)
    bed = 0.3 * np.sin(i / 6.0) * np.cos(j / 5.0) + 0.004 * i
    h = 1.2 - bed + 0.05 * np.sin(i / 4.0 + j / 7.0)
    hu = h * (0.35 + 0.1 * np.sin(j / 4.0))
    hv = h * (-0.2 + 0.1 * np.cos(i / 3.0))
    cfg = Config(bc_x="extrapolate", bc_y="extrapolate", dtype=dtype, friction="manning", manning_n=0.03,
                 h_min=1e-6, rainfall=20.0 / 3.6e6)
    mesh = Mesh2D(n, n, 2.0, 2.0, ngh=4)

    def solver():
        s = Solver2D(mesh, cfg, np.stack([h, hu, hv]), bed)
        cls = np.zeros((n + 8, n + 8), np.uint8)
        cls[4:-4, 4:-4] = ((i + 2 * j) % 8).astype(np.uint8)
        s.set_manning_table(cls, np.linspace(0.02, 0.09, 8))
        return CompressedSolver.from_dense(s, cfl_linf=True, track_max_depth=False)

    ref = solver()
    dts = np.array([ref.step() for _ in range(S)])
    f = solver()
    rng = np.random.default_rng(seed)
    gauges = np.sort(rng.choice(np.flatnonzero(f.fm.active), size=n_gauges, replace=False))
    prob = GaugeMisfit(f, gauges, dts, checkpoint_every=K)
    n0 = prob.params()[0]
    prob.set_params(ntab=n0 * (1.0 + 0.25 * rng.uniform(-1.0, 1.0, len(n0))))
    prob.h_star = prob.simulate(keep_checkpoints=False).copy()
    prob.set_params(ntab=n0)
    return prob
