#!/usr/bin/env python3
"""Cell-updates/s and HBM roofline of the fused finite-volume step on one MI355X.

Workload (default): the everywhere-wet rain-on-slope grid, 25600 x 25000 = 640 M cells,
dense layout, float32, dx = 3 m, bed = 1e-3 slope + seeded fBm (Hurst 0.8, 0.5 m),
h0 = 0.1 m, 50 mm/h uniform rain, Manning n = 0.03, free-outflow edges, CFL 0.5. A step
is one full time step (residual, rain, update, friction, wet/dry, next-step CFL with dt
on the device). Prints one JSON line.

  python bench.py [--steps 200] [--warmup 20] [--layout dense|flat] [--nx 25600 --ny 25000]
                  [--no-fp64] [--no-cpu] [--traffic profiles/<file>.json]

Alongside the throughput it reports the roofline of the step kernel (algorithmic bytes
per launch / measured launch time vs 8 TB/s), the max depth error against the same
workload run by the float64 build of the kernel, and the CPU NumPy oracle timed on a
bounded crop of the same workload on this host.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
# q read 12 + bed 4 + q write 12 (uniform n, uniform rain, no flags plane in either layout)
BYTES_PER_CELL = {"dense": 28.0, "flat": 28.0}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def timed_steps(s, steps):
    """Run `steps` fused steps; return (wall seconds, event milliseconds)."""
    from geoswe_amd import _lib as L
    e0, e1 = C.c_void_p(), C.c_void_p()
    L.call("gs_event_create", C.byref(e0))
    L.call("gs_event_create", C.byref(e1))
    L.sync(s.stream)
    t0 = time.perf_counter()
    L.call("gs_event_record", e0, s.stream)
    s.advance(steps, sync=False)
    L.call("gs_event_record", e1, s.stream)
    L.sync(s.stream)
    wall = time.perf_counter() - t0
    ms = C.c_float()
    L.call("gs_event_elapsed_ms", C.byref(ms), e0, e1)
    L.call("gs_event_destroy", e0)
    L.call("gs_event_destroy", e1)
    return wall, float(ms.value)


def fp64_error(nx, ny, bed, steps):
    """Max |h - h64| after `steps` steps of the rain-on-slope case, h64 from the float64 build
    of the kernel on the float32 run's inputs and wet/dry thresholds: the float32 run (h32)
    against the float64 run with its own CFL steps and against the float64 build replaying
    the float32 run's dt sequence (which takes the dt schedule out of the comparison), and the
    float32 planes stepped with float64 arithmetic against the float64 run."""
    from geoswe_amd.cases import rain_on_slope_c2
    log(f"float64 comparison: float32, float64 and mixed runs of {steps} steps")

    def err(h, h64, name):
        e = np.abs(h - h64)
        k = int(np.argmax(e))
        i, j = np.unravel_index(k, e.shape)
        r = {"max": float(e.max()), "at": [int(i), int(j)], "p99999": float(np.percentile(e, 99.999)),
             "median": float(np.median(e)), "cells_above_1e-5": int((e > 1e-5).sum())}
        log(f"{name}: max |h - h64| {e.max():.3e} m at ({i}, {j}); 99.999th percentile {r['p99999']:.3e} m, "
            f"median {r['median']:.3e} m")
        return r

    s = rain_on_slope_c2(nx, ny, "float32", bed=bed)
    dts = [s.step() for _ in range(steps)]
    h32 = s.depth().astype(np.float64)
    t32 = s.t
    del s
    s = rain_on_slope_c2(nx, ny, "float32", bed=bed, compute_dtype="float64")
    s.advance(steps)
    hmx = s.depth().astype(np.float64)
    del s
    out = {"steps": steps, "t32": t32}
    s64 = rain_on_slope_c2(nx, ny, "float64", bed=bed, f32_inputs=True)
    s64.advance(steps)
    h64 = s64.depth()
    out["own_dt"] = err(h32, h64, "float32 vs float64")
    out["own_dt"]["t64"] = s64.t
    out["mixed_vs_fp64"] = err(hmx, h64, "float32 planes with float64 arithmetic vs float64")
    del s64, h64, hmx
    s64 = rain_on_slope_c2(nx, ny, "float64", bed=bed, f32_inputs=True)
    for dt in dts:
        s64.step(dt)
    out["frozen_dt"] = err(h32, s64.depth(), "float32 vs float64 replaying the float32 dt sequence")
    out["frozen_dt"]["t64"] = s64.t
    del s64
    out["max_depth_err_vs_fp64_m"] = out["own_dt"]["max"]
    return out


def mixed_timing(nx, ny, bed, steps, warmup):
    """The rain-on-slope step with float32 planes and float64 arithmetic (GS_F32X64): wall
    and HIP-event time over `steps` steps after `warmup`."""
    from geoswe_amd.cases import rain_on_slope_c2
    s = rain_on_slope_c2(nx, ny, "float32", bed=bed, compute_dtype="float64")
    s.advance(warmup)
    wall, ev_ms = timed_steps(s, steps)
    cells = nx * ny
    out = {"value": cells * steps / wall, "unit": "cell-updates/s", "steps": steps, "ms_per_step": 1e3 * wall / steps,
           "kernel_ms": ev_ms / steps, "kernel": "gs::k_dense_sweep<double, 1, float>",
           "hbm_frac": BYTES_PER_CELL["dense"] * cells / (ev_ms / steps * 1e-3) / 1e9 / HBM_PEAK_GBS}
    log(f"float32 planes, float64 arithmetic: {out['ms_per_step']:.3f} ms/step, {out['value'] / 1e9:.2f} G "
        f"cell-updates/s")
    return out


def flat_layout_leg(nx, ny, bed, steps, warmup, dense_depth=None) -> dict:
    """The same rain-on-slope case in the flat layout (every cell active, row-sweep tiles):
    wall and HIP-event time over `steps` steps after `warmup`, and the depth after them
    against the dense run's depth after as many steps (bit for bit)."""
    from geoswe_amd.cases import rain_on_slope_c2
    from geoswe_amd.compressed import CompressedSolver
    s = rain_on_slope_c2(nx, ny, "float32", bed=bed)
    s = CompressedSolver.from_dense(s, cfl_linf=True, track_max_depth=False)
    s.advance(warmup)
    wall, ev_ms = timed_steps(s, steps)
    cells = nx * ny
    kern_ms = ev_ms / steps
    achieved = BYTES_PER_CELL["flat"] * cells / (kern_ms * 1e-3) / 1e9
    out = {"value": cells * steps / wall, "unit": "cell-updates/s", "steps": steps, "ms_per_step": 1e3 * wall / steps,
           "kernel_ms": kern_ms, "tiles": int(len(s.fm.tile_e0)),
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": achieved / HBM_PEAK_GBS, "kernel": "gs::k_flat_sweep<float, false, true>",
                        "bytes_per_cell": BYTES_PER_CELL["flat"]}}
    if dense_depth is not None:
        out["equals_dense"] = bool(np.array_equal(s.depth(), dense_depth))
    log(f"flat layout: {out['ms_per_step']:.3f} ms/step, {out['value'] / 1e9:.2f} G cell-updates/s"
        + (f", depth equal to the dense run: {out['equals_dense']}" if dense_depth is not None else ""))
    return out


def cpu_baseline(bed_spec, crop=(1024, 1000), steps=80):
    """The NumPy oracle (float32, one thread) on a crop of the same workload."""
    from oracle import generators as GN
    from oracle import swe_oracle as O
    nx, ny = crop
    b = GN.bed_crop(0, 0, nx, ny, bed_spec.seed, bed_spec.octaves, bed_spec.L0, bed_spec.hurst, bed_spec.mu,
                    bed_spec.scale, bed_spec.slope_x, bed_spec.dx, bed_spec.offset, np.float32)
    h = np.full((nx, ny), 0.1)
    case = O.DenseCase(h=h, hu=0 * h, hv=0 * h, bed=b, dx=3.0, dy=3.0, bc_x="fall", bc_y="fall",
                       manning=np.float32(0.03), rain=50.0 / 3.6e6)
    o = O.DenseOracle(case, O.Physics(friction=True), np.float32)
    o.run_steps(1)                     # warm-up (allocations, page faults)
    t0 = time.perf_counter()
    o.run_steps(steps)
    el = time.perf_counter() - t0
    return {"value": nx * ny * steps / el, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "sample": f"NumPy oracle (oracle/swe_oracle.py), float32, {nx}x{ny} crop of the same bed, "
                      f"{steps} steps in {el:.1f} s, one thread; host has {os.cpu_count()} logical CPUs"}


def valu_calibration(reps: int = 2) -> dict:
    """VALU issue rate of the device (gs_bench_valu): wave64 instructions per second for
    independent float32 multiply/add chains (mode 0) and for chains in which half the
    instructions are hardware square roots and reciprocals (mode 1); from the two, the issue
    cost of one transcendental instruction in simple-instruction slots. Each figure is the
    best of `reps` bursts of 10 back-to-back launches (~1-3 ms each) after a warm-up burst."""
    from geoswe_amd import _lib as L
    blocks, iters = 8 * 256, 4096                  # 8 workgroups per CU of the MI355X's 256
    out = L.DeviceArray((blocks * 256,), np.float32)
    e0, e1 = C.c_void_p(), C.c_void_p()
    L.call("gs_event_create", C.byref(e0))
    L.call("gs_event_create", C.byref(e1))
    rate = {}
    burst = 10
    for mode, per_step in ((0, 16), (1, 32)):
        # a sustained burst, as the step runs: back-to-back launches after a warm-up burst
        # (the shader clock ramps up under load; isolated launches run at a lower clock)
        best = None
        for _ in range(reps):
            for _ in range(burst):
                L.call("gs_bench_valu", C.c_void_p(out.ptr), blocks, iters, mode, None)
            L.call("gs_event_record", e0, None)
            for _ in range(burst):
                L.call("gs_bench_valu", C.c_void_p(out.ptr), blocks, iters, mode, None)
            L.call("gs_event_record", e1, None)
            L.sync()
            ms = C.c_float()
            L.call("gs_event_elapsed_ms", C.byref(ms), e0, e1)
            per = ms.value / burst
            best = per if best is None else min(best, per)
        rate[mode] = blocks * 4 * (iters - iters % 8) * per_step / (best * 1e-3)
    # mode 1 per step: 16 simple + 16 transcendental; time = (16 + 16 x) / rate0 = 32 / rate1
    slots = (32.0 * rate[0] / rate[1] - 16.0) / 16.0
    return {"simple_wave_insts_per_s": rate[0], "mixed_wave_insts_per_s": rate[1],
            "transcendental_slots": slots, "workgroups": blocks, "iters": iters}


def copy_calibration(nbytes: int, reps: int, quiet: bool = False) -> dict:
    """Time gs_bench_copy (read n + write n bytes per launch); prints one JSON line unless quiet."""
    from geoswe_amd import _lib as L
    nbytes -= nbytes % 16
    a = L.DeviceArray((nbytes,), np.uint8, zero=True)
    b = L.DeviceArray((nbytes,), np.uint8, zero=True)
    L.call("gs_bench_copy", C.c_void_p(b.ptr), C.c_void_p(a.ptr), C.c_size_t(nbytes), None)
    e0, e1 = C.c_void_p(), C.c_void_p()
    L.call("gs_event_create", C.byref(e0))
    L.call("gs_event_create", C.byref(e1))
    L.sync()
    L.call("gs_event_record", e0, None)
    for _ in range(reps):
        L.call("gs_bench_copy", C.c_void_p(b.ptr), C.c_void_p(a.ptr), C.c_size_t(nbytes), None)
    L.call("gs_event_record", e1, None)
    L.sync()
    ms = C.c_float()
    L.call("gs_event_elapsed_ms", C.byref(ms), e0, e1)
    per = ms.value / reps
    out = {"copy_bytes": nbytes, "reps": reps, "ms_per_copy": per,
           "GBps_read_plus_write": 2 * nbytes / (per * 1e-3) / 1e9}
    if not quiet:
        print(json.dumps(out), flush=True)
    return out


def coastal(args):
    """--workload coastal: the coastal leg as its own JSON line."""
    out = coastal_leg(args.n, args.steps, args.sweep_rows, args.forcings)
    print(json.dumps(out), flush=True)
    if not out["dense_equals_flat"]:
        sys.exit(1)


def coastal_leg(n: int, steps: int, sweep_rows=None, forcings: bool = False) -> dict:
    """The compressed coastal case: flat and dense layouts stepped side by side, depth and
    residual compared bit for bit at steps 1, 500 and the last; both timed with HIP events.
    The line is the flat layout's (the production layout of this case)."""
    from geoswe_amd import _lib as L
    from geoswe_amd.cases import coastal_c3
    t0 = time.perf_counter()
    log(f"building the coastal case {n}x{n}")
    case = coastal_c3(n=n)
    d = case.dense(track_max_depth=False)
    f = case.flat(track_max_depth=False, sweep_rows=sweep_rows)
    fm = f.fm
    act = fm.active
    ij = fm.ij[act] - case.mesh.ngh
    log(f"built in {time.perf_counter() - t0:.0f} s: {fm.n_active / 1e6:.1f} M active cells "
        f"({fm.n_active / n / n:.1%} of the box), {fm.n_stored / 1e6:.1f} M stored, {len(fm.run_k0)} runs")

    def compare(tag):
        hd = d.depth()[ij[:, 0], ij[:, 1]]
        hf = f.q_flat()[0][act]
        same_h = bool(np.array_equal(hd, hf))
        rd = d.residual()
        rf = f.residual()
        same_r = bool(np.array_equal(rf[:, act], rd[:, ij[:, 0], ij[:, 1]]))
        del rd, rf
        same_clock = d.nsteps == f.nsteps and d.t == f.t
        log(f"{tag}: steps {d.nsteps}/{f.nsteps}, t {d.t:.6f}/{f.t:.6f} s, depth equal {same_h}, "
            f"residual equal {same_r}, wet cells {int((hf > 1e-6).sum())}")
        return same_h and same_r and same_clock

    ok = compare("t=0")
    marks = [1, 500, steps] if steps > 500 else [1, steps]
    tf = td = 0.0
    done = 0
    for m in marks:
        k = m - done
        wf, ef = timed_steps(f, k)
        wd, ed = timed_steps(d, k)
        if m > 1:
            tf += ef
            td += ed
        done = m
        ok = compare(f"step {m}") and ok
    timed = done - 1
    for s_ in (d, f):
        if s_._clockv().status & L.ST_NONFINITE:
            raise FloatingPointError("non-finite state in the coastal run")
    ms_f, ms_d = tf / timed, td / timed
    nact = fm.n_active
    bpc = 33.0      # q 12 + bed 4 + flags 1 + rain frame 4 + q' 12 per active cell
    achieved = bpc * nact / (ms_f * 1e-3) / 1e9
    traffic = None
    tp = ROOT / "profiles" / "flat_sweep_coastal_16384_traffic.json"
    if n == 16384 and tp.exists():
        traffic = json.loads(tp.read_text()).get("bytes_per_launch")
    log(f"flat {ms_f:.3f} ms/step ({nact / (ms_f * 1e-3) / 1e9:.1f} G active cell-updates/s), dense "
        f"{ms_d:.3f} ms/step over the whole box; {achieved:.0f} GB/s algorithmic (flat)")
    out = {
        "metric": "active cell-updates/s (compressed coastal step, flat layout, 1 MI355X)",
        "value": nact / (ms_f * 1e-3),
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": timed,
        "warmup": 1,
        "ms_per_step": ms_f,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded fBm DEM, Voronoi Manning classes, fBm rain frames, M2 tide)",
        "config": {"workload": f"coastal {n}x{n}, {nact} active cells ({nact / n / n:.1%}), 8 Manning classes, "
                               f"gridded rain every 300 s, tidal stage on the sea edge, fall edges, CFL 0.5 (L-inf)",
                   "layout": "flat", "active_cells": int(nact), "stored_cells": int(fm.n_stored),
                   "sweep_tiles": int(len(fm.tile_e0)), "sweep_rows": int(fm.tile_irc[:, 1].max(initial=0))},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": "gs::k_flat_sweep<float, false, true>",
                     "kernel_ms": ms_f, "bytes_per_cell": bpc},
        "dense_ms_per_step": ms_d,
        "dense_equals_flat": ok,
    }
    if forcings:
        out["forcing_tail"] = coastal_forcing_tail(case, steps, sweep_rows, ms_f)
    return out


def coastal_forcing_tail(case, steps, sweep_rows, plain_ms):
    """The coastal step with a driver's whole forcing tail attached (infiltration, sponge,
    IDW gauge ring, Green-Ampt with drain-tau, clamp, karst drain): a fresh flat solver, one
    warm-up step, then the same number of timed steps as the plain run."""
    from geoswe_amd import _lib as L
    from geoswe_amd.cases import attach_coastal_forcings
    t0 = time.perf_counter()
    f = case.flat(track_max_depth=False, sweep_rows=sweep_rows)
    cells = attach_coastal_forcings(case, f)["cells"]
    log(f"forcing tail attached in {time.perf_counter() - t0:.0f} s: {cells}")
    timed_steps(f, 1)
    k = max(steps - 1, 1)
    _, ms = timed_steps(f, k)
    if f._clockv().status & L.ST_NONFINITE:
        raise FloatingPointError("non-finite state in the forced coastal run")
    ms /= k
    log(f"with the forcing tail: {ms:.3f} ms/step against {plain_ms:.3f} ms/step plain "
        f"({(ms / plain_ms - 1) * 100:+.1f} %)")
    return {"ms_per_step": ms, "plain_ms_per_step": plain_ms, "overhead": ms / plain_ms - 1.0, "cells": cells,
            "kernels": ["gs::k_flat_ring_force<float>", "gs::k_flat_sweep_ext<float>"]}


def ensemble(args):
    """--workload ensemble: the ensemble leg as its own JSON line."""
    out = ensemble_leg(args.members, args.steps, args.graph_batch)
    print(json.dumps(out), flush=True)
    if not out["members_bitwise_equal_plain"]:
        sys.exit(1)


def ensemble_leg(members: int, steps: int, graph_batch: int) -> dict:
    """The batched ensemble: seeded uneven coastal members stepped one member after another
    (plain CompressedSolver runs) and together (one launch per step, hipGraph replay, the
    autotuned launch shape); every member compared bit for bit."""
    from geoswe_amd import _lib as L
    from geoswe_amd.cases import ensemble_c4, ensemble_sides
    from geoswe_amd.ensemble import Ensemble
    sides = ensemble_sides(members)
    t0 = time.perf_counter()
    log(f"building {len(sides)} members, sides {min(sides)}..{max(sides)}, {sum(x * x for x in sides) / 1e6:.1f} M "
        f"box cells")
    cases = ensemble_c4(sides=sides)
    plain = [c.flat(track_max_depth=False) for c in cases]
    mem = [c.flat(track_max_depth=False) for c in cases]
    del cases
    act = sum(m.n_active for m in mem)
    log(f"built in {time.perf_counter() - t0:.0f} s; {act / 1e6:.1f} M active cells")
    for m in plain:                    # first dt and set-up outside the timed region
        m._ensure_lam()
        m._open_mode()
    L.sync()
    e0, e1 = C.c_void_p(), C.c_void_p()
    L.call("gs_event_create", C.byref(e0))
    L.call("gs_event_create", C.byref(e1))
    w0 = time.perf_counter()
    L.call("gs_event_record", e0, None)
    for m in plain:
        m.advance(steps, sync=False)
    L.call("gs_event_record", e1, None)
    L.sync()
    w_plain = time.perf_counter() - w0
    ms = C.c_float()
    L.call("gs_event_elapsed_ms", C.byref(ms), e0, e1)
    ev_plain = ms.value
    log(f"sequential plain runs: {w_plain:.2f} s wall for {steps} steps x {len(plain)} members")
    ens = Ensemble(mem, graph_batch=graph_batch)
    tune = ens.autotune(trial_steps=3, heights=(16, 32, 64))
    log(f"autotuned launch: {tune}")
    ens.advance(1, graph_batch=0)       # first launch outside the timed region
    L.sync()
    w0 = time.perf_counter()
    L.call("gs_event_record", e0, None)
    ens.advance(steps - 1, sync=False)
    L.call("gs_event_record", e1, None)
    L.sync()
    w_ens = (time.perf_counter() - w0) * steps / (steps - 1)
    L.call("gs_event_elapsed_ms", C.byref(ms), e0, e1)
    ev_ens = ms.value * steps / (steps - 1)
    same = all(a.nsteps == b.nsteps and a.t == b.t and np.array_equal(a.q_flat(), b.q_flat())
               for a, b in zip(plain, mem))
    dts = sorted({round(m.t / m.nsteps, 6) for m in mem})
    v_plain = act * steps / w_plain
    v_ens = act * steps / w_ens
    log(f"ensemble: {w_ens:.2f} s wall; {v_ens / 1e9:.2f} vs {v_plain / 1e9:.2f} G cell-updates/s "
        f"({v_ens / v_plain:.2f}x); members bitwise equal to their plain runs: {same}")
    # the reference's plain loop: one host read-back of the step's dt per step
    # (compressed_solver.py _step_loop: lambda read-back after every fused step)
    w0 = time.perf_counter()
    for m in plain:
        for _ in range(steps):
            m.step()
    L.sync()
    w_sync = time.perf_counter() - w0
    v_sync = act * steps / w_sync
    log(f"sequential runs with a per-step dt read-back: {w_sync:.2f} s wall, {v_sync / 1e9:.2f} G cell-updates/s "
        f"(ensemble {v_ens / v_sync:.2f}x)")
    achieved = 33.0 * act * steps / (ev_ens * 1e-3) / 1e9
    traffic = None
    tp = ROOT / "profiles" / "ens_step_64_members_traffic.json"
    if members == 64 and tp.exists():
        traffic = json.loads(tp.read_text()).get("bytes_per_launch")
    out = {
        "metric": "cell-updates/s (batched ensemble of flat domains, 1 MI355X)",
        "value": v_ens, "unit": "cell-updates/s", "n_gpus": 1, "steps": steps, "warmup": 1,
        "ms_per_step": 1e3 * w_ens / steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded coastal members: fBm DEM, Voronoi classes, fBm rain, tide)",
        "config": {"workload": f"{len(sides)} coastal members, sides {min(sides)}..{max(sides)} (log-uniform), "
                               f"{act} active cells, rain x U[0.5,2], n x (1+U[-0.3,0.3]), per-member CFL dt",
                   "members": len(sides), "active_cells": int(act), "box_cells": int(sum(x * x for x in sides))},
        "sequential_plain_value": v_plain, "speedup_vs_sequential": v_ens / v_plain,
        "sequential_readback_value": v_sync, "speedup_vs_sequential_readback": v_ens / v_sync,
        "sequential_plain_s": w_plain, "ensemble_s": w_ens, "event_ms_plain": ev_plain, "event_ms_ensemble": ev_ens,
        "launch": tune, "members_bitwise_equal_plain": same, "distinct_member_dt": len(dts),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": "gs::k_ens_step<float, 2>",
                     "bytes_per_cell": 33.0},
    }
    L.call("gs_event_destroy", e0)
    L.call("gs_event_destroy", e1)
    return out


def adjoint(args):
    """--workload adjoint: the adjoint leg as its own JSON line."""
    out = adjoint_leg(args.n_adjoint, args.steps, args.checkpoint, args.sweep_rows, fd=not args.no_fd,
                      fd_bed_cells=args.fd_bed_cells)
    print(json.dumps(out), flush=True)


def adjoint_leg(n: int, S: int, K: int, sweep_rows=None, fd: bool = True, fd_bed_cells: int = 64) -> dict:
    """Gauge-misfit calibration gradients on the coastal case: forward and adjoint cost,
    central finite differences (eps 1e-3 and 1e-4) on every class n and on seeded bed cells
    near the gauges, and float64 torch autograd on a small coastal case."""
    from geoswe_amd import _lib as L
    from geoswe_amd.adjoint import GaugeMisfit
    from geoswe_amd.cases import coastal_c3
    t0 = time.perf_counter()
    case = coastal_c3(n=n)
    ref = case.flat(track_max_depth=False)
    ref._ensure_lam()
    ref._open_mode()
    dts = []
    for _ in range(S):
        dts.append(ref.step())
    del ref
    f = case.flat(track_max_depth=False, sweep_rows=sweep_rows)
    fm = f.fm
    h0 = f.q_flat()[0]
    rng = np.random.default_rng(20261205)
    gauges = np.sort(rng.choice(np.flatnonzero(fm.active & (h0 > 0.05)), size=32, replace=False))
    prob = GaugeMisfit(f, gauges, np.array(dts), checkpoint_every=K)
    n0, b0 = prob.params()
    n_true = n0 * (1.0 + 0.3 * rng.uniform(-1.0, 1.0, len(n0)))
    prob.set_params(ntab=n_true)
    prob.h_star = prob.simulate(keep_checkpoints=False).copy()
    prob.set_params(ntab=n0)
    log(f"built {n}x{n} ({fm.n_active / 1e6:.2f} M active), {S} steps to t={sum(dts):.1f} s in "
        f"{time.perf_counter() - t0:.0f} s")
    prob.simulate()
    L.sync()
    w0 = time.perf_counter()
    prob.simulate()
    t_fwd = time.perf_counter() - w0
    J0 = prob.loss()
    w0 = time.perf_counter()
    gn, gb = prob.gradient()
    t_adj = time.perf_counter() - w0
    # the production forward of the same solver: CFL steps, one launch per step, no gauges
    prob.s.advance(1)
    w0 = time.perf_counter()
    prob.s.advance(S)
    t_plain = time.perf_counter() - w0
    log(f"J = {J0:.6e}; calibration forward {1e3 * t_fwd / S:.3f} ms/step, plain forward "
        f"{1e3 * t_plain / S:.3f} ms/step, gradient {1e3 * t_adj / S:.3f} ms/step ({t_adj / t_fwd:.2f}x the "
        f"calibration forward, {t_adj / t_plain:.2f}x the plain forward)")

    def J(nn=None, bb=None):
        prob.set_params(ntab=n0 if nn is None else nn, bed=b0 if bb is None else bb)
        prob.simulate(keep_checkpoints=False)
        return prob.loss()

    # central differences at eps = 1e-3 and 1e-4: the loss is only piecewise smooth (wet/dry
    # switches of rain-fed land cells, the still-water ties of HLLC's wave-speed min/max), so
    # the spread between the two is the finite differences' own uncertainty
    def fd_at(eps, nn_dir=None, cell=None):
        if cell is None:
            e = np.zeros_like(n0)
            e[nn_dir] = eps
            return (J(nn=n0 + e) - J(nn=n0 - e)) / (2 * eps)
        e = np.zeros_like(b0)
        e[cell] = eps
        return (J(bb=b0 + e) - J(bb=b0 - e)) / (2 * eps)

    if not fd:
        return {"forward_ms_per_step": 1e3 * t_fwd / S, "adjoint_ms_per_step": 1e3 * t_adj / S,
                "plain_forward_ms_per_step": 1e3 * t_plain / S, "value": t_adj / t_fwd}
    fd_n = np.array([[fd_at(eps, nn_dir=k) for k in range(len(n0))] for eps in (1e-3, 1e-4)])
    gij = fm.ij[gauges]
    near = np.flatnonzero(fm.active & (np.abs(fm.ij[:, None, :] - gij[None]).max(axis=2).min(axis=1) <= 4))
    cells = np.sort(rng.choice(near, size=min(fd_bed_cells, len(near)), replace=False))
    fd_b = np.array([[fd_at(eps, cell=c) for c in cells] for eps in (1e-3, 1e-4)])
    prob.set_params(ntab=n0, bed=b0)

    def stats(adj, fd):
        big = np.abs(fd[0]) >= 1e-2 * np.abs(fd[0]).max()
        rel = np.abs(adj - fd[0])[big] / np.abs(fd[0])[big]
        spread = np.abs(fd[0] - fd[1])[big] / np.abs(fd[0])[big]
        near_ = np.abs(adj - fd[1])[big] / np.abs(fd[1])[big]
        return {"compared": int(big.sum()), "rel_err_eps1e-3_median": float(np.median(rel)),
                "rel_err_eps1e-3_max": float(rel.max()), "within_1e-3": float(np.mean(rel <= 1e-3)),
                "fd_spread_median": float(np.median(spread)),
                "within_fd_spread": float(np.mean(rel <= np.maximum(spread, 1e-3))),
                "rel_err_eps1e-4_median": float(np.median(near_))}

    st_n = stats(gn, fd_n)
    st_b = stats(gb[cells], fd_b)
    log(f"FD (eps 1e-3 / 1e-4) vs adjoint, class n: {st_n}")
    log(f"FD (eps 1e-3 / 1e-4) vs adjoint, bed cells near the gauges: {st_b}")
    del prob, f
    horizon = fd64_horizon_check(n, 500)
    log(f"float64 central differences vs adjoint at 500 steps: {horizon}")
    torch_err = None
    try:
        torch_err = torch_check()
        log(f"float64 torch autograd check (128x128, 200 steps): {torch_err}")
    except ImportError:
        log("torch not importable: float64 autograd check skipped")
    out = {
        "metric": "adjoint step cost / forward step cost (gauge-misfit gradient, 1 MI355X)",
        "value": t_adj / t_fwd, "unit": "x forward", "n_gpus": 1, "steps": S, "warmup": 1,
        "ms_per_step": 1e3 * t_adj / S, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (coastal generator, truth run with perturbed Manning classes)",
        "config": {"workload": f"coastal {n}x{n} flat, {fm.n_active} active cells, {S} steps, checkpoint every "
                               f"{K}, 32 gauges, gradients w.r.t. {len(n0)} class n and {fm.n_stored} bed cells"},
        "forward_ms_per_step": 1e3 * t_fwd / S, "adjoint_ms_per_step": 1e3 * t_adj / S,
        "plain_forward_ms_per_step": 1e3 * t_plain / S, "adjoint_over_plain_forward": t_adj / t_plain,
        "forward_cell_updates_per_s": fm.n_active * S / t_fwd,
        "grad_n": gn.tolist(), "fd_n_eps1e-3": fd_n[0].tolist(), "fd_n_eps1e-4": fd_n[1].tolist(),
        "fd_vs_adjoint_n": st_n, "fd_vs_adjoint_bed": st_b,
        "torch_f64_autograd": torch_err,
        "fd64_500_steps": horizon,
    }
    return out


def fd64_horizon_check(n: int, S: int) -> dict:
    """The adjoint against float64 central differences on the calibration case over a
    horizon where the loss is smooth: the float64 build of the same coastal problem (its own
    CFL dt sequence, gauges and truth run, as the float32 leg builds them) over S steps;
    per class, differences at eps = 1e-5, 1e-6, 1e-7; a class counts as smooth where two
    consecutive eps agree to 1e-3, and the adjoint is compared with the finer of the two."""
    from geoswe_amd.adjoint import GaugeMisfit
    from geoswe_amd.cases import coastal_c3
    case = coastal_c3(n=n)
    case.cfg.dtype = "float64"
    ref = case.flat(track_max_depth=False)
    dts = [ref.step() for _ in range(S)]
    del ref
    f = case.flat(track_max_depth=False)
    fm = f.fm
    rng = np.random.default_rng(20261205)
    gauges = np.sort(rng.choice(np.flatnonzero(fm.active & (f.q_flat()[0] > 0.05)), size=32, replace=False))
    prob = GaugeMisfit(f, gauges, np.array(dts), checkpoint_every=100)
    n0, _ = prob.params()
    prob.set_params(ntab=n0 * (1.0 + 0.3 * rng.uniform(-1.0, 1.0, len(n0))))
    prob.h_star = prob.simulate(keep_checkpoints=False).copy()
    prob.set_params(ntab=n0)
    prob.simulate()
    gn, _ = prob.gradient()

    def J(nn):
        prob.set_params(ntab=nn)
        prob.simulate(keep_checkpoints=False)
        return prob.loss()

    epss = (1e-5, 1e-6, 1e-7)
    rows, rel = [], []
    for c in range(len(n0)):
        fd = []
        for eps in epss:
            e = np.zeros_like(n0)
            e[c] = eps
            fd.append((J(n0 + e) - J(n0 - e)) / (2 * eps))
        smooth = None
        for a_, b_ in zip(fd[:-1], fd[1:]):
            if abs(a_ - b_) <= 1e-3 * abs(b_):
                smooth = b_
        r = None if smooth is None else abs(gn[c] - smooth) / abs(smooth)
        if r is not None:
            rel.append(r)
        rows.append({"class": c, "adjoint": float(gn[c]), "fd": [float(x) for x in fd], "rel_err": r})
    prob.set_params(ntab=n0)
    return {"steps": S, "eps": list(epss), "classes": rows, "smooth_classes": len(rel),
            "max_rel_err_smooth": max(rel) if rel else None,
            "within_1e-3": float(np.mean(np.array(rel) <= 1e-3)) if rel else None}


def torch_check():
    """Relative gradient differences of the float64 device adjoint against torch autograd
    through the float64 torch restatement of the step (tests/torch_ref.py), 128x128, 200 steps."""
    sys.path.insert(0, str(ROOT / "tests"))
    import torch
    import adjoint_cases as AC
    case, prob = AC.coastal_problem(n=128, S=200, K=50, dtype="float64", n_gauges=32)
    n0, b0 = prob.params()
    prob.simulate()
    gn, gb = prob.gradient()
    nt = torch.tensor(n0, requires_grad=True)
    bt = torch.tensor(b0, requires_grad=True)
    Jt, _ = AC.torch_misfit(case, prob, nt, bt)
    Jt.backward()
    tn, tb = nt.grad.numpy(), bt.grad.numpy()
    return {"n_max_rel": float(np.abs(gn - tn).max() / np.abs(tn).max()),
            "bed_normwise_rel": float(np.linalg.norm(gb - tb) / np.linalg.norm(tb)),
            "bed_max_rel": float(np.abs(gb - tb).max() / np.abs(tb).max())}


def dambreak(args):
    """--workload dambreak: the dam-break leg as its own JSON line."""
    print(json.dumps(dambreak_leg(args.steps)), flush=True)


def dambreak_leg(steps: int) -> dict:
    """The 256x256 circular dam break (walls, no friction, no rain), 2000 steps of
    step(cfl_dt()): the float32 GPU run against the float64 GPU run and against the float64
    NumPy oracle run; max |dh|, relative L2 of h, hu, hv at steps 1, 100 and the last, the
    final time and the dt sequences. The float32 oracle is timed on the same run (one
    thread) as the CPU baseline."""
    from geoswe_amd import Solver2D
    from geoswe_amd.cases import dam_break_c1
    from oracle import swe_oracle as O
    marks = sorted({1, min(100, steps), steps})
    runs = {}
    for dt_name in ("float32", "float64"):
        mesh, cfg, q0, b = dam_break_c1(dtype=dt_name)
        s = Solver2D(mesh, cfg, q0, b)
        snaps, dts = {}, []
        w0 = time.perf_counter()
        for k in range(1, steps + 1):
            dts.append(s.step())
            if k in marks:
                snaps[k] = s.q_interior.astype(np.float64)
        runs[dt_name] = (snaps, np.array(dts), s.t, time.perf_counter() - w0)
        log(f"GPU {dt_name}: {steps} steps to t={s.t:.6f} s in {runs[dt_name][3]:.2f} s")
    mesh, cfg, q0, b = dam_break_c1(dtype="float64")
    orc = {}
    for dt_name, T in (("float64", np.float64), ("float32", np.float32)):
        case = O.DenseCase(h=q0[0], hu=q0[1], hv=q0[2], bed=b, dx=1.0, dy=1.0, bc_x="wall", bc_y="wall")
        o = O.DenseOracle(case, O.Physics(cfl=0.5), T)
        snaps, dts = {}, []
        w0 = time.perf_counter()
        for k in range(1, steps + 1):
            dt = o.next_dt()
            dts.append(dt)
            o.step(dt)
            if k in marks:
                snaps[k] = o.interior.astype(np.float64)
        orc[dt_name] = (snaps, np.array(dts), o.t, time.perf_counter() - w0)
        log(f"oracle {dt_name}: {steps} steps in {orc[dt_name][3]:.1f} s")

    def compare(a, b_):
        out = {}
        for k in marks:
            qa, qb = a[0][k], b_[0][k]
            out[str(k)] = {"max_abs_dh": float(np.abs(qa[0] - qb[0]).max()),
                           "rel_l2": [float(np.linalg.norm(qa[m] - qb[m]) / max(np.linalg.norm(qb[m]), 1e-300))
                                      for m in range(3)]}
        out["t_final_diff"] = float(a[2] - b_[2])
        out["max_rel_dt_diff"] = float(np.abs(a[1] - b_[1]).max() / b_[1].max())
        return out

    gpu32_vs_gpu64 = compare(runs["float32"], runs["float64"])
    gpu64_vs_oracle64 = compare(runs["float64"], orc["float64"])
    gpu32_vs_oracle32 = compare(runs["float32"], orc["float32"])
    exact32 = all(np.array_equal(runs["float32"][0][k], orc["float32"][0][k]) for k in marks)
    log(f"float32 vs float64 after {steps} steps: max |dh| {gpu32_vs_gpu64[str(steps)]['max_abs_dh']:.3e} m; "
        f"GPU float32 == oracle float32 bit for bit: {exact32}")
    cells = 256 * 256
    w32 = runs["float32"][3]
    cpu = orc["float32"][3]
    out = {
        "metric": "cell-updates/s (circular dam break 256x256, step(cfl_dt()) loop, 1 MI355X)",
        "value": cells * steps / w32, "unit": "cell-updates/s", "n_gpus": 1, "steps": steps, "warmup": 0,
        "ms_per_step": 1e3 * w32 / steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (h=10 m inside r<40 cells, 1 m outside, flat bed, walls)",
        "config": {"workload": "circular dam break 256x256, dx=1 m, CFL 0.5, no friction, no rain, walls"},
        "float32_vs_float64_gpu": gpu32_vs_gpu64, "float64_gpu_vs_float64_oracle": gpu64_vs_oracle64,
        "float32_gpu_vs_float32_oracle": gpu32_vs_oracle32, "float32_gpu_bitwise_oracle": exact32,
        "cpu_baseline": {"value": cells * steps / cpu, "unit": "cell-updates/s", "cores": 1, "kind": "port",
                         "sample": f"NumPy oracle float32, the full run ({steps} steps) in {cpu:.1f} s"},
    }
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--nx", type=int, default=25600)
    ap.add_argument("--ny", type=int, default=25000)
    ap.add_argument("--layout", choices=["dense", "flat"], default="dense")
    ap.add_argument("--sweep-rows", type=int, default=None,
                    help="flat layout: rows per row-sweep tile (default: chosen from the domain size)")
    ap.add_argument("--tile-rows", type=int, default=0)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--traffic", default=str(ROOT / "profiles" / "dense_sweep_rain_on_slope_640M_traffic.json"))
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="instead of the step, time the plain copy kernel over this many bytes (HBM calibration)")
    ap.add_argument("--copy-reps", type=int, default=10)
    ap.add_argument("--valu-peak", action="store_true",
                    help="instead of the step, time the VALU calibration kernel (issue rate of simple and "
                         "transcendental instructions)")
    ap.add_argument("--workload", choices=["rain-on-slope", "coastal", "ensemble", "adjoint", "dambreak"], default="rain-on-slope",
                    help="coastal: the compressed coastal case (flat and dense layouts, bitwise cross-check); "
                         "ensemble: batched uneven members vs sequential runs; adjoint: gauge-misfit gradients")
    ap.add_argument("--members", type=int, default=64)
    ap.add_argument("--graph-batch", type=int, default=50)
    ap.add_argument("--n-adjoint", type=int, default=2048)
    ap.add_argument("--checkpoint", type=int, default=100)
    ap.add_argument("--no-fd", action="store_true", help="adjoint: timings only (no finite differences)")
    ap.add_argument("--fd-bed-cells", type=int, default=64, help="adjoint: bed cells checked by finite differences")
    ap.add_argument("--no-legs", action="store_true",
                    help="default workload: leave out the ensemble and adjoint legs of the line")
    ap.add_argument("--ens-steps", type=int, default=1000, help="ensemble leg: steps (the configuration's 1000)")
    ap.add_argument("--adj-steps", type=int, default=4000, help="adjoint leg: steps (the configuration's 4000)")
    ap.add_argument("--err-steps", type=int, default=220,
                    help="float64 comparison: steps of both runs (the configuration's 20 warm-up + 200 timed)")
    ap.add_argument("--n", type=int, default=16384, help="coastal box side")
    ap.add_argument("--forcings", action="store_true",
                    help="coastal: also time the step with a driver's whole forcing tail attached")
    args = ap.parse_args()
    if args.copy_bytes:
        copy_calibration(args.copy_bytes, args.copy_reps)
        return
    if args.valu_peak:
        v = valu_calibration()
        v["simple_wave_insts_per_cu_per_ns"] = v["simple_wave_insts_per_s"] / 256 / 1e9   # 256 CUs
        print(json.dumps(v), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported (a multi-GPU run "
              f"would split the domain into strips along y with a 2-cell halo exchange per step)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "coastal":
        return coastal(args)
    if args.workload == "ensemble":
        return ensemble(args)
    if args.workload == "adjoint":
        return adjoint(args)
    if args.workload == "dambreak":
        return dambreak(args)

    from geoswe_amd import _lib as L
    from geoswe_amd.cases import FbmBed, C2_SEED, rain_on_slope_c2
    from geoswe_amd.compressed import CompressedSolver

    nx, ny = args.nx, args.ny
    cells = nx * ny
    log(f"building the rain-on-slope case {nx}x{ny} ({cells / 1e6:.1f} M cells), layout {args.layout}")
    bed = FbmBed(seed=C2_SEED).normalise(nx, ny)
    s = rain_on_slope_c2(nx, ny, "float32", tile_rows=args.tile_rows, bed=bed)
    if args.layout == "flat":
        log("packing the flat layout (every cell active)")
        s = CompressedSolver.from_dense(s, cfl_linf=True, track_max_depth=False, sweep_rows=args.sweep_rows)
        log(f"flat layout: {s.n_active / 1e6:.1f} M active, {s.n_stored / 1e6:.1f} M stored, "
            f"{len(s.fm.tile_e0)} sweep tiles")
    log(f"warm-up {args.warmup} steps")
    s.advance(args.warmup)
    log(f"timing {args.steps} steps")
    wall, ev_ms = timed_steps(s, args.steps)
    c = s._clockv()
    if c.status & L.ST_NONFINITE:
        raise FloatingPointError("non-finite state during the benchmark")
    ms_step = 1e3 * wall / args.steps
    kern_ms = ev_ms / args.steps
    value = cells * args.steps / wall
    achieved = BYTES_PER_CELL[args.layout] * cells / (kern_ms * 1e-3) / 1e9
    traffic = None
    valu = None
    tp = Path(args.traffic)
    if tp.exists() and args.layout == "dense" and (nx, ny) == (25600, 25000):
        traffic = json.loads(tp.read_text()).get("bytes_per_launch")
        vp = ROOT / "profiles" / "dense_sweep_rain_on_slope_640M_valu.json"
        if vp.exists():
            v = json.loads(vp.read_text())
            # the step is instruction-bound: VALU wave-instructions per launch (SQ_INSTS_VALU)
            # over this run's launch time
            valu = {"insts_per_cell": v["valu_insts_per_cell"], "wave_insts_per_launch": v["valu_wave_insts_per_launch"],
                    "wave_insts_per_s": v["valu_wave_insts_per_launch"] / (kern_ms * 1e-3), "source": v["source"]}
            # the device's measured float32 multiply/add issue rate (gs_bench_valu) and the
            # step's VALU rate against it: at or above 1 the step is bound by instruction issue
            # (its compares, selects and moves issue faster than float32 arithmetic)
            cal = valu_calibration()
            valu.update({"fp32_peak_wave_insts_per_s": cal["simple_wave_insts_per_s"],
                         "rate_over_fp32_peak": valu["wave_insts_per_s"] / cal["simple_wave_insts_per_s"],
                         "transcendental_slots": cal["transcendental_slots"],
                         "peak_source": "gs_bench_valu (bench.py --valu-peak)"})
    # measured copy-kernel bandwidth of this device (2 GiB read + 2 GiB written per launch)
    copy = copy_calibration(1 << 31, 10, quiet=True)
    log(f"{ms_step:.3f} ms/step wall, {kern_ms:.3f} ms/launch (events), {value / 1e9:.2f} G cell-updates/s, "
        f"{achieved:.0f} GB/s algorithmic; copy kernel {copy['GBps_read_plus_write']:.0f} GB/s")
    err64 = None
    mixed = None
    flat = None
    if not args.no_legs and args.layout == "dense":
        log("flat layout leg: the same case, every cell active")
        flat = flat_layout_leg(nx, ny, bed, args.steps, args.warmup, dense_depth=s.depth())
    if not args.no_fp64 and args.layout == "dense":
        del s
        mixed = mixed_timing(nx, ny, bed, args.steps, args.warmup)
        err64 = fp64_error(nx, ny, bed, args.err_steps)
        mixed["max_depth_err_vs_fp64_m"] = err64["mixed_vs_fp64"]["max"]
    cpu = None
    if not args.no_cpu:
        log("CPU baseline (NumPy oracle on a crop)")
        cpu = cpu_baseline(bed)
        log(f"CPU {cpu['value'] / 1e6:.2f} M cell-updates/s")
    out = {
        "metric": "cell-updates/s & %HBM roofline, 1 MI355X: plain vs batched ensemble vs adjoint",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded fBm bed generated on the device)",
        "config": {"workload": f"rain-on-slope {args.layout} {nx}x{ny} ({cells / 1e6:.0f} M cells), dx=3 m, "
                               f"fBm H=0.8 0.5 m + slope 1e-3, h0=0.1 m, 50 mm/h, n=0.03, fall edges, CFL 0.5",
                   "layout": args.layout, "cells": cells},
        # the step is bound by instruction issue (VALU), not by HBM: "valu" carries its issue rate
        # against the device's measured float32 issue rate; achieved/peak/frac are the HBM figures
        "roofline": {"bound": "valu" if valu else "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "gs::k_dense_sweep<float, 1, float>" if args.layout == "dense" else "gs::k_flat_sweep<float, false, true>",
                     "kernel_ms": kern_ms, "bytes_per_cell": BYTES_PER_CELL[args.layout], "valu": valu,
                     "copy_GBps": copy["GBps_read_plus_write"],
                     "frac_of_copy": achieved / copy["GBps_read_plus_write"]},
        "max_depth_err_vs_fp64_m": err64["max_depth_err_vs_fp64_m"] if err64 else None,
        "fp64_comparison": err64,
        # the same step with float32 planes and float64 arithmetic (Config.compute_dtype):
        # the configuration meeting the float64 depth tolerance, timed over the same steps
        "float64_arithmetic": mixed,
        "flat_layout": flat,
        "cpu_baseline": cpu,
    }
    if not args.no_legs and (nx, ny) == (25600, 25000) and args.layout == "dense":
        log(f"ensemble leg: 64 coastal members, {args.ens_steps} steps")
        e = ensemble_leg(64, args.ens_steps, args.graph_batch)
        out["ensemble"] = {k: e[k] for k in ("value", "unit", "steps", "ms_per_step", "config", "sequential_plain_value",
                                             "speedup_vs_sequential", "sequential_readback_value",
                                             "speedup_vs_sequential_readback", "launch", "members_bitwise_equal_plain",
                                             "distinct_member_dt", "roofline")}
        log(f"adjoint leg: coastal 2048x2048, {args.adj_steps} steps, checkpoint every {args.checkpoint}")
        a = adjoint_leg(2048, args.adj_steps, args.checkpoint, fd=True, fd_bed_cells=16)
        log("coastal leg: 16384x16384 flat and dense, 2000 steps")
        c = coastal_leg(16384, 2000)
        out["coastal"] = {k: c[k] for k in ("value", "unit", "steps", "ms_per_step", "config", "roofline",
                                            "dense_ms_per_step", "dense_equals_flat")}
        log("dam-break leg: 256x256, 2000 steps, GPU float32/float64 and the NumPy oracle")
        d = dambreak_leg(2000)
        out["dambreak"] = {k: d[k] for k in ("value", "unit", "steps", "ms_per_step", "config", "float32_vs_float64_gpu",
                                             "float64_gpu_vs_float64_oracle", "float32_gpu_bitwise_oracle",
                                             "cpu_baseline")}
        out["adjoint"] = {k: a[k] for k in ("value", "unit", "steps", "config", "forward_ms_per_step",
                                            "adjoint_ms_per_step", "plain_forward_ms_per_step",
                                            "adjoint_over_plain_forward", "forward_cell_updates_per_s", "grad_n",
                                            "fd_n_eps1e-3", "fd_n_eps1e-4", "fd_vs_adjoint_n", "fd_vs_adjoint_bed",
                                            "fd64_500_steps", "torch_f64_autograd") if k in a}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
