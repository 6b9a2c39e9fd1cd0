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
BYTES_PER_CELL = {"dense": 28.0, "flat": 28.0}   # q read 12 + bed 4 + q write 12 (uniform n, uniform rain)


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


def cpu_baseline(bed_spec, crop=(1024, 1000), steps=30):
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


def copy_calibration(nbytes: int, reps: int):
    """Time gs_bench_copy (read n + write n bytes per launch); prints one JSON line."""
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
    print(json.dumps({"copy_bytes": nbytes, "reps": reps, "ms_per_copy": per,
                      "GBps_read_plus_write": 2 * nbytes / (per * 1e-3) / 1e9}), flush=True)


def coastal(args):
    """The compressed coastal case: flat and dense layouts stepped side by side, depth and
    residual compared bit for bit at steps 1, 500 and the last; both timed with HIP events.
    Prints one JSON line for the flat layout (the production layout of this case)."""
    from geoswe_amd import _lib as L
    from geoswe_amd.cases import coastal_c3
    n = args.n
    t0 = time.perf_counter()
    log(f"building the coastal case {n}x{n}")
    case = coastal_c3(n=n)
    d = case.dense(track_max_depth=False)
    f = case.flat(track_max_depth=False)
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
    marks = [1, 500, args.steps] if args.steps > 500 else [1, args.steps]
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
                   "layout": "flat", "active_cells": int(nact), "stored_cells": int(fm.n_stored)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "gs::k_flat_step<float, 1>",
                     "kernel_ms": ms_f, "bytes_per_cell": bpc},
        "dense_ms_per_step": ms_d,
        "dense_equals_flat": ok,
    }
    print(json.dumps(out), flush=True)
    if not ok:
        sys.exit(1)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--nx", type=int, default=25600)
    ap.add_argument("--ny", type=int, default=25000)
    ap.add_argument("--layout", choices=["dense", "flat"], default="dense")
    ap.add_argument("--tile-rows", type=int, default=0)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--traffic", default=str(ROOT / "profiles" / "dense_sweep_rain_on_slope_640M_traffic.json"))
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="instead of the step, time the plain copy kernel over this many bytes (HBM calibration)")
    ap.add_argument("--copy-reps", type=int, default=10)
    ap.add_argument("--workload", choices=["rain-on-slope", "coastal"], default="rain-on-slope",
                    help="coastal: the compressed coastal case (flat and dense layouts, bitwise cross-check)")
    ap.add_argument("--n", type=int, default=16384, help="coastal box side")
    args = ap.parse_args()
    if args.copy_bytes:
        return copy_calibration(args.copy_bytes, args.copy_reps)
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported by this build (a multi-GPU run "
              f"would split the domain into strips along y with a 2-cell halo exchange per step)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "coastal":
        return coastal(args)

    from geoswe_amd import _lib as L
    from geoswe_amd.cases import FbmBed, C2_SEED, rain_on_slope_c2
    from geoswe_amd.compressed import CompressedSolver

    nx, ny = args.nx, args.ny
    cells = nx * ny
    log(f"building the rain-on-slope case {nx}x{ny} ({cells / 1e6:.1f} M cells), layout {args.layout}")
    bed = FbmBed(seed=C2_SEED).normalise(nx, ny)
    s = rain_on_slope_c2(nx, ny, "float32", tile_rows=args.tile_rows, bed=bed)
    if args.layout == "flat":
        s = CompressedSolver.from_dense(s, cfl_linf=True, track_max_depth=False)
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
    tp = Path(args.traffic)
    if tp.exists() and args.layout == "dense" and (nx, ny) == (25600, 25000):
        traffic = json.loads(tp.read_text()).get("bytes_per_launch")
    log(f"{ms_step:.3f} ms/step wall, {kern_ms:.3f} ms/launch (events), {value / 1e9:.2f} G cell-updates/s, "
        f"{achieved:.0f} GB/s algorithmic")
    h32 = None
    err64 = None
    if not args.no_fp64:
        log("float64 run of the same workload (FP64 oracle build of the kernel)")
        h32 = s.depth() if args.layout == "dense" else None
        if h32 is not None:
            del s
            s64 = rain_on_slope_c2(nx, ny, "float64", bed=bed, f32_inputs=True)
            s64.advance(args.warmup + args.steps)
            h64 = s64.depth()
            e = np.abs(h32.astype(np.float64) - h64)
            err64 = float(e.max())
            k = int(np.argmax(e))
            log(f"max error at cell {np.unravel_index(k, e.shape)}, h64 there {h64.flat[k]:.4g} m; "
                f"99.999th percentile {np.percentile(e, 99.999):.3e} m, median {np.median(e):.3e} m")
            log(f"max |h32 - h64| after {args.warmup + args.steps} steps: {err64:.3e} m")
            del s64, h64
    cpu = None
    if not args.no_cpu:
        log("CPU baseline (NumPy oracle on a crop)")
        cpu = cpu_baseline(bed)
        log(f"CPU {cpu['value'] / 1e6:.2f} M cell-updates/s")
    out = {
        "metric": "cell-updates/s (fused FV step, 1 MI355X)",
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
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "gs::k_dense_sweep<float, 1>" if args.layout == "dense" else "gs::k_flat_step<float, 1>",
                     "kernel_ms": kern_ms, "bytes_per_cell": BYTES_PER_CELL[args.layout]},
        "max_depth_err_vs_fp64_m": err64,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
