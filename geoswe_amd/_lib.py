"""ctypes binding of libgeoswe_hip.so (the C ABI declared in include/geoswe_abi.h).

The library is the only compute path of this package: there is no CPU fallback. If the
shared object is missing, ``lib()`` raises with the build command; if no GPU is
visible, device calls fail with the HIP error the library reports.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "_lib" / "libgeoswe_hip.so"

GS_BC = {"extrapolate": 0, "periodic": 1, "wall": 2, "fall": 3}
GS_F32, GS_F64, GS_F32X64 = 0, 1, 2
FLAG_CLASS_MASK, FLAG_CFL_SKIP, FLAG_INACTIVE, FLAG_STAGE = 0x1F, 0x20, 0x40, 0x80
FLAG_FORCE = 0x40          # flat layout: sponge / clamp cell (the dense layout's inactive bit)
MAX_CLASSES = 32        # class ids the flags byte holds
MAX_CLASSES_PLANE = 256  # class ids a class plane holds (GsDense.cls, GsFlat.cls)
ST_DONE, ST_NONFINITE = 1, 2
ADJ_MAX_BLOCKS = 8192


class GsClock(C.Structure):
    _fields_ = [("t", C.c_double), ("dt", C.c_double), ("t_end", C.c_double), ("dt_max", C.c_double),
                ("eta64", C.c_double), ("nsteps", C.c_int64), ("max_steps", C.c_int64),
                ("lam_acc", C.c_uint64), ("lam_bits", C.c_uint64), ("blocks_done", C.c_uint32),
                ("status", C.c_int32), ("dt32", C.c_float), ("rain32", C.c_float), ("eta32", C.c_float),
                ("rain_idx", C.c_int32), ("use_host_dt", C.c_int32), ("run_mode", C.c_int32),
                ("pad_", C.c_int32 * 6)]


class GsPhys(C.Structure):
    _fields_ = [("g", C.c_double), ("h_min", C.c_double), ("h_cfl", C.c_double), ("cfl", C.c_double),
                ("dmin", C.c_double), ("lam_floor", C.c_double), ("vcap", C.c_double),
                ("friction", C.c_int32), ("quadratic", C.c_int32), ("keep_h", C.c_int32),
                ("cfl_l2", C.c_int32), ("n_classes", C.c_int32), ("pad_", C.c_int32)]


class GsForcing(C.Structure):
    _fields_ = [("n_rain", C.c_int32), ("rain_gridded", C.c_int32), ("rain_t", C.c_void_p),
                ("rain_rate", C.c_void_p), ("rain_cap", C.c_void_p), ("rain_frames", C.c_void_p),
                ("rain_frame_stride", C.c_int64), ("rain_row_pitch", C.c_int64),
                ("n_stage", C.c_int32), ("pad_", C.c_int32), ("stage_t", C.c_void_p),
                ("stage_v", C.c_void_p)]


class GsDense(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("pitch", C.c_int64), ("bc_x", C.c_int32),
                ("bc_y", C.c_int32), ("dtype", C.c_int32), ("tile_rows", C.c_int32), ("dx", C.c_double),
                ("dy", C.c_double), ("q", (C.c_void_p * 3) * 2), ("bed", C.c_void_p), ("flags", C.c_void_p),
                ("ntab", C.c_void_p), ("n_field", C.c_void_p), ("max_h", C.c_void_p), ("cls", C.c_void_p)]


class GsFlat(C.Structure):
    _fields_ = [("n_stored", C.c_int64), ("n_runs", C.c_int32), ("dtype", C.c_int32), ("nyp", C.c_int32),
                ("ngh", C.c_int32), ("dx", C.c_double), ("dy", C.c_double), ("q", (C.c_void_p * 3) * 2),
                ("bed", C.c_void_p), ("flags", C.c_void_p), ("ntab", C.c_void_p), ("n_field", C.c_void_p),
                ("max_h", C.c_void_p), ("run_k0", C.c_void_p), ("run_len", C.c_void_p),
                ("run_i", C.c_void_p), ("run_j0", C.c_void_p), ("run_d", C.c_void_p),
                ("chunk_run", C.c_void_p), ("n_tiles", C.c_int64), ("tile_e0", C.c_void_p),
                ("tile_irc", C.c_void_p), ("ent", C.c_void_p), ("sponge", C.c_void_p), ("clamp", C.c_void_p), ("infil", C.c_void_p), ("drain", C.c_void_p),
                ("ga_cls", C.c_void_p), ("ga_ks", C.c_void_p), ("ga_psi", C.c_void_p), ("ga_dth", C.c_void_p),
                ("ga_F", C.c_void_p), ("ga_fmax", C.c_void_p), ("ga_inv_tau", C.c_void_p),
                ("rain_lookup", C.c_void_p), ("n_ring_mode", C.c_int64), ("ring_mode_ids", C.c_void_p),
                ("ring_mode_nb", C.c_void_p), ("ring_mode_kind", C.c_void_p), ("ring_mode_eta", C.c_double),
                ("n_force_ring", C.c_int64), ("force_ring", C.c_void_p), ("force_ring_row", C.c_void_p),
                ("ring_ng", C.c_int32), ("ring_nt", C.c_int32), ("ring_bed", C.c_void_p), ("ring_w", C.c_void_p),
                ("ring_t", C.c_void_p), ("ring_stage", C.c_void_p), ("cls", C.c_void_p)]


class GsFlatMeshHost(C.Structure):
    _fields_ = [("n_stored", C.c_int64), ("n_active", C.c_int64), ("n_runs", C.c_int64), ("n_chunks", C.c_int64),
                ("n_tiles", C.c_int64), ("n_entries", C.c_int64), ("sweep_rows", C.c_int32), ("pad_", C.c_int32),
                ("ij", C.c_void_p), ("active", C.c_void_p), ("run_k0", C.c_void_p), ("run_len", C.c_void_p),
                ("run_i", C.c_void_p), ("run_j0", C.c_void_p), ("run_d", C.c_void_p), ("chunk_run", C.c_void_p),
                ("tile_e0", C.c_void_p), ("tile_irc", C.c_void_p), ("ent", C.c_void_p)]


class GsAdjFlat(C.Structure):
    _fields_ = [("n_stored", C.c_int64), ("dtype", C.c_int32), ("n_classes", C.c_int32), ("dx", C.c_double),
                ("dy", C.c_double), ("active", C.c_void_p), ("nbr", C.c_void_p), ("bed", C.c_void_p),
                ("flags", C.c_void_p), ("ntab", C.c_void_p), ("q", C.c_void_p * 3), ("h_next", C.c_void_p),
                ("pre", C.c_void_p * 3), ("lam", C.c_void_p * 3), ("w", C.c_void_p * 3), ("xf", C.c_void_p),
                ("yf", C.c_void_p), ("gz", C.c_void_p), ("gz_prev", C.c_void_p), ("bbar", C.c_void_p),
                ("nbar", C.c_void_p), ("npart", C.c_void_p)]


# symbol -> (restype, argtypes); every entry point returns int
_P = C.c_void_p
_SIGS = {
    "gs_version": (C.c_int, []),
    "gs_last_error": (C.c_char_p, []),
    "gs_get_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "gs_set_device": (C.c_int, [C.c_int]),
    "gs_malloc": (C.c_int, [C.POINTER(C.c_void_p), C.c_size_t]),
    "gs_free": (C.c_int, [_P]),
    "gs_memcpy_h2d": (C.c_int, [_P, _P, C.c_size_t, _P]),
    "gs_memcpy_d2h": (C.c_int, [_P, _P, C.c_size_t, _P]),
    "gs_memcpy_d2d": (C.c_int, [_P, _P, C.c_size_t, _P]),
    "gs_memcpy2d": (C.c_int, [_P, C.c_size_t, _P, C.c_size_t, C.c_size_t, C.c_size_t, C.c_int, _P]),
    "gs_memset": (C.c_int, [_P, C.c_int, C.c_size_t, _P]),
    "gs_stream_create": (C.c_int, [C.POINTER(C.c_void_p)]),
    "gs_stream_destroy": (C.c_int, [_P]),
    "gs_stream_synchronize": (C.c_int, [_P]),
    "gs_event_create": (C.c_int, [C.POINTER(C.c_void_p)]),
    "gs_event_destroy": (C.c_int, [_P]),
    "gs_event_record": (C.c_int, [_P, _P]),
    "gs_event_elapsed_ms": (C.c_int, [C.POINTER(C.c_float), _P, _P]),
    "gs_clock_reset": (C.c_int, [_P, C.c_double, _P]),
    "gs_clock_set_target": (C.c_int, [_P, C.c_double, C.c_double, C.c_int64, C.POINTER(GsPhys), GsForcing, _P]),
    "gs_clock_set_dt": (C.c_int, [_P, C.c_double, C.POINTER(GsPhys), GsForcing, _P]),
    "gs_dense_step": (C.c_int, [C.POINTER(GsDense), C.POINTER(GsPhys), GsForcing, _P, _P]),
    "gs_dense_residual": (C.c_int, [C.POINTER(GsDense), C.POINTER(GsPhys), C.c_int, _P, _P, _P, _P]),
    "gs_dense_cfl": (C.c_int, [C.POINTER(GsDense), C.POINTER(GsPhys), GsForcing, _P, _P]),
    "gs_flat_step": (C.c_int, [C.POINTER(GsFlat), C.POINTER(GsPhys), GsForcing, _P, _P]),
    "gs_flat_residual": (C.c_int, [C.POINTER(GsFlat), C.POINTER(GsPhys), C.c_int, _P, _P, _P, _P]),
    "gs_flat_cfl": (C.c_int, [C.POINTER(GsFlat), C.POINTER(GsPhys), GsForcing, _P, _P]),
    "gs_flat_launch_info": (C.c_int, [C.POINTER(GsFlat), C.POINTER(GsPhys), GsForcing, C.POINTER(C.c_int),
                                      C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "gs_flat_step_pre": (C.c_int, [C.POINTER(GsFlat), C.POINTER(GsPhys), GsForcing, _P, _P, _P, _P, _P]),
    "gs_flat_sample": (C.c_int, [C.POINTER(GsFlat), _P, _P, C.c_int, _P, _P]),
    "gs_adj_flat_step": (C.c_int, [C.POINTER(GsAdjFlat), C.POINTER(GsPhys), C.c_double, _P]),
    "gs_adj_add": (C.c_int, [_P, C.c_int, _P, _P, C.c_int, _P]),
    "gs_flat_run_frozen": (C.c_int, [C.POINTER(GsFlat), C.POINTER(GsPhys), GsForcing, _P, C.c_int, _P, _P, C.c_int,
                                     _P, C.c_int, _P, _P]),
    "gs_flat_replay_pre": (C.c_int, [C.POINTER(GsFlat), C.POINTER(GsPhys), GsForcing, _P, C.c_int, _P, _P, _P, _P]),
    "gs_adj_flat_sweep": (C.c_int, [C.POINTER(GsAdjFlat), C.POINTER(GsPhys), C.c_int, _P, _P, _P, _P, _P, C.c_int,
                                    _P, C.c_int64, _P]),
    "gs_adj_flat_bedgrad": (C.c_int, [C.POINTER(GsAdjFlat), C.POINTER(GsPhys), _P]),
    "gs_adj_flat_nsum": (C.c_int, [C.POINTER(GsAdjFlat), _P]),
    "gs_ens_create": (C.c_int, [C.POINTER(GsFlat), C.POINTER(GsPhys), C.POINTER(GsForcing), C.POINTER(_P), C.c_int,
                                C.POINTER(_P)]),
    "gs_ens_destroy": (C.c_int, [_P]),
    "gs_ens_set_launch": (C.c_int, [_P, C.c_int, C.c_int]),
    "gs_ens_get_launch": (C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int64)]),
    "gs_ens_step": (C.c_int, [_P, _P]),
    "gs_ens_run": (C.c_int, [_P, C.c_int64, C.c_int, _P]),
    "gs_ens_autotune": (C.c_int, [_P, C.c_int, _P, C.POINTER(C.c_double)]),
    "gs_ens_run_noalloc": (C.c_int, [C.POINTER(GsFlat), C.POINTER(GsPhys), C.POINTER(GsForcing), C.POINTER(_P),
                                     C.c_int, C.c_int64, _P]),
    "gs_fill": (C.c_int, [_P, C.c_int, C.c_int64, C.c_int64, C.c_int64, C.c_double, _P]),
    "gs_gen_bed": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.c_uint32,
                             C.c_int,
                             C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                             C.c_int, _P]),
    "gs_bench_copy": (C.c_int, [_P, _P, C.c_size_t, _P]),
    "gs_bench_valu": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, _P]),
    "gs_crmath_check": (C.c_int, [C.c_int, _P, _P]),
    "gs_flat_mesh_build": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _P]),
    "gs_flat_mesh_free": (C.c_int, [_P]),
    "gs_gen_voronoi": (C.c_int, [_P, C.c_int, C.c_int, C.c_int64, C.c_int64, C.c_int64, _P, _P, _P, C.c_int, _P]),
    "gs_graph_begin": (C.c_int, [_P]),
    "gs_graph_end": (C.c_int, [_P, C.POINTER(_P)]),
    "gs_graph_nodes": (C.c_int, [_P, C.POINTER(C.c_int64)]),
    "gs_graph_launch": (C.c_int, [_P, _P]),
    "gs_graph_destroy": (C.c_int, [_P]),
}

_LIB = None


class GsError(RuntimeError):
    """A failing libgeoswe_hip call (usage error code or HIP error)."""


def lib():
    """Load libgeoswe_hip.so once; raise if it has not been built."""
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise ImportError(
                f"{LIB_PATH} is missing: build the HIP library first "
                f"(python -c 'import __graft_entry__ as g; g.build()' or make -C geoswe_amd/csrc)")
        h = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = h
    return _LIB


def call(name: str, *args) -> None:
    """Call an int-returning entry point and raise GsError on a non-zero code."""
    rc = getattr(lib(), name)(*args)
    if rc != 0:
        msg = lib().gs_last_error().decode(errors="replace")
        raise GsError(f"{name} failed ({rc}): {msg}")


def device_count() -> int:
    n = C.c_int(0)
    try:
        call("gs_get_device_count", C.byref(n))
    except GsError:
        return 0
    return int(n.value)


class DeviceArray:
    """A device allocation owned by Python (freed on garbage collection)."""

    def __init__(self, shape, dtype, zero: bool = False, stream=None):
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        p = C.c_void_p()
        call("gs_malloc", C.byref(p), C.c_size_t(max(self.nbytes, 1)))
        self.ptr = int(p.value)
        self._lib = lib()
        if zero:
            call("gs_memset", C.c_void_p(self.ptr), 0, C.c_size_t(self.nbytes), stream)

    @classmethod
    def from_host(cls, a: np.ndarray, stream=None) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a, stream)
        return d

    def upload(self, a: np.ndarray, stream=None) -> None:
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.nbytes != self.nbytes:
            raise ValueError(f"upload of {a.nbytes} bytes into a {self.nbytes}-byte array")
        call("gs_memcpy_h2d", C.c_void_p(self.ptr), C.c_void_p(a.ctypes.data), C.c_size_t(a.nbytes), stream)
        if stream is not None:
            sync(stream)

    def download(self, stream=None) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        call("gs_memcpy_d2h", C.c_void_p(out.ctypes.data), C.c_void_p(self.ptr), C.c_size_t(self.nbytes), stream)
        sync(stream)
        return out

    def offset(self, elems: int) -> int:
        return self.ptr + int(elems) * self.dtype.itemsize

    def __del__(self):
        p = getattr(self, "ptr", 0)
        if p and getattr(self, "_lib", None) is not None:
            try:
                self._lib.gs_free(C.c_void_p(p))
            except Exception:
                pass
            self.ptr = 0


def sync(stream=None) -> None:
    call("gs_stream_synchronize", stream)


class StepGraph:
    """Steps recorded into an executable HIP graph (gs_graph_*). ``replay()`` enqueues the
    recorded launches on the owner's stream; ``close()`` releases the graph."""

    def __init__(self, record, stream=None):
        """Run ``record(capture_stream)`` under stream capture: every entry point it calls on
        that stream is recorded instead of run."""
        cap = C.c_void_p()
        call("gs_stream_create", C.byref(cap))
        self._h = None
        self.stream = stream
        try:
            call("gs_graph_begin", cap)
            err = None
            try:
                record(cap)
            except BaseException as e:       # end the capture in any case, then re-raise
                err = e
            h = C.c_void_p()
            rc = lib().gs_graph_end(cap, C.byref(h))
            if err is not None:
                if h.value:
                    call("gs_graph_destroy", h)
                raise err
            if rc != 0:
                raise GsError(f"gs_graph_end failed ({rc}): {lib().gs_last_error().decode(errors='replace')}")
            self._h = h
        finally:
            call("gs_stream_destroy", cap)
        n = C.c_int64(0)
        call("gs_graph_nodes", self._h, C.byref(n))
        self.nodes = int(n.value)

    def replay(self, times: int = 1) -> None:
        if self._h is None:
            raise RuntimeError("StepGraph: closed")
        for _ in range(int(times)):
            call("gs_graph_launch", self._h, self.stream)

    def close(self) -> None:
        if self._h is not None:
            call("gs_graph_destroy", self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def phys_struct(g, h_min, h_cfl, cfl, dmin, vcap, friction, quadratic, keep_h, cfl_l2, n_classes,
                floor_depth=None) -> GsPhys:
    """The step's scalars. ``floor_depth`` is the depth whose gravity-wave speed floors the
    CFL wave speed: the dense solver's ``h_cfl`` (solver.py:2663-2665, the default), the
    flat solver's ``h_min`` (compressed_solver.py:713-719, 741-745)."""
    p = GsPhys()
    p.g, p.h_min, p.h_cfl, p.cfl, p.dmin = float(g), float(h_min), float(h_cfl), float(cfl), float(dmin)
    # the dry-partition floor, with the reference's own host expression
    p.lam_floor = (float(g) * float(h_cfl if floor_depth is None else floor_depth)) ** 0.5
    p.vcap = float(vcap) if vcap is not None else math.inf
    p.friction, p.quadratic, p.keep_h, p.cfl_l2 = int(friction), int(quadratic), int(keep_h), int(cfl_l2)
    p.n_classes = int(n_classes)
    return p


def read_clock(dev: DeviceArray, stream=None) -> GsClock:
    raw = dev.download(stream)
    c = GsClock()
    C.memmove(C.byref(c), raw.ctypes.data, C.sizeof(GsClock))
    return c


def new_clock(t0: float = 0.0, stream=None) -> DeviceArray:
    assert C.sizeof(GsClock) == 128, C.sizeof(GsClock)
    d = DeviceArray((C.sizeof(GsClock),), np.uint8)
    call("gs_clock_reset", C.c_void_p(d.ptr), C.c_double(t0), stream)
    return d
