"""Gradients of a gauge-depth misfit with respect to per-class Manning n and per-cell bed.

The forward model is a ``CompressedSolver`` stepped with a frozen dt sequence (the CFL
steps of a reference run, replayed with ``gs_clock_set_dt``), so dt carries no
derivative. The loss is

    J = sum_{s=1..S} sum_g (h_s(g) - h*_s(g))^2

over the depths h_s at gauge cells after every step. Its gradient comes from the
hand-written reverse step ``gs_adj_flat_step`` (geoswe_amd/csrc/gs_adjoint.hip) with
two-level checkpointing: the forward pass keeps the state every K steps; the reverse pass
recomputes each K-step segment from its checkpoint (``gs_flat_step_pre``, which also keeps
the pre-friction state the update's reverse needs), stores that segment's states on the
device, and runs the K reverse steps over them. Memory is O(S/K + K) states; compute is
two forward passes plus the reverse steps.

``GaugeMisfitFn`` wraps this in a ``torch.autograd.Function`` of (ntab, bed).

The reference (src/geoswe) has no adjoint; the forward it differentiates is the flat
fused step of compressed_rhs.py:673-865 as built here.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _lib as L
from .compressed import CompressedSolver


def flat_neighbours(fm) -> np.ndarray:
    """(N, 4) int32 flat ids of (i+1, j), (i-1, j), (i, j+1), (i, j-1) of every stored
    cell, -1 where that cell is not stored. Checks that every active cell has its four
    neighbours and their neighbours along the same axis (the ring-2 stored set)."""
    if fm.n_stored >= 2 ** 31:
        raise ValueError("flat_neighbours: more than 2^31 stored cells")
    idx = np.full((fm.nxp, fm.nyp), -1, np.int32)
    i, j = fm.ij[:, 0].astype(np.int64), fm.ij[:, 1].astype(np.int64)
    idx[i, j] = np.arange(fm.n_stored, dtype=np.int32)

    def at(di, dj):
        ii, jj = i + di, j + dj
        ok = (ii >= 0) & (ii < fm.nxp) & (jj >= 0) & (jj < fm.nyp)
        out = np.full(fm.n_stored, -1, np.int32)
        out[ok] = idx[ii[ok], jj[ok]]
        return out

    nb = np.stack([at(1, 0), at(-1, 0), at(0, 1), at(0, -1)], 1)
    act = fm.active
    if (nb[act] < 0).any():
        raise ValueError("flat_neighbours: an active cell lacks a stored neighbour")
    for ax, (u, d) in enumerate(((0, 1), (2, 3))):
        for k in (u, d):
            m = nb[act, k]
            if (nb[m, u] < 0).any() or (nb[m, d] < 0).any():
                raise ValueError("flat_neighbours: the stored set is not two cells deep around the active set")
    return nb.astype(np.int32)


class GaugeMisfit:
    """J(ntab, bed) for one CompressedSolver, a frozen dt sequence and gauge cells."""

    def __init__(self, solver: CompressedSolver, gauges: np.ndarray, dts: np.ndarray,
                 h_star: Optional[np.ndarray] = None, checkpoint_every: int = 100):
        if solver._cls is not None:
            raise NotImplementedError("GaugeMisfit: the reverse kernels read the Manning class from the flags "
                                      "byte (at most 32 classes); this solver has a class plane")
        if solver.has_coastal_forcing:
            raise NotImplementedError("GaugeMisfit: the coastal forcings (sponge, gauge ring, clamp, infiltration, "
                                      "drain) have no reverse kernels; build the misfit on a solver without them")
        self.s = solver
        solver._tables_taken = True
        fm = solver.fm
        self.dtype = solver.dtype
        self.N = fm.n_stored
        self.gauges = np.asarray(gauges, np.int64)
        if self.gauges.ndim != 1 or not len(self.gauges):
            raise ValueError("gauges must be a non-empty 1-D array of flat ids")
        if (self.gauges < 0).any() or (self.gauges >= self.N).any() or not fm.active[self.gauges].all():
            raise ValueError("gauges must be active flat ids")
        self.dts = np.asarray(dts, np.float64)
        if self.dts.ndim != 1 or not len(self.dts) or not np.all(self.dts > 0):
            raise ValueError("dts must be a non-empty 1-D array of positive steps")
        self.S = len(self.dts)
        self._dts_c = np.ascontiguousarray(self.dts, np.float64)     # read by the C loops on the host
        self.K = int(checkpoint_every)
        if self.K < 1:
            raise ValueError("checkpoint_every must be >= 1")
        G = len(self.gauges)
        self.h_star = np.zeros((self.S, G)) if h_star is None else np.asarray(h_star, np.float64)
        if self.h_star.shape != (self.S, G):
            raise ValueError(f"h_star must have shape ({self.S}, {G})")
        self._d_gauges = L.DeviceArray.from_host(self.gauges)
        self._d_samples = L.DeviceArray((self.S, G), self.dtype, zero=True)
        self._nbr = L.DeviceArray.from_host(flat_neighbours(fm))
        self._active = L.DeviceArray.from_host(fm.active.astype(np.uint8))
        cur = solver.nsteps & 1
        self._q0 = [L.DeviceArray((self.N,), self.dtype) for _ in range(3)]
        for k in range(3):
            L.call("gs_memcpy_d2d", C.c_void_p(self._q0[k].ptr), C.c_void_p(solver._bufs[cur][k].ptr),
                   C.c_size_t(self._q0[k].nbytes), None)
        self.t0 = solver.t
        self._ckpt = {}
        self._ckpt_t = {}
        self.samples = None

    # ------------------------------------------------------------------ parameters
    @property
    def n_classes(self) -> int:
        return self.s._n_classes

    def set_params(self, ntab=None, bed=None) -> None:
        """Manning n per class (n_classes,) and/or the flat bed (N,), host arrays."""
        if ntab is not None:
            t = np.zeros(L.MAX_CLASSES_PLANE, self.dtype)
            nt = np.asarray(ntab, np.float64).ravel()
            if len(nt) != self.n_classes:
                raise ValueError(f"ntab has {len(nt)} entries, the solver {self.n_classes} classes")
            t[:len(nt)] = nt
            self.s._ntab.upload(t)
        if bed is not None:
            b = np.asarray(bed).ravel()
            if len(b) != self.N:
                raise ValueError(f"bed has {len(b)} entries, the flat mesh {self.N}")
            self.s._bed.upload(b.astype(self.dtype))

    def params(self):
        return (self.s._ntab.download()[:self.n_classes].astype(np.float64),
                self.s._bed.download().astype(np.float64))

    # ------------------------------------------------------------------ forward
    def _reset(self, q_src, t0: float) -> None:
        s = self.s
        for k in range(3):
            L.call("gs_memcpy_d2d", C.c_void_p(s._bufs[0][k].ptr), C.c_void_p(q_src[k].ptr),
                   C.c_size_t(self._q0[k].nbytes), None)
        L.call("gs_clock_reset", C.c_void_p(s._clock.ptr), C.c_double(t0), None)
        s._lam_valid = False
        s._target = None

    def _set_dt(self, dt: float, p) -> None:
        s = self.s
        L.call("gs_clock_set_dt", C.c_void_p(s._clock.ptr), C.c_double(dt), p, s._fo.s, None)

    def simulate(self, keep_checkpoints: bool = True) -> np.ndarray:
        """S frozen-dt steps from the initial state; gauge depths (S, G) after each step."""
        s = self.s
        self._reset(self._q0, self.t0)
        K = self.K
        nseg = (self.S + K - 1) // K
        ck = None
        if keep_checkpoints:
            for k in range(nseg):
                if k not in self._ckpt:
                    self._ckpt[k] = [L.DeviceArray((self.N,), self.dtype) for _ in range(3)]
            ck = (C.c_void_p * (3 * nseg))(*[self._ckpt[k][m].ptr for k in range(nseg) for m in range(3)])
            t = self.t0                     # the clock's own float64 sums, in the same order
            for st in range(self.S):
                if st % K == 0:
                    self._ckpt_t[st // K] = t
                t = t + float(self.dts[st])
        L.call("gs_flat_run_frozen", C.byref(s._desc()), C.byref(s._phys()), s._fo.s, C.c_void_p(s._clock.ptr),
               self.S, C.c_void_p(self._dts_c.ctypes.data), C.c_void_p(self._d_gauges.ptr), len(self.gauges),
               C.c_void_p(self._d_samples.ptr), K, ck, None)
        L.sync(None)
        c = s._clockv()
        if c.status & L.ST_NONFINITE:
            raise FloatingPointError("GaugeMisfit: non-finite state in the forward run")
        if c.nsteps != self.S:
            raise RuntimeError(f"GaugeMisfit: {c.nsteps} steps taken, {self.S} expected")
        self.samples = self._d_samples.download().astype(np.float64)
        return self.samples

    def loss(self, samples: Optional[np.ndarray] = None) -> float:
        h = self.samples if samples is None else samples
        return float(np.sum((h - self.h_star) ** 2))

    # ------------------------------------------------------------------ reverse
    def gradient(self):
        """(dJ/dn per class (n_classes,), dJ/dbed (N,)) at the parameters of the last
        simulate() (float64 host arrays). Leaves the solver at the initial state and time."""
        if self.samples is None:
            raise RuntimeError("gradient: call simulate() first")
        s, N, K, S = self.s, self.N, self.K, self.S
        dt_ = self.dtype
        G = len(self.gauges)
        nseg = (S + K - 1) // K
        if not hasattr(self, "_ws"):     # device workspace, allocated once per problem
            nblk = min(max(1, (N + 255) // 256), L.ADJ_MAX_BLOCKS)
            self._ws = dict(
                seg_states=[[L.DeviceArray((N,), dt_) for _ in range(3)] for _ in range(K + 1)],
                seg_pre=[[L.DeviceArray((N,), dt_) for _ in range(3)] for _ in range(K)],
                lam=[L.DeviceArray((N,), dt_) for _ in range(3)],
                w=[L.DeviceArray((N,), dt_) for _ in range(3)],
                xf=L.DeviceArray((10, N), dt_), yf=L.DeviceArray((10, N), dt_),
                gz=L.DeviceArray((2, 2, N), dt_),          # double-buffered bed-gradient cotangents
                bbar=L.DeviceArray((N,), np.float64), nbar=L.DeviceArray((L.MAX_CLASSES,), np.float64),
                npart=L.DeviceArray((nblk, 32), np.float64))
        ws = self._ws
        seg_states, seg_pre = ws["seg_states"], ws["seg_pre"]
        lam, w, xf, yf, gz = ws["lam"], ws["w"], ws["xf"], ws["yf"], ws["gz"]
        bbar, nbar, npart = ws["bbar"], ws["nbar"], ws["npart"]
        for arr in lam + w + [xf, yf, gz, bbar, nbar, npart]:   # accumulators and scratch start at zero
            L.call("gs_memset", C.c_void_p(arr.ptr), 0, C.c_size_t(arr.nbytes), None)
        coef = (2.0 * (self.samples - self.h_star)).astype(dt_)          # (S, G): dJ/dh_s at gauges
        d_coef = L.DeviceArray.from_host(coef)
        es = d_coef.dtype.itemsize
        dcode = L.GS_F32 if dt_ == np.float32 else L.GS_F64
        a = L.GsAdjFlat()
        a.n_stored, a.dtype, a.n_classes = N, dcode, self.n_classes
        a.dx, a.dy = float(s.mesh.dx), float(s.mesh.dy)
        a.active, a.nbr, a.bed = self._active.ptr, self._nbr.ptr, s._bed.ptr
        a.flags, a.ntab = (s._flags.ptr if s._flags is not None else None), s._ntab.ptr
        for k in range(3):
            a.lam[k], a.w[k] = lam[k].ptr, w[k].ptr
        a.xf, a.yf, a.gz, a.bbar, a.nbar, a.npart = xf.ptr, yf.ptr, gz.ptr, bbar.ptr, nbar.ptr, npart.ptr
        phys = s._phys()
        p = C.byref(phys)
        desc = s._desc()
        clk = C.c_void_p(s._clock.ptr)
        # cotangent of q_S: the last step's gauge term
        L.call("gs_adj_add", C.c_void_p(lam[0].ptr), dcode, C.c_void_p(self._d_gauges.ptr),
               C.c_void_p(d_coef.ptr + (S - 1) * G * es), G, None)
        nbytes = self._q0[0].nbytes
        # every slot of the segment store starts as the initial state: the step writes only
        # active cells, and the ring cells are frozen at their initial values
        for st_ in seg_states:
            for m in range(3):
                L.call("gs_memcpy_d2d", C.c_void_p(st_[m].ptr), C.c_void_p(self._q0[m].ptr), C.c_size_t(nbytes),
                       None)
        state_ptrs = (C.c_void_p * (3 * (K + 1)))(*[x[m].ptr for x in seg_states for m in range(3)])
        pre_ptrs = (C.c_void_p * (3 * K))(*[x[m].ptr for x in seg_pre for m in range(3)])
        for seg in range(nseg - 1, -1, -1):
            s0 = seg * K
            n_in = min(K, S - s0)
            # recompute the segment from its checkpoint straight into the store (step j reads
            # slot j, writes slot j + 1 and keeps its pre-friction state), then sweep it back
            self._reset(self._ckpt[seg], self._ckpt_t[seg])
            for m in range(3):
                L.call("gs_memcpy_d2d", C.c_void_p(seg_states[0][m].ptr), C.c_void_p(self._ckpt[seg][m].ptr),
                       C.c_size_t(nbytes), None)
            dptr = C.c_void_p(self._dts_c.ctypes.data + 8 * s0)
            L.call("gs_flat_replay_pre", C.byref(desc), p, s._fo.s, clk, n_in, dptr, state_ptrs, pre_ptrs, None)
            # the loss term of the state after step s0 + j, for s0 + j >= 1
            coef_ptrs = (C.c_void_p * n_in)(*[(d_coef.ptr + (s0 + j - 1) * G * es) if s0 + j >= 1 else None
                                              for j in range(n_in)])
            L.call("gs_adj_flat_sweep", C.byref(a), p, n_in, dptr, state_ptrs, pre_ptrs,
                   C.c_void_p(self._d_gauges.ptr), coef_ptrs, G, C.c_void_p(gz.ptr), C.c_int64(s0), None)
        # the bed-gradient term of the last reverse step (step 0 wrote half 0), the class sums
        a.gz, a.gz_prev = gz.ptr, None
        L.call("gs_adj_flat_bedgrad", C.byref(a), p, None)
        L.call("gs_adj_flat_nsum", C.byref(a), None)
        # leave the solver at the initial state and time (the replay moved its clock)
        self._reset(self._q0, self.t0)
        L.sync(None)
        return nbar.download()[:self.n_classes].copy(), bbar.download().copy()


try:
    import torch

    class GaugeMisfitFn(torch.autograd.Function):
        """J = GaugeMisfitFn.apply(ntab, bed, problem): the misfit as a torch scalar whose
        backward runs the device adjoint (ntab (n_classes,), bed (N,) in flat order)."""

        @staticmethod
        def forward(ctx, ntab, bed, problem):
            problem.set_params(ntab.detach().cpu().double().numpy(), bed.detach().cpu().double().numpy())
            problem.simulate()
            ctx.problem = problem
            ctx.dev = (ntab.device, bed.device)
            return torch.tensor(problem.loss(), dtype=torch.float64, device=ntab.device)

        @staticmethod
        def backward(ctx, g):
            gn, gb = ctx.problem.gradient()
            gs = float(g)
            return (torch.as_tensor(gn * gs, device=ctx.dev[0]), torch.as_tensor(gb * gs, device=ctx.dev[1]),
                    None)
except ImportError:     # torch is optional for the solver path
    GaugeMisfitFn = None
