"""NumPy restatement of GeoSWE's production finite-volume time step (test oracle).

Test infrastructure only: imported by ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py``; never by the product package.

The arithmetic follows the reference's fused CUDA kernel text term by term and in the
same evaluation order (no FMA contraction), so in float32 it is the bit pattern the
production kernel is specified to produce, and in float64 it is the FP64 oracle. The
array-op structure (face arrays, ``np.where`` chains) mirrors the reference's own NumPy
path, which makes its timing representative of that CPU path.

Sources followed (paths relative to the reference checkout):

* primitives with dry gate ............ src/geoswe/rhs_cuda.py:931-938 (swe.py:12-27)
* central bed gradients ............... rhs_cuda.py:1644-1650 (well_balanced.py:234-241)
* SRM face reconstruction ............. rhs_cuda.py:1042-1184 (well_balanced.py:244-269)
* HLLC flux ........................... rhs_cuda.py:986-1036 (flux.py:56-151)
* one-sided SRM bed source ............ rhs_cuda.py:1060-1183 (well_balanced.py:303-362)
* fused update tail ................... rhs_cuda.py:1757-1796; flat form compressed_rhs.py:716-748
* CFL wave speed ...................... solver.py:152-205, 2569-2666; compressed_solver.py:200-240
* dt rule and rain step limit ......... solver.py:2657-2666, 2975-3027
* ghost-cell fill ..................... bc.py:50-98; ghost bed solver.py:1466-1484
* rain lookup / stage boundary ........ forcing.py:170-216, 315-377
* open-boundary sponge, stage clamp .... compressed_solver.py:1320-1326, 424-440
* IDW gauge ring ....................... compressed_solver.py:287-307, 2338-2348
* infiltration, karst drain ............ compressed_solver.py:143-158, 269-286
* Green-Ampt infiltration, drain-tau ... compressed_solver.py:320-421
* ring modes of the cached replay ...... compressed_solver.py:450-465, 486-575, 2134-2144

Documented departures (shared with the HIP kernels, which implement the same rule):

* Each face is evaluated with elevations relative to its lower cell's bed (a per-face
  shift of the vertical datum). The expressions are the reference's in exact arithmetic;
  in float32 the face depths and the bed-slope source no longer lose |b|-sized ulps (on a
  77 m slope one ulp of b is 7.6e-6 m, a 0.25 % error in the source of a 3 m cell).

* ``h ** (-4/3)`` in the friction coefficient uses ``pow_m43`` (an inverse-cube-root
  Newton iteration built from IEEE +,-,* only) instead of ``powf(h, -4/3)``, so CPU and
  GPU results are reproducible bit for bit. Its error against the exact power is below
  13 ulp, the same order as ``powf(h, (float)(-4/3))`` itself (which carries the
  rounding of the exponent).
* HLLC star celerity: the reference's ``a* = sqrt(g * (h_rt^2 / g))`` is taken as ``h_rt``
  (for ``h_rt > 0``, else 0), its exact value; it differs from the reference's rounding by
  at most a few ulp and saves a division and a square root per face.
* The CFL reduction keeps NaN (IEEE-bit maximum), as the reference flat kernels do
  (compressed_solver.py:200-240); the dense reference kernel drops a NaN lambda.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

G = 9.81

BC_KINDS = ("extrapolate", "periodic", "wall", "fall")


# ---------------------------------------------------------------------------
# boundary fill (bc.py:50-98) and ghost bed (solver.py:1466-1484)
# ---------------------------------------------------------------------------

def bc_fill(q: np.ndarray, ngh: int, bc_x: str, bc_y: str) -> None:
    """Fill the ghost layers of a padded state ``q`` (3, nxp, nyp) in place, x then y.

    Restates bc.py:50-98: extrapolate copies the last interior layer, periodic wraps,
    wall mirrors with the normal momentum negated, fall zeroes the ghosts.
    Pinned by tests/test_oracle_invariants.py::test_ghost_fill_by_hand (hand-built cases per
    kind) and by the reference's golden residuals under wall, extrapolate and periodic edges.
    """
    for axis, kind in ((1, bc_x), (2, bc_y)):
        n = q.shape[axis]
        sl = lambda a, b: (slice(None),) * axis + (slice(a, b),)
        if kind == "periodic":
            q[sl(0, ngh)] = q[sl(n - 2 * ngh, n - ngh)]
            q[sl(n - ngh, n)] = q[sl(ngh, 2 * ngh)]
        elif kind == "extrapolate":
            q[sl(0, ngh)] = q[sl(ngh, ngh + 1)]
            q[sl(n - ngh, n)] = q[sl(n - ngh - 1, n - ngh)]
        elif kind == "wall":
            lo = np.flip(q[sl(ngh, 2 * ngh)], axis=axis).copy()
            hi = np.flip(q[sl(n - 2 * ngh, n - ngh)], axis=axis).copy()
            lo[axis] = -lo[axis]          # normal momentum: hu for x (axis 1), hv for y (axis 2)
            hi[axis] = -hi[axis]
            q[sl(0, ngh)] = lo
            q[sl(n - ngh, n)] = hi
        elif kind == "fall":
            q[sl(0, ngh)] = 0
            q[sl(n - ngh, n)] = 0
        else:
            raise ValueError(f"unknown BC kind {kind!r}")


def pad_bed(b_int: np.ndarray, ngh: int, bc_x: str, bc_y: str, dtype=np.float32) -> np.ndarray:
    """Padded bed with ghosts edge-replicated (periodic axes wrap), x first (solver.py:1466-1484)."""
    nx, ny = b_int.shape
    b = np.zeros((nx + 2 * ngh, ny + 2 * ngh), dtype=dtype)
    b[ngh:ngh + nx, ngh:ngh + ny] = b_int.astype(dtype, copy=False)
    if bc_x == "periodic":
        b[:ngh, :] = b[-2 * ngh:-ngh, :]
        b[-ngh:, :] = b[ngh:2 * ngh, :]
    else:
        b[:ngh, :] = b[ngh:ngh + 1, :]
        b[-ngh:, :] = b[-ngh - 1:-ngh, :]
    if bc_y == "periodic":
        b[:, :ngh] = b[:, -2 * ngh:-ngh]
        b[:, -ngh:] = b[:, ngh:2 * ngh]
    else:
        b[:, :ngh] = b[:, ngh:ngh + 1]
        b[:, -ngh:] = b[:, -ngh - 1:-ngh]
    return b


# ---------------------------------------------------------------------------
# scalar helpers
# ---------------------------------------------------------------------------

def kernel_consts(dx: float, dy: float, g: float, h_min: float, dtype):
    """The kernel's scalar arguments as the reference casts them (solver.py:2275-2279):
    ``inv_dx = (T)(1/dx)`` and, inside the kernel, ``dx = 1/inv_dx`` in T (rhs_cuda.py:928-929)."""
    T = np.dtype(dtype).type
    inv_dx = T(1.0 / dx)
    inv_dy = T(1.0 / dy)
    return dict(T=T, inv_dx=inv_dx, inv_dy=inv_dy, dxk=T(1) / inv_dx, dyk=T(1) / inv_dy,
                g=T(g), h_min=T(h_min))


def icbrt_f32(h: np.ndarray) -> np.ndarray:
    """``h ** (-1/3)`` in float32 from IEEE +,-,* only (bit-reproducible on the GPU).

    Newton's method on f(y) = y^-3 - h, ``y <- y * (4/3 - (h/3) * y^3)``, three iterations
    from the magic-constant seed ``bits(y0) = 0x54a2fa8c - bits(h)/3`` (Moroz et al.,
    "Fast calculation of cube and inverse cube roots using a magic constant", 2021).
    """
    T = np.float32
    h = np.asarray(h).astype(T, copy=False)
    bits = h.view(np.uint32)
    y = (np.uint32(0x54A2FA8C) - bits // np.uint32(3)).view(T)
    xt = h * T(1.0 / 3.0)
    c43 = T(4.0 / 3.0)
    for _ in range(3):
        y = y * (c43 - xt * ((y * y) * y))
    return y


def pow_m43(h: np.ndarray) -> np.ndarray:
    """``h ** (-4/3)`` for ``h > 0`` from IEEE +,-,* only, so CPU and GPU agree bit for bit.

    float32: ``y = icbrt_f32(h)``, then ``(y*y)*(y*y)``. float64: the float32 root refined
    by two more Newton steps in float64, then ``(y*y)*(y*y)``.
    Pinned by tests/test_oracle_invariants.py against ``h**(-4/3)`` (float32 < 13 ulp,
    float64 <= 16 ulp).
    """
    h = np.asarray(h)
    if h.dtype == np.float64:
        y = icbrt_f32(h.astype(np.float32)).astype(np.float64)
        xt = h * (1.0 / 3.0)
        for _ in range(2):
            y = y * (4.0 / 3.0 - xt * ((y * y) * y))
        y2 = y * y
        return y2 * y2
    y = icbrt_f32(h)
    y2 = y * y
    return y2 * y2


# ---------------------------------------------------------------------------
# HLLC (rhs_cuda.py:986-1036) and SRM face reconstruction (rhs_cuda.py:1042-1184)
# ---------------------------------------------------------------------------

def hllc(hL, uNL, uTL, hR, uNR, uTR, g, h_min):
    """HLLC flux (Fh, FN, FT) in the face frame, kernel form.

    Differs from flux.py:80-151 exactly as the production kernel does: cell velocities
    gated on the cell depth, no SM clamp (rhs_cuda.py:1001-1007), Sigma = 0.
    Pinned by tests/test_oracle_invariants.py: consistency F(q,q) = physical flux, mirror
    antisymmetry, both-dry zero, rest pressure; and through the residual by the reference's
    golden NumPy residuals and steps (tests/test_oracle_golden.py).
    """
    T = type(g)
    z = T(0)
    aL = np.sqrt(g * np.where(hL > z, hL, z))
    aR = np.sqrt(g * np.where(hR > z, hR, z))
    u_st = (T(0.5) * (uNL + uNR) + aL) - aR
    h_rt = T(0.5) * (aL + aR) + T(0.25) * (uNL - uNR)
    # star celerity sqrt(g * h_rt^2/g) taken as h_rt itself (exact arithmetic identity)
    a_st = np.where(h_rt > z, h_rt, z)
    SL_w = np.minimum(uNL - aL, u_st - a_st)
    SR_w = np.maximum(uNR + aR, u_st + a_st)
    dryL = hL <= h_min
    dryR = hR <= h_min
    SL = np.where(dryL, uNR - T(2.0) * aR, SL_w)
    SR = np.where(dryR, uNL + T(2.0) * aL, SR_w)
    both = dryL & dryR
    SL = np.where(both, z, SL)
    SR = np.where(both, z, SR)
    denom = hR * (uNR - SR) - hL * (uNL - SL)
    big = np.abs(denom) > T(1.0e-14)
    with np.errstate(divide="ignore", invalid="ignore"):
        SM = np.where(big, ((SL * hR) * (uNR - SR) - (SR * hL) * (uNL - SL)) / np.where(big, denom, T(1)),
                      T(0.5) * (uNL + uNR))
    huNL = hL * uNL
    huNR = hR * uNR
    huTL = hL * uTL
    huTR = hR * uTR
    pL = (T(0.5) * g) * hL * hL
    pR = (T(0.5) * g) * hR * hR
    FL_h, FL_N, FL_T = huNL, huNL * uNL + pL, huNL * uTL
    FR_h, FR_N, FR_T = huNR, huNR * uNR + pR, huNR * uTR
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        facL = (SL - uNL) / ((SL - SM) + T(1.0e-30))
        facR = (SR - uNR) / ((SR - SM) + T(1.0e-30))
    hLs = hL * facL
    hRs = hR * facR
    FLs_h = FL_h + SL * (hLs - hL)
    FLs_N = FL_N + SL * (hLs * SM - huNL)
    FLs_T = FL_T + SL * (hLs * uTL - huTL)
    FRs_h = FR_h + SR * (hRs - hR)
    FRs_N = FR_N + SR * (hRs * SM - huNR)
    FRs_T = FR_T + SR * (hRs * uTR - huTR)
    c1, c2, c3 = SL >= z, SM >= z, SR >= z
    Fh = np.where(c1, FL_h, np.where(c2, FLs_h, np.where(c3, FRs_h, FR_h)))
    FN = np.where(c1, FL_N, np.where(c2, FLs_N, np.where(c3, FRs_N, FR_N)))
    FT = np.where(c1, FL_T, np.where(c2, FLs_T, np.where(c3, FRs_T, FR_T)))
    return Fh, FN, FT


def srm_face(bL, bR, gzL, gzR, hL, hR, half_d, g, h_min):
    """SRM face depths and the two one-sided bed sources of one face (rhs_cuda.py:1042-1115).

    Returns (h_Lf, h_Rf, gate, srcL, srcR): ``srcL`` is the lower-index cell's source in
    its own frame (it subtracts ``srcL*inv_d``), ``srcR`` the upper cell's (it adds
    ``srcR*inv_d``), each with the dry gate on the *other* cell's depth.

    Elevations are measured from the lower cell's bed bL (eta_L - bL = hL,
    eta_R - bL = hR + (bR - bL), z - bL): the reference's expressions in exact arithmetic,
    without the cancellation of (z_f - b_c) against float32 ulps of |b|.
    Pinned by lake-at-rest (tests/test_oracle_invariants.py) and by the reference NumPy
    residual golden vectors (tests/test_oracle_golden.py).
    """
    T = type(g)
    z = T(0)
    db = bR - bL
    zL = half_d * gzL
    zR = db - half_d * gzR
    zf0 = np.where(db > z, db, z)
    dzc = zR - zL
    dz = db - dzc
    etaR = hR + db
    detaL = np.maximum(z, np.minimum(dz, etaR - hL))
    detaR = np.maximum(z, np.minimum(-dz, hL - etaR))
    hLf = np.maximum(z, (hL + detaL) - zf0)
    hRf = np.maximum(z, (etaR + detaR) - zf0)
    gate = ~((hL < h_min) & (hR < h_min))
    # owner L: gate on the neighbour R being dry
    headL = zf0 - hL
    dzL = np.where(hR < h_min, np.maximum(z, headL), np.maximum(z, np.minimum(dzc, headL)))
    srcL = ((T(0.5) * g) * (hLf + hL)) * (zf0 - dzL)
    # owner R (its left face): dz_clip in R's frame is zL - zR
    headR = zf0 - etaR
    dzR = np.where(hL < h_min, np.maximum(z, headR), np.maximum(z, np.minimum(zL - zR, headR)))
    srcR = ((T(0.5) * g) * (hRf + hR)) * ((zf0 - dzR) - db)
    return hLf, hRf, gate, srcL, srcR


def primitives(q0, q1, q2, h_min):
    """Depth, velocities with the dry gate, as the kernel's PRIMS macro (rhs_cuda.py:931-938)."""
    T = type(h_min)
    wet = q0 > h_min
    hs = np.where(wet, q0, h_min)
    with np.errstate(divide="ignore", invalid="ignore"):
        u = np.where(wet, q1 / hs, T(0))
        v = np.where(wet, q2 / hs, T(0))
    return u, v


# ---------------------------------------------------------------------------
# residual on a padded grid (the kernel window [2, n-3]^2)
# ---------------------------------------------------------------------------

def residual(q: np.ndarray, b: np.ndarray, dx: float, dy: float, g: float = G,
             h_min: float = 1e-6, active: Optional[np.ndarray] = None) -> np.ndarray:
    """SRM-HLLC residual of a padded, ghost-filled state (rhs_cuda.py:900-1208).

    ``q`` (3, nxp, nyp) and ``b`` (nxp, nyp) share a dtype (float32 or float64). The
    residual is evaluated on the kernel window [2, nxp-3] x [2, nyp-3] intersected with
    ``active`` (padded uint8/bool, None = all); every other cell gets 0, like the
    compact launch plus the ``inside_mask == 0`` branch. Accumulation order per cell is
    x right face (h, hu, hv, hu source), x left, y top, y bottom (Appendix A.2 order).
    """
    dt_ = q.dtype
    c = kernel_consts(dx, dy, g, h_min, dt_)
    T, ix, iy, gk, hm = c["T"], c["inv_dx"], c["inv_dy"], c["g"], c["h_min"]
    z = T(0)
    h, hu, hv = q[0], q[1], q[2]
    nxp, nyp = h.shape
    u, v = primitives(h, hu, hv, hm)
    gzx = np.zeros_like(b)
    gzy = np.zeros_like(b)
    gzx[1:-1, :] = (T(0.5) * (b[2:, :] - b[:-2, :])) * ix
    gzy[:, 1:-1] = (T(0.5) * (b[:, 2:] - b[:, :-2])) * iy
    hdx = T(0.5) * c["dxk"]
    hdy = T(0.5) * c["dyk"]
    # x faces between rows i and i+1 (shape nxp-1, nyp)
    xhL, xhR, xg, xsL, xsR = srm_face(b[:-1], b[1:], gzx[:-1], gzx[1:], h[:-1], h[1:], hdx, gk, hm)
    xFh, xFN, xFT = hllc(xhL, u[:-1], v[:-1], xhR, u[1:], v[1:], gk, hm)
    # y faces between columns j and j+1 (shape nxp, nyp-1): normal v, transverse u
    yhL, yhR, yg, ysL, ysR = srm_face(b[:, :-1], b[:, 1:], gzy[:, :-1], gzy[:, 1:], h[:, :-1], h[:, 1:],
                                      hdy, gk, hm)
    yFh, yFN, yFT = hllc(yhL, v[:, :-1], u[:, :-1], yhR, v[:, 1:], u[:, 1:], gk, hm)

    W = (slice(2, nxp - 2), slice(2, nyp - 2))          # the kernel window
    # face views per window cell
    R = (slice(2, nxp - 2), slice(2, nyp - 2))          # x face i+1/2 has index i
    L = (slice(1, nxp - 3), slice(2, nyp - 2))          # x face i-1/2 has index i-1
    Tt = (slice(2, nxp - 2), slice(2, nyp - 2))         # y face j+1/2 has index j
    B = (slice(2, nxp - 2), slice(1, nyp - 3))          # y face j-1/2 has index j-1
    shp = (nxp - 4, nyp - 4)
    rh = np.zeros(shp, dt_)
    rhu = np.zeros(shp, dt_)
    rhv = np.zeros(shp, dt_)
    gR, gL, gT, gB = xg[R], xg[L], yg[Tt], yg[B]
    with np.errstate(over="ignore", invalid="ignore"):
        rh = np.where(gR, rh - xFh[R] * ix, rh)
        rhu = np.where(gR, rhu - xFN[R] * ix, rhu)
        rhv = np.where(gR, rhv - xFT[R] * ix, rhv)
        rhu = np.where(gR, rhu - xsL[R] * ix, rhu)
        rh = np.where(gL, rh + xFh[L] * ix, rh)
        rhu = np.where(gL, rhu + xFN[L] * ix, rhu)
        rhv = np.where(gL, rhv + xFT[L] * ix, rhv)
        rhu = np.where(gL, rhu + xsR[L] * ix, rhu)
        rh = np.where(gT, rh - yFh[Tt] * iy, rh)
        rhu = np.where(gT, rhu - yFT[Tt] * iy, rhu)
        rhv = np.where(gT, rhv - yFN[Tt] * iy, rhv)
        rhv = np.where(gT, rhv - ysL[Tt] * iy, rhv)
        rh = np.where(gB, rh + yFh[B] * iy, rh)
        rhu = np.where(gB, rhu + yFT[B] * iy, rhu)
        rhv = np.where(gB, rhv + yFN[B] * iy, rhv)
        rhv = np.where(gB, rhv + ysR[B] * iy, rhv)
    out = np.zeros_like(q)
    out[0][W], out[1][W], out[2][W] = rh, rhu, rhv
    if active is not None:
        out *= (np.asarray(active) != 0)
        out[out == 0] = 0   # keep +0 (masked cells are written as +0 in the kernel)
    return out


# ---------------------------------------------------------------------------
# fused update tail (rhs_cuda.py:1757-1796) and the CFL wave speed
# ---------------------------------------------------------------------------

@dataclass
class Physics:
    """Scalars of the step (Config production defaults, solver.py:629-860)."""
    g: float = G
    h_min: float = 1e-6
    h_cfl: float = 1e-6          # CFL floor (h_min_cfl, floored at 1e-6 for float32)
    cfl: float = 0.5
    friction: bool = False       # Config.friction == 'manning_implicit'
    vcap: float = 15.0           # friction_velocity_cap_ms
    quadratic: bool = True       # friction_quadratic_alpha
    keep_h: bool = True          # SWE_WETDRY_ZERO_H unset (solver.py:428)
    cfl_l2: bool = False         # velocity norm of the CFL: L-inf (dense) or L2 (flat default)
    flat_cfl: bool = False       # the flat layout's wave-speed rule and floor (see lam_max, dt_from_lam)


def update(q_int, r_int, dt: float, ph: Physics, n_cell=None, rain=None, active=None):
    """One cell update from the residual, interior cells only (rhs_cuda.py:1757-1796).

    ``q_int``/``r_int`` are (3, ...) arrays of the same dtype; ``n_cell`` Manning n per
    cell (array or scalar; used when ``ph.friction``); ``rain`` m/s per cell or scalar
    (already in the state dtype), added to ``rhs_h`` in the state precision; ``active``
    marks cells whose residual and rain are kept (inactive cells update with a zero
    residual, like the reference's dense carry kernel).
    Pinned by the reference's own step (tests/golden/step_vcap_f64.npz): from its residual,
    a class table with a fifth of the cells above the 15 m/s velocity cap, the quadratic and
    the linear root, equal to 1e-14 of each field's scale (tests/test_oracle_golden.py); by
    the bowl step golden; by the friction root and dry-cell closed forms
    (tests/test_oracle_invariants.py).
    """
    T = q_int.dtype.type
    h_min = T(ph.h_min)
    g = T(ph.g)
    dtT = T(dt)
    z = T(0)
    rr0, r1, r2 = r_int[0], r_int[1], r_int[2]
    if rain is not None:
        rr0 = rr0 + (np.asarray(rain, dtype=q_int.dtype) if np.ndim(rain) else T(rain))
    if active is not None:
        act = np.asarray(active) != 0
        rr0 = np.where(act, rr0, z)
        r1 = np.where(act, r1, z)
        r2 = np.where(act, r2, z)
    with np.errstate(over="ignore", invalid="ignore", divide="ignore"):
        h = q_int[0] + dtT * rr0
        hu = q_int[1] + dtT * r1
        hv = q_int[2] + dtT * r2
        dry = h < h_min
        if ph.friction:
            hs = np.where(h > h_min, h, h_min)
            u = hu / hs
            v = hv / hs
            modU = np.sqrt(u * u + v * v)
            n = (np.asarray(n_cell, dtype=q_int.dtype) if np.ndim(n_cell) else
                 np.full(h.shape, T(n_cell)))
            h43 = pow_m43(np.where(dry, T(1), hs))
            if math.isfinite(ph.vcap):
                n_cri = np.sqrt(T(1) / ((((T(1.0e-10) + dtT) * g) * h43) * (modU + T(1.0e-30))))
                n = np.where((modU > T(ph.vcap)) & (n_cri > n), n_cri, n)
            Cf = ((g * n) * n) * h43
            if ph.quadratic:
                two = ((T(2.0) * dtT) * Cf) * modU
                alpha = T(2.0) / (np.sqrt(T(1.0) + T(2.0) * two) + T(1.0))
            else:
                alpha = T(1.0) / (T(1.0) + ((dtT * Cf) * modU))
            hu = np.where(dry, hu, hu * alpha)
            hv = np.where(dry, hv, hv * alpha)
        if ph.keep_h:
            h = np.where(dry, np.where(h > z, h, z), h)
        else:
            h = np.where(dry, z, h)
        hu = np.where(dry, z, hu)
        hv = np.where(dry, z, hv)
    return np.stack([h, hu, hv])


def lam_max(q_int, ph: Physics, skip=None) -> float:
    """Max over cells of the CFL wave speed, as a Python float, NaN kept.

    Dense rule (lammax_fp32_lean, solver.py:152-205): cells with ``h < h_cfl`` are skipped
    and ``u = hu/h``. Flat rule (``ph.flat_cfl``; cfl_lammax_flat_blk and the fused step's
    lambda, compressed_solver.py:200-240, compressed_rhs.py:1041-1053): cells with
    ``h < h_min`` are skipped and ``u = hu/max(h, h_cfl)`` (``h_cfl`` is the flat solver's
    ``h_min_cfl``). Speed = velocity norm (L-inf ``max(|u|,|v|)`` or L2 ``sqrt(u*u+v*v)``)
    + ``sqrt(g*h)``, evaluated in the state dtype and maximised as float64 IEEE bits, the
    order-independent reduction the device performs; a NaN depth passes the skip test and
    yields a NaN lambda, which dominates the bit maximum.
    Pinned by closed forms on hand-built states (tests/test_oracle_invariants.py), by the
    reference's C1 dt sequence (golden) for the dense L-inf rule, and by the reference's
    rain-on-dry-bed run (golden) for the floor.
    """
    T = q_int.dtype.type
    h, hu, hv = q_int[0], q_int[1], q_int[2]
    with np.errstate(over="ignore", invalid="ignore", divide="ignore"):
        if ph.flat_cfl:
            take = ~(h < T(ph.h_min))
            hs = np.where(h > T(ph.h_cfl), h, T(ph.h_cfl))
        else:
            take = ~(h < T(ph.h_cfl))
            hs = h
        if skip is not None:
            take &= ~(np.asarray(skip) != 0)
        u = hu / hs
        v = hv / hs
        if ph.cfl_l2:
            vel = np.sqrt(u * u + v * v)
        else:
            au, av = np.abs(u), np.abs(v)
            vel = np.where(au > av, au, av)
        lam = vel + np.sqrt(T(ph.g) * h)
    lam = lam[take].astype(np.float64)
    if lam.size == 0:
        return 0.0
    if np.isnan(lam).any():
        return math.nan
    return float(lam.view(np.uint64).max().view(np.float64))


def dt_from_lam(lam_max: float, ph: Physics, dmin: float) -> float:
    """dt = cfl*min(dx,dy)/(max(lam, floor) + 1e-12) in float64, with the floor
    sqrt(g*h_cfl) of the dense solver (solver.py:2657-2666) or sqrt(g*h_min) of the flat one
    (compressed_solver.py:713-719, 741-745)."""
    if not math.isfinite(lam_max):
        raise FloatingPointError(f"non-finite max wave speed {lam_max}")
    floor = (ph.g * (ph.h_min if ph.flat_cfl else ph.h_cfl)) ** 0.5
    if lam_max < floor:
        lam_max = floor
    return ph.cfl * dmin / (lam_max + 1.0e-12)


def rain_step_cap(rmax: float, ph: Physics, dmin: float) -> float:
    """The rain-film step bound (cfl*dmin)^(2/3)/(g*R)^(1/3) (solver.py:3005-3027); inf if dry.
    Pinned by its fixed point dt = cfl*dmin/sqrt(g*R*dt) (tests/test_oracle_invariants.py) and
    by the reference's rain-on-dry-bed run, whose first step it sets
    (tests/golden/rain_dry_run_f64.npz)."""
    if rmax <= 0.0:
        return math.inf
    return (ph.cfl * dmin) ** (2.0 / 3.0) / (ph.g * rmax) ** (1.0 / 3.0)


# ---------------------------------------------------------------------------
# forcings (forcing.py:170-216, 315-377)
# ---------------------------------------------------------------------------

def rain_index(time_s, t: float) -> int:
    """Piecewise-constant segment index: bisect_right(time_s, t) - 1, clamped (forcing.py:198-199)."""
    import bisect
    tl = list(np.asarray(time_s, dtype=np.float64).tolist())
    idx = max(0, bisect.bisect_right(tl, float(t)) - 1)
    return min(idx, len(tl) - 1)


def stage_at_time(time_s, stage_m, t: float) -> float:
    """Linear interpolation of the stage, clamped to the end values (forcing.py:315-347)."""
    import bisect
    tl = list(np.asarray(time_s, dtype=np.float64).tolist())
    sl = list(np.asarray(stage_m, dtype=np.float64).tolist())
    t = float(t)
    if t <= tl[0]:
        return sl[0]
    if t >= tl[-1]:
        return sl[-1]
    i = bisect.bisect_left(tl, t)
    t0, t1 = tl[i - 1], tl[i]
    s0, s1 = sl[i - 1], sl[i]
    w = (t - t0) / (t1 - t0)
    return s0 + w * (s1 - s0)


# ---------------------------------------------------------------------------
# a dense run (Solver2D.run/step, solver.py:2695-3003)
# ---------------------------------------------------------------------------

# exp(r) on |r| <= ln2/2 by Horner's rule on the Taylor coefficients 1/n!, n = 13 .. 0
# (the truncation term r^14/14! is below 4e-18 relative)
_EXP_C = (1.6059043836821613e-10, 2.08767569878681e-09, 2.505210838544172e-08, 2.755731922398589e-07,
          2.7557319223985893e-06, 2.48015873015873e-05, 0.0001984126984126984, 0.001388888888888889,
          0.008333333333333333, 0.041666666666666664, 0.16666666666666666, 0.5, 1.0, 1.0)
_LN2_HI, _LN2_LO, _INV_LN2 = 6.93147180369123816490e-01, 1.90821492927058770002e-10, 1.44269504088896338700e+00


def exp_ieee(x) -> np.ndarray:
    """``exp(x)`` in float64 from IEEE +, -, *, round-to-nearest-integer and ``ldexp`` only, so
    CPU and GPU agree bit for bit: Cody-Waite reduction ``x = k ln2 + r`` with ln2 split in a
    32-bit head and a tail (k ln2_hi is exact), ``exp(r)`` by a degree-13 Taylor polynomial in
    Horner form, scaled by ``2^k``. 0 below -745.2, +inf above 709.8, NaN kept. Stands in for
    the reference's fast ``__expf`` in the drain-tau decay (compressed_solver.py:355-356,
    417), which differs by a few ulp. Pinned against ``numpy.exp`` (<= 2 ulp).
    """
    x = np.asarray(x, np.float64)
    out = np.empty_like(x)
    big, small, nan = x > 709.8, x < -745.2, x != x
    mid = ~(big | small | nan)
    xm = x[mid]
    kf = np.rint(xm * _INV_LN2)
    r = (xm - kf * _LN2_HI) - kf * _LN2_LO
    p = np.full_like(r, _EXP_C[0])
    for c in _EXP_C[1:]:
        p = p * r + c
    out[mid] = np.ldexp(p, kf.astype(np.int64))
    out[big] = np.inf
    out[small] = 0.0
    out[nan] = x[nan]
    return out


_EXPF_C = tuple(np.float32(c) for c in (0.00019841270113829523, 0.0013888889225199819, 0.008333333767950535,
                                          0.0416666679084301, 0.1666666716337204, 0.5, 1.0, 1.0))
_LN2_HI_F, _LN2_LO_F, _INV_LN2_F = np.float32(0.693145751953125), np.float32(1.428606765330187e-06), \
    np.float32(1.4426950216293335)


def exp_ieee_f32(x) -> np.ndarray:
    """``exp(x)`` in float32 from IEEE float32 +, -, *, round-to-nearest-integer and ``ldexp``
    only (the float32 form of :func:`exp_ieee`): Cody-Waite reduction with ln2 split in a
    15-bit head (k ln2_hi exact for |k| < 512) and a tail, a degree-7 Taylor polynomial in
    Horner form (truncation below 5e-9 relative), scaled by 2^k. 0 below -103.98, +inf above
    88.73, NaN kept. Pinned against ``numpy.exp`` (<= 2 ulp)."""
    x = np.asarray(x, np.float32)
    out = np.empty_like(x)
    big, small, nan = x > np.float32(88.73), x < np.float32(-103.98), x != x
    mid = ~(big | small | nan)
    xm = x[mid]
    kf = np.rint(xm * _INV_LN2_F)
    r = (xm - kf * _LN2_HI_F) - kf * _LN2_LO_F
    p = np.full_like(r, _EXPF_C[0])
    for c in _EXPF_C[1:]:
        p = p * r + c
    out[mid] = np.ldexp(p, kf.astype(np.int32))
    out[big] = np.inf
    out[small] = 0.0
    out[nan] = x[nan]
    return out


def green_ampt_drain(q_int: np.ndarray, F: np.ndarray, ga: dict, dt: float, active=None) -> None:
    """Fused Green-Ampt infiltration and linear-reservoir drain (ga_drain_flat,
    compressed_solver.py:320-367), in place on a (3, nx, ny) state and the cumulative
    infiltration ``F``. ``ga``: ``cls`` (nx, ny) soil class, tables ``Ks``, ``psi``, ``dth``
    (None: no Green-Ampt), ``Fmax`` (nx, ny) or None (no cap), ``inv_tau`` (nx, ny) or None (no
    drain). On active cells with ``h > 0``: with ``K = Ks[c] > 0``, ``a = F0 + K dt``,
    ``disc = a*a + ((4 K dt) (psi[c] + h)) dth[c]``, ``F1 = (a + sqrt(max(disc, 0)))/2``,
    ``dF = min(max(F1 - F0, 0), h)`` capped by the room ``Fmax - F0`` (0 when none left),
    ``h <- h - dF``, ``F <- F0 + dF``, ``scale = h'/h``; then with ``it = inv_tau > 0`` and
    ``h > 0``: ``decay = exp(-dt*it)`` by :func:`exp_ieee_f32` in float32 or :func:`exp_ieee` in
    float64, ``scale *= decay``,
    ``h *= decay``; momentum times ``scale``. The ``ga`` and ``drain`` modes of the reference
    (ga_flat, drain_tau_flat, 369-421) are this with one part left out, bit for bit."""
    dtp = q_int.dtype
    T = dtp.type
    h, hu, hv = q_int[0], q_int[1], q_int[2]
    on = ~(h <= 0)
    if active is not None:
        on &= np.asarray(active, bool)
    dtv = T(dt)
    hh = h[on]
    scale = np.ones_like(hh)
    if ga.get("Ks") is not None:
        c = np.asarray(ga["cls"])[on]
        K = np.asarray(ga["Ks"]).astype(dtp)[c]
        psi = np.asarray(ga["psi"]).astype(dtp)[c]
        dth = np.asarray(ga["dth"]).astype(dtp)[c]
        F0 = F[on]
        g = K > 0
        KsDt = K * dtv
        head = psi + hh
        a = F0 + KsDt
        disc = a * a + ((T(4.0) * KsDt) * head) * dth
        F1 = T(0.5) * (a + np.sqrt(np.fmax(disc, T(0))))
        dF = F1 - F0
        dF = np.where(dF > 0, dF, T(0))
        dF = np.where(dF > hh, hh, dF)
        Fmax = ga.get("Fmax")
        if Fmax is not None:
            room = np.asarray(Fmax).astype(dtp)[on] - F0
            dF = np.where(dF > room, np.where(room > 0, room, T(0)), dF)
        h_new = hh - dF
        with np.errstate(invalid="ignore", divide="ignore"):
            scale = np.where(g, scale * (h_new / hh), scale)
        hh = np.where(g, h_new, hh)
        Fn = F0.copy()
        Fn[g] = F0[g] + dF[g]
        F[on] = Fn
    if ga.get("inv_tau") is not None:
        it = np.asarray(ga["inv_tau"]).astype(dtp)[on]
        d = (it > 0) & (hh > 0)
        arg = (-dtv) * it[d]
        decay = exp_ieee_f32(arg) if dtp == np.float32 else exp_ieee(arg.astype(np.float64))
        scale[d] = scale[d] * decay
        hh[d] = hh[d] * decay
    h[on] = hh
    hu[on] = hu[on] * scale
    hv[on] = hv[on] * scale


def ring_mode_apply(q_int: np.ndarray, b_int: np.ndarray, rm: dict, h_min: float) -> None:
    """The ring modes of the cached replay at the top of a step, in place on a (3, nx, ny)
    state (compressed_solver.py:2134-2144): kind 0 copies the neighbour's h, hu, hv; kind 1
    sets still water ``max(eta - b, 0)`` (state dtype) with zero momentum; kind 2 does the same
    and then takes the neighbour's velocity, ``hu = h*(hu_nb*r)`` with ``r = 1/h_nb`` where
    ``h_nb > h_min``, else 0 (ghost_uv, 450-465)."""
    dt = q_int.dtype
    T = dt.type
    c = np.asarray(rm["cells"])
    nbc = np.asarray(rm["nb"])
    kind = np.asarray(rm["kind"])
    k0 = kind == 0
    q_int[:, c[k0, 0], c[k0, 1]] = q_int[:, nbc[k0, 0], nbc[k0, 1]]
    ks = kind > 0
    d = T(rm["eta"]) - np.asarray(b_int).astype(dt)[c[ks, 0], c[ks, 1]]
    h = np.where(d > 0, d, T(0))
    q_int[0, c[ks, 0], c[ks, 1]] = h
    q_int[1, c[ks, 0], c[ks, 1]] = 0
    q_int[2, c[ks, 0], c[ks, 1]] = 0
    k2 = kind == 2
    if k2.any():
        hn = q_int[0, nbc[k2, 0], nbc[k2, 1]]
        with np.errstate(divide="ignore"):
            inv = np.where(hn > T(h_min), T(1) / hn, T(0))
        hg = q_int[0, c[k2, 0], c[k2, 1]]
        q_int[1, c[k2, 0], c[k2, 1]] = hg * (q_int[1, nbc[k2, 0], nbc[k2, 1]] * inv)
        q_int[2, c[k2, 0], c[k2, 1]] = hg * (q_int[2, nbc[k2, 0], nbc[k2, 1]] * inv)


def ring_depth(ring: dict, t: float, dtype) -> np.ndarray:
    """Still-water depth of each gauge-ring cell at time ``t`` (the step's start time).

    Stage interpolation (compressed_solver.py:2338-2348): ``ti = searchsorted(t_common, t) -
    1`` clamped to ``[0, nt - 2]``; ``wt = (t - t_common[ti]) / (t_common[ti+1] - t_common[ti]
    + 1e-12)`` in float64, clamped to ``[0, 1]``; per gauge ``s = S[ti]*(1 - wt) + S[ti+1]*wt``
    in the state dtype. IDW depth (ring_bc_flat, 287-307): ``eta = sum_g w_g s_g`` from 0 in
    gauge order, ``h = max(0, eta - bed)`` (``fmax``: a NaN gives 0).
    """
    dt = np.dtype(dtype)
    tc = np.asarray(ring["t_common"], np.float64)
    ti = int(np.searchsorted(tc, t) - 1)
    ti = max(0, min(len(tc) - 2, ti))
    wt = float((t - tc[ti]) / (tc[ti + 1] - tc[ti] + 1e-12))
    wt = min(1.0, max(0.0, wt))
    S = np.asarray(ring["stage_all"]).astype(dt)
    s = S[:, ti] * dt.type(1.0 - wt) + S[:, ti + 1] * dt.type(wt)
    w = np.asarray(ring["w"]).astype(dt).reshape(len(ring["cells"]), S.shape[0])
    eta = np.zeros(w.shape[0], dt)
    for g in range(w.shape[1]):
        eta = eta + w[:, g] * s[g]
    return np.fmax(dt.type(0), eta - np.asarray(ring["bed"]).astype(dt))


def coastal_force(q_int: np.ndarray, sponge=None, clamp=None, ring=None, t: float = 0.0, drain=None,
                  infil=None, ga=None) -> None:
    """Infiltration, open-boundary sponge, gauge ring, stage clamp, karst drain, in place on a
    (3, nx, ny) state.

    Infiltration (infiltrate_flat, compressed_solver.py:143-158), ``infil = (cls, rate, dt,
    h_min, active)``: on active cells with ``h >= h_min``, ``h' = max(h - rate[cls]*dt, 0)``
    and ``hu, hv`` scaled by ``h'/h`` (``dt`` and ``h_min`` in the state dtype). Karst drain
    (drain_cap, 269-286), ``drain`` an ``h_tgt`` field (+inf where none), last: where
    ``h > h_tgt``, ``r = h_tgt/h`` (0 when ``h <= 1e-12``), ``h = h_tgt``, momentum scaled by r.

    Sponge (compressed_solver.py:1320-1326): ``h <- h*keep + amb`` as a multiply then an
    add (two roundings), ``hu, hv <- hu*keep, hv*keep``, on the band (``keep != 1`` or
    ``amb != 0``). Clamp (compressed_solver.py:424-440): where ``h > h_max``,
    ``r = h_max/h``, ``h <- h_max``, ``hu, hv <- hu*r, hv*r``; clamp cells are those with an
    ``h_max`` that is neither +inf nor NaN. A cell named by either list gets both (the
    identity sponge ``h*1 + 0`` included), as the fused step applies them. ``ring``
    (``{"cells": (n, 2) ij, "bed", "w": (n, NG), "t_common", "stage_all": (NG, nt)}``) sets
    its cells to still water at :func:`ring_depth` of time ``t`` between the two; ``ga`` =
    ``(spec, F, dt, active)`` runs :func:`green_ampt_drain` after the ring. The cells
    named by the sponge band, the clamp or the drain are the forced cells: each gets the
    sponge (if any), the clamp (if any) and the drain (if any).
    """
    dt = q_int.dtype
    h, hu, hv = q_int[0], q_int[1], q_int[2]
    if infil is not None:
        cls, rate, idt, hmin, act = infil
        rate = np.asarray(rate).astype(dt)[np.asarray(cls)]
        on = ~(h < dt.type(hmin))
        if act is not None:
            on &= np.asarray(act, bool)
        hn = h[on] - rate[on] * dt.type(idt)
        hn = np.where(hn < 0, dt.type(0), hn)
        r = hn / h[on]
        h[on] = hn
        hu[on] = hu[on] * r
        hv[on] = hv[on] * r
    cells = np.zeros(q_int.shape[1:], bool)
    if sponge is not None:
        kp = np.asarray(sponge[0]).astype(dt)
        am = np.asarray(sponge[1]).astype(dt)
        cells |= (kp != 1) | (am != 0)
    if clamp is not None:
        hx = np.asarray(clamp).astype(dt)
        cells |= ~(np.isnan(hx) | (hx == np.inf))
    if drain is not None:
        ht = np.asarray(drain).astype(dt)
        cells |= ~(np.isnan(ht) | (ht == np.inf))
    if sponge is not None:
        h[cells] = h[cells] * kp[cells] + am[cells]
        hu[cells] = hu[cells] * kp[cells]
        hv[cells] = hv[cells] * kp[cells]
    if ring is not None:
        ii, jj = np.asarray(ring["cells"])[:, 0], np.asarray(ring["cells"])[:, 1]
        h[ii, jj] = ring_depth(ring, t, dt)
        hu[ii, jj] = 0
        hv[ii, jj] = 0
    if ga is not None:
        green_ampt_drain(q_int, ga[1], ga[0], ga[2], ga[3])
    if clamp is not None:
        over = cells & (h > hx)
        r = hx[over] / h[over]
        h[over] = hx[over]
        hu[over] = hu[over] * r
        hv[over] = hv[over] * r
    if drain is not None:
        over = cells & (h > ht)
        r = np.where(h[over] > dt.type(1e-12), ht[over] / h[over], dt.type(0))
        h[over] = ht[over]
        hu[over] = hu[over] * r
        hv[over] = hv[over] * r


@dataclass
class DenseCase:
    """Inputs of a dense run on an unpadded ``(nx, ny)`` grid."""
    h: np.ndarray
    hu: np.ndarray
    hv: np.ndarray
    bed: np.ndarray
    dx: float
    dy: float
    bc_x: str = "wall"
    bc_y: str = "wall"
    ngh: int = 4
    manning: object = 0.0              # scalar or (nx, ny) n
    rain: object = None                # None, scalar m/s, or (time_s, rate_mm_h[, frames]) tuple
    active: Optional[np.ndarray] = None   # (nx, ny) bool inside mask
    stage: Optional[dict] = None       # {"cells": (M,2) unpadded ij, "time_s", "stage_m"}
    sponge: Optional[tuple] = None     # (keep, amb), (nx, ny) each
    clamp: Optional[np.ndarray] = None  # (nx, ny) h_max, +inf where none
    ring: Optional[dict] = None         # gauge ring: see coastal_force
    infil: Optional[tuple] = None       # (class (nx, ny) ints, rate per class m/s)
    drain: Optional[np.ndarray] = None  # (nx, ny) karst target depth, +inf where none
    ga: Optional[dict] = None           # Green-Ampt / drain-tau: see green_ampt_drain (+ "F" initial)
    ring_mode: Optional[dict] = None    # {"cells", "nb" (n, 2) ij (-1: none), "kind", "eta"}: see ring_mode_apply


class DenseOracle:
    """Dense time stepping in the reference order: BC fill -> residual -> update -> t += dt
    -> stage -> infiltration -> sponge -> gauge ring (at the step's start time) -> Green-Ampt
    and drain-tau -> clamp -> karst drain -> (next) CFL. Works in float32 or float64."""

    def __init__(self, case: DenseCase, ph: Physics, dtype=np.float32, store_dtype=None):
        """``store_dtype`` narrower than ``dtype`` (float32 with float64): the state is kept in
        ``store_dtype``, i.e. rounded to it after every step and read back exactly, and the
        arithmetic is ``dtype`` (the library's float32-planes, float64-arithmetic mode)."""
        self.case, self.ph = case, ph
        self.dtype = np.dtype(dtype)
        self.store = None if store_dtype is None else np.dtype(store_dtype)
        nx, ny = case.h.shape
        ngh = case.ngh
        self.nx, self.ny, self.ngh = nx, ny, ngh
        self.q = np.zeros((3, nx + 2 * ngh, ny + 2 * ngh), dtype=self.dtype)
        for k, a in enumerate((case.h, case.hu, case.hv)):
            self.q[k, ngh:ngh + nx, ngh:ngh + ny] = np.asarray(a).astype(self.dtype)
        self.b = pad_bed(np.asarray(case.bed), ngh, case.bc_x, case.bc_y, self.dtype)
        self.act_pad = None
        if case.active is not None:
            self.act_pad = np.zeros(self.q.shape[1:], np.uint8)
            self.act_pad[ngh:ngh + nx, ngh:ngh + ny] = np.asarray(case.active).astype(np.uint8)
        self.t = 0.0
        self.nsteps = 0
        self.dmin = min(case.dx, case.dy)
        self.ga_F = None
        if case.ga is not None:
            F0 = case.ga.get("F")
            self.ga_F = (np.zeros((nx, ny), self.dtype) if F0 is None else np.asarray(F0).astype(self.dtype))
        self.dts: list = []

    @property
    def interior(self):
        g = self.ngh
        return self.q[:, g:g + self.nx, g:g + self.ny]

    def rain_now(self):
        """(rate array/scalar in the state dtype, rmax float64) at self.t, or (None, 0)."""
        r = self.case.rain
        if r is None:
            return None, 0.0
        if np.isscalar(r):
            return (self.dtype.type(r), float(r)) if r > 0.0 else (None, 0.0)
        time_s, rate = r[0], np.asarray(r[1])
        idx = rain_index(time_s, self.t)
        if rate.ndim == 1:
            m = float(rate[idx]) / 3.6e6
            return self.dtype.type(m), m
        frame = (rate[idx] / 3.6e6).astype(self.dtype)
        return frame, float(frame.max())

    cfl_skip = None       # (nx, ny) bool: cells left out of the CFL (set_cfl_ghost_mask)

    def cfl_dt(self) -> float:
        return dt_from_lam(lam_max(self.interior, self.ph, skip=self.cfl_skip), self.ph, self.dmin)

    def step(self, dt: float) -> None:
        c, ph = self.case, self.ph
        if c.ring_mode is not None:
            g = self.ngh
            ring_mode_apply(self.q[:, g:g + self.nx, g:g + self.ny], np.asarray(c.bed), c.ring_mode, ph.h_min)
        bc_fill(self.q, self.ngh, c.bc_x, c.bc_y)
        r = residual(self.q, self.b, c.dx, c.dy, ph.g, ph.h_min, active=self.act_pad)
        g = self.ngh
        I = (slice(None), slice(g, g + self.nx), slice(g, g + self.ny))
        rain, _ = self.rain_now()
        n = c.manning
        if np.ndim(n):
            n = np.asarray(n, dtype=np.float32)
        act = None if c.active is None else np.asarray(c.active)
        self.q[I] = update(self.q[I], r[I], dt, ph, n_cell=n, rain=rain, active=act)
        t0 = self.t
        self.t += dt
        if c.stage is not None:
            eta = self.dtype.type(stage_at_time(c.stage["time_s"], c.stage["stage_m"], self.t))
            ii = np.asarray(c.stage["cells"])[:, 0]
            jj = np.asarray(c.stage["cells"])[:, 1]
            bed = self.b[ii + g, jj + g]
            self.q[0, ii + g, jj + g] = np.maximum(eta - bed, self.dtype.type(0))
            self.q[1, ii + g, jj + g] = 0
            self.q[2, ii + g, jj + g] = 0
        if any(f is not None for f in (c.sponge, c.clamp, c.ring, c.drain, c.infil, c.ga)):
            inf = None
            if c.infil is not None:
                inf = (c.infil[0], c.infil[1], dt, ph.h_min, act)
            ga = None if c.ga is None else (c.ga, self.ga_F, dt, act)
            coastal_force(self.q[I], c.sponge, c.clamp, c.ring, t0, c.drain, inf, ga)
        if self.store is not None:
            self.q[I] = self.q[I].astype(self.store).astype(self.dtype)
        self.nsteps += 1

    def rain_max_all(self) -> float:
        """The largest rain rate of the whole table as the step reads it (state dtype, m/s)."""
        r = self.case.rain
        if r is None:
            return 0.0
        if np.isscalar(r):
            return float(self.dtype.type(r))
        rate = np.asarray(r[1])
        return float((rate / 3.6e6).astype(self.dtype).max())

    def next_dt(self, t_end=None, dt_max=None) -> float:
        """The step a run takes next: CFL, rain-film bound, dt_max, clip to t_end. The dense
        solver bounds each step by the current segment's largest rate (solver.py:3005-3027);
        the flat one (``ph.flat_cfl``) by the largest rate of the whole table, rounded to the
        state dtype (compressed_solver.py:3429-3436)."""
        dt = self.cfl_dt()
        rmax = self.rain_max_all() if self.ph.flat_cfl else self.rain_now()[1]
        dt = min(dt, rain_step_cap(rmax, self.ph, self.dmin))
        if dt_max is not None and dt > dt_max:
            dt = dt_max
        if t_end is not None and self.t + dt > t_end:
            dt = t_end - self.t
        return dt

    def run_steps(self, nsteps: int, dt_max=None) -> None:
        for _ in range(nsteps):
            dt = self.next_dt(dt_max=dt_max)
            self.dts.append(dt)
            self.step(dt)

    def run(self, t_end: float, dt_max=None, max_steps=10 ** 7) -> int:
        steps = 0
        while self.t < t_end - 1.0e-12 and steps < max_steps:
            dt = self.next_dt(t_end=t_end, dt_max=dt_max)
            self.dts.append(dt)
            self.step(dt)
            steps += 1
        return steps
