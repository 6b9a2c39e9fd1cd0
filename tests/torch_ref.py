"""Float64 torch restatement of the flat fused step, for autograd gradients (test only).

The step of the oracle (oracle/swe_oracle.py: hllc, srm_face, primitives, residual,
update) written with torch operations on the padded box: the active cells are updated,
every other stored cell (the frozen ring and ghosts) keeps its value, stage cells take
h = max(eta - b, 0) with zero momentum. Branches are torch.where over both sides, with
square roots and divisions guarded so that an unselected side can neither produce nor
propagate a NaN under autograd (a guarded sqrt has zero derivative at 0). The gradient
of a loss through S such steps is the independent check of the hand-written device
adjoint (geoswe_amd/csrc/gs_adjoint.hip).
"""
from __future__ import annotations

import torch

Tn = torch.Tensor


def _sqrt0(x: Tn) -> Tn:
    pos = x > 0
    return torch.where(pos, torch.sqrt(torch.where(pos, x, torch.ones_like(x))), torch.zeros_like(x))


def _min(a: Tn, b: Tn) -> Tn:       # a <= b selects a (the device reverse's choice at a tie)
    return torch.where(a <= b, a, b)


def _max(a: Tn, b: Tn) -> Tn:       # a >= b selects a
    return torch.where(a >= b, a, b)


def _pos(x: Tn) -> Tn:              # max(0, x) with no derivative at 0
    return torch.where(x > 0, x, torch.zeros_like(x))


def _div(a: Tn, b: Tn, m: Tn) -> Tn:
    bb = torch.where(m, b, torch.ones_like(b))
    return torch.where(m, a / bb, torch.zeros_like(a))


def hllc(hL, uNL, uTL, hR, uNR, uTR, g, hm):
    z = torch.zeros_like(hL)
    aL = _sqrt0(g * _pos(hL))
    aR = _sqrt0(g * _pos(hR))
    u_st = (0.5 * (uNL + uNR) + aL) - aR
    h_rt = 0.5 * (aL + aR) + 0.25 * (uNL - uNR)
    a_st = torch.where(h_rt > 0, h_rt, z)
    SL_w = _min(uNL - aL, u_st - a_st)
    SR_w = _max(uNR + aR, u_st + a_st)
    dryL = hL <= hm
    dryR = hR <= hm
    both = dryL & dryR
    SL = torch.where(both, z, torch.where(dryL, uNR - 2.0 * aR, SL_w))
    SR = torch.where(both, z, torch.where(dryR, uNL + 2.0 * aL, SR_w))
    denom = hR * (uNR - SR) - hL * (uNL - SL)
    big = torch.abs(denom) > 1.0e-14
    SM = torch.where(big, _div((SL * hR) * (uNR - SR) - (SR * hL) * (uNL - SL), denom, big), 0.5 * (uNL + uNR))
    c1, c2, c3 = SL >= 0, SM >= 0, SR >= 0
    mL = ~c1 & c2
    mR = ~c1 & ~c2 & c3
    huNL, huNR = hL * uNL, hR * uNR
    pL = (0.5 * g * hL) * hL
    pR = (0.5 * g * hR) * hR
    FL = (huNL, huNL * uNL + pL, huNL * uTL)
    FR = (huNR, huNR * uNR + pR, huNR * uTR)
    facL = _div(SL - uNL, (SL - SM) + 1.0e-30, mL)
    facR = _div(SR - uNR, (SR - SM) + 1.0e-30, mR)
    hLs, hRs = hL * facL, hR * facR
    FLs = (huNL + SL * (hLs - hL), FL[1] + SL * (hLs * SM - huNL), huNL * uTL + SL * (hLs * uTL - hL * uTL))
    FRs = (huNR + SR * (hRs - hR), FR[1] + SR * (hRs * SM - huNR), huNR * uTR + SR * (hRs * uTR - hR * uTR))
    return tuple(torch.where(c1, a, torch.where(c2, b, torch.where(c3, c, d))) for a, b, c, d in zip(FL, FLs, FRs, FR))


def srm_face(bL, bR, gzL, gzR, hL, hR, half_d, g, hm):
    z = torch.zeros_like(hL)
    db = bR - bL
    zL = half_d * gzL
    zR = db - half_d * gzR
    zf0 = torch.where(db > 0, db, z)
    dzc = zR - zL
    dz = db - dzc
    etaR = hR + db
    detaL = _pos(_min(dz, etaR - hL))
    detaR = _pos(_min(-dz, hL - etaR))
    hLf = _pos((hL + detaL) - zf0)
    hRf = _pos((etaR + detaR) - zf0)
    gate = ~((hL < hm) & (hR < hm))
    headL = zf0 - hL
    dzL = torch.where(hR < hm, _pos(headL), _pos(_min(dzc, headL)))
    srcL = ((0.5 * g) * (hLf + hL)) * (zf0 - dzL)
    headR = zf0 - etaR
    dzR = torch.where(hL < hm, _pos(headR), _pos(_min(zL - zR, headR)))
    srcR = ((0.5 * g) * (hRf + hR)) * ((zf0 - dzR) - db)
    return hLf, hRf, gate, srcL, srcR


def primitives(h, hu, hv, hm):
    wet = h > hm
    return _div(hu, h, wet), _div(hv, h, wet)


def residual(q, b, ix, iy, hdx, hdy, g, hm):
    """(3, nxp, nyp) residual on the window [2, n-3]^2 (zeros outside)."""
    h, hu, hv = q[0], q[1], q[2]
    u, v = primitives(h, hu, hv, hm)
    zx = torch.zeros_like(b[:1])
    zy = torch.zeros_like(b[:, :1])
    gzx = torch.cat([zx, (0.5 * (b[2:] - b[:-2])) * ix, zx], 0)
    gzy = torch.cat([zy, (0.5 * (b[:, 2:] - b[:, :-2])) * iy, zy], 1)
    xhL, xhR, xg, xsL, xsR = srm_face(b[:-1], b[1:], gzx[:-1], gzx[1:], h[:-1], h[1:], hdx, g, hm)
    xFh, xFN, xFT = hllc(xhL, u[:-1], v[:-1], xhR, u[1:], v[1:], g, hm)
    yhL, yhR, yg, ysL, ysR = srm_face(b[:, :-1], b[:, 1:], gzy[:, :-1], gzy[:, 1:], h[:, :-1], h[:, 1:], hdy, g, hm)
    yFh, yFN, yFT = hllc(yhL, v[:, :-1], u[:, :-1], yhR, v[:, 1:], u[:, 1:], g, hm)
    nxp, nyp = h.shape
    R = (slice(2, nxp - 2), slice(2, nyp - 2))
    Lf = (slice(1, nxp - 3), slice(2, nyp - 2))
    Tf = (slice(2, nxp - 2), slice(2, nyp - 2))
    Bf = (slice(2, nxp - 2), slice(1, nyp - 3))
    gR, gL, gT, gB = xg[R], xg[Lf], yg[Tf], yg[Bf]

    def m(gate, x):
        return torch.where(gate, x, torch.zeros_like(x))

    rh = (-m(gR, xFh[R] * ix) + m(gL, xFh[Lf] * ix) - m(gT, yFh[Tf] * iy) + m(gB, yFh[Bf] * iy))
    rhu = (-m(gR, xFN[R] * ix) - m(gR, xsL[R] * ix) + m(gL, xFN[Lf] * ix) + m(gL, xsR[Lf] * ix)
           - m(gT, yFT[Tf] * iy) + m(gB, yFT[Bf] * iy))
    rhv = (-m(gR, xFT[R] * ix) + m(gL, xFT[Lf] * ix) - m(gT, yFN[Tf] * iy) - m(gT, ysL[Tf] * iy)
           + m(gB, yFN[Bf] * iy) + m(gB, ysR[Bf] * iy))
    pad = torch.nn.functional.pad
    return torch.stack([pad(rh, (2, 2, 2, 2)), pad(rhu, (2, 2, 2, 2)), pad(rhv, (2, 2, 2, 2))])


def update(q, r, dt, n_cell, rain, hm, g, vcap, quadratic=True):
    h1 = q[0] + dt * (r[0] + rain)
    hu1 = q[1] + dt * r[1]
    hv1 = q[2] + dt * r[2]
    dry = h1 < hm
    hs = torch.where(h1 > hm, h1, torch.full_like(h1, hm))
    u = hu1 / hs
    v = hv1 / hs
    modU = _sqrt0(u * u + v * v)
    h43 = torch.where(dry, torch.ones_like(hs), hs) ** (-4.0 / 3.0)
    D = (((1.0e-10 + dt) * g) * h43) * (modU + 1.0e-30)
    n_cri = _sqrt0(1.0 / D)
    n = torch.where((modU > vcap) & (n_cri > n_cell), n_cri, n_cell)
    Cf = ((g * n) * n) * h43
    if quadratic:
        two = ((2.0 * dt) * Cf) * modU
        alpha = 2.0 / (torch.sqrt(1.0 + 2.0 * two) + 1.0)
    else:
        alpha = 1.0 / (1.0 + ((dt * Cf) * modU))
    z = torch.zeros_like(h1)
    h = torch.where(dry, _pos(h1), h1)
    hu = torch.where(dry, z, hu1 * alpha)
    hv = torch.where(dry, z, hv1 * alpha)
    return torch.stack([h, hu, hv])


def step(q, b, n_cell, rain, dt, active, stage, eta, c):
    """One flat step on the padded box: only `active` cells change; `stage` cells take
    h = max(eta - b, 0), hu = hv = 0."""
    r = residual(q, b, c["ix"], c["iy"], c["hdx"], c["hdy"], c["g"], c["hm"])
    qn = update(q, r, dt, n_cell, rain, c["hm"], c["g"], c["vcap"])
    hs = _pos(eta - b)
    st = torch.stack([hs, torch.zeros_like(hs), torch.zeros_like(hs)])
    qn = torch.where(stage[None], st, qn)
    return torch.where(active[None], qn, q)
