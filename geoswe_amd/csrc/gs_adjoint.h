// gs_adjoint.h -- hand-written reverse mode (vector-Jacobian products) of the step
// arithmetic in gs_device.h: the HLLC flux, the SRM face (reconstruction, flux, both
// one-sided bed sources), the dry-gated primitive velocities and the fused update with
// Manning friction. Each function recomputes the forward values it needs and then runs
// the reverse sweep statement by statement in the exact reverse order of the forward
// text. At a max/min the cotangent follows the argument the forward selected (a <= b
// selects a in min, a >= b selects a in max); sqrt(0) and the dry cut-offs pass no
// cotangent. Host and device: the CPU check builds these same functions.
#pragma once

#include "gs_device.h"

namespace gs {

// A quotient that only the reverse sweep uses (no recomputed forward value and no branch
// depends on it): the float32 device build takes the hardware reciprocal (1 ulp) times the
// numerator; float64 and the host build divide. The recomputed forward values keep the
// forward's correctly rounded operations (vdiv, vsqrt), so the reverse follows exactly the
// branches the forward took.
template <typename T>
GS_HD T rdiv(T a, T b) {
#if defined(__HIP_DEVICE_COMPILE__)
  if constexpr (sizeof(T) == 4) return a * __builtin_amdgcn_rcpf(b);
#endif
  return a / b;
}

// Cotangents of the 10 face inputs.
template <typename T>
struct FaceGrad {
  T bL, bR, gzL, gzR, hL, hR, uNL, uTL, uNR, uTR;
};

// Reverse of hllc(): given cotangents of (Fh, FN, FT), ACCUMULATES into the six input
// cotangents.
template <typename T>
GS_HD void hllc_vjp(T hL, T uNL, T uTL, T hR, T uNR, T uTR, T g, T h_min, T bFh, T bFN, T bFT, T& ghL,
                    T& guNL, T& guTL, T& ghR, T& guNR, T& guTR) {
  const T z = T(0);
  // ---- forward (as hllc)
  const T aL = vsqrt(g * (hL > z ? hL : z));
  const T aR = vsqrt(g * (hR > z ? hR : z));
  const T u_st = (T(0.5) * (uNL + uNR) + aL) - aR;
  const T h_rt = T(0.5) * (aL + aR) + T(0.25) * (uNL - uNR);
  const T a_st = (h_rt > z) ? h_rt : z;
  const T xl = uNL - aL, yl = u_st - a_st;
  const T xr = uNR + aR, yr = u_st + a_st;
  const T SL_w = vmin(xl, yl);
  const T SR_w = vmax(xr, yr);
  const bool dryL = (hL <= h_min);
  const bool dryR = (hR <= h_min);
  const bool both = dryL && dryR;
  T SL = dryL ? (uNR - T(2) * aR) : SL_w;
  T SR = dryR ? (uNL + T(2) * aL) : SR_w;
  if (both) { SL = z; SR = z; }
  const T denom = hR * (uNR - SR) - hL * (uNL - SL);
  const bool big = vabs(denom) > T(1.0e-14);
  const T SM = big ? vdiv((SL * hR) * (uNR - SR) - (SR * hL) * (uNL - SL), denom) : T(0.5) * (uNL + uNR);
  int br;   // 0 left flux, 1 right flux, 2 left star, 3 right star
  if (SL >= z) br = 0;
  else if (!(SM >= z) && !(SR >= z)) br = 1;
  else br = (SM >= z) ? 2 : 3;

  // ---- reverse, branch-free: the one-sided and the star contributions are both formed and
  // combined with value selects (the unselected side adds an exact zero). A branch that picks
  // which accumulator to update makes the compiler select between their addresses and keep
  // them in memory.
  const bool one = br <= 1;
  // one-sided flux of side L (br 0) or R (br 1): Fh = h uN, FN = h uN uN + (0.5 g h) h,
  // FT = h uN uT
  const bool l = (br == 0);
  const T h1 = l ? hL : hR, uN1 = l ? uNL : uNR, uT1 = l ? uTL : uTR;
  const T huN1 = h1 * uN1;
  const T ghuN1 = bFh + bFN * uN1 + bFT * uT1;
  const T guN_1 = bFN * huN1 + ghuN1 * h1;
  const T guT_1 = bFT * huN1;
  const T gh_1 = bFN * (g * h1) + ghuN1 * uN1;
  // star state of side lft (br 2: left, br 3: right)
  const bool lft = (br == 2);
  const T S = lft ? SL : SR;
  const T uN = lft ? uNL : uNR;
  const T uT = lft ? uTL : uTR;
  const T hh = lft ? hL : hR;
  const T huN = hh * uN;
  const T den = (S - SM) + T(1.0e-30);
  const T fac = one ? z : rdiv(S - uN, den);
  const T hs = hh * fac;
  // Fh = huN + S (hs - hh); FN = F_N + S (hs SM - huN); FT = huN uT + S (hs uT - hh uT)
  T gS = bFh * (hs - hh) + bFN * (hs * SM - huN) + bFT * (hs * uT - hh * uT);
  const T ghs = bFh * S + bFN * (S * SM) + bFT * (S * uT);
  T ghh = -(bFh * S) - bFT * (S * uT);
  T ghuN = bFh - bFN * S + bFT * uT;
  const T gSMs = bFN * (S * hs);
  const T guT = bFT * huN + bFT * (S * (hs - hh));
  const T gF_N = bFN;
  ghh += ghs * fac;
  const T gfac = ghs * hh;
  const T gnum = one ? z : rdiv(gfac, den);
  const T gden = one ? z : rdiv(-(gfac * fac), den);
  gS += gnum + gden;
  T guN = -gnum;
  ghuN += gF_N * uN;
  guN += gF_N * huN;
  ghh += gF_N * (g * hh);
  ghh += ghuN * uN;
  guN += ghuN * hh;
  const bool sl = !one && lft, sr = !one && !lft;
  ghL += l && one ? gh_1 : (sl ? ghh : z);
  guNL += l && one ? guN_1 : (sl ? guN : z);
  guTL += l && one ? guT_1 : (sl ? guT : z);
  ghR += !l && one ? gh_1 : (sr ? ghh : z);
  guNR += !l && one ? guN_1 : (sr ? guN : z);
  guTR += !l && one ? guT_1 : (sr ? guT : z);
  T gSL = sl ? gS : z, gSR = sr ? gS : z;
  const T gSM = one ? z : gSMs - gden;
  // SM = Nn / denom (|denom| > 1e-14) or 0.5 (uNL + uNR)
  const T gNn = big ? rdiv(gSM, denom) : z;
  const T gdn = big ? rdiv(-(gSM * SM), denom) : z;
  gSL += gNn * (hR * (uNR - SR) + SR * hL) + gdn * hL;
  gSR += gNn * (-(SL * hR) - hL * (uNL - SL)) - gdn * hR;
  ghR += gNn * (SL * (uNR - SR)) + gdn * (uNR - SR);
  ghL += gNn * (-(SR * (uNL - SL))) - gdn * (uNL - SL);
  guNR += gNn * (SL * hR) + gdn * hR + (big ? z : T(0.5) * gSM);
  guNL += gNn * (-(SR * hL)) - gdn * hL + (big ? z : T(0.5) * gSM);
  // SL, SR (both dry: constants)
  const T gsl = both ? z : gSL, gsr = both ? z : gSR;
  const bool l1 = !dryL && (xl <= yl), l2 = !dryL && !(xl <= yl);
  const bool r1 = !dryR && (xr >= yr), r2 = !dryR && !(xr >= yr);
  guNR += dryL ? gsl : z;
  guNL += l1 ? gsl : z;
  guNL += dryR ? gsr : z;
  guNR += r1 ? gsr : z;
  T gaL = (dryR ? T(2) * gsr : z) - (l1 ? gsl : z);
  T gaR = (r1 ? gsr : z) - (dryL ? T(2) * gsl : z);
  T gust = (l2 ? gsl : z) + (r2 ? gsr : z);
  const T gast = (r2 ? gsr : z) - (l2 ? gsl : z);
  const T ghrt = (h_rt > z) ? gast : z;
  gaL += T(0.5) * ghrt;
  gaR += T(0.5) * ghrt;
  guNL += T(0.25) * ghrt;
  guNR -= T(0.25) * ghrt;
  guNL += T(0.5) * gust;
  guNR += T(0.5) * gust;
  gaL += gust;
  gaR -= gust;
  if (hL > z && aL > z) ghL += gaL * rdiv(g, T(2) * aL);
  if (hR > z && aR > z) ghR += gaR * rdiv(g, T(2) * aR);
}

// Reverse of face(): cotangents of (Fh, FN, FT, srcL, srcR) -> cotangents of the 10
// inputs (overwritten). The gate is not differentiated (a dry face contributes nothing).
template <typename T>
GS_HD void face_vjp(T bL, T bR, T gzL, T gzR, T hL, T hR, T uNL, T uTL, T uNR, T uTR, T half_d, T g, T h_min,
                    T bFh, T bFN, T bFT, T bsL, T bsR, FaceGrad<T>& G) {
  const T z = T(0);
  // ---- forward (as face)
  const T db = bR - bL;
  const T zL = half_d * gzL;
  const T zR = db - half_d * gzR;
  const T zf0 = (db > z) ? db : z;
  const T dzc = zR - zL;
  const T dz = db - dzc;
  const T etaR = hR + db;
  const T mL = vmin(dz, etaR - hL);
  const T detaL = vmax(z, mL);
  const T mR = vmin(-dz, hL - etaR);
  const T detaR = vmax(z, mR);
  const T aLf = (hL + detaL) - zf0;
  const T aRf = (etaR + detaR) - zf0;
  const T hLf = vmax(z, aLf);
  const T hRf = vmax(z, aRf);
  const T headL = zf0 - hL;
  const T mdl = vmin(dzc, headL);
  const T dzL = (hR < h_min) ? vmax(z, headL) : vmax(z, mdl);
  const T AL = (T(0.5) * g) * (hLf + hL);
  const T BL = zf0 - dzL;
  const T headR = zf0 - etaR;
  const T zlr = zL - zR;
  const T mdr = vmin(zlr, headR);
  const T dzR = (hL < h_min) ? vmax(z, headR) : vmax(z, mdr);
  const T AR = (T(0.5) * g) * (hRf + hR);
  const T BR = (zf0 - dzR) - db;

  // ---- reverse
  T gdb = z, gzLv = z, gzRv = z, gzf0 = z, gdzc = z, gdz = z, getaR = z, gdetaL = z, gdetaR = z;
  T ghLf = z, ghRf = z, ghL = z, ghR = z, gheadL = z, gheadR = z;
  T guNL = z, guTL = z, guNR = z, guTR = z;
  // srcR = AR * BR
  const T gAR = bsR * BR, gBR = bsR * AR;
  ghRf += (T(0.5) * g) * gAR;
  ghR += (T(0.5) * g) * gAR;
  gzf0 += gBR;
  const T gdzR = -gBR;
  gdb -= gBR;
  {
    const bool pr = (hL < h_min) ? (headR > z) : false;
    const bool pz = !(hL < h_min) && (mdr > z) && (zlr <= headR);
    const bool ph = !(hL < h_min) && (mdr > z) && !(zlr <= headR);
    gheadR += (pr || ph) ? gdzR : z;
    gzLv += pz ? gdzR : z;
    gzRv -= pz ? gdzR : z;
  }
  gzf0 += gheadR;
  getaR -= gheadR;
  // srcL = AL * BL
  const T gAL = bsL * BL, gBL = bsL * AL;
  ghLf += (T(0.5) * g) * gAL;
  ghL += (T(0.5) * g) * gAL;
  gzf0 += gBL;
  const T gdzL = -gBL;
  {
    const bool ph = (hR < h_min) ? (headL > z) : ((mdl > z) && !(dzc <= headL));
    const bool pc = !(hR < h_min) && (mdl > z) && (dzc <= headL);
    gheadL += ph ? gdzL : z;
    gdzc += pc ? gdzL : z;
  }
  gzf0 += gheadL;
  ghL -= gheadL;
  // flux
  hllc_vjp<T>(hLf, uNL, uTL, hRf, uNR, uTR, g, h_min, bFh, bFN, bFT, ghLf, guNL, guTL, ghRf, guNR, guTR);
  // face depths
  const T gR_ = (aRf > z) ? ghRf : z, gL_ = (aLf > z) ? ghLf : z;
  getaR += gR_;
  gdetaR += gR_;
  gzf0 -= gR_;
  ghL += gL_;
  gdetaL += gL_;
  gzf0 -= gL_;
  {
    const bool a1 = (mR > z) && (-dz <= hL - etaR), a2 = (mR > z) && !(-dz <= hL - etaR);
    gdz -= a1 ? gdetaR : z;
    ghL += a2 ? gdetaR : z;
    getaR -= a2 ? gdetaR : z;
    const bool b1 = (mL > z) && (dz <= etaR - hL), b2 = (mL > z) && !(dz <= etaR - hL);
    gdz += b1 ? gdetaL : z;
    getaR += b2 ? gdetaL : z;
    ghL -= b2 ? gdetaL : z;
  }
  ghR += getaR;
  gdb += getaR;
  gdb += gdz;
  gdzc -= gdz;
  gzRv += gdzc;
  gzLv -= gdzc;
  if (db > z) gdb += gzf0;
  gdb += gzRv;
  G.gzR = -(half_d * gzRv);
  G.gzL = half_d * gzLv;
  G.bR = gdb;
  G.bL = -gdb;
  G.hL = ghL;
  G.hR = ghR;
  G.uNL = guNL;
  G.uTL = guTL;
  G.uNR = guNR;
  G.uTR = guTR;
}

// Reverse of prims(): cotangents of (u, v) ACCUMULATED into those of (h, hu, hv).
template <typename T>
GS_HD void prims_vjp(T h, T hu, T hv, T h_min, T gu, T gv, T& gh, T& ghu, T& ghv) {
  if (!(h > h_min)) return;
  const T u = vdiv(hu, h), v = vdiv(hv, h);
  ghu += rdiv(gu, h);
  ghv += rdiv(gv, h);
  gh -= rdiv(gu * u + gv * v, h);
}

// Reverse of update() from its pre-friction values h1 = q0 + dt*rr0, hu1, hv1:
// cotangents of the outputs (h, hu, hv) -> cotangents of (h1, hu1, hv1) and of the cell's
// Manning n (overwritten). dt is not differentiated (the step sequence is frozen).
template <typename T>
GS_HD void update_vjp(T h1, T hu1, T hv1, T dt, T n, const Phys<T>& P, T bh, T bhu, T bhv, T& g1, T& gu1, T& gv1,
                      T& gn) {
  const T z = T(0);
  g1 = z;
  gu1 = z;
  gv1 = z;
  gn = z;
  if (h1 < P.h_min) {
    if (P.keep_h && h1 > z) g1 = bh;
    return;
  }
  if (!P.friction) {
    g1 = bh;
    gu1 = bhu;
    gv1 = bhv;
    return;
  }
  const T hs = (h1 > P.h_min) ? h1 : P.h_min;
  const T u = vdiv(hu1, hs);
  const T v = vdiv(hv1, hs);
  const T modU = vsqrt(u * u + v * v);
  const T h43 = pow_m43(hs);
  T ne = n, n_cri = z, D = z;
  bool capped = false;
  if (P.has_vcap && modU > P.vcap) {
    D = (((T(1.0e-10) + dt) * P.g) * h43) * (modU + T(1.0e-30));
    n_cri = vsqrt(vdiv(T(1), D));
    if (n_cri > n) { ne = n_cri; capped = true; }
  }
  const T Cf = ((P.g * ne) * ne) * h43;
  T alpha, s = z;
  if (P.quadratic) {
    const T two = ((T(2) * dt) * Cf) * modU;
    s = vsqrt(T(1) + T(2) * two);
    alpha = vdiv(T(2), s + T(1));
  } else {
    alpha = vdiv(T(1), T(1) + ((dt * Cf) * modU));
  }
  // hu = hu1 alpha, hv = hv1 alpha, h = h1
  const T galpha = bhu * hu1 + bhv * hv1;
  gu1 = bhu * alpha;
  gv1 = bhv * alpha;
  g1 = bh;
  T gCf, gmodU = z;
  if (P.quadratic) {
    const T gs = galpha * (-(alpha * alpha) * T(0.5));   // d(2/(s+1))/ds = -alpha^2/2
    const T gtwo = rdiv(gs, s);                          // ds/dtwo = 1/s
    gCf = gtwo * ((T(2) * dt) * modU);
    gmodU += gtwo * ((T(2) * dt) * Cf);
  } else {
    const T gx = -(galpha * alpha) * alpha;
    gCf = gx * (dt * modU);
    gmodU += gx * (dt * Cf);
  }
  const T gne = gCf * ((T(2) * P.g) * ne * h43);
  T gh43 = gCf * ((P.g * ne) * ne);
  if (capped) {
    const T gD = gne * rdiv(-(T(0.5) * n_cri), D);
    gh43 += gD * (((T(1.0e-10) + dt) * P.g) * (modU + T(1.0e-30)));
    gmodU += gD * (((T(1.0e-10) + dt) * P.g) * h43);
  } else {
    gn = gne;
  }
  T ghs = gh43 * rdiv((T(-4) / T(3)) * h43, hs);
  if (modU > z) {
    const T gu = gmodU * rdiv(u, modU), gv = gmodU * rdiv(v, modU);
    gu1 += rdiv(gu, hs);
    gv1 += rdiv(gv, hs);
    ghs -= rdiv(gu * u + gv * v, hs);
  }
  if (h1 > P.h_min) g1 += ghs;
}

}  // namespace gs
