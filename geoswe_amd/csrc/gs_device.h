// gs_device.h -- device arithmetic of the SRM-HLLC finite-volume step, shared by every
// layout (dense sweep, flat runs, ensemble) so that all of them produce the same bits.
//
// Evaluation order follows the reference kernel text src/geoswe/rhs_cuda.py:900-1208 and
// the update tail rhs_cuda.py:1757-1796 term by term. The library is compiled with
// -ffp-contract=off and correctly rounded fp32 division/sqrt, so every operation below is
// one IEEE-754 rounding and the NumPy oracle reproduces it exactly.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/geoswe_abi.h"

// Host and device: the same arithmetic also builds for the CPU (host-side checks of the
// hand-written adjoint against finite differences of these functions).
#define GS_HD __host__ __device__ __forceinline__

#include "gs_crmath.h"

namespace gs {

template <typename T> GS_HD T vmax(T a, T b) { return fmax(a, b); }
template <typename T> GS_HD T vmin(T a, T b) { return fmin(a, b); }
template <> GS_HD float vmax<float>(float a, float b) { return fmaxf(a, b); }
template <> GS_HD float vmin<float>(float a, float b) { return fminf(a, b); }
// Correctly rounded float32 division and square root. On the device they are the short
// refinements of gs_crmath.h (the IEEE results bit for bit, in 9 instructions where the
// compiler's general expansions take 11 and 16); on the host, the IEEE operators.
GS_HD float vsqrt(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return sqrt_cr(x);
#else
  return sqrtf(x);
#endif
}
GS_HD double vsqrt(double x) { return sqrt(x); }
// sqrt of an operand that is >= 2^-96, +-0, +inf or NaN (no range branch on the device)
GS_HD float vsqrt_big(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return sqrt_cr_big(x);
#else
  return sqrtf(x);
#endif
}
GS_HD double vsqrt_big(double x) { return sqrt(x); }
GS_HD float vdiv(float a, float b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return div_cr(a, b);
#else
  return a / b;
#endif
}
GS_HD double vdiv(double a, double b) { return a / b; }
GS_HD float vabs(float x) { return fabsf(x); }
GS_HD double vabs(double x) { return fabs(x); }

// Per-face SRM reconstruction + HLLC flux + the two one-sided bed sources.
// L is the lower-index cell, R the upper one, along the face normal.
template <typename T>
struct Face {
  T Fh, FN, FT;   // HLLC mass, normal-momentum and transverse-momentum fluxes
  T srcL, srcR;   // bed source of L (subtracted) and of R (added), rhs_cuda.py:1060-1115
  bool gate;      // false when both cells are dry: the face contributes nothing
};

// HLLC in the face frame, rhs_cuda.py:986-1036. Only the branch the cascade selects is
// evaluated for the star state; the selected value is computed with the reference's
// expressions, so the result is the reference's bit pattern.
template <typename T>
GS_HD void hllc(T hL, T uNL, T uTL, T hR, T uNR, T uTR, T g, T h_min,
                                     T& Fh, T& FN, T& FT) {
  const T z = T(0);
  const T aL = vsqrt(g * (hL > z ? hL : z));
  const T aR = vsqrt(g * (hR > z ? hR : z));
  const T u_st = (T(0.5) * (uNL + uNR) + aL) - aR;
  const T h_rt = T(0.5) * (aL + aR) + T(0.25) * (uNL - uNR);
  // star celerity: the reference forms h* = h_rt^2/g and a* = sqrt(g h*), which is h_rt
  // itself in exact arithmetic; taking it directly saves a division and a square root
  const T a_st = (h_rt > z) ? h_rt : z;
  const T SL_w = vmin(uNL - aL, u_st - a_st);
  const T SR_w = vmax(uNR + aR, u_st + a_st);
  const bool dryL = (hL <= h_min);
  const bool dryR = (hR <= h_min);
  T SL = dryL ? (uNR - T(2) * aR) : SL_w;
  T SR = dryR ? (uNL + T(2) * aL) : SR_w;
  if (dryL && dryR) { SL = z; SR = z; }
  const T denom = hR * (uNR - SR) - hL * (uNL - SL);
  T SM;
  if (vabs(denom) > T(1.0e-14)) {
    SM = vdiv((SL * hR) * (uNR - SR) - (SR * hL) * (uNL - SL), denom);
  } else {
    SM = T(0.5) * (uNL + uNR);
  }
  const T huNL = hL * uNL, huNR = hR * uNR;
  const T pL = (T(0.5) * g * hL) * hL;
  const T pR = (T(0.5) * g * hR) * hR;
  const T FL_N = huNL * uNL + pL;
  const T FR_N = huNR * uNR + pR;
  if (SL >= z) {
    Fh = huNL; FN = FL_N; FT = huNL * uTL;
  } else if (!(SM >= z) && !(SR >= z)) {
    Fh = huNR; FN = FR_N; FT = huNR * uTR;
  } else {
    // star state of the side the cascade picks: SM >= 0 -> left, else right
    const bool lft = (SM >= z);
    const T S = lft ? SL : SR;
    const T uN = lft ? uNL : uNR;
    const T uT = lft ? uTL : uTR;
    const T hh = lft ? hL : hR;
    const T huN = lft ? huNL : huNR;
    const T F_N = lft ? FL_N : FR_N;
    const T fac = vdiv(S - uN, (S - SM) + T(1.0e-30));
    const T hs = hh * fac;
    Fh = huN + S * (hs - hh);
    FN = F_N + S * (hs * SM - huN);
    FT = huN * uT + S * (hs * uT - hh * uT);
  }
}

// One face: SRM face depths (rhs_cuda.py:1042-1056), flux, and both cells' sources with
// the gate on the other cell's depth (rhs_cuda.py:1060-1073, 1101-1115). The reference
// evaluates each face twice, once in each cell's frame; the two evaluations are equal
// bit for bit (IEEE negation is exact), so one evaluation serves both cells.
//
// Elevations are taken relative to the lower cell's bed bL (a per-face shift of the
// vertical datum): eta_L -> hL, eta_R -> hR + (bR - bL), z -> z - bL. The expressions
// are the reference's term by term in exact arithmetic, but (z_f - b_c) and the face
// depths no longer cancel against |b|: with beds of tens of metres one float32 ulp of b
// is several micrometres, which the absolute-elevation form adds to every face depth
// and to the bed-slope source.
template <typename T>
GS_HD void face(T bL, T bR, T gzL, T gzR, T hL, T hR, T uNL, T uTL, T uNR, T uTR,
                                     T half_d, T g, T h_min, Face<T>& f) {
  const T z = T(0);
  f.gate = !(hL < h_min && hR < h_min);
  const T db = bR - bL;                   // bed of R in L's datum
  const T zL = half_d * gzL;              // reconstructed face bed from L
  const T zR = db - half_d * gzR;         // reconstructed face bed from R
  const T zf0 = (db > z) ? db : z;        // max(bL, bR) - bL
  const T dzc = zR - zL;
  const T dz = db - dzc;
  const T etaR = hR + db;                 // eta_R - bL; eta_L - bL is hL
  const T detaL = vmax(z, vmin(dz, etaR - hL));
  const T detaR = vmax(z, vmin(-dz, hL - etaR));
  const T hLf = vmax(z, (hL + detaL) - zf0);
  const T hRf = vmax(z, (etaR + detaR) - zf0);
  hllc<T>(hLf, uNL, uTL, hRf, uNR, uTR, g, h_min, f.Fh, f.FN, f.FT);
  const T headL = zf0 - hL;
  const T dzL = (hR < h_min) ? vmax(z, headL) : vmax(z, vmin(dzc, headL));
  f.srcL = ((T(0.5) * g) * (hLf + hL)) * (zf0 - dzL);
  const T headR = zf0 - etaR;
  const T dzR = (hL < h_min) ? vmax(z, headR) : vmax(z, vmin(zL - zR, headR));
  f.srcR = ((T(0.5) * g) * (hRf + hR)) * ((zf0 - dzR) - db);
}

// Primitive velocities with the dry gate (rhs_cuda.py:931-938).
template <typename T>
GS_HD void prims(T h, T hu, T hv, T h_min, T& u, T& v) {
  const bool wet = h > h_min;
  const T hs = wet ? h : h_min;
  u = wet ? vdiv(hu, hs) : T(0);
  v = wet ? vdiv(hv, hs) : T(0);
}

// h^(-1/3) from IEEE +,-,* only: Newton on f(y) = y^-3 - h, y <- y*(4/3 - (h/3)*y^3),
// three iterations from the magic-constant seed of Moroz et al. (2021),
// bits(y0) = 0x54a2fa8c - bits(h)/3.
GS_HD float icbrt(float h) {
  const uint32_t bits = __builtin_bit_cast(uint32_t, h);
  float y = __builtin_bit_cast(float, 0x54A2FA8Cu - bits / 3u);
  const float xt = h * (float)(1.0 / 3.0);
  const float c43 = (float)(4.0 / 3.0);
  y = y * (c43 - xt * ((y * y) * y));
  y = y * (c43 - xt * ((y * y) * y));
  y = y * (c43 - xt * ((y * y) * y));
  return y;
}
// h^(-4/3) = (y*y)*(y*y); bit-reproducible on the CPU (numpy float32 ops).
GS_HD float pow_m43(float h) {
  const float y = icbrt(h);
  const float y2 = y * y;
  return y2 * y2;
}
// float64: the float32 root refined by two more Newton steps in float64.
GS_HD double pow_m43(double h) {
  double y = (double)icbrt((float)h);
  const double xt = h * (1.0 / 3.0);
  const double c43 = 4.0 / 3.0;
  y = y * (c43 - xt * ((y * y) * y));
  y = y * (c43 - xt * ((y * y) * y));
  const double y2 = y * y;
  return y2 * y2;
}

// Scalars of the cell update, converted once per launch.
template <typename T>
struct Phys {
  T g, h_min, h_cfl, vcap, inv_dx, inv_dy, half_dx, half_dy;
  int friction, quadratic, keep_h, cfl_l2, has_vcap;
};

// Fused update tail, rhs_cuda.py:1757-1796 (flat form compressed_rhs.py:716-748):
// Euler update, wet/dry rule, point-implicit Manning friction with the velocity cap.
template <typename T>
GS_HD void update(T q0, T q1, T q2, T rr0, T r1, T r2, T dt, T n,
                                       const Phys<T>& P, T& h, T& hu, T& hv) {
  h = q0 + dt * rr0;
  hu = q1 + dt * r1;
  hv = q2 + dt * r2;
  if (h < P.h_min) {
    h = P.keep_h ? (h > T(0) ? h : T(0)) : T(0);
    hu = T(0);
    hv = T(0);
  } else if (P.friction) {
    const T hs = (h > P.h_min) ? h : P.h_min;
    const T u = vdiv(hu, hs);
    const T v = vdiv(hv, hs);
    const T modU = vsqrt(u * u + v * v);
    const T h43 = pow_m43(hs);
    if (P.has_vcap && modU > P.vcap) {
      const T n_cri = vsqrt(vdiv(T(1), (((T(1.0e-10) + dt) * P.g) * h43) * (modU + T(1.0e-30))));
      if (n_cri > n) n = n_cri;
    }
    const T Cf = ((P.g * n) * n) * h43;
    T alpha;
    if (P.quadratic) {
      const T two = ((T(2) * dt) * Cf) * modU;
      // the root's operand is >= 1
      alpha = vdiv(T(2), vsqrt_big(T(1) + T(2) * two) + T(1));
    } else {
      alpha = vdiv(T(1), T(1) + ((dt * Cf) * modU));
    }
    hu = hu * alpha;
    hv = hv * alpha;
  }
}

// CFL wave speed of one cell as an unsigned key whose maximum is the maximum speed: the
// IEEE bits of the speed in the state's precision (non-negative floats order as their
// bits, and a NaN of either sign lies above +inf, so NaN dominates), 0 for a skipped
// cell. One integer max per cell; lam_key_bits64 turns the reduced key into the float64
// bits the clock holds (canonical NaN). The two layouts follow their own reference kernels:
//   dense (FLAT = false), lammax_fp32_lean (solver.py:152-205): cells with h < h_cfl are
//     skipped and the velocity divides by h;
//   flat (FLAT = true), cfl_lammax_flat_blk and the fused step's lambda
//     (compressed_solver.py:200-240, compressed_rhs.py:1041-1053): cells with h < h_min are
//     skipped and the velocity divides by max(h, h_cfl) (h_min_cfl there).
// Both keep a NaN speed (compressed_solver.py:200-240 NaN rule).
template <typename T> struct LamKey { using type = uint64_t; };
template <> struct LamKey<float> { using type = uint32_t; };
template <typename T> using lam_t = typename LamKey<T>::type;

__device__ __forceinline__ uint64_t lam_key_bits64(uint32_t k) {
  if ((k & 0x7FFFFFFFu) > 0x7F800000u) return 0x7FF8000000000000ull;
  return (uint64_t)__double_as_longlong((double)__uint_as_float(k));
}
__device__ __forceinline__ uint64_t lam_key_bits64(uint64_t k) {
  if ((k & 0x7FFFFFFFFFFFFFFFull) > 0x7FF0000000000000ull) return 0x7FF8000000000000ull;
  return k;
}

template <typename T, bool FLAT = false>
__device__ __forceinline__ lam_t<T> lam_key(T h, T hu, T hv, const Phys<T>& P) {
  if (h < (FLAT ? P.h_min : P.h_cfl)) return 0;
  const T hs = FLAT ? (h > P.h_cfl ? h : P.h_cfl) : h;
  T vel;
  if (P.cfl_l2) {
    const T u = vdiv(hu, hs);
    const T v = vdiv(hv, hs);
    vel = vsqrt(u * u + v * v);
  } else {
    // max(|hu/hs|, |hv/hs|) = max(|hu|, |hv|)/hs bit for bit: correctly rounded division
    // by the same positive hs is monotonic, so one division serves both components
    const T au = vabs(hu), av = vabs(hv);
    vel = vdiv((au > av) ? au : av, hs);
  }
  // h >= h_min > 0 (or NaN) here, so sqrt(g*max(h, 0)) of the flat kernels is sqrt(g*h);
  // the speed is +0 or more, or NaN
  const T lam = vel + vsqrt(P.g * h);
  if (sizeof(T) == 4) return (lam_t<T>)__float_as_uint((float)lam);
  return (lam_t<T>)__double_as_longlong((double)lam);
}

}  // namespace gs
