// gs_crmath.h -- correctly rounded float32 division and square root in fewer instructions
// than the compiler's general expansions, for the operands the step produces.
//
// Both are refinements of the hardware's 1-ulp approximations (v_rcp_f32, v_sqrt_f32) by
// fused multiply-adds. The refinement steps are written here once, as functions of the
// approximation, so that the exhaustive host check (tests/native/cr_check.cpp) runs the
// very code the kernels run, fed with every value the 1-ulp hardware result can take.
//
// Division, after Markstein (IBM J. Res. Dev. 34(1), 1990; Theorem 8.10 of Muller et al.,
// Handbook of Floating-Point Arithmetic): with y = RN(1/b) and q0 = RN(a*y), the remainder
// r = a - b*q0 is exact and RN(q0 + r*y) = RN(a/b). y comes from one Newton step on the
// hardware reciprocal; that this step returns RN(1/b) for every significand of b and every
// admissible hardware result is what the host check proves exhaustively. The operands are
// brought into range by v_div_scale and the last fused multiply-add is v_div_fmas, which
// undoes the scaling with a single rounding (so subnormal quotients round once), and
// v_div_fixup supplies the IEEE results for zeros, infinities and NaNs: the sequence is the
// compiler's own expansion with its second correction step removed (9 instructions instead
// of 11) where the quotient is normal. Where v_div_scale flags the quotient for rescaling
// (a subnormal or near-overflow quotient), the final rounding is on a coarser grid than the
// theorem's and the second correction is kept.
//
// Square root, the compiler's own neighbour test without its range scaling and its class
// check (9 instructions instead of 16): s = v_sqrt_f32(x) is within 1 ulp of sqrt(x); the
// signs of x - s_dn*s and x - s_up*s (one rounding each, by fma) pick RN(sqrt(x)) among
// s - 1 ulp, s and s + 1 ulp. Valid for x = +-0, +-inf, NaN and |x| >= 2^-96, where the
// residuals cannot underflow to zero (and the hardware square root sees no subnormal, which
// it takes as zero: sqrt of a negative subnormal must be NaN, not -0); the callers that can
// meet 0 < |x| < 2^-96 take the general expansion there.
#pragma once

#include <stdint.h>
#include <string.h>

#ifndef GS_HD
#define GS_HD __host__ __device__ __forceinline__
#endif

namespace gs {

GS_HD float cr_bits_f(uint32_t u) {
  float f;
  memcpy(&f, &u, 4);
  return f;
}
GS_HD uint32_t cr_f_bits(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
}

// One Newton step on an approximate reciprocal y0 of b: RN(1/b) when y0 is within 1 ulp.
GS_HD float cr_recip_refine(float b, float y0) {
  const float e = fmaf(-b, y0, 1.0f);
  return fmaf(e, y0, y0);
}

// Markstein's correction of q0 = RN(a*y) with y = RN(1/b), unscaled form (the host check;
// the device form scales a and b with v_div_scale and ends in v_div_fmas).
GS_HD float cr_div_refine(float a, float b, float y) {
  const float q0 = a * y;
  const float r = fmaf(-b, q0, a);
  return fmaf(r, y, q0);
}

// Selection of RN(sqrt(x)) among the neighbours of an approximation s within 1 ulp.
GS_HD float cr_sqrt_refine(float x, float s) {
  const uint32_t sb = cr_f_bits(s);
  const float s_dn = cr_bits_f(sb - 1u);
  const float s_up = cr_bits_f(sb + 1u);
  const float r_dn = fmaf(-s_dn, s, x);
  const float r_up = fmaf(-s_up, s, x);
  float t = (r_dn <= 0.0f) ? s_dn : s;
  t = (r_up > 0.0f) ? s_up : t;
  return t;
}

#if defined(__HIPCC__)
// Correctly rounded a/b for every float32 a, b (the IEEE quotient, bit for bit).
__device__ __forceinline__ float div_cr(float a, float b) {
  bool scale_den, scale_num;
  const float bs = __builtin_amdgcn_div_scalef(a, b, false, &scale_den);   // scaled denominator
  const float as = __builtin_amdgcn_div_scalef(a, b, true, &scale_num);    // scaled numerator
  (void)scale_den;
  const float y = cr_recip_refine(bs, __builtin_amdgcn_rcpf(bs));
  float q0 = as * y;
  float r = fmaf(-bs, q0, as);
  if (__builtin_expect(scale_num, 0)) {
    // a quotient div_fmas scales into the subnormal (or top) range is rounded on a coarser
    // grid, where one correction can miss a near-tie: take the second, as the general
    // expansion does (the branch no lane takes on the states the step produces)
    q0 = fmaf(r, y, q0);
    r = fmaf(-bs, q0, as);
  }
  const float q = __builtin_amdgcn_div_fmasf(r, y, q0, scale_num);
  return __builtin_amdgcn_div_fixupf(q, b, a);
}

// Correctly rounded sqrt(x) for |x| >= 2^-96 (and +-0, +-inf, NaN).
__device__ __forceinline__ float sqrt_cr_big(float x) {
  return cr_sqrt_refine(x, __builtin_amdgcn_sqrtf(x));
}

// Correctly rounded sqrt(x) for every float32 x: the short form, and the general expansion
// for 0 < |x| < 2^-96 (a branch no lane takes on the states the step produces).
__device__ __forceinline__ float sqrt_cr(float x) {
  if (__builtin_expect(x != 0.0f && __builtin_fabsf(x) < 0x1p-96f, 0)) return __builtin_sqrtf(x);
  return sqrt_cr_big(x);
}
#endif

}  // namespace gs
