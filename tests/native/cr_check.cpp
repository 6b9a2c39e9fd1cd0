// Host proof of the refinement steps in geoswe_amd/csrc/gs_crmath.h.
//
// The kernels feed those steps with the hardware's 1-ulp approximations (v_rcp_f32,
// v_sqrt_f32). This program feeds them with every value such an approximation can take
// (the correctly rounded result and its two neighbours) and compares with the IEEE result
// of the host (correctly rounded 1/b, a/b and sqrtf):
//   recip  every significand of b (one binade suffices: the step commutes with scaling by
//          powers of two while 1/b is normal, and the device scales b into that range);
//   div    Markstein's correction from y = RN(1/b), for 256 divisors (1, the all-ones
//          significand, 254 seeded ones) against every numerator significand, quotients in
//          two binades;
//   sqrt   every x in [1, 4) (two binades cover all significands and both exponent
//          parities; the selection commutes with scaling x by 4), the six binades above the
//          2^-96 threshold, the top binades and the special values.
// Prints one line per check and "failures N"; exit status 1 on any failure.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>

#define GS_HD inline
#include "../../geoswe_amd/csrc/gs_crmath.h"

using gs::cr_bits_f;
using gs::cr_f_bits;

static uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static bool same(float a, float b) {
  if (std::isnan(a) && std::isnan(b)) return true;
  return cr_f_bits(a) == cr_f_bits(b);
}

int main() {
  long long fail = 0;

  // 1. one Newton step on the reciprocal. Approximations within 1 ulp of 1/b (what
  //    v_rcp_f32 returns) must refine to RN(1/b), with one exception that is known and
  //    recorded: for the all-ones significand, 1/b lies just above a midpoint and its lower
  //    neighbour refines to itself (a tie broken to even). What the hardware returns there is
  //    pinned on the GPU (every float32 b through the device path).
  {
    long long bad = 0, checked = 0, outside = 0, allones_lower = 0;
#pragma omp parallel for reduction(+ : bad, checked, outside, allones_lower) schedule(static)
    for (int64_t m = 0; m < (1 << 23); ++m) {
      const float b = cr_bits_f(0x3F800000u | (uint32_t)m);
      const float t = 1.0f / b;
      const double exact = 1.0 / (double)b;
      const double ulp = std::ldexp(1.0, std::ilogb((double)t) - 23);
      for (int d = -1; d <= 1; ++d) {
        const float y0 = cr_bits_f(cr_f_bits(t) + (uint32_t)d);
        checked++;
        if (same(gs::cr_recip_refine(b, y0), t)) continue;
        if (std::fabs((double)y0 - exact) >= ulp) outside++;          // not a 1-ulp approximation
        else if (m == 0x7FFFFF && d == -1) allones_lower++;           // the recorded exception
        else bad++;
      }
    }
    printf("recip  %lld approximations of 1/b over every significand: %lld wrong within 1 ulp "
           "(%lld beyond 1 ulp, %lld the all-ones significand's lower neighbour)\n",
           checked, bad, outside, allones_lower);
    fail += bad + (allones_lower != 1);
  }

  // 2. Markstein's division correction from y = RN(1/b)
  {
    float divisors[256];
    divisors[0] = 1.0f;
    divisors[1] = cr_bits_f(0x3FFFFFFFu);   // significand all ones
    uint64_t s = 12345;
    for (int k = 2; k < 256; ++k) divisors[k] = cr_bits_f(0x3F800000u | (uint32_t)(splitmix(s) & 0x7FFFFF));
    long long bad = 0, checked = 0;
    for (int k = 0; k < 256; ++k) {
      const float b = divisors[k];
      const float y = 1.0f / b;
#pragma omp parallel for reduction(+ : bad, checked) schedule(static)
      for (int64_t m = 0; m < (1 << 23); ++m) {
        const float a = cr_bits_f(0x3F800000u | (uint32_t)m);
        checked++;
        if (!same(gs::cr_div_refine(a, b, y), a / b)) bad++;
      }
    }
    printf("div    %lld quotients (256 divisors x every numerator significand): %lld wrong\n", checked, bad);
    fail += bad;
  }

  // 3. square-root neighbour selection
  {
    long long bad = 0, checked = 0;
    auto run = [&](uint32_t lo, uint32_t hi) {
      long long b_ = 0, c_ = 0;
#pragma omp parallel for reduction(+ : b_, c_) schedule(static)
      for (int64_t u = lo; u < (int64_t)hi; ++u) {
        const float x = cr_bits_f((uint32_t)u);
        const float t = sqrtf(x);
        for (int d = -1; d <= 1; ++d) {
          const float s0 = cr_bits_f(cr_f_bits(t) + (uint32_t)d);
          c_++;
          if (!same(gs::cr_sqrt_refine(x, s0), t)) b_++;
        }
      }
      bad += b_;
      checked += c_;
    };
    run(cr_f_bits(1.0f), cr_f_bits(4.0f));                 // [1, 4)
    run(cr_f_bits(0x1p-96f), cr_f_bits(0x1p-90f));         // above the threshold
    run(cr_f_bits(0x1p124f), 0x7F800000u);                 // the top binades
    // special values: the approximation is exact there (v_sqrt_f32 returns them as they are)
    const float sp[] = {0.0f, -0.0f, INFINITY, NAN, -1.0f, -INFINITY};
    for (float x : sp) {
      checked++;
      if (!same(gs::cr_sqrt_refine(x, sqrtf(x)), sqrtf(x))) bad++;
    }
    printf("sqrt   %lld approximations: %lld wrong\n", checked, bad);
    fail += bad;
  }

  printf("failures %lld\n", fail);
  return fail ? 1 : 0;
}
