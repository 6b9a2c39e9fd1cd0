// gs_common.h -- pieces shared by the layout kernels: error reporting, the device-side
// step scheduler (the reference's host dt arithmetic moved onto the GPU) and the
// last-workgroup grid reduction of the CFL wave speed.
#pragma once

#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "gs_device.h"

namespace gs {

// ---------------------------------------------------------------------------
// error reporting
// ---------------------------------------------------------------------------
extern thread_local char g_err[512];

static inline int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

static inline int hip_fail(hipError_t e, const char* what) {
  snprintf(g_err, sizeof g_err, "%s: %s", what, hipGetErrorString(e));
  return -(int)e;
}

#define GS_HIP(call, what)                       \
  do {                                           \
    hipError_t _e = (call);                      \
    if (_e != hipSuccess) return hip_fail(_e, what); \
  } while (0)

// ---------------------------------------------------------------------------
// device-side scheduling (Solver2D.run / cfl_dt host arithmetic, on the device)
// ---------------------------------------------------------------------------
struct Plan {
  double cfl, dmin, lam_floor;
  int n_rain, n_stage;
  const double* rain_t;
  const double* rain_rate;
  const double* rain_cap;
  const double* stage_t;
  const double* stage_v;
};

static inline Plan make_plan(const GsPhys& p, const GsForcing& f) {
  Plan pl;
  pl.cfl = p.cfl;
  pl.dmin = p.dmin;
  pl.lam_floor = p.lam_floor;
  pl.n_rain = f.n_rain;
  pl.n_stage = f.n_stage;
  pl.rain_t = f.rain_t;
  pl.rain_rate = f.rain_rate;
  pl.rain_cap = f.rain_cap;
  pl.stage_t = f.stage_t;
  pl.stage_v = f.stage_v;
  return pl;
}

// bisect_right(a, x) - 1, clamped to [0, n-1] (forcing.py:198-199)
__device__ static int seg_index(const double* a, int n, double x) {
  int lo = 0, hi = n;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (x < a[mid]) hi = mid; else lo = mid + 1;
  }
  int idx = lo - 1;
  if (idx < 0) idx = 0;
  if (idx > n - 1) idx = n - 1;
  return idx;
}

// StageBoundary.stage_at_time (forcing.py:315-347)
__device__ static double stage_at(const double* ts, const double* sv, int n, double t) {
  if (t <= ts[0]) return sv[0];
  if (t >= ts[n - 1]) return sv[n - 1];
  int lo = 0, hi = n;  // bisect_left
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (ts[mid] < t) lo = mid + 1; else hi = mid;
  }
  const int i = lo;
  const double t0 = ts[i - 1], t1 = ts[i];
  const double s0 = sv[i - 1], s1 = sv[i];
  const double w = (t - t0) / (t1 - t0);
  return s0 + w * (s1 - s0);
}

// Forcing values of the step that starts at c->t with step c->dt.
__device__ static void plan_forcing(GsClock* c, const Plan& p) {
  int ridx = 0;
  float r32 = 0.0f;
  if (p.n_rain > 0) {
    ridx = seg_index(p.rain_t, p.n_rain, c->t);
    r32 = (float)p.rain_rate[ridx];
  }
  c->rain_idx = ridx;
  c->rain32 = r32;
  if (p.n_stage > 0) {
    const double e = stage_at(p.stage_t, p.stage_v, p.n_stage, c->t + c->dt);
    c->eta64 = e;
    c->eta32 = (float)e;
  }
}

// dt from the wave speed of the current state (solver.py:2657-2666) and, in run mode,
// the rain-film cap, dt_max and the clip to t_end (solver.py:2990-2999, 3005-3027).
__device__ static void plan_dt(GsClock* c, const Plan& p) {
  int st = c->status & ~GS_ST_DONE;
  double lam = __longlong_as_double((long long)c->lam_bits);
  const double t = c->t;
  if (!(lam == lam) || isinf(lam)) {
    c->status = st | GS_ST_NONFINITE | GS_ST_DONE;
    return;
  }
  if (lam < p.lam_floor) lam = p.lam_floor;
  double dt = p.cfl * p.dmin / (lam + 1.0e-12);
  if (c->run_mode) {
    if (p.n_rain > 0) {
      const double cap = p.rain_cap[seg_index(p.rain_t, p.n_rain, t)];
      if (cap < dt) dt = cap;
    }
    if (c->dt_max > 0.0 && dt > c->dt_max) dt = c->dt_max;
    if (t + dt > c->t_end) dt = c->t_end - t;
    if (!(t < c->t_end - 1.0e-12)) st |= GS_ST_DONE;
  }
  if (c->nsteps >= c->max_steps) st |= GS_ST_DONE;
  c->dt = dt;
  c->dt32 = (float)dt;
  c->use_host_dt = 0;
  c->status = st;
  plan_forcing(c, p);
}

// Called by the last workgroup of a step (advance) or of a CFL pass.
__device__ static void plan_finish(GsClock* c, const Plan& p, uint64_t lam, bool advance) {
  if (advance) {
    c->t = c->t + c->dt;
    c->nsteps = c->nsteps + 1;
  }
  c->lam_bits = lam;
  c->lam_acc = 0ull;
  c->blocks_done = 0u;
  plan_dt(c, p);
}

// Neighbour-lane values for the y-faces: lane l receives lane l+1's value (lane_upper)
// or lane l-1's (lane_lower); the end lanes keep their own. The shuffle runs on the LDS
// crossbar, beside the VALU work of the faces: DPP wavefront shifts (one VALU move per 32
// bits plus hazard waits) made the VALU-bound sweeps 1-6 % slower on MI355X.
template <typename T> __device__ __forceinline__ T lane_upper(T x) { return __shfl_down(x, 1, 64); }
template <typename T> __device__ __forceinline__ T lane_lower(T x) { return __shfl_up(x, 1, 64); }

__device__ __forceinline__ uint64_t wave_max_u64(uint64_t v) {
  for (int off = 32; off > 0; off >>= 1) {
    const uint64_t o = __shfl_xor(v, off, 64);
    v = o > v ? o : v;
  }
  return v;
}

// Use an atomic's returned value: the wave waits until the atomic has been performed, and
// the compiler moves no memory access across this point.
__device__ __forceinline__ void await_atomic(uint64_t v) { __asm__ volatile("" ::"v"(v) : "memory"); }

// Block reduction of a per-thread wave-speed (IEEE bits, order independent) and the
// last-workgroup-done hand-off of the grid maximum to plan_finish. Only atomics cross
// workgroups here (the state a workgroup stores is read by the next launch), so the
// hand-off orders atomics and nothing else: thread 0 folds the workgroup's maximum into
// the clock and waits until that atomic has been performed before it counts the
// workgroup done, so the count that reveals the last workgroup follows every maximum, and
// the last workgroup reads the total with a device-scope atomic load. A release on the
// count would also write back the XCD's L2 (its dirty state lines) once per workgroup,
// which cost 10 % of the coastal step at one workgroup per four tiles.
__device__ static void grid_reduce_and_plan(uint64_t lam, GsClock* c, const Plan& p, bool advance) {
  __shared__ uint64_t s_w[16];
  __shared__ int s_last;
  lam = wave_max_u64(lam);
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) s_w[wave] = lam;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t m = 0ull;
    const int nw = (blockDim.x + 63) >> 6;
    for (int w = 0; w < nw; ++w) m = s_w[w] > m ? s_w[w] : m;
    if (m) await_atomic(__hip_atomic_fetch_max(&c->lam_acc, m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    const unsigned int prev =
        __hip_atomic_fetch_add(&c->blocks_done, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s_last = (prev == gridDim.x - 1) ? 1 : 0;
  }
  __syncthreads();
  if (s_last && threadIdx.x == 0) {
    const uint64_t total = __hip_atomic_load(&c->lam_acc, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    plan_finish(c, p, total, advance);
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
template <typename T>
static inline Phys<T> make_phys(const GsPhys& p, double dx, double dy) {
  Phys<T> P;
  P.g = (T)p.g;
  P.h_min = (T)p.h_min;
  P.h_cfl = (T)p.h_cfl;
  P.vcap = (T)p.vcap;
  P.has_vcap = isfinite(p.vcap) && p.vcap > 0.0;
  P.inv_dx = (T)(1.0 / dx);
  P.inv_dy = (T)(1.0 / dy);
  // the reference kernel recomputes dx = 1/inv_dx in T (rhs_cuda.py:928-929)
  P.half_dx = T(0.5) * (T(1) / P.inv_dx);
  P.half_dy = T(0.5) * (T(1) / P.inv_dy);
  P.friction = p.friction;
  P.quadratic = p.quadratic;
  P.keep_h = p.keep_h;
  P.cfl_l2 = p.cfl_l2;
  return P;
}

static inline int check_phys(const GsPhys* p) {
  if (!p) return fail(GS_E_ARG, "phys is NULL");
  if (!(p->g > 0.0) || !(p->h_min > 0.0) || !(p->cfl > 0.0) || !(p->dmin > 0.0))
    return fail(GS_E_ARG, "phys: g, h_min, cfl and dmin must be positive");
  if (p->friction && (p->n_classes < 1 || p->n_classes > GS_MAX_CLASSES))
    return fail(GS_E_ARG, "phys: n_classes must be in [1, %d] with friction on (got %d)", GS_MAX_CLASSES,
                p->n_classes);
  return GS_OK;
}

static inline int check_forcing(const GsForcing& f) {
  if (f.n_rain < 0 || f.n_stage < 0) return fail(GS_E_ARG, "forcing: negative table size");
  if (f.n_rain > 0 && (!f.rain_t || !f.rain_rate || !f.rain_cap))
    return fail(GS_E_ARG, "forcing: rain tables are NULL");
  if (f.rain_gridded && (f.n_rain < 1 || !f.rain_frames))
    return fail(GS_E_ARG, "forcing: gridded rain needs frames");
  if (f.n_stage > 0 && (!f.stage_t || !f.stage_v)) return fail(GS_E_ARG, "forcing: stage tables are NULL");
  return GS_OK;
}


}  // namespace gs
