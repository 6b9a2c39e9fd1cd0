// gs_kernels.hip -- MI355X (gfx950) kernels and C ABI of the fused finite-volume step.
//
// Dense layout: one wavefront owns a strip of 64 contiguous columns (60 written, two
// halo lanes on each side) and sweeps it down `tile_rows` rows. Each x-face is computed
// once and carried in registers from one row to the next; each y-face is computed once
// by the lane left of it and handed to the lane on its right with a cross-lane shuffle.
// Ghost cells are never stored: the stencil loads synthesise them from the boundary
// kinds. The residual, rain, Euler update, friction, wet/dry rule, stage boundary,
// running max and the next step's CFL wave speed are one pass; the last workgroup turns
// the wave-speed maximum into the next dt on the device (no host synchronisation).
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "gs_common.h"

namespace gs {

thread_local char g_err[512] = "";

// clock set-up kernels (one thread)
__global__ void k_clock_reset(GsClock* c, double t0) {
  c->t = t0;
  c->dt = 0.0;
  c->t_end = INFINITY;
  c->dt_max = 0.0;
  c->eta64 = 0.0;
  c->nsteps = 0;
  c->max_steps = INT64_MAX;
  c->lam_acc = 0ull;
  c->lam_bits = 0ull;
  c->blocks_done = 0u;
  c->status = 0;
  c->dt32 = 0.0f;
  c->rain32 = 0.0f;
  c->eta32 = 0.0f;
  c->rain_idx = 0;
  c->use_host_dt = 0;
  c->run_mode = 0;
  // the padding too, so that a clock's bytes are a function of its history alone
  for (int i = 0; i < 6; ++i) c->pad_[i] = 0;
}

__global__ void k_clock_target(GsClock* c, Plan p, double t_end, double dt_max, long long max_steps,
                               int run_mode) {
  c->t_end = t_end;
  c->dt_max = dt_max;
  c->max_steps = max_steps < 0 ? INT64_MAX : c->nsteps + max_steps;
  c->run_mode = run_mode;
  c->status &= ~GS_ST_DONE;
  plan_dt(c, p);
}

__global__ void k_clock_dt(GsClock* c, Plan p, double dt) {
  c->dt = dt;
  c->dt32 = (float)dt;
  c->use_host_dt = 1;
  c->status &= ~GS_ST_DONE;
  if (c->status & GS_ST_NONFINITE) c->status |= GS_ST_DONE;
  if (c->nsteps >= c->max_steps) c->status |= GS_ST_DONE;
  plan_forcing(c, p);
}


// ---------------------------------------------------------------------------
// dense layout
// ---------------------------------------------------------------------------
constexpr int kOut = 60;      // columns written per strip (lanes 2..61)
constexpr int kBlock = 256;   // 4 strips per workgroup

// T is the arithmetic type, S the type of the planes in memory (S = float with T = double:
// float32 planes stepped with float64 arithmetic, each value rounded once when it is stored).
template <typename T, typename S = T>
struct DenseK {
  int nx, ny, bc_x, bc_y, tile_rows, ncs, nrs, ntiles;
  int64_t pitch;
  S* buf[2][3];
  const S* bed;
  const uint8_t* flags;
  const S* ntab;
  int n_classes;
  const S* n_field;
  S* max_h;
  const uint8_t* cls;   // per-cell class plane (ids 0..255), or NULL: the flags' class bits
  T* rhs[3];
  int which;
  Phys<T> P;
  int rain_mode;  // 0 none, 1 uniform, 2 gridded
  const S* rain_frames;
  int64_t frame_stride, rain_pitch;
  int stage;
  GsClock* clock;
  Plan plan;
};

// Ghost synthesis (bc.py:50-98 state, solver.py:1466-1484 bed). Rows/columns outside the
// interior map to the interior cell the boundary kind copies; `zero` for 'fall';
// `flip` negates the normal momentum for 'wall'.
__device__ __forceinline__ int map_state(int i, int n, int bc, bool& zero, bool& flip) {
  if (i >= 0 && i < n) return i;
  switch (bc) {
    case GS_BC_PERIODIC: return i < 0 ? i + n : i - n;
    case GS_BC_WALL: flip = true; return i < 0 ? -1 - i : 2 * n - 1 - i;
    case GS_BC_FALL: zero = true; return i < 0 ? 0 : n - 1;
    default: return i < 0 ? 0 : n - 1;
  }
}
__device__ __forceinline__ int map_bed(int i, int n, int bc) {
  if (i >= 0 && i < n) return i;
  if (bc == GS_BC_PERIODIC) return i < 0 ? i + n : i - n;
  return i < 0 ? 0 : n - 1;
}

// Column part of the ghost map, fixed for a lane over the whole sweep.
struct ColMap {
  int mj, mjb;     // state column, bed column
  bool zero, flip;
};
__device__ __forceinline__ ColMap col_map(int c, int ny, int bc) {
  ColMap m;
  m.zero = false;
  m.flip = false;
  m.mj = map_state(c, ny, bc, m.zero, m.flip);
  m.mjb = map_bed(c, ny, bc);
  return m;
}

// h, hu, hv of cell (i, column map cm); the row map is uniform across the wavefront.
template <typename T, typename S>
__device__ __forceinline__ void load_q(const DenseK<T, S>& a, S* const* q, int i, const ColMap& cm, T& h, T& hu,
                                       T& hv) {
  bool zx = false, fx = false;
  const int mi = map_state(i, a.nx, a.bc_x, zx, fx);
  const int64_t idx = (int64_t)mi * a.pitch + cm.mj;
  h = (T)q[0][idx];
  hu = (T)q[1][idx];
  hv = (T)q[2][idx];
  if (fx) hu = -hu;
  if (cm.flip) hv = -hv;
  if (zx || cm.zero) { h = T(0); hu = T(0); hv = T(0); }
}

template <typename T, typename S>
__device__ __forceinline__ T load_b(const DenseK<T, S>& a, int i, const ColMap& cm) {
  const int mi = map_bed(i, a.nx, a.bc_x);
  return (T)a.bed[(int64_t)mi * a.pitch + cm.mjb];
}


// MODE 0: residual into a.rhs; MODE 1: fused step.
template <typename T, int MODE, typename S = T>
__global__ __launch_bounds__(kBlock) void k_dense_sweep(DenseK<T, S> a) {
  int par = 0;
  T dt = T(0), rain_u = T(0), eta = T(0);
  int ridx = 0;
  if (MODE == 1) {
    const GsClock* c = a.clock;
    if (c->status & GS_ST_DONE) return;  // the whole grid sees the same clock
    par = (int)(c->nsteps & 1);
    ridx = c->rain_idx;
    if (sizeof(T) == 4) {
      dt = (T)c->dt32;
      rain_u = (T)c->rain32;
      eta = (T)c->eta32;
    } else {
      dt = (T)c->dt;
      rain_u = a.rain_mode == 1 ? (T)a.plan.rain_rate[ridx] : T(0);
      eta = (T)c->eta64;
    }
  }
  __shared__ T s_ntab[GS_MAX_CLASSES];
  if (MODE == 1)
    for (int k = (int)threadIdx.x; k < a.n_classes; k += (int)blockDim.x) s_ntab[k] = (T)a.ntab[k];
  __syncthreads();

  S* const* qin = MODE == 1 ? a.buf[par] : a.buf[a.which];
  S* const* qout = a.buf[par ^ 1];
  const Phys<T>& P = a.P;
  const T ix = P.inv_dx, iy = P.inv_dy, hdx = P.half_dx, hdy = P.half_dy, g = P.g, hm = P.h_min;
  const S* frame = (MODE == 1 && a.rain_mode == 2) ? a.rain_frames + (int64_t)ridx * a.frame_stride : nullptr;

  // XCD-aware order: workgroups are dealt round-robin to the 8 XCDs; give each XCD a
  // contiguous range of tiles so neighbouring strips share that XCD's L2.
  const unsigned nb = gridDim.x, b = blockIdx.x;
  const unsigned qd = nb / 8u, rm = nb % 8u, xcd = b % 8u, kk = b / 8u;
  const unsigned lb = (xcd < rm) ? xcd * (qd + 1u) + kk : rm * (qd + 1u) + (xcd - rm) * qd + kk;
  const int lane = threadIdx.x & 63;
  // the wavefront index as a scalar: tile, rows and row addresses are then uniform values
  const int tile = (int)(lb * 4u) + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  lam_t<T> lam = 0;

  if (tile < a.ntiles) {
    const int rs = tile / a.ncs, cs = tile - rs * a.ncs;
    const int c = cs * kOut + lane - 2;
    const int cl = c < -2 ? -2 : (c > a.ny + 1 ? a.ny + 1 : c);   // loads stay within one ghost pair
    const int i0 = rs * a.tile_rows;
    const int i_end = min(i0 + a.tile_rows, a.nx);
    const bool col_ok = (lane >= 2) && (lane < 2 + kOut) && (c < a.ny);

    const ColMap cm = col_map(cl, a.ny, a.bc_y);
    T bm2 = load_b(a, i0 - 2, cm), bm1 = load_b(a, i0 - 1, cm), b0 = load_b(a, i0, cm),
      bp1 = load_b(a, i0 + 1, cm);
    T hP, huP, hvP, uP, vP;
    load_q(a, qin, i0 - 1, cm, hP, huP, hvP);
    prims(hP, huP, hvP, hm, uP, vP);
    const T gzxP = (T(0.5) * (b0 - bm2)) * ix;
    T hC, huC, hvC, uC, vC;
    load_q(a, qin, i0, cm, hC, huC, hvC);
    prims(hC, huC, hvC, hm, uC, vC);
    T gzxC = (T(0.5) * (bp1 - bm1)) * ix;
    Face<T> fx;
    face<T>(bm1, b0, gzxP, gzxC, hP, hC, uP, vP, uC, vC, hdx, g, hm, fx);
    // the x-face below the current row, as the products the current cell adds (zero when
    // the face is gated off: see the accumulation below)
    const bool xg = fx.gate;
    T xFh = xg ? fx.Fh * ix : T(0), xFN = xg ? fx.FN * ix : T(0), xFT = xg ? fx.FT * ix : T(0),
      xsR = xg ? fx.srcR * ix : T(0);

    for (int i = i0; i < i_end; ++i) {
      const T bp2 = load_b(a, i + 2, cm);
      T hN, huN, hvN, uN, vN;
      load_q(a, qin, i + 1, cm, hN, huN, hvN);
      prims(hN, huN, hvN, hm, uN, vN);
      const T gzxN = (T(0.5) * (bp2 - b0)) * ix;
      Face<T> fn;
      face<T>(b0, bp1, gzxC, gzxN, hC, hN, uC, vC, uN, vN, hdx, g, hm, fn);
      const bool ng = fn.gate;
      const T nFh = ng ? fn.Fh * ix : T(0), nFN = ng ? fn.FN * ix : T(0), nFT = ng ? fn.FT * ix : T(0);
      const T nsL = ng ? fn.srcL * ix : T(0), nsR = ng ? fn.srcR * ix : T(0);

      // y faces of row i: lane computes its upper face (c, c+1)
      const T bl = lane_lower(b0), br = lane_upper(b0);
      const T gzy = (T(0.5) * (br - bl)) * iy;
      const T hR = lane_upper(hC), uR = lane_upper(uC), vR = lane_upper(vC);
      const T gzyR = lane_upper(gzy);
      Face<T> fy;
      face<T>(b0, br, gzy, gzyR, hC, hR, vC, uC, vR, uR, hdy, g, hm, fy);
      const bool yg = fy.gate;
      const T yFh = yg ? fy.Fh * iy : T(0), yFN = yg ? fy.FN * iy : T(0), yFT = yg ? fy.FT * iy : T(0);
      const T ysL = yg ? fy.srcL * iy : T(0);
      const T yFhB = lane_lower(yFh), yFNB = lane_lower(yFN), yFTB = lane_lower(yFT);
      const T ysRB = lane_lower(yg ? fy.srcR * iy : T(0));

      // The four faces in the reference's order. A face gated off (both cells dry) adds
      // zeros: the sums start at +0 and IEEE addition yields -0 only from two -0 terms, so
      // they never hold -0, and x + 0 = x - 0 = x for every other x; the gate costs one
      // select per face term instead of one per use.
      T rh = T(0), rhu = T(0), rhv = T(0);
      rh = rh - nFh;
      rhu = rhu - nFN;
      rhv = rhv - nFT;
      rhu = rhu - nsL;
      rh = rh + xFh;
      rhu = rhu + xFN;
      rhv = rhv + xFT;
      rhu = rhu + xsR;
      rh = rh - yFh;
      rhu = rhu - yFT;
      rhv = rhv - yFN;
      rhv = rhv - ysL;
      rh = rh + yFhB;
      rhu = rhu + yFTB;
      rhv = rhv + yFNB;
      rhv = rhv + ysRB;

      if (col_ok) {
        const int64_t idx = (int64_t)i * a.pitch + c;
        const uint8_t fl = a.flags ? a.flags[idx] : (uint8_t)0;
        const bool inactive = (fl & GS_FLAG_INACTIVE) != 0;
        if (inactive) { rh = T(0); rhu = T(0); rhv = T(0); }
        if (MODE == 0) {
          a.rhs[0][idx] = rh;
          a.rhs[1][idx] = rhu;
          a.rhs[2][idx] = rhv;
        } else {
          T rr0 = rh;
          if (!inactive) {
            if (a.rain_mode == 1) rr0 = rr0 + rain_u;
            else if (a.rain_mode == 2) rr0 = rr0 + (T)frame[(int64_t)i * a.rain_pitch + c];
          }
          T n = T(0);
          if (P.friction)
            n = a.n_field ? (T)a.n_field[idx] : s_ntab[a.cls ? a.cls[idx] : (fl & GS_FLAG_CLASS_MASK)];
          T h, hu, hv;
          update<T>(hC, huC, hvC, rr0, rhu, rhv, dt, n, P, h, hu, hv);
          if (fl & GS_FLAG_STAGE) {
            const T hs = eta - b0;
            h = hs > T(0) ? hs : T(0);
            hu = T(0);
            hv = T(0);
          }
          // the stored values (one rounding each when S is narrower than T) are the state the
          // running max and the next step's wave speed see
          const S hS = (S)h, huS = (S)hu, hvS = (S)hv;
          if (a.max_h && hS > a.max_h[idx]) a.max_h[idx] = hS;
          qout[0][idx] = hS;
          qout[1][idx] = huS;
          qout[2][idx] = hvS;
          if (!(fl & GS_FLAG_CFL_SKIP)) {
            const lam_t<T> lb_ = lam_key<T>((T)hS, (T)huS, (T)hvS, P);
            lam = lb_ > lam ? lb_ : lam;
          }
        }
      }

      // slide the window one row down
      hC = hN; huC = huN; hvC = hvN; uC = uN; vC = vN; gzxC = gzxN;
      bm1 = b0; b0 = bp1; bp1 = bp2;
      xFh = nFh; xFN = nFN; xFT = nFT; xsR = nsR;
    }
  }
  if (MODE == 1) grid_reduce_and_plan(lam_key_bits64(lam), a.clock, a.plan, true);
}

// CFL wave speed of the current dense state, then plan dt (no time advance).
template <typename T, typename S = T>
__global__ __launch_bounds__(kBlock) void k_dense_cfl(DenseK<T, S> a) {
  const int par = (int)(a.clock->nsteps & 1);
  S* const* q = a.buf[par];
  lam_t<T> lam = 0;
  const int64_t n = (int64_t)a.nx * a.ny;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = k / a.ny, j = k - i * a.ny;
    const int64_t idx = i * a.pitch + j;
    if (a.flags && (a.flags[idx] & GS_FLAG_CFL_SKIP)) continue;
    const lam_t<T> l = lam_key<T>((T)q[0][idx], (T)q[1][idx], (T)q[2][idx], a.P);
    lam = l > lam ? l : lam;
  }
  grid_reduce_and_plan(lam_key_bits64(lam), a.clock, a.plan, false);
}

// Rows per strip: 64 when that still gives 8192 strips, else fewer. On the 640 M-cell grid
// 64-row strips (167 K wavefronts) ran the step in 7.93 ms against 8.11 ms for 128 rows
// and 8.30 ms for 256 (the extra halo rows of shorter strips cost less than the better
// balance of more wavefronts gains); 32 rows took 8.00 ms.
static int pick_tile_rows(int nx, int ncs, int req) {
  if (req > 0) return req;
  const int cand[3] = {64, 32, 16};
  for (int k = 0; k < 3; ++k) {
    const int64_t nrs = (nx + cand[k] - 1) / cand[k];
    if ((int64_t)ncs * nrs >= 8192) return cand[k];
  }
  return 16;
}

template <typename T, typename S = T>
static int build_dense(const GsDense* d, const GsPhys* p, const GsForcing& f, GsClock* clk, DenseK<T, S>& a,
                       bool need_clock) {
  if (!d) return fail(GS_E_ARG, "dense descriptor is NULL");
  if (d->nx < 4 || d->ny < 4) return fail(GS_E_RANGE, "dense grid needs nx, ny >= 4 (got %d x %d)", d->nx, d->ny);
  if (d->pitch < d->ny) return fail(GS_E_ARG, "pitch %lld < ny %d", (long long)d->pitch, d->ny);
  if ((int64_t)d->nx * d->pitch > ((int64_t)1 << 40)) return fail(GS_E_RANGE, "grid too large");
  if (d->bc_x < 0 || d->bc_x > 3 || d->bc_y < 0 || d->bc_y > 3) return fail(GS_E_ARG, "bad boundary kind");
  if (!d->bed) return fail(GS_E_ARG, "bed is NULL");
  for (int s = 0; s < 2; ++s)
    for (int k = 0; k < 3; ++k)
      if (!d->q[s][k]) return fail(GS_E_ARG, "state plane q[%d][%d] is NULL", s, k);
  int rc = check_phys(p);
  if (rc) return rc;
  rc = check_forcing(f);
  if (rc) return rc;
  if (p->friction && !d->ntab && !d->n_field) return fail(GS_E_ARG, "friction needs ntab or n_field");
  if (need_clock && !clk) return fail(GS_E_ARG, "clock is NULL");
  memset(&a, 0, sizeof a);
  a.nx = d->nx;
  a.ny = d->ny;
  a.bc_x = d->bc_x;
  a.bc_y = d->bc_y;
  a.pitch = d->pitch;
  a.ncs = (d->ny + kOut - 1) / kOut;
  a.tile_rows = pick_tile_rows(d->nx, a.ncs, d->tile_rows);
  a.nrs = (d->nx + a.tile_rows - 1) / a.tile_rows;
  if ((int64_t)a.ncs * a.nrs > (int64_t)INT32_MAX / 2) return fail(GS_E_RANGE, "too many tiles");
  a.ntiles = a.ncs * a.nrs;
  for (int s = 0; s < 2; ++s)
    for (int k = 0; k < 3; ++k) a.buf[s][k] = (S*)d->q[s][k];
  a.bed = (const S*)d->bed;
  a.flags = d->flags;
  a.ntab = (const S*)d->ntab;
  a.n_classes = d->ntab ? (p->n_classes > 0 ? p->n_classes : 1) : 0;
  a.n_field = (const S*)d->n_field;
  a.cls = d->cls;
  a.max_h = (S*)d->max_h;
  a.P = make_phys<T>(*p, d->dx, d->dy);
  a.rain_mode = f.n_rain > 0 ? (f.rain_gridded ? 2 : 1) : 0;
  a.rain_frames = (const S*)f.rain_frames;
  a.frame_stride = f.rain_frame_stride;
  a.rain_pitch = f.rain_row_pitch;
  a.stage = f.n_stage > 0;
  a.clock = clk;
  a.plan = make_plan(*p, f);
  return GS_OK;
}

template <typename T, typename S = T>
static int dense_step_t(const GsDense* d, const GsPhys* p, GsForcing f, GsClock* clk, hipStream_t s) {
  DenseK<T, S> a;
  int rc = build_dense<T, S>(d, p, f, clk, a, true);
  if (rc) return rc;
  const unsigned blocks = (unsigned)((a.ntiles + 3) / 4);
  hipLaunchKernelGGL((k_dense_sweep<T, 1, S>), dim3(blocks), dim3(kBlock), 0, s, a);
  GS_HIP(hipGetLastError(), "k_dense_sweep launch");
  return GS_OK;
}

template <typename T, typename S = T>
static int dense_residual_t(const GsDense* d, const GsPhys* p, int which, void* r0, void* r1, void* r2,
                            hipStream_t s) {
  GsForcing f;
  memset(&f, 0, sizeof f);
  DenseK<T, S> a;
  int rc = build_dense<T, S>(d, p, f, nullptr, a, false);
  if (rc) return rc;
  if (which < 0 || which > 1) return fail(GS_E_ARG, "which must be 0 or 1");
  if (!r0 || !r1 || !r2) return fail(GS_E_ARG, "residual planes are NULL");
  a.which = which;
  a.rhs[0] = (T*)r0;
  a.rhs[1] = (T*)r1;
  a.rhs[2] = (T*)r2;
  // MODE 0 writes qout = buf[which ^ 1] never; buf[par ^ 1] is unused there
  const unsigned blocks = (unsigned)((a.ntiles + 3) / 4);
  hipLaunchKernelGGL((k_dense_sweep<T, 0, S>), dim3(blocks), dim3(kBlock), 0, s, a);
  GS_HIP(hipGetLastError(), "k_dense_sweep(residual) launch");
  return GS_OK;
}

template <typename T, typename S = T>
static int dense_cfl_t(const GsDense* d, const GsPhys* p, GsForcing f, GsClock* clk, hipStream_t s) {
  DenseK<T, S> a;
  int rc = build_dense<T, S>(d, p, f, clk, a, true);
  if (rc) return rc;
  const int64_t n = (int64_t)a.nx * a.ny;
  int64_t blocks = (n + kBlock - 1) / kBlock;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL((k_dense_cfl<T, S>), dim3((unsigned)blocks), dim3(kBlock), 0, s, a);
  GS_HIP(hipGetLastError(), "k_dense_cfl launch");
  return GS_OK;
}

}  // namespace gs

using namespace gs;

extern "C" {

int gs_version(void) { return GS_ABI_VERSION; }
const char* gs_last_error(void) { return g_err; }

int gs_get_device_count(int* n) {
  if (!n) return fail(GS_E_ARG, "n is NULL");
  GS_HIP(hipGetDeviceCount(n), "hipGetDeviceCount");
  return GS_OK;
}
int gs_set_device(int dev) {
  GS_HIP(hipSetDevice(dev), "hipSetDevice");
  return GS_OK;
}

int gs_malloc(void** ptr, size_t bytes) {
  if (!ptr) return fail(GS_E_ARG, "ptr is NULL");
  *ptr = nullptr;
  if (bytes == 0) bytes = 256;
  GS_HIP(hipMalloc(ptr, bytes), "hipMalloc");
  return GS_OK;
}
int gs_free(void* ptr) {
  if (ptr) GS_HIP(hipFree(ptr), "hipFree");
  return GS_OK;
}
int gs_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream) {
  if (!bytes) return GS_OK;
  GS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream), "hipMemcpyAsync h2d");
  return GS_OK;
}
int gs_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream) {
  if (!bytes) return GS_OK;
  GS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream), "hipMemcpyAsync d2h");
  return GS_OK;
}
int gs_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream) {
  if (!bytes) return GS_OK;
  GS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream), "hipMemcpyAsync d2d");
  return GS_OK;
}
int gs_memcpy2d(void* dst, size_t dpitch, const void* src, size_t spitch, size_t width, size_t height, int kind,
                void* stream) {
  hipMemcpyKind k = kind == 1 ? hipMemcpyHostToDevice : (kind == 2 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
  if (kind < 1 || kind > 3) return fail(GS_E_ARG, "memcpy2d kind must be 1, 2 or 3");
  if (!width || !height) return GS_OK;
  GS_HIP(hipMemcpy2DAsync(dst, dpitch, src, spitch, width, height, k, (hipStream_t)stream), "hipMemcpy2DAsync");
  return GS_OK;
}
int gs_memset(void* dst, int value, size_t bytes, void* stream) {
  if (!bytes) return GS_OK;
  GS_HIP(hipMemsetAsync(dst, value, bytes, (hipStream_t)stream), "hipMemsetAsync");
  return GS_OK;
}
int gs_stream_create(void** stream) {
  if (!stream) return fail(GS_E_ARG, "stream is NULL");
  hipStream_t s;
  GS_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
  *stream = (void*)s;
  return GS_OK;
}
int gs_stream_destroy(void* stream) {
  if (stream) GS_HIP(hipStreamDestroy((hipStream_t)stream), "hipStreamDestroy");
  return GS_OK;
}
int gs_stream_synchronize(void* stream) {
  GS_HIP(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
  return GS_OK;
}
int gs_event_create(void** ev) {
  if (!ev) return fail(GS_E_ARG, "ev is NULL");
  hipEvent_t e;
  GS_HIP(hipEventCreate(&e), "hipEventCreate");
  *ev = (void*)e;
  return GS_OK;
}
int gs_event_destroy(void* ev) {
  if (ev) GS_HIP(hipEventDestroy((hipEvent_t)ev), "hipEventDestroy");
  return GS_OK;
}
int gs_event_record(void* ev, void* stream) {
  GS_HIP(hipEventRecord((hipEvent_t)ev, (hipStream_t)stream), "hipEventRecord");
  return GS_OK;
}
int gs_event_elapsed_ms(float* ms, void* start, void* stop) {
  if (!ms) return fail(GS_E_ARG, "ms is NULL");
  GS_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop), "hipEventElapsedTime");
  return GS_OK;
}

int gs_clock_reset(GsClock* clock_dev, double t0, void* stream) {
  if (!clock_dev) return fail(GS_E_ARG, "clock is NULL");
  hipLaunchKernelGGL(k_clock_reset, dim3(1), dim3(1), 0, (hipStream_t)stream, clock_dev, t0);
  GS_HIP(hipGetLastError(), "k_clock_reset launch");
  return GS_OK;
}

int gs_clock_set_target(GsClock* clock_dev, double t_end, double dt_max, int64_t max_steps, const GsPhys* phys,
                        GsForcing forc, void* stream) {
  if (!clock_dev) return fail(GS_E_ARG, "clock is NULL");
  int rc = check_phys(phys);
  if (rc) return rc;
  rc = check_forcing(forc);
  if (rc) return rc;
  const int run_mode = isfinite(t_end) ? 1 : 0;
  hipLaunchKernelGGL(k_clock_target, dim3(1), dim3(1), 0, (hipStream_t)stream, clock_dev, make_plan(*phys, forc),
                     t_end, dt_max, (long long)max_steps, run_mode);
  GS_HIP(hipGetLastError(), "k_clock_target launch");
  return GS_OK;
}

int gs_clock_set_dt(GsClock* clock_dev, double dt, const GsPhys* phys, GsForcing forc, void* stream) {
  if (!clock_dev) return fail(GS_E_ARG, "clock is NULL");
  if (!(dt >= 0.0) || isinf(dt)) return fail(GS_E_ARG, "dt must be finite and >= 0");
  int rc = check_phys(phys);
  if (rc) return rc;
  rc = check_forcing(forc);
  if (rc) return rc;
  hipLaunchKernelGGL(k_clock_dt, dim3(1), dim3(1), 0, (hipStream_t)stream, clock_dev, make_plan(*phys, forc), dt);
  GS_HIP(hipGetLastError(), "k_clock_dt launch");
  return GS_OK;
}

int gs_dense_step(const GsDense* d, const GsPhys* phys, GsForcing forc, GsClock* clock_dev, void* stream) {
  if (!d) return fail(GS_E_ARG, "dense descriptor is NULL");
  if (d->dtype == GS_F32) return dense_step_t<float>(d, phys, forc, clock_dev, (hipStream_t)stream);
  if (d->dtype == GS_F64) return dense_step_t<double>(d, phys, forc, clock_dev, (hipStream_t)stream);
  if (d->dtype == GS_F32X64) return dense_step_t<double, float>(d, phys, forc, clock_dev, (hipStream_t)stream);
  return fail(GS_E_DTYPE, "unsupported dtype %d", d->dtype);
}

int gs_dense_residual(const GsDense* d, const GsPhys* phys, int which, void* rhs0, void* rhs1, void* rhs2,
                      void* stream) {
  if (!d) return fail(GS_E_ARG, "dense descriptor is NULL");
  if (d->dtype == GS_F32) return dense_residual_t<float>(d, phys, which, rhs0, rhs1, rhs2, (hipStream_t)stream);
  if (d->dtype == GS_F64) return dense_residual_t<double>(d, phys, which, rhs0, rhs1, rhs2, (hipStream_t)stream);
  if (d->dtype == GS_F32X64)
    return dense_residual_t<double, float>(d, phys, which, rhs0, rhs1, rhs2, (hipStream_t)stream);
  return fail(GS_E_DTYPE, "unsupported dtype %d", d->dtype);
}

int gs_dense_cfl(const GsDense* d, const GsPhys* phys, GsForcing forc, GsClock* clock_dev, void* stream) {
  if (!d) return fail(GS_E_ARG, "dense descriptor is NULL");
  if (d->dtype == GS_F32) return dense_cfl_t<float>(d, phys, forc, clock_dev, (hipStream_t)stream);
  if (d->dtype == GS_F64) return dense_cfl_t<double>(d, phys, forc, clock_dev, (hipStream_t)stream);
  if (d->dtype == GS_F32X64) return dense_cfl_t<double, float>(d, phys, forc, clock_dev, (hipStream_t)stream);
  return fail(GS_E_DTYPE, "unsupported dtype %d", d->dtype);
}

}  // extern "C"
