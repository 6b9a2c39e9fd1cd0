// gs_flat.hip -- compressed active-cell ("flat") layout of the fused finite-volume step.
//
// Stored cells are the active set dilated by two cells (4-neighbour cross), laid out in
// row-major padded (i, j) order like the reference's CompressedMesh2D
// (src/geoswe/compressed_mesh.py:139-157, compressed_rhs.py:241-369). Because of the
// ring-2 dilation, the +-1 and +-2 neighbours along j of an active cell are the flat ids
// k+-1, k+-2; the rows i+-1, i+-2 sit at offsets that are constant over a run of
// consecutive active cells, so the kernel reads a small run table instead of a per-cell
// neighbour table. Cells in no run are the frozen ring: never written, identical in both
// state buffers.
//
// A wavefront pass covers 64 consecutive flat ids and writes the middle 60 (the chunk
// stride of the run index). Each lane evaluates its two x-faces with gathered neighbours and its
// upper y-face with the lane above; the lower y-face comes from the lane below by a
// cross-lane shuffle. The face arithmetic is gs_device.h's, shared with the dense sweep,
// so the residual is bitwise the dense one.
#include <atomic>
#include <mutex>
#include <vector>

#include "gs_common.h"

namespace gs {

constexpr int kFOut = 60;     // flat ids written per wavefront (lanes 2..61)
constexpr int kFBlock = 256;

template <typename T>
struct FlatK {
  int64_t n;
  int n_runs, nyp, ngh;
  int64_t nchunks;
  T* buf[2][3];
  const T* bed;
  const uint8_t* flags;
  const T* ntab;
  int n_classes;
  const T* n_field;
  T* max_h;
  const uint8_t* cls;   // per-cell class plane (ids 0..255), or NULL: the flags' class bits
  T* rhs[3];
  int which;
  const int64_t* run_k0;
  const int32_t* run_len;
  const int32_t* run_i;
  const int32_t* run_j0;
  const int32_t* run_d;
  const int32_t* chunk_run;
  Phys<T> P;
  int rain_mode;
  const T* rain_frames;
  int64_t frame_stride, rain_pitch;
  GsClock* clock;
  Plan plan;
  T* pre[3];        // nullable: pre-friction h, hu, hv of active cells (adjoint recomputation)
  int64_t n_tiles;  // row-sweep tiles (0: chunk traversal)
  const int64_t* tile_e0;
  const int32_t* tile_irc;
  const uint64_t* ent;
  const T* sponge;  // nullable [N][2] (keep, amb) of GS_FLAG_FORCE cells
  const T* clamp;   // nullable [N] h_max of GS_FLAG_FORCE cells
  const T* infil;   // nullable [32] infiltration rate per Manning class
  const T* drain;   // nullable [N] karst target depth of GS_FLAG_FORCE cells
  const uint8_t* ga_cls;  // Green-Ampt / drain-tau (nullable tables, see the ABI header)
  const T* ga_ks;
  const T* ga_psi;
  const T* ga_dth;
  T* ga_F;
  const T* ga_fmax;
  const T* ga_inv_tau;
  const int32_t* rain_lookup;  // rain_mode 3: native-table column of each cell
  int64_t n_rmode;             // ring modes of the cached replay (top of the step)
  const int64_t* rmode_ids;
  const int64_t* rmode_nb;
  const uint8_t* rmode_kind;
  T rmode_eta;
  int64_t n_fring;  // forced cells of the ring (stored, not active): sponge, clamp, gauge ring
  const int64_t* fring;
  const int32_t* fring_row;  // nullable: gauge-ring row of each entry, -1 for none
  int ring_ng, ring_nt;
  const T* ring_bed;
  const T* ring_w;
  const double* ring_t;
  const T* ring_stage;
};


// Everything one chunk pass reads, as values: the state buffers picked for this step (a
// runtime index into the argument's buffer table would put the table in private memory and
// every state access behind a generic pointer), the mesh arrays, the clock's per-step
// values and the physics scalars. Filled from the kernel argument (plain step) or from
// the member table (ensemble).
#define GS_G __attribute__((address_space(1)))

template <typename T>
struct FlatCtx {
  const GS_G T* qi0;
  const GS_G T* qi1;
  const GS_G T* qi2;
  GS_G T* qo0;
  GS_G T* qo1;
  GS_G T* qo2;
  const GS_G T* bed;
  const GS_G uint8_t* flags;
  const GS_G T* n_field;
  GS_G T* max_h;
  const GS_G uint8_t* cls;
  GS_G T* rhs0;
  GS_G T* rhs1;
  GS_G T* rhs2;
  GS_G T* pre0;
  GS_G T* pre1;
  GS_G T* pre2;
  const GS_G int64_t* run_k0;
  const GS_G int32_t* run_len;
  const GS_G int32_t* run_i;
  const GS_G int32_t* run_j0;
  const GS_G int32_t* run_d;
  const GS_G int32_t* chunk_run;
  const GS_G int64_t* tile_e0;
  const GS_G int32_t* tile_irc;
  const GS_G uint64_t* ent;
  const GS_G T* sponge;
  const GS_G T* clamp;
  const GS_G T* infil;
  const GS_G T* drain;
  const GS_G uint8_t* ga_cls;
  const GS_G T* ga_ks;
  const GS_G T* ga_psi;
  const GS_G T* ga_dth;
  GS_G T* ga_F;
  const GS_G T* ga_fmax;
  const GS_G T* ga_inv_tau;
  const GS_G int32_t* rain_lookup;
  const GS_G T* frame;   // gridded rain frame of this step (rain_mode 2)
  const T* ntab;         // Manning n per class (LDS copy or the member's table)
  int64_t n, rain_pitch;
  int n_runs, ngh, rain_mode;
  T dt, rain_u, eta;
  Phys<T> P;
};

// Device buffers are global memory; the explicit address space lets the compiler use global
// (not generic) accesses for pointers that come from device memory (the member table).
template <typename P>
__device__ __forceinline__ GS_G P* gp(P* p) {
  return (GS_G P*)p;
}

// The clock's per-step values for one domain (read once per launch, or once per member
// switch in the ensemble).
template <typename T>
struct FlatClk {
  bool live;        // false when the clock says the run is done
  int par;          // input buffer (MODE 1: nsteps & 1; MODE 0: the residual selector)
  T dt, rain_u, eta;
  const GS_G T* frame;
};

template <typename T, int MODE>
__device__ __forceinline__ FlatClk<T> flat_clock(const FlatK<T>& a) {
  FlatClk<T> k;
  k.live = true;
  k.par = MODE == 1 ? 0 : a.which;
  k.dt = T(0);
  k.rain_u = T(0);
  k.eta = T(0);
  k.frame = nullptr;
  if (MODE == 1) {
    const GS_G GsClock* c = gp(a.clock);
    if (c->status & GS_ST_DONE) {
      k.live = false;
      return k;
    }
    k.par = (int)(c->nsteps & 1);
    const int ridx = c->rain_idx;
    if (sizeof(T) == 4) {
      k.dt = (T)c->dt32;
      k.rain_u = (T)c->rain32;
      k.eta = (T)c->eta32;
    } else {
      k.dt = (T)c->dt;
      k.rain_u = a.rain_mode == 1 ? (T)gp(a.plan.rain_rate)[ridx] : T(0);
      k.eta = (T)c->eta64;
    }
    if (a.rain_mode >= 2) k.frame = gp(a.rain_frames) + (int64_t)ridx * a.frame_stride;
  }
  return k;
}

// The step context: the layout of `a` with the buffers the clock values pick.
template <typename T, int MODE>
__device__ __forceinline__ void flat_ctx(const FlatK<T>& a, const FlatClk<T>& k, FlatCtx<T>& x) {
  const bool in1 = k.par != 0;
  const bool out1 = MODE == 1 ? !in1 : true;   // MODE 0 writes no state
  x.qi0 = gp(in1 ? a.buf[1][0] : a.buf[0][0]);
  x.qi1 = gp(in1 ? a.buf[1][1] : a.buf[0][1]);
  x.qi2 = gp(in1 ? a.buf[1][2] : a.buf[0][2]);
  x.qo0 = gp(out1 ? a.buf[1][0] : a.buf[0][0]);
  x.qo1 = gp(out1 ? a.buf[1][1] : a.buf[0][1]);
  x.qo2 = gp(out1 ? a.buf[1][2] : a.buf[0][2]);
  x.bed = gp(a.bed);
  x.flags = gp(a.flags);
  x.n_field = gp(a.n_field);
  x.cls = gp(a.cls);
  x.max_h = gp(a.max_h);
  x.rhs0 = gp(a.rhs[0]);
  x.rhs1 = gp(a.rhs[1]);
  x.rhs2 = gp(a.rhs[2]);
  x.pre0 = gp(a.pre[0]);
  x.pre1 = gp(a.pre[1]);
  x.pre2 = gp(a.pre[2]);
  x.run_k0 = gp(a.run_k0);
  x.run_len = gp(a.run_len);
  x.run_i = gp(a.run_i);
  x.run_j0 = gp(a.run_j0);
  x.run_d = gp(a.run_d);
  x.chunk_run = gp(a.chunk_run);
  x.tile_e0 = gp(a.tile_e0);
  x.tile_irc = gp(a.tile_irc);
  x.ent = gp(a.ent);
  x.sponge = gp(a.sponge);
  x.clamp = gp(a.clamp);
  x.infil = gp(a.infil);
  x.drain = gp(a.drain);
  x.ga_cls = gp(a.ga_cls);
  x.ga_ks = gp(a.ga_ks);
  x.ga_psi = gp(a.ga_psi);
  x.ga_dth = gp(a.ga_dth);
  x.ga_F = gp(a.ga_F);
  x.ga_fmax = gp(a.ga_fmax);
  x.ga_inv_tau = gp(a.ga_inv_tau);
  x.rain_lookup = gp(a.rain_lookup);
  x.frame = k.frame;
  x.ntab = a.ntab;
  x.n = a.n;
  x.rain_pitch = a.rain_pitch;
  x.n_runs = a.n_runs;
  x.ngh = a.ngh;
  x.rain_mode = a.rain_mode;
  x.dt = k.dt;
  x.rain_u = k.rain_u;
  x.eta = k.eta;
  x.P = a.P;
}

// One wavefront pass over flat ids chunk*60-2 .. chunk*60+61: the residual (MODE 0) or the
// fused update (MODE 1) of the 60 middle ids, folding their CFL speed bits into lam.
// Every lane of the wavefront must call it (cross-lane shuffles).
// Element at byte offset `off` of an array segment that starts at `row` (uniform).
template <typename P>
__device__ __forceinline__ const GS_G P& gat(const GS_G P* row, uint32_t off) {
  return *(const GS_G P*)((const GS_G char*)row + off);
}
template <typename P>
__device__ __forceinline__ GS_G P& gat(GS_G P* row, uint32_t off) {
  return *(GS_G P*)((GS_G char*)row + off);
}

// Sponge band then stage clamp on a GS_FLAG_FORCE cell, after the update, the stage and the
// running max (sponge_band, compressed_solver.py:1320-1326: h*keep then + amb, two roundings;
// stage_clamp_flat, 424-440). keep = 1, amb = 0, h_max = +inf leave a cell unchanged.
template <typename T>
__device__ __forceinline__ void sponge_cell(T kp, T am, T& h, T& hu, T& hv) {
  h = h * kp + am;
  hu = hu * kp;
  hv = hv * kp;
}
template <typename T>
__device__ __forceinline__ void clamp_cell(T hx, T& h, T& hu, T& hv) {
  if (h > hx) {
    const T ratio = hx / h;
    h = hx;
    hu = hu * ratio;
    hv = hv * ratio;
  }
}
template <typename T>
__device__ __forceinline__ void drain_cell(T ht, T& h, T& hu, T& hv) {
  if (h > ht) {
    const T r = h > T(1e-12) ? ht / h : T(0);
    h = ht;
    hu = hu * r;
    hv = hv * r;
  }
}
// Constant-rate infiltration of one active cell (infiltrate_flat): nothing below h_min.
template <typename T>
__device__ __forceinline__ void infil_cell(T rate, T dt, T hmin, T& h, T& hu, T& hv) {
  if (!(h < hmin)) {
    const T dh = rate * dt;
    T hn = h - dh;
    hn = hn < T(0) ? T(0) : hn;
    const T r = hn / h;
    h = hn;
    hu = hu * r;
    hv = hv * r;
  }
}
// exp(x) in float64 from IEEE +, -, *, rint and ldexp only, so that it is the oracle's
// exp_ieee bit for bit: Cody-Waite reduction x = k ln2 + r (ln2 split in a 32-bit head, so
// k*LN2_HI is exact, and a tail), exp(r) on |r| <= ln2/2 by the degree-13 Taylor polynomial in
// Horner form, scaled by 2^k. It replaces the fast exponential of the reference's drain decay
// (float64 state; float32 state uses exp_ieee_f below).
__device__ __forceinline__ double exp_ieee(double x) {
  if (!(x == x)) return x;
  if (x > 709.8) return __builtin_inf();
  if (x < -745.2) return 0.0;
  const double kf = rint(x * 1.44269504088896338700e+00);
  const double r = (x - kf * 6.93147180369123816490e-01) - kf * 1.90821492927058770002e-10;
  double p = 1.6059043836821613e-10;
  p = p * r + 2.08767569878681e-09;
  p = p * r + 2.505210838544172e-08;
  p = p * r + 2.755731922398589e-07;
  p = p * r + 2.7557319223985893e-06;
  p = p * r + 2.48015873015873e-05;
  p = p * r + 0.0001984126984126984;
  p = p * r + 0.001388888888888889;
  p = p * r + 0.008333333333333333;
  p = p * r + 0.041666666666666664;
  p = p * r + 0.16666666666666666;
  p = p * r + 0.5;
  p = p * r + 1.0;
  p = p * r + 1.0;
  return ldexp(p, (int)kf);
}

// The float32 form: ln2 split in a 15-bit head (k*LN2_HI exact for |k| < 512) and a tail, a
// degree-7 Taylor polynomial (truncation below 5e-9 relative), every operation in float32.
__device__ __forceinline__ float exp_ieee_f(float x) {
  if (!(x == x)) return x;
  if (x > 88.73f) return __builtin_inff();
  if (x < -103.98f) return 0.0f;
  const float kf = rintf(x * 1.4426950216293335f);
  const float r = (x - kf * 0.693145751953125f) - kf * 1.428606765330187e-06f;
  float p = 0.00019841270113829523f;
  p = p * r + 0.0013888889225199819f;
  p = p * r + 0.008333333767950535f;
  p = p * r + 0.0416666679084301f;
  p = p * r + 0.1666666716337204f;
  p = p * r + 0.5f;
  p = p * r + 1.0f;
  p = p * r + 1.0f;
  return ldexpf(p, (int)kf);
}

// Green-Ampt infiltration then drain-tau of one active cell (ga_drain_flat); `ob` is the
// cell's offset in a one-byte plane.
template <typename T>
__device__ __forceinline__ void ga_cell(const FlatCtx<T>& x, int64_t base, uint32_t off, uint32_t ob, T& h, T& hu,
                                        T& hv) {
  if (h <= T(0)) return;
  T scale = T(1);
  if (x.ga_ks) {
    const int c = gat(x.ga_cls + base, ob);
    const T K = x.ga_ks[c];
    if (K > T(0)) {
      const T KsDt = K * x.dt;
      const T head = x.ga_psi[c] + h;
      GS_G T& Fr = gat(x.ga_F + base, off);
      const T F0 = Fr;
      const T a = F0 + KsDt;
      const T disc = a * a + ((T(4) * KsDt) * head) * x.ga_dth[c];
      const T F1 = T(0.5) * (a + sqrt(disc > T(0) ? disc : T(0)));   // fmax(disc, 0): NaN gives 0
      T dF = F1 - F0;
      dF = dF > T(0) ? dF : T(0);
      if (dF > h) dF = h;
      if (x.ga_fmax) {
        const T room = gat(x.ga_fmax + base, off) - F0;
        if (dF > room) dF = room > T(0) ? room : T(0);
      }
      const T hn = h - dF;
      scale = scale * (hn / h);
      h = hn;
      Fr = F0 + dF;
    }
  }
  if (x.ga_inv_tau) {
    const T it = gat(x.ga_inv_tau + base, off);
    if (it > T(0) && h > T(0)) {
      const T arg = (-x.dt) * it;
      T decay;
      if constexpr (sizeof(T) == 4) decay = exp_ieee_f(arg);
      else decay = exp_ieee(arg);
      scale = scale * decay;
      h = h * decay;
    }
  }
  hu = hu * scale;
  hv = hv * scale;
}

// The forcings after the running max, in the reference's order: infiltration (every active
// cell, by Manning class), the sponge of a GS_FLAG_FORCE cell, (the gauge ring: ring cells
// only, in the ring pass), Green-Ampt and drain-tau (every active cell), then the clamp and
// the karst drain of a GS_FLAG_FORCE cell. GA = false leaves Green-Ampt out of the code (the
// plain sweep's instantiation, which has no register to spare for it).
template <typename T, bool GA>
__device__ __forceinline__ void active_force(const FlatCtx<T>& x, uint8_t fg, int64_t base, uint32_t off,
                                             uint32_t ob, T& h, T& hu, T& hv) {
  if (x.infil)
    infil_cell<T>(x.infil[x.cls ? gat(x.cls + base, ob) : (fg & GS_FLAG_CLASS_MASK)], x.dt, x.P.h_min, h, hu, hv);
  const bool fc = (fg & GS_FLAG_FORCE) != 0;
  if (fc && x.sponge)
    sponge_cell<T>(gat(x.sponge + 2 * base, 2 * off), gat(x.sponge + 2 * base + 1, 2 * off), h, hu, hv);
  if (GA && (x.ga_ks || x.ga_inv_tau)) ga_cell<T>(x, base, off, ob, h, hu, hv);
  if (fc && x.clamp) clamp_cell<T>(gat(x.clamp + base, off), h, hu, hv);
  if (fc && x.drain) drain_cell<T>(gat(x.drain + base, off), h, hu, hv);
}

template <typename T, int MODE>
__device__ __forceinline__ void flat_chunk(const FlatCtx<T>& x, int64_t chunk, int lane, lam_t<T>& lam) {
  const FlatCtx<T>& a = x;
  const Phys<T>& P = a.P;
  const T ix = P.inv_dx, iy = P.inv_dy, hdx = P.half_dx, hdy = P.half_dy, g = P.g, hm = P.h_min;
  const int64_t k = chunk * kFOut + lane - 2;
  const int64_t kl = k < 0 ? 0 : (k >= a.n ? a.n - 1 : k);
  // this lane's cell (needed by every lane for the y-face shuffles)
  const T hC = x.qi0[kl], huC = x.qi1[kl], hvC = x.qi2[kl];
  const T bC = a.bed[kl];
  T uC, vC;
  prims(hC, huC, hvC, hm, uC, vC);
  const T bl = lane_lower(bC), br = lane_upper(bC);
  const T gzy = (T(0.5) * (br - bl)) * iy;
  // upper y-face (k, k+1)
  const T hR = lane_upper(hC), uR = lane_upper(uC), vR = lane_upper(vC);
  const T gzyR = lane_upper(gzy);
  Face<T> fy;
  face<T>(bC, br, gzy, gzyR, hC, hR, vC, uC, vR, uR, hdy, g, hm, fy);
  const T yFhB = lane_lower(fy.Fh), yFNB = lane_lower(fy.FN), yFTB = lane_lower(fy.FT);
  const T ysRB = lane_lower(fy.srcR);
  const int ygB = lane_lower((int)fy.gate);

  // run of this cell
  int r = -1;
  if (lane >= 2 && lane < 2 + kFOut && k < a.n) {
    int rr = a.chunk_run[chunk];
    while (rr < a.n_runs && a.run_k0[rr] + a.run_len[rr] <= k) ++rr;
    if (rr < a.n_runs && a.run_k0[rr] <= k) r = rr;
  }
  if (r < 0) return;
  const GS_G int32_t* d = a.run_d + 4 * (int64_t)r;
  const int64_t kE = k + d[0], kW = k + d[1], kEE = k + d[2], kWW = k + d[3];
  const T bE = a.bed[kE], bW = a.bed[kW], bEE = a.bed[kEE], bWW = a.bed[kWW];
  const T hE = x.qi0[kE], huE = x.qi1[kE], hvE = x.qi2[kE];
  const T hW = x.qi0[kW], huW = x.qi1[kW], hvW = x.qi2[kW];
  T uE, vE, uW, vW;
  prims(hE, huE, hvE, hm, uE, vE);
  prims(hW, huW, hvW, hm, uW, vW);
  const T gzxC = (T(0.5) * (bE - bW)) * ix;
  const T gzxE = (T(0.5) * (bEE - bC)) * ix;
  const T gzxW = (T(0.5) * (bC - bWW)) * ix;
  Face<T> fr, fl;
  face<T>(bC, bE, gzxC, gzxE, hC, hE, uC, vC, uE, vE, hdx, g, hm, fr);
  face<T>(bW, bC, gzxW, gzxC, hW, hC, uW, vW, uC, vC, hdx, g, hm, fl);
  T rh = T(0), rhu = T(0), rhv = T(0);
  if (fr.gate) {
    rh = rh - fr.Fh * ix;
    rhu = rhu - fr.FN * ix;
    rhv = rhv - fr.FT * ix;
    rhu = rhu - fr.srcL * ix;
  }
  if (fl.gate) {
    rh = rh + fl.Fh * ix;
    rhu = rhu + fl.FN * ix;
    rhv = rhv + fl.FT * ix;
    rhu = rhu + fl.srcR * ix;
  }
  if (fy.gate) {
    rh = rh - fy.Fh * iy;
    rhu = rhu - fy.FT * iy;
    rhv = rhv - fy.FN * iy;
    rhv = rhv - fy.srcL * iy;
  }
  if (ygB) {
    rh = rh + yFhB * iy;
    rhu = rhu + yFTB * iy;
    rhv = rhv + yFNB * iy;
    rhv = rhv + ysRB * iy;
  }
  if (MODE == 0) {
    a.rhs0[k] = rh;
    a.rhs1[k] = rhu;
    a.rhs2[k] = rhv;
    return;
  }
  const uint8_t fg = a.flags ? a.flags[k] : (uint8_t)0;
  T rr0 = rh;
  if (a.rain_mode == 1) {
    rr0 = rr0 + x.rain_u;
  } else if (a.rain_mode == 2) {
    const int64_t ii = (int64_t)a.run_i[r] - a.ngh;
    const int64_t jj = (int64_t)a.run_j0[r] + (k - a.run_k0[r]) - a.ngh;
    rr0 = rr0 + x.frame[ii * a.rain_pitch + jj];
  } else if (a.rain_mode == 3) {
    rr0 = rr0 + x.frame[x.rain_lookup[k]];
  }
  if (a.pre0) {   // the first three statements of update(), kept for its reverse
    a.pre0[k] = hC + x.dt * rr0;
    a.pre1[k] = huC + x.dt * rhu;
    a.pre2[k] = hvC + x.dt * rhv;
  }
  T n = T(0);
  if (P.friction) n = a.n_field ? a.n_field[k] : x.ntab[a.cls ? a.cls[k] : (fg & GS_FLAG_CLASS_MASK)];
  T h, hu, hv;
  update<T>(hC, huC, hvC, rr0, rhu, rhv, x.dt, n, P, h, hu, hv);
  if (fg & GS_FLAG_STAGE) {
    const T hs = x.eta - bC;
    h = hs > T(0) ? hs : T(0);
    hu = T(0);
    hv = T(0);
  }
  if (a.max_h && h > a.max_h[k]) a.max_h[k] = h;
  active_force<T, true>(x, fg, k, 0u, 0u, h, hu, hv);
  x.qo0[k] = h;
  x.qo1[k] = hu;
  x.qo2[k] = hv;
  if (!(fg & GS_FLAG_CFL_SKIP)) {
    const lam_t<T> lb_ = lam_key<T, true>(h, hu, hv, P);
    lam = lb_ > lam ? lb_ : lam;
  }
}

// Lane `lane`'s cell in one window row of a sweep tile (entry e = first stored id, stored
// mask, active mask): the stored cells of a row are consecutive flat ids, so the lane's cell
// is the row's first id (uniform across the wavefront) plus the number of stored lanes below
// it. The byte offset of that count is returned; accesses add it to the row's start.
template <typename T>
__device__ __forceinline__ uint32_t sweep_lane(const GS_G uint64_t* e, int lane, int64_t& base, bool& st) {
  const uint64_t sm = e[1];
  base = (int64_t)e[0];
  const unsigned below =
      __builtin_amdgcn_mbcnt_hi((unsigned)(sm >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)sm, 0u));
  st = ((sm >> lane) & 1ull) != 0ull;
  return below * (uint32_t)sizeof(T);
}

// State and bed of a lane's cell; lanes on unstored cells read nothing and hold zeros
// (their faces are never used: every face an active cell needs has its stencil stored).
template <typename T>
__device__ __forceinline__ void sweep_q(const FlatCtx<T>& x, int64_t base, uint32_t off, bool st, T& h, T& hu,
                                        T& hv) {
  h = T(0);
  hu = T(0);
  hv = T(0);
  if (st) {
    h = gat(x.qi0 + base, off);
    hu = gat(x.qi1 + base, off);
    hv = gat(x.qi2 + base, off);
  }
}

template <typename T>
__device__ __forceinline__ T sweep_b(const FlatCtx<T>& x, int64_t base, uint32_t off, bool st) {
  T b = T(0);
  if (st) b = gat(x.bed + base, off);
  return b;
}

// One row-sweep tile: the fused update of the active lanes of rows i0 .. i0+nr-1 of one
// 64-column window. As in the dense sweep, every face is evaluated once: the x-face below
// a row is carried from the row above, the y-face below a lane comes from the lane below
// by shuffle, and the residual terms are added in the same order as there (and as in
// flat_chunk), so all three traversals give the same bits. Every lane of the wavefront
// must call it. PRE: also write the pre-friction state (the adjoint's recomputation); a
// separate instantiation keeps those pointers out of the plain step's scalar registers.
// LEAN: the domain has no per-cell Manning field or class plane, no running maximum and no
// active-cell forcing, so none of their pointers is held (the host checks that they are all
// null).
template <typename T, bool PRE, bool EXT, bool LEAN = false>
__device__ __forceinline__ void flat_tile(const FlatCtx<T>& x, int64_t t, int lane, lam_t<T>& lam) {
  const Phys<T>& P = x.P;
  const T ix = P.inv_dx, iy = P.inv_dy, hdx = P.half_dx, hdy = P.half_dy, g = P.g, hm = P.h_min;
  const GS_G int32_t* irc = x.tile_irc + 3 * t;
  const int i0 = irc[0], nr = irc[1], c0 = irc[2];
  const GS_G uint64_t* E = x.ent + 3 * x.tile_e0[t];   // entry of row i0 - 2
  bool sm2, sm1, sC, sN;
  int64_t gm2, gm1, gC, gN;   // first flat id of rows i-2, i-1, i, i+1 (uniform)
  const uint32_t om2 = sweep_lane<T>(E, lane, gm2, sm2);
  const uint32_t om1 = sweep_lane<T>(E + 3, lane, gm1, sm1);
  uint32_t oC = sweep_lane<T>(E + 6, lane, gC, sC);
  uint32_t oN = sweep_lane<T>(E + 9, lane, gN, sN);
  const T bm2 = sweep_b(x, gm2, om2, sm2);
  T bm1 = sweep_b(x, gm1, om1, sm1), b0 = sweep_b(x, gC, oC, sC), bp1 = sweep_b(x, gN, oN, sN);
  T hP, huP, hvP, uP, vP;
  sweep_q(x, gm1, om1, sm1, hP, huP, hvP);
  prims(hP, huP, hvP, hm, uP, vP);
  const T gzxP = (T(0.5) * (b0 - bm2)) * ix;
  T hC, huC, hvC, uC, vC;
  sweep_q(x, gC, oC, sC, hC, huC, hvC);
  prims(hC, huC, hvC, hm, uC, vC);
  T gzxC = (T(0.5) * (bp1 - bm1)) * ix;
  Face<T> fx;
  face<T>(bm1, b0, gzxP, gzxC, hP, hC, uP, vP, uC, vC, hdx, g, hm, fx);
  // the x-face below the current row, as the products the current cell adds (zero when
  // the face is gated off: see the accumulation below)
  const bool xg = fx.gate;
  T xFh = xg ? fx.Fh * ix : T(0), xFN = xg ? fx.FN * ix : T(0), xFT = xg ? fx.FT * ix : T(0),
    xsR = xg ? fx.srcR * ix : T(0);

  for (int r = 0; r < nr; ++r) {
    const GS_G uint64_t* Er = E + 3 * (r + 2);   // this row's entry
    bool sNN;
    int64_t gNN;
    const uint32_t oNN = sweep_lane<T>(Er + 6, lane, gNN, sNN);
    const T bp2 = sweep_b(x, gNN, oNN, sNN);
    T hN, huN, hvN, uN, vN;
    sweep_q(x, gN, oN, sN, hN, huN, hvN);
    prims(hN, huN, hvN, hm, uN, vN);
    const T gzxN = (T(0.5) * (bp2 - b0)) * ix;
    Face<T> fn;
    face<T>(b0, bp1, gzxC, gzxN, hC, hN, uC, vC, uN, vN, hdx, g, hm, fn);
    const bool ng = fn.gate;
    const T nFh = ng ? fn.Fh * ix : T(0), nFN = ng ? fn.FN * ix : T(0), nFT = ng ? fn.FT * ix : T(0);
    const T nsL = ng ? fn.srcL * ix : T(0), nsR = ng ? fn.srcR * ix : T(0);

    // y faces of the row: each lane computes its upper face (j, j+1)
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

    if ((Er[2] >> lane) & 1ull) {   // an active cell: lanes 2..61 only
      const uint32_t ob = oC / (uint32_t)sizeof(T);   // byte offset in a 1-byte plane
      const uint8_t fg = x.flags ? gat(x.flags + gC, ob) : (uint8_t)0;
      T rr0 = rh;
      if (x.rain_mode == 1) {
        rr0 = rr0 + x.rain_u;
      } else if (x.rain_mode == 2) {
        rr0 = rr0 + x.frame[(int64_t)(i0 + r - x.ngh) * x.rain_pitch + (c0 + lane - x.ngh)];
      } else if (EXT && x.rain_mode == 3) {
        rr0 = rr0 + x.frame[gat(x.rain_lookup + gC, (oC / (uint32_t)sizeof(T)) * 4u)];
      }
      if (PRE) {   // the first three statements of update(), kept for its reverse
        gat(x.pre0 + gC, oC) = hC + x.dt * rr0;
        gat(x.pre1 + gC, oC) = huC + x.dt * rhu;
        gat(x.pre2 + gC, oC) = hvC + x.dt * rhv;
      }
      T n = T(0);
      if (P.friction)
        n = (!LEAN && x.n_field) ? gat(x.n_field + gC, oC)
                                 : x.ntab[(!LEAN && x.cls) ? gat(x.cls + gC, ob) : (fg & GS_FLAG_CLASS_MASK)];
      T h, hu, hv;
      update<T>(hC, huC, hvC, rr0, rhu, rhv, x.dt, n, P, h, hu, hv);
      if (fg & GS_FLAG_STAGE) {
        const T hs = x.eta - b0;
        h = hs > T(0) ? hs : T(0);
        hu = T(0);
        hv = T(0);
      }
      if (!LEAN) {
        if (x.max_h && h > gat(x.max_h + gC, oC)) gat(x.max_h + gC, oC) = h;
        active_force<T, EXT>(x, fg, gC, oC, ob, h, hu, hv);
      }
      gat(x.qo0 + gC, oC) = h;
      gat(x.qo1 + gC, oC) = hu;
      gat(x.qo2 + gC, oC) = hv;
      if (!(fg & GS_FLAG_CFL_SKIP)) {
        const lam_t<T> lb_ = lam_key<T, true>(h, hu, hv, P);
        lam = lb_ > lam ? lb_ : lam;
      }
    }

    // slide the window one row down
    hC = hN; huC = huN; hvC = hvN; uC = uN; vC = vN; gzxC = gzxN;
    bm1 = b0; b0 = bp1; bp1 = bp2;
    xFh = nFh; xFN = nFN; xFT = nFT; xsR = nsR;
    gC = gN; oC = oN; gN = gNN; oN = oNN; sN = sNN;
  }
}

// Workgroup b's place in an XCD-major order: workgroups are dealt round-robin to the 8
// XCDs, so giving each XCD a contiguous range of wavefronts (hence of chunks in every
// pass) keeps the rows i+-1 that a chunk gathers mostly in that XCD's L2.
__device__ __forceinline__ unsigned xcd_rank(unsigned b, unsigned nb) {
  const unsigned qd = nb / 8u, rm = nb % 8u, xcd = b % 8u, kk = b / 8u;
  return (xcd < rm) ? xcd * (qd + 1u) + kk : rm * (qd + 1u) + (xcd - rm) * qd + kk;
}

// One step of one flat domain on a resident grid, each wavefront striding over chunks.
template <typename T, int MODE>
__global__ __launch_bounds__(kFBlock) void k_flat_step(FlatK<T> a) {
  const FlatClk<T> ck = flat_clock<T, MODE>(a);
  if (!ck.live) return;   // the whole grid sees the same clock
  FlatCtx<T> x;
  flat_ctx<T, MODE>(a, ck, x);
  __shared__ T s_ntab[GS_MAX_CLASSES];
  if (MODE == 1)
    for (int k = (int)threadIdx.x; k < a.n_classes; k += (int)blockDim.x) s_ntab[k] = a.ntab[k];
  __syncthreads();
  x.ntab = s_ntab;
  const int lane = threadIdx.x & 63;
  const int64_t nwaves = (int64_t)gridDim.x * (kFBlock / 64);
  lam_t<T> lam = 0;
  for (int64_t chunk = (int64_t)xcd_rank(blockIdx.x, gridDim.x) * (kFBlock / 64) + (threadIdx.x >> 6);
       chunk < a.nchunks; chunk += nwaves)
    flat_chunk<T, MODE>(x, chunk, lane, lam);
  if (MODE == 1) grid_reduce_and_plan(lam_key_bits64(lam), a.clock, a.plan, true);
}

// One step of one flat domain by row-sweep tiles, one wavefront per tile in band-major tile
// order and XCD-major wavefront order (the loop also serves a grid smaller than that). The float32 plain
// step is held to 72 VGPRs (7 wavefronts per SIMD instead of 6; a few scalar registers are
// then reloaded in the row loop), which measured 2-3 % faster. LEAN (a domain with no
// per-cell Manning field or class plane, running maximum or active-cell forcing) drops those pointers:
// 34 scalar registers spilled to lanes instead of 75, 4 % faster on the coastal case.
template <typename T, bool PRE, bool LEAN = false>
__global__ __launch_bounds__(kFBlock) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 && !PRE ? 7 : 1))) void k_flat_sweep(FlatK<T> a) {
  const FlatClk<T> ck = flat_clock<T, 1>(a);
  if (!ck.live) return;   // the whole grid sees the same clock
  FlatCtx<T> x;
  flat_ctx<T, 1>(a, ck, x);
  __shared__ T s_ntab[GS_MAX_CLASSES];
  for (int k = (int)threadIdx.x; k < a.n_classes; k += (int)blockDim.x) s_ntab[k] = a.ntab[k];
  __syncthreads();
  x.ntab = s_ntab;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int64_t nwaves = (int64_t)gridDim.x * (kFBlock / 64);
  lam_t<T> lam = 0;
  // PRE (the adjoint's recomputation) also compiles the extended forcings in, so that a
  // replay reads rain through the per-cell lookup exactly as the forward run's extended sweep
  for (int64_t t = (int64_t)xcd_rank(blockIdx.x, gridDim.x) * (kFBlock / 64) + wv; t < a.n_tiles; t += nwaves)
    flat_tile<T, PRE, PRE, LEAN>(x, t, lane, lam);
  grid_reduce_and_plan(lam_key_bits64(lam), a.clock, a.plan, true);
}

// The sweep with the extended forcings compiled in: Green-Ampt infiltration, drain-tau and
// rain read through a per-cell lookup into a native-grid table (launched only when one of
// them is set; the plain sweep has no register to spare for them). The float32 build is
// held to 72 VGPRs (7 wavefronts per SIMD instead of 6, a few registers then spilled):
// 2.62 against 2.74 ms per step with the coastal forcing tail on the 16384^2 case.
template <typename T>
__global__ __launch_bounds__(kFBlock) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 7 : 1))) void k_flat_sweep_ext(FlatK<T> a) {
  const FlatClk<T> ck = flat_clock<T, 1>(a);
  if (!ck.live) return;
  FlatCtx<T> x;
  flat_ctx<T, 1>(a, ck, x);
  __shared__ T s_ntab[GS_MAX_CLASSES];
  for (int k = (int)threadIdx.x; k < a.n_classes; k += (int)blockDim.x) s_ntab[k] = a.ntab[k];
  __syncthreads();
  x.ntab = s_ntab;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int64_t nwaves = (int64_t)gridDim.x * (kFBlock / 64);
  lam_t<T> lam = 0;
  for (int64_t t = (int64_t)xcd_rank(blockIdx.x, gridDim.x) * (kFBlock / 64) + wv; t < a.n_tiles; t += nwaves)
    flat_tile<T, false, true>(x, t, lane, lam);
  grid_reduce_and_plan(lam_key_bits64(lam), a.clock, a.plan, true);
}

// The forcings on the ring cells of the band (stored, not active): out = force(in), one
// thread per listed id. Launched just before the step kernel on the same stream, it reads
// the clock the step will read (its input buffer, its done state); the step kernel only
// reads these cells, from the input buffer, so the two never touch the same word.
template <typename T>
__global__ __launch_bounds__(kFBlock) void k_flat_ring_force(FlatK<T> a) {
  const FlatClk<T> ck = flat_clock<T, 1>(a);
  if (!ck.live) return;
  FlatCtx<T> x;
  flat_ctx<T, 1>(a, ck, x);
  const bool sp = x.sponge != nullptr, cl = x.clamp != nullptr;
  const GS_G int64_t* ids = gp(a.fring);
  const GS_G int32_t* rows = gp(a.fring_row);
  // gauge stage weights at the step's start time (uniform: every thread takes the same path)
  T wa = T(0), wb = T(0);
  int ti = 0;
  if (rows) {
    const GS_G double* tc = gp(a.ring_t);
    const double t = gp(a.clock)->t;
    int lo = 0, hi = a.ring_nt;   // bisect_left: the number of samples below t
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (tc[mid] < t) lo = mid + 1; else hi = mid;
    }
    ti = lo - 1;
    if (ti > a.ring_nt - 2) ti = a.ring_nt - 2;
    if (ti < 0) ti = 0;
    double wt = (t - tc[ti]) / (tc[ti + 1] - tc[ti] + 1.0e-12);
    wt = wt < 1.0 ? wt : 1.0;
    wt = wt > 0.0 ? wt : 0.0;
    wa = (T)(1.0 - wt);
    wb = (T)wt;
  }
  for (int64_t i = (int64_t)blockIdx.x * kFBlock + threadIdx.x; i < a.n_fring; i += (int64_t)gridDim.x * kFBlock) {
    const int64_t k = ids[i];
    T h = x.qi0[k], hu = x.qi1[k], hv = x.qi2[k];
    const bool fc = x.flags != nullptr && (x.flags[k] & GS_FLAG_FORCE);
    if (fc && sp) sponge_cell<T>(x.sponge[2 * k], x.sponge[2 * k + 1], h, hu, hv);
    const int r = rows ? rows[i] : -1;
    if (r >= 0) {
      const GS_G T* w = gp(a.ring_w) + (int64_t)r * a.ring_ng;
      const GS_G T* S = gp(a.ring_stage);
      T eta = T(0);
      for (int g = 0; g < a.ring_ng; ++g) {
        const T sg = S[(int64_t)g * a.ring_nt + ti] * wa + S[(int64_t)g * a.ring_nt + ti + 1] * wb;
        eta = eta + w[g] * sg;
      }
      const T d = eta - gp(a.ring_bed)[r];
      h = d > T(0) ? d : T(0);   // fmaxf(0, d): NaN gives 0
      hu = T(0);
      hv = T(0);
    }
    if (fc && cl) clamp_cell<T>(x.clamp[k], h, hu, hv);
    if (fc && x.drain) drain_cell<T>(x.drain[k], h, hu, hv);
    x.qo0[k] = h;
    x.qo1[k] = hu;
    x.qo2[k] = hv;
  }
}

// The ring modes of the cached replay on the step's input state, before the step reads
// it: each listed ring cell is overwritten from its active neighbour (extrapolate), with
// still water (stage), or with still-water depth and the neighbour's velocity. Only ring
// cells are written and only active cells (which nothing writes here) are read from.
template <typename T>
__global__ __launch_bounds__(kFBlock) void k_flat_ring_mode(FlatK<T> a) {
  const FlatClk<T> ck = flat_clock<T, 1>(a);
  if (!ck.live) return;
  FlatCtx<T> x;
  flat_ctx<T, 1>(a, ck, x);
  GS_G T* q0 = (GS_G T*)x.qi0;
  GS_G T* q1 = (GS_G T*)x.qi1;
  GS_G T* q2 = (GS_G T*)x.qi2;
  const GS_G int64_t* ids = gp(a.rmode_ids);
  const GS_G int64_t* nbs = gp(a.rmode_nb);
  const GS_G uint8_t* kind = gp(a.rmode_kind);
  for (int64_t i = (int64_t)blockIdx.x * kFBlock + threadIdx.x; i < a.n_rmode; i += (int64_t)gridDim.x * kFBlock) {
    const int64_t k = ids[i];
    const int kd = kind[i];
    if (kd == 0) {
      const int64_t nb = nbs[i];
      q0[k] = q0[nb];
      q1[k] = q1[nb];
      q2[k] = q2[nb];
      continue;
    }
    const T d = a.rmode_eta - x.bed[k];
    const T h = d > T(0) ? d : T(0);
    T hu = T(0), hv = T(0);
    if (kd == 2) {
      const int64_t nb = nbs[i];
      const T hn = q0[nb];
      const T inv = hn > a.P.h_min ? T(1) / hn : T(0);
      hu = h * (q1[nb] * inv);
      hv = h * (q2[nb] * inv);
    }
    q0[k] = h;
    q1[k] = hu;
    q2[k] = hv;
  }
}

// CFL wave speed of the current flat state over active cells, then plan dt.
template <typename T>
__global__ __launch_bounds__(kFBlock) void k_flat_cfl(FlatK<T> a) {
  const int par = (int)(a.clock->nsteps & 1);
  T* const* q = a.buf[par];
  lam_t<T> lam = 0;
  for (int64_t r = (int64_t)blockIdx.x * (kFBlock / 64) + (threadIdx.x >> 6); r < a.n_runs;
       r += (int64_t)gridDim.x * (kFBlock / 64)) {
    const int64_t k0 = a.run_k0[r];
    const int len = a.run_len[r];
    for (int t = threadIdx.x & 63; t < len; t += 64) {
      const int64_t k = k0 + t;
      if (a.flags && (a.flags[k] & GS_FLAG_CFL_SKIP)) continue;
      const lam_t<T> l = lam_key<T, true>(q[0][k], q[1][k], q[2][k], a.P);
      lam = l > lam ? l : lam;
    }
  }
  grid_reduce_and_plan(lam_key_bits64(lam), a.clock, a.plan, false);
}

template <typename T>
static int build_flat(const GsFlat* f, const GsPhys* p, const GsForcing& fo, GsClock* clk, FlatK<T>& a,
                      bool need_clock) {
  if (!f) return fail(GS_E_ARG, "flat descriptor is NULL");
  if (f->n_stored < 1 || f->n_stored > ((int64_t)1 << 40)) return fail(GS_E_RANGE, "n_stored out of range");
  if (f->n_runs < 0) return fail(GS_E_ARG, "n_runs < 0");
  if (f->n_runs > 0 && (!f->run_k0 || !f->run_len || !f->run_i || !f->run_j0 || !f->run_d || !f->chunk_run))
    return fail(GS_E_ARG, "run tables are NULL");
  if (f->n_tiles < 0) return fail(GS_E_ARG, "n_tiles < 0");
  if (f->n_tiles > 0 && (!f->tile_e0 || !f->tile_irc || !f->ent)) return fail(GS_E_ARG, "tile tables are NULL");
  if (f->n_force_ring < 0 || f->n_force_ring > f->n_stored) return fail(GS_E_RANGE, "n_force_ring out of range");
  if (f->n_ring_mode < 0 || f->n_ring_mode > f->n_stored) return fail(GS_E_RANGE, "n_ring_mode out of range");
  if (f->n_ring_mode > 0 && (!f->ring_mode_ids || !f->ring_mode_nb || !f->ring_mode_kind))
    return fail(GS_E_ARG, "ring mode arrays are NULL");
  if (f->ga_ks && (!f->ga_cls || !f->ga_psi || !f->ga_dth || !f->ga_F))
    return fail(GS_E_ARG, "Green-Ampt needs the soil class, psi, dth and F arrays");
  if (f->n_force_ring > 0 && !f->force_ring) return fail(GS_E_ARG, "force_ring is NULL");
  if (f->n_force_ring > 0 && !f->sponge && !f->clamp && !f->drain && !f->force_ring_row)
    return fail(GS_E_ARG, "force_ring lists cells but no sponge, clamp, drain or gauge ring is set");
  if (f->force_ring_row && (f->ring_ng < 1 || f->ring_nt < 2 || !f->ring_bed || !f->ring_w || !f->ring_t ||
                            !f->ring_stage))
    return fail(GS_E_ARG, "gauge ring: needs >= 1 gauge, >= 2 time samples and its bed, weight, time and stage "
                "arrays");
  if (!f->bed) return fail(GS_E_ARG, "bed is NULL");
  for (int s = 0; s < 2; ++s)
    for (int k = 0; k < 3; ++k)
      if (!f->q[s][k]) return fail(GS_E_ARG, "state q[%d][%d] is NULL", s, k);
  int rc = check_phys(p);
  if (rc) return rc;
  rc = check_forcing(fo);
  if (rc) return rc;
  if (p->friction && !f->ntab && !f->n_field) return fail(GS_E_ARG, "friction needs ntab or n_field");
  if (need_clock && !clk) return fail(GS_E_ARG, "clock is NULL");
  memset(&a, 0, sizeof a);
  a.n = f->n_stored;
  a.n_runs = f->n_runs;
  a.nyp = f->nyp;
  a.ngh = f->ngh;
  a.nchunks = (f->n_stored + kFOut - 1) / kFOut;
  for (int s = 0; s < 2; ++s)
    for (int k = 0; k < 3; ++k) a.buf[s][k] = (T*)f->q[s][k];
  a.bed = (const T*)f->bed;
  a.flags = f->flags;
  a.ntab = (const T*)f->ntab;
  a.n_classes = f->ntab ? (p->n_classes > 0 ? p->n_classes : 1) : 0;
  a.n_field = (const T*)f->n_field;
  a.max_h = (T*)f->max_h;
  a.cls = f->cls;
  a.run_k0 = f->run_k0;
  a.run_len = f->run_len;
  a.run_i = f->run_i;
  a.run_j0 = f->run_j0;
  a.run_d = f->run_d;
  a.chunk_run = f->chunk_run;
  a.n_tiles = f->n_tiles;
  a.tile_e0 = f->tile_e0;
  a.tile_irc = f->tile_irc;
  a.ent = f->ent;
  a.sponge = (const T*)f->sponge;
  a.clamp = (const T*)f->clamp;
  a.infil = (const T*)f->infil;
  a.drain = (const T*)f->drain;
  a.ga_cls = f->ga_cls;
  a.ga_ks = (const T*)f->ga_ks;
  a.ga_psi = (const T*)f->ga_psi;
  a.ga_dth = (const T*)f->ga_dth;
  a.ga_F = (T*)f->ga_F;
  a.ga_fmax = (const T*)f->ga_fmax;
  a.ga_inv_tau = (const T*)f->ga_inv_tau;
  a.n_fring = f->n_force_ring;
  a.fring = f->force_ring;
  a.fring_row = f->force_ring_row;
  a.ring_ng = f->ring_ng;
  a.ring_nt = f->ring_nt;
  a.ring_bed = (const T*)f->ring_bed;
  a.ring_w = (const T*)f->ring_w;
  a.ring_t = f->ring_t;
  a.ring_stage = (const T*)f->ring_stage;
  a.P = make_phys<T>(*p, f->dx, f->dy);
  a.rain_mode = fo.n_rain > 0 ? (fo.rain_gridded ? (f->rain_lookup ? 3 : 2) : 1) : 0;
  a.rain_lookup = f->rain_lookup;
  a.n_rmode = f->n_ring_mode;
  a.rmode_ids = f->ring_mode_ids;
  a.rmode_nb = f->ring_mode_nb;
  a.rmode_kind = f->ring_mode_kind;
  a.rmode_eta = (T)f->ring_mode_eta;
  a.rain_frames = (const T*)fo.rain_frames;
  a.frame_stride = fo.rain_frame_stride;
  a.rain_pitch = fo.rain_row_pitch;
  a.clock = clk;
  a.plan = make_plan(*p, fo);
  return GS_OK;
}

// Workgroups of a one-wavefront-per-item kernel that are resident at once on the current
// device (cached per device and kernel); the grid never exceeds the work.
template <typename F>
static unsigned resident_grid(F kern, std::atomic<int>* cache, int64_t items) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) dev = -1;
  int per = (dev >= 0 && dev < 64) ? cache[dev].load(std::memory_order_relaxed) : 0;
  if (per <= 0) {
    int occ = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, kFBlock, 0) != hipSuccess) occ = 4;
    if (dev < 0 || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) cus = 256;
    per = (occ > 0 ? occ : 1) * (cus > 0 ? cus : 1);
    if (dev >= 0 && dev < 64) cache[dev].store(per, std::memory_order_relaxed);
  }
  const int64_t need = (items + kFBlock / 64 - 1) / (kFBlock / 64);
  return (unsigned)(need < per ? (need > 0 ? need : 1) : per);
}

template <typename T, int MODE>
static unsigned flat_grid(int64_t nchunks) {
  static std::atomic<int> cache[64];
  return resident_grid(k_flat_step<T, MODE>, cache, nchunks);
}

// The sweeps launch one wavefront per tile (four tiles per workgroup) when the tiles fill
// the resident grid more than four times: the hardware then keeps every SIMD fed until the
// last tile, where a resident grid striding over the tiles left a partial last round (and
// measured 10-14 % slower on the coastal case, 4.9 rounds, and the full 640 M-cell grid,
// 23). Fewer tiles run on the resident grid: one wavefront per tile from one resident grid
// up measured 9 % faster on a 2048^2 domain (1.3 rounds) but 4.5 % slower summed over the
// 64 ensemble members' plain runs (mostly shorter tiles of smaller domains).
template <typename K>
static unsigned sweep_grid(K kern, std::atomic<int>* cache, int64_t ntiles) {
  const unsigned resident = resident_grid(kern, cache, ntiles);
  const int64_t need = (ntiles + kFBlock / 64 - 1) / (kFBlock / 64);
  if (need <= 4 * (int64_t)resident) return resident;
  return (unsigned)need;
}

template <typename T, bool PRE, bool LEAN = false>
static unsigned sweep_grid_for(int64_t ntiles) {
  static std::atomic<int> cache[64];
  return sweep_grid(k_flat_sweep<T, PRE, LEAN>, cache, ntiles);
}

// The domain may take the LEAN sweep: nothing it would skip is set.
template <typename T>
static bool flat_lean(const FlatK<T>& a) {
  return !a.n_field && !a.cls && !a.max_h && !a.sponge && !a.clamp && !a.infil && !a.drain;
}

template <typename T>
static unsigned sweep_ext_grid(int64_t ntiles) {
  static std::atomic<int> cache[64];
  return sweep_grid(k_flat_sweep_ext<T>, cache, ntiles);
}

// The launch gs_flat_step makes for `a` (without the pre-friction state): which kernel, its
// grid and the kernel's resident grid on this device.
template <typename T>
static void flat_launch_choice(const FlatK<T>& a, int* kernel, int64_t* blocks, int64_t* resident) {
  static std::atomic<int> c_plain[64], c_ext[64];
  if (a.n_tiles <= 0) {
    *kernel = 0;
    *blocks = flat_grid<T, 1>(a.nchunks);
    *resident = *blocks;
  } else if (a.ga_ks || a.ga_inv_tau || a.rain_mode == 3) {
    *kernel = 2;
    *blocks = sweep_ext_grid<T>(a.n_tiles);
    *resident = resident_grid(k_flat_sweep_ext<T>, c_ext, a.n_tiles);
  } else if (flat_lean(a)) {
    static std::atomic<int> c_lean[64];
    *kernel = 1;
    *blocks = sweep_grid_for<T, false, true>(a.n_tiles);
    *resident = resident_grid(k_flat_sweep<T, false, true>, c_lean, a.n_tiles);
  } else {
    *kernel = 1;
    *blocks = sweep_grid_for<T, false>(a.n_tiles);
    *resident = resident_grid(k_flat_sweep<T, false>, c_plain, a.n_tiles);
  }
}

// The fused step: row-sweep tiles when the descriptor has them, else chunks of the run table.
template <typename T>
static int flat_step_launch(const FlatK<T>& a, hipStream_t s) {
  if (a.n_rmode > 0) {
    const int64_t nb = (a.n_rmode + kFBlock - 1) / kFBlock;
    hipLaunchKernelGGL((k_flat_ring_mode<T>), dim3((unsigned)(nb < 1024 ? nb : 1024)), dim3(kFBlock), 0, s, a);
    GS_HIP(hipGetLastError(), "k_flat_ring_mode launch");
  }
  if (a.n_fring > 0) {
    const int64_t nb = (a.n_fring + kFBlock - 1) / kFBlock;
    hipLaunchKernelGGL((k_flat_ring_force<T>), dim3((unsigned)(nb < 1024 ? nb : 1024)), dim3(kFBlock), 0, s, a);
    GS_HIP(hipGetLastError(), "k_flat_ring_force launch");
  }
  if (a.n_tiles > 0 && a.pre[0]) {
    hipLaunchKernelGGL((k_flat_sweep<T, true>), dim3(sweep_grid_for<T, true>(a.n_tiles)), dim3(kFBlock), 0, s, a);
    GS_HIP(hipGetLastError(), "k_flat_sweep(pre) launch");
  } else if (a.n_tiles > 0 && (a.ga_ks || a.ga_inv_tau || a.rain_mode == 3)) {
    hipLaunchKernelGGL((k_flat_sweep_ext<T>), dim3(sweep_ext_grid<T>(a.n_tiles)), dim3(kFBlock), 0, s, a);
    GS_HIP(hipGetLastError(), "k_flat_sweep_ext launch");
  } else if (a.n_tiles > 0 && flat_lean(a)) {
    hipLaunchKernelGGL((k_flat_sweep<T, false, true>), dim3(sweep_grid_for<T, false, true>(a.n_tiles)), dim3(kFBlock), 0, s, a);
    GS_HIP(hipGetLastError(), "k_flat_sweep(lean) launch");
  } else if (a.n_tiles > 0) {
    hipLaunchKernelGGL((k_flat_sweep<T, false>), dim3(sweep_grid_for<T, false>(a.n_tiles)), dim3(kFBlock), 0, s, a);
    GS_HIP(hipGetLastError(), "k_flat_sweep launch");
  } else {
    hipLaunchKernelGGL((k_flat_step<T, 1>), dim3(flat_grid<T, 1>(a.nchunks)), dim3(kFBlock), 0, s, a);
    GS_HIP(hipGetLastError(), "k_flat_step launch");
  }
  return GS_OK;
}

template <typename T>
static int flat_step_t(const GsFlat* f, const GsPhys* p, GsForcing fo, GsClock* clk, hipStream_t s,
                       void* pre0 = nullptr, void* pre1 = nullptr, void* pre2 = nullptr) {
  FlatK<T> a;
  int rc = build_flat<T>(f, p, fo, clk, a, true);
  if (rc) return rc;
  if ((pre0 || pre1 || pre2) && !(pre0 && pre1 && pre2)) return fail(GS_E_ARG, "pre: all three planes or none");
  if (pre0 && a.n_tiles == 0)
    return fail(GS_E_ARG, "the pre-friction state is written by the row-sweep traversal only (n_tiles > 0)");
  a.pre[0] = (T*)pre0;
  a.pre[1] = (T*)pre1;
  a.pre[2] = (T*)pre2;
  return flat_step_launch<T>(a, s);
}

template <typename T>
static int flat_residual_t(const GsFlat* f, const GsPhys* p, int which, void* r0, void* r1, void* r2,
                           hipStream_t s) {
  GsForcing fo;
  memset(&fo, 0, sizeof fo);
  FlatK<T> a;
  int rc = build_flat<T>(f, p, fo, nullptr, a, false);
  if (rc) return rc;
  if (which < 0 || which > 1) return fail(GS_E_ARG, "which must be 0 or 1");
  if (!r0 || !r1 || !r2) return fail(GS_E_ARG, "residual arrays are NULL");
  a.which = which;
  a.rhs[0] = (T*)r0;
  a.rhs[1] = (T*)r1;
  a.rhs[2] = (T*)r2;
  hipLaunchKernelGGL((k_flat_step<T, 0>), dim3(flat_grid<T, 0>(a.nchunks)), dim3(kFBlock), 0, s, a);
  GS_HIP(hipGetLastError(), "k_flat_step(residual) launch");
  return GS_OK;
}

template <typename T>
static int flat_cfl_t(const GsFlat* f, const GsPhys* p, GsForcing fo, GsClock* clk, hipStream_t s) {
  FlatK<T> a;
  int rc = build_flat<T>(f, p, fo, clk, a, true);
  if (rc) return rc;
  int64_t blocks = ((int64_t)a.n_runs + 3) / 4;
  if (blocks < 1) blocks = 1;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL((k_flat_cfl<T>), dim3((unsigned)blocks), dim3(kFBlock), 0, s, a);
  GS_HIP(hipGetLastError(), "k_flat_cfl launch");
  return GS_OK;
}


// Depth of the current state at listed flat ids (gauges), read after a step.
template <typename T>
__global__ void k_flat_sample(FlatK<T> a, const int64_t* ids, int n, T* out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int par = (int)(a.clock->nsteps & 1);
  out[t] = a.buf[par][0][ids[t]];
}

// ---------------------------------------------------------------------------
// batched ensemble: every member's sweep tiles in one launch, per-member clocks
// ---------------------------------------------------------------------------
template <typename T>
struct EnsK {
  const FlatK<T>* __restrict__ mem;   // member views (device)
  const int64_t* __restrict__ tile0;  // [n_members + 1] first global tile of each member
  int n_members;
  int tpp;                            // consecutive tiles per wavefront pass
  int64_t ntiles;                     // total tiles
  unsigned* done;                     // workgroups finished in this launch
};

// Read a value that the launch never writes through the constant address space: a
// uniform address then gives scalar loads. The host pass sees an ordinary read.
template <typename U>
__device__ __forceinline__ U const_read(const U* p) {
#if defined(__HIP_DEVICE_COMPILE__)
  return *(const __attribute__((address_space(4))) U*)p;
#else
  return *p;
#endif
}

// Fold a wavefront's CFL speed bits of member m into that member's clock (lane 0 waits
// until its atomic has been performed, as grid_reduce_and_plan's thread 0 does).
template <typename T>
__device__ __forceinline__ void ens_flush(const EnsK<T>& e, int m, uint64_t lam) {
  lam = wave_max_u64(lam);
  if (m >= 0 && (threadIdx.x & 63) == 0 && lam)
    await_atomic(__hip_atomic_fetch_max(&e.mem[m].clock->lam_acc, lam, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

// The member that owns global tile gc (tile0: the members' tile prefix, n + 1 entries).
__device__ __forceinline__ int ens_member_of(const int64_t* tile0, int n, int64_t gc) {
  int lo = 0, hi = n;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (const_read(tile0 + mid) <= gc) lo = mid;
    else hi = mid;
  }
  return lo;
}

// One step of every live member. The global tile space is the members' sweep tiles end to
// end; a wavefront takes `tpp` consecutive tiles per pass and strides by the grid (the
// autotuner usually picks one wavefront per tile, as the plain sweep launches), finding a
// tile's member by bisection of the tile prefix and reading the member's view and clock
// values from the member table with scalar loads. A wavefront folds its speed maximum into
// a member's clock when it leaves the member; its last maximum goes through the
// workgroup's thread 0, which folds the workgroup's maxima (one atomic per member) before
// counting the workgroup done. The last workgroup to finish plans each member's next dt
// from its own maximum (plan_finish), as the plain step's last workgroup does for its
// single domain; the hand-off orders atomics only, as grid_reduce_and_plan's does.
// KIND kEnsExt compiles the extended forcings (Green-Ampt, drain-tau, rain through a
// per-cell lookup) in; kEnsPlain leaves them out; kEnsLean also leaves out the per-cell
// Manning field, the running maximum and the active-cell forcings (every member has none
// of them), and its float32 build is held to 72 VGPRs (7 wavefronts per SIMD: 1.81 against
// 1.93 ms per step on the 64-member case at 82 VGPRs, though 16 VGPRs then spill, a few of
// them reloaded in the row loop).
enum { kEnsPlain = 0, kEnsExt = 1, kEnsLean = 2 };

template <typename T, int KIND>
__global__ __launch_bounds__(kFBlock) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 && KIND == kEnsLean ? 7 : 1)))
void k_ens_step(EnsK<T> e, const FlatK<T>* __restrict__ mem, const int64_t* __restrict__ tile0) {
  const int lane = threadIdx.x & 63;
  const int64_t nwaves = (int64_t)gridDim.x * (kFBlock / 64);
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int64_t wave = (int64_t)xcd_rank(blockIdx.x, gridDim.x) * (kFBlock / 64) + wv;
  const int64_t tpp = e.tpp;
  int m = -1;          // the member whose maximum `lam` holds
  lam_t<T> lam = 0;
  for (int64_t base = wave * tpp; base < e.ntiles; base += nwaves * tpp) {
    const int64_t stop = base + tpp < e.ntiles ? base + tpp : e.ntiles;
    for (int64_t gc = base; gc < stop; ++gc) {
      // the tile's member, its view and clock values, looked up per tile so that none of
      // them is held across tiles
      const int mt = ens_member_of(tile0, e.n_members, gc);
      if (mt != m) {
        ens_flush(e, m, lam_key_bits64(lam));
        lam = 0;
        m = mt;
      }
      const FlatK<T> v = const_read(mem + mt);
      const FlatClk<T> ck = flat_clock<T, 1>(v);
      if (!ck.live) continue;
      FlatCtx<T> x;
      flat_ctx<T, 1>(v, ck, x);
      flat_tile<T, false, KIND == kEnsExt, KIND == kEnsLean>(x, gc - const_read(tile0 + mt), lane, lam);
    }
  }
  __shared__ int s_m[kFBlock / 64];
  __shared__ uint64_t s_l[kFBlock / 64];
  __shared__ int s_last;
  const uint64_t lw = wave_max_u64(lam_key_bits64(lam));
  if (lane == 0) {
    s_m[wv] = m;
    s_l[wv] = lw;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 0; w < kFBlock / 64; ++w) {
      const int mw = s_m[w];
      const uint64_t l = s_l[w];
      if (mw < 0 || !l) continue;
      if (w + 1 < kFBlock / 64 && s_m[w + 1] == mw) {   // the next wave's member: fold into its slot
        s_l[w + 1] = s_l[w + 1] > l ? s_l[w + 1] : l;
        continue;
      }
      await_atomic(__hip_atomic_fetch_max(&e.mem[mw].clock->lam_acc, l, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    }
    const unsigned prev = __hip_atomic_fetch_add(e.done, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s_last = (prev == gridDim.x - 1) ? 1 : 0;
  }
  __syncthreads();
  if (!s_last) return;
  for (int mm = threadIdx.x; mm < e.n_members; mm += blockDim.x) {
    GsClock* c = e.mem[mm].clock;
    if (c->status & GS_ST_DONE) continue;
    const uint64_t tot = __hip_atomic_load(&c->lam_acc, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    plan_finish(c, e.mem[mm].plan, tot, true);
  }
  if (threadIdx.x == 0) __hip_atomic_store(e.done, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename T>
static int ens_occupancy(int kind) {
  int occ = 0;
  const hipError_t h =
      kind == kEnsExt    ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_ens_step<T, kEnsExt>, kFBlock, 0)
      : kind == kEnsLean ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_ens_step<T, kEnsLean>, kFBlock, 0)
                         : hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_ens_step<T, kEnsPlain>, kFBlock, 0);
  if (h != hipSuccess || occ < 1) occ = 1;
  return occ;
}

// The instantiation the members need: extended if any member has an extended forcing, lean
// if every member may take the lean sweep, else plain.
template <typename T>
static int ens_kind(const FlatK<T>* views, int n) {
  bool lean = true;
  for (int m = 0; m < n; ++m) {
    if (views[m].ga_ks || views[m].ga_inv_tau || views[m].rain_mode == 3) return kEnsExt;
    lean = lean && flat_lean(views[m]);
  }
  return lean ? kEnsLean : kEnsPlain;
}

template <typename T>
static void ens_launch_kind(int kind, int64_t blocks, hipStream_t s, const EnsK<T>& k) {
  if (kind == kEnsExt)
    hipLaunchKernelGGL((k_ens_step<T, kEnsExt>), dim3((unsigned)blocks), dim3(kFBlock), 0, s, k, k.mem, k.tile0);
  else if (kind == kEnsLean)
    hipLaunchKernelGGL((k_ens_step<T, kEnsLean>), dim3((unsigned)blocks), dim3(kFBlock), 0, s, k, k.mem, k.tile0);
  else
    hipLaunchKernelGGL((k_ens_step<T, kEnsPlain>), dim3((unsigned)blocks), dim3(kFBlock), 0, s, k, k.mem, k.tile0);
}

}  // namespace gs

using namespace gs;

extern "C" {

int gs_flat_step(const GsFlat* f, const GsPhys* phys, GsForcing forc, GsClock* clock_dev, void* stream) {
  if (!f) return fail(GS_E_ARG, "flat descriptor is NULL");
  if (f->dtype == GS_F32) return flat_step_t<float>(f, phys, forc, clock_dev, (hipStream_t)stream);
  if (f->dtype == GS_F64) return flat_step_t<double>(f, phys, forc, clock_dev, (hipStream_t)stream);
  return fail(GS_E_DTYPE, "unsupported dtype %d", f->dtype);
}

int gs_flat_residual(const GsFlat* f, const GsPhys* phys, int which, void* rhs0, void* rhs1, void* rhs2,
                     void* stream) {
  if (!f) return fail(GS_E_ARG, "flat descriptor is NULL");
  if (f->dtype == GS_F32) return flat_residual_t<float>(f, phys, which, rhs0, rhs1, rhs2, (hipStream_t)stream);
  if (f->dtype == GS_F64) return flat_residual_t<double>(f, phys, which, rhs0, rhs1, rhs2, (hipStream_t)stream);
  return fail(GS_E_DTYPE, "unsupported dtype %d", f->dtype);
}

int gs_flat_cfl(const GsFlat* f, const GsPhys* phys, GsForcing forc, GsClock* clock_dev, void* stream) {
  if (!f) return fail(GS_E_ARG, "flat descriptor is NULL");
  if (f->dtype == GS_F32) return flat_cfl_t<float>(f, phys, forc, clock_dev, (hipStream_t)stream);
  if (f->dtype == GS_F64) return flat_cfl_t<double>(f, phys, forc, clock_dev, (hipStream_t)stream);
  return fail(GS_E_DTYPE, "unsupported dtype %d", f->dtype);
}

int gs_flat_launch_info(const GsFlat* f, const GsPhys* phys, GsForcing forc, int* kernel, int64_t* blocks,
                        int64_t* resident) {
  using namespace gs;
  if (!f || !kernel || !blocks || !resident) return fail(GS_E_ARG, "gs_flat_launch_info: NULL argument");
  if (f->dtype == GS_F32) {
    FlatK<float> a;
    const int rc = build_flat<float>(f, phys, forc, nullptr, a, false);
    if (rc) return rc;
    flat_launch_choice<float>(a, kernel, blocks, resident);
  } else if (f->dtype == GS_F64) {
    FlatK<double> a;
    const int rc = build_flat<double>(f, phys, forc, nullptr, a, false);
    if (rc) return rc;
    flat_launch_choice<double>(a, kernel, blocks, resident);
  } else {
    return fail(GS_E_DTYPE, "unsupported dtype %d", f->dtype);
  }
  return GS_OK;
}

int gs_flat_step_pre(const GsFlat* f, const GsPhys* phys, GsForcing forc, GsClock* clock_dev, void* pre0,
                     void* pre1, void* pre2, void* stream) {
  using namespace gs;
  if (!pre0 || !pre1 || !pre2) return fail(GS_E_ARG, "pre arrays are NULL");
  if (!f) return fail(GS_E_ARG, "flat descriptor is NULL");
  if (f->dtype == GS_F64)
    return flat_step_t<double>(f, phys, forc, clock_dev, (hipStream_t)stream, pre0, pre1, pre2);
  if (f->dtype == GS_F32)
    return flat_step_t<float>(f, phys, forc, clock_dev, (hipStream_t)stream, pre0, pre1, pre2);
  return fail(GS_E_DTYPE, "unsupported dtype %d", f->dtype);
}

int gs_flat_sample(const GsFlat* f, const GsClock* clock_dev, const int64_t* ids, int n, void* out, void* stream) {
  using namespace gs;
  if (!f || !clock_dev || (n > 0 && (!ids || !out))) return fail(GS_E_ARG, "NULL argument");
  if (n < 0) return fail(GS_E_ARG, "n < 0");
  if (n == 0) return GS_OK;
  GsPhys p;
  memset(&p, 0, sizeof p);
  p.g = 9.81; p.h_min = 1e-6; p.h_cfl = 1e-6; p.cfl = 0.5; p.dmin = 1.0; p.n_classes = 1;
  GsForcing fo;
  memset(&fo, 0, sizeof fo);
  const unsigned nb = (unsigned)((n + 255) / 256);
  if (f->dtype == GS_F64) {
    FlatK<double> a;
    int rc = build_flat<double>(f, &p, fo, (GsClock*)clock_dev, a, true);
    if (rc) return rc;
    hipLaunchKernelGGL(k_flat_sample<double>, dim3(nb), dim3(256), 0, (hipStream_t)stream, a, ids, n, (double*)out);
  } else {
    FlatK<float> a;
    int rc = build_flat<float>(f, &p, fo, (GsClock*)clock_dev, a, true);
    if (rc) return rc;
    hipLaunchKernelGGL(k_flat_sample<float>, dim3(nb), dim3(256), 0, (hipStream_t)stream, a, ids, n, (float*)out);
  }
  GS_HIP(hipGetLastError(), "k_flat_sample launch");
  return GS_OK;
}


}  // extern "C"

// ---------------------------------------------------------------------------
// ensemble host object
// ---------------------------------------------------------------------------
struct GsEns {
  int dtype = 0;
  int n = 0;
  void* mem = nullptr;          // device FlatK<T>[n]
  int64_t* tile0 = nullptr;     // device [n + 1]
  unsigned* done = nullptr;     // device counter
  int64_t ntiles = 0;
  int blocks = 0;               // 0: one occupancy wave of the device
  int tpp = 1;
  int cus = 256;
  int occ = 1;                  // resident k_ens_step workgroups per CU
  int kind = 0;                 // k_ens_step instantiation (kEnsPlain / kEnsExt / kEnsLean)
  std::vector<GsFlat> flat;
  std::vector<GsClock*> clocks;
  hipStream_t gstream = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int g_batch = 0, g_blocks = -1, g_tpp = -1;
};

namespace gs {

// Member m's view for k_ens_step, with the checks of what the ensemble step supports.
template <typename T>
static int ens_member_view(const GsFlat* members, const GsPhys* phys, const GsForcing* forcings,
                           GsClock* const* clocks, int m, FlatK<T>& v) {
  const int rc = build_flat<T>(&members[m], &phys[m], forcings[m], clocks[m], v, true);
  if (rc) {
    char msg[512];
    snprintf(msg, sizeof msg, "%s", g_err);
    return fail(rc, "member %d: %s", m, msg);
  }
  if (v.n_runs > 0 && v.n_tiles < 1) return fail(GS_E_ARG, "member %d has active cells but no row-sweep tiles", m);
  if (v.n_rmode > 0) return fail(GS_E_ARG, "member %d has ring modes; the ensemble step does not apply them", m);
  if (v.n_fring > 0)
    return fail(GS_E_ARG, "member %d forces ring cells (gauge ring, or sponge or clamp outside the active set); the ensemble "
                "step updates active cells only", m);
  return GS_OK;
}

// Grid of one k_ens_step launch: `blocks` resident workgroups (0: one occupancy wave of the
// device), never more than the tiles need.
static int64_t ens_grid(int64_t blocks, int occ, int cus, int64_t ntiles, int tpp) {
  if (blocks <= 0) blocks = (int64_t)occ * cus;
  const int64_t need = (ntiles + 4LL * tpp - 1) / (4LL * tpp);
  if (blocks > need) blocks = need;
  return blocks < 1 ? 1 : blocks;
}

template <typename T>
static int ens_build_views(GsEns* e, const GsFlat* members, const GsPhys* phys, const GsForcing* forcings) {
  std::vector<FlatK<T>> views((size_t)e->n);
  std::vector<int64_t> c0((size_t)e->n + 1, 0);
  for (int m = 0; m < e->n; ++m) {
    const int rc = ens_member_view<T>(members, phys, forcings, e->clocks.data(), m, views[m]);
    if (rc) return rc;
    c0[m + 1] = c0[m] + views[m].n_tiles;
  }
  e->ntiles = c0[e->n];
  e->kind = ens_kind<T>(views.data(), e->n);
  e->occ = ens_occupancy<T>(e->kind);
  GS_HIP(hipMalloc(&e->mem, sizeof(FlatK<T>) * (size_t)e->n), "ensemble views");
  GS_HIP(hipMalloc((void**)&e->tile0, sizeof(int64_t) * ((size_t)e->n + 1)), "ensemble tile table");
  GS_HIP(hipMalloc((void**)&e->done, 256), "ensemble counter");
  GS_HIP(hipMemcpy(e->mem, views.data(), sizeof(FlatK<T>) * (size_t)e->n, hipMemcpyHostToDevice), "views h2d");
  GS_HIP(hipMemcpy(e->tile0, c0.data(), sizeof(int64_t) * ((size_t)e->n + 1), hipMemcpyHostToDevice),
         "tile table h2d");
  GS_HIP(hipMemset(e->done, 0, 256), "counter clear");
  return GS_OK;
}

template <typename T>
static int ens_launch(GsEns* e, hipStream_t s) {
  EnsK<T> k;
  k.mem = (const FlatK<T>*)e->mem;
  k.tile0 = e->tile0;
  k.n_members = e->n;
  k.tpp = e->tpp;
  k.ntiles = e->ntiles;
  k.done = e->done;
  const int64_t blocks = ens_grid(e->blocks, e->occ, e->cus, e->ntiles, e->tpp);
  ens_launch_kind<T>(e->kind, blocks, s, k);
  GS_HIP(hipGetLastError(), "k_ens_step launch");
  return GS_OK;
}

static int ens_step(GsEns* e, hipStream_t s) {
  return e->dtype == GS_F64 ? ens_launch<double>(e, s) : ens_launch<float>(e, s);
}

static void ens_drop_graph(GsEns* e) {
  if (e->exec) (void)hipGraphExecDestroy(e->exec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  e->exec = nullptr;
  e->graph = nullptr;
  e->g_batch = 0;
}

static int ens_capture(GsEns* e, int batch) {
  ens_drop_graph(e);
  GS_HIP(hipStreamBeginCapture(e->gstream, hipStreamCaptureModeThreadLocal), "begin capture");
  int rc = GS_OK;
  for (int i = 0; i < batch && rc == GS_OK; ++i) rc = ens_step(e, e->gstream);
  hipGraph_t g = nullptr;
  const hipError_t ee = hipStreamEndCapture(e->gstream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (ee != hipSuccess) return hip_fail(ee, "end capture");
  e->graph = g;
  GS_HIP(hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0), "graph instantiate");
  e->g_batch = batch;
  e->g_blocks = e->blocks;
  e->g_tpp = e->tpp;
  return GS_OK;
}

}  // namespace gs

extern "C" {

int gs_ens_create(const GsFlat* members, const GsPhys* phys, const GsForcing* forcings, GsClock* const* clocks,
                  int n_members, GsEns** out) {
  using namespace gs;
  if (!out) return fail(GS_E_ARG, "out is NULL");
  *out = nullptr;
  if (n_members < 1 || n_members > (1 << 20)) return fail(GS_E_RANGE, "n_members must be in 1..2^20");
  if (!members || !phys || !forcings || !clocks) return fail(GS_E_ARG, "member arrays are NULL");
  const int dt = members[0].dtype;
  if (dt != GS_F32 && dt != GS_F64) return fail(GS_E_DTYPE, "unsupported dtype %d", dt);
  for (int m = 0; m < n_members; ++m) {
    if (members[m].dtype != dt) return fail(GS_E_DTYPE, "member %d dtype differs from member 0", m);
    if (!clocks[m]) return fail(GS_E_ARG, "clock of member %d is NULL", m);
  }
  GsEns* e = new GsEns();
  e->dtype = dt;
  e->n = n_members;
  e->flat.assign(members, members + n_members);
  e->clocks.assign(clocks, clocks + n_members);
  int dev = 0;
  if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&e->cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e->cus < 1) e->cus = 256;
  int rc = dt == GS_F64 ? ens_build_views<double>(e, members, phys, forcings)
                        : ens_build_views<float>(e, members, phys, forcings);
  if (rc == GS_OK) {
    hipError_t h = hipStreamCreateWithFlags(&e->gstream, hipStreamNonBlocking);
    if (h == hipSuccess) h = hipEventCreateWithFlags(&e->ev_in, hipEventDisableTiming);
    if (h == hipSuccess) h = hipEventCreateWithFlags(&e->ev_out, hipEventDisableTiming);
    if (h != hipSuccess) rc = hip_fail(h, "ensemble stream/events");
  }
  if (rc) {
    gs_ens_destroy(e);
    return rc;
  }
  *out = e;
  return GS_OK;
}

int gs_ens_destroy(GsEns* e) {
  if (!e) return GS_OK;
  gs::ens_drop_graph(e);
  if (e->gstream) (void)hipStreamSynchronize(e->gstream);
  if (e->ev_in) (void)hipEventDestroy(e->ev_in);
  if (e->ev_out) (void)hipEventDestroy(e->ev_out);
  if (e->gstream) (void)hipStreamDestroy(e->gstream);
  if (e->mem) (void)hipFree(e->mem);
  if (e->tile0) (void)hipFree(e->tile0);
  if (e->done) (void)hipFree(e->done);
  delete e;
  return GS_OK;
}

int gs_ens_set_launch(GsEns* e, int blocks, int tiles_per_pass) {
  using namespace gs;
  if (!e) return fail(GS_E_ARG, "ensemble is NULL");
  if (blocks < 0 || blocks > (1 << 24)) return fail(GS_E_RANGE, "blocks must be in 0..2^24");
  if (tiles_per_pass < 1 || tiles_per_pass > 64) return fail(GS_E_RANGE, "tiles_per_pass must be in 1..64");
  e->blocks = blocks;
  e->tpp = tiles_per_pass;
  return GS_OK;
}

int gs_ens_get_launch(const GsEns* e, int* blocks, int* tiles_per_pass, int64_t* total_tiles) {
  using namespace gs;
  if (!e) return fail(GS_E_ARG, "ensemble is NULL");
  if (blocks) {
    int b = e->blocks;
    if (b == 0) b = e->occ * e->cus;
    *blocks = b;
  }
  if (tiles_per_pass) *tiles_per_pass = e->tpp;
  if (total_tiles) *total_tiles = e->ntiles;
  return GS_OK;
}

int gs_ens_step(GsEns* e, void* stream) {
  using namespace gs;
  if (!e) return fail(GS_E_ARG, "ensemble is NULL");
  return ens_step(e, (hipStream_t)stream);
}

int gs_ens_run(GsEns* e, int64_t nsteps, int graph_batch, void* stream) {
  using namespace gs;
  if (!e) return fail(GS_E_ARG, "ensemble is NULL");
  if (nsteps < 0) return fail(GS_E_ARG, "nsteps < 0");
  hipStream_t s = (hipStream_t)stream;
  if (graph_batch <= 0 || nsteps < graph_batch) {
    for (int64_t i = 0; i < nsteps; ++i) {
      const int rc = ens_step(e, s);
      if (rc) return rc;
    }
    return GS_OK;
  }
  if (graph_batch > 100000) return fail(GS_E_RANGE, "graph_batch too large");
  GS_HIP(hipEventRecord(e->ev_in, s), "event record");
  GS_HIP(hipStreamWaitEvent(e->gstream, e->ev_in, 0), "stream wait");
  if (!e->exec || e->g_batch != graph_batch || e->g_blocks != e->blocks || e->g_tpp != e->tpp) {
    const int rc = ens_capture(e, graph_batch);
    if (rc) return rc;
  }
  const int64_t reps = nsteps / graph_batch, rest = nsteps % graph_batch;
  for (int64_t r = 0; r < reps; ++r) GS_HIP(hipGraphLaunch(e->exec, e->gstream), "graph launch");
  for (int64_t i = 0; i < rest; ++i) {
    const int rc = ens_step(e, e->gstream);
    if (rc) return rc;
  }
  GS_HIP(hipEventRecord(e->ev_out, e->gstream), "event record");
  GS_HIP(hipStreamWaitEvent(s, e->ev_out, 0), "stream wait");
  return GS_OK;
}

int gs_ens_autotune(GsEns* e, int trial_steps, void* stream, double* best_ms) {
  using namespace gs;
  if (!e) return fail(GS_E_ARG, "ensemble is NULL");
  if (trial_steps < 1 || trial_steps > 10000) return fail(GS_E_RANGE, "trial_steps must be in 1..10000");
  hipStream_t s = (hipStream_t)stream;
  const size_t es = e->dtype == GS_F64 ? 8 : 4;
  GS_HIP(hipStreamSynchronize(s), "sync");
  // snapshot: the current state buffer, the running max, the Green-Ampt cumulative
  // infiltration and the clock of every member (everything a step writes)
  constexpr int kSnap = 5;
  auto plane = [&](int m, int par, int k) -> void* {
    return k < 3 ? e->flat[m].q[par][k] : (k == 3 ? e->flat[m].max_h : e->flat[m].ga_F);
  };
  std::vector<GsClock> clk((size_t)e->n);
  std::vector<void*> snap;
  int rc = GS_OK;
  auto cleanup = [&]() {
    for (void* p : snap) (void)hipFree(p);
    snap.clear();
  };
  for (int m = 0; m < e->n && rc == GS_OK; ++m) {
    hipError_t h = hipMemcpy(&clk[m], e->clocks[m], sizeof(GsClock), hipMemcpyDeviceToHost);
    if (h != hipSuccess) { rc = hip_fail(h, "clock d2h"); break; }
    const int par = (int)(clk[m].nsteps & 1);
    const size_t bytes = (size_t)e->flat[m].n_stored * es;
    for (int k = 0; k < kSnap && rc == GS_OK; ++k) {
      const void* src = plane(m, par, k);
      void* p = nullptr;
      if (src) {
        h = hipMalloc(&p, bytes);
        if (h == hipSuccess) h = hipMemcpy(p, src, bytes, hipMemcpyDeviceToDevice);
        if (h != hipSuccess) rc = hip_fail(h, "autotune snapshot");
      }
      snap.push_back(p);
    }
  }
  if (rc) { cleanup(); return rc; }
  auto restore = [&]() -> int {
    for (int m = 0; m < e->n; ++m) {
      const int par = (int)(clk[m].nsteps & 1);
      const size_t bytes = (size_t)e->flat[m].n_stored * es;
      for (int k = 0; k < kSnap; ++k) {
        void* dst = plane(m, par, k);
        void* p = snap[(size_t)m * kSnap + k];
        if (dst && p) GS_HIP(hipMemcpyAsync(dst, p, bytes, hipMemcpyDeviceToDevice, s), "restore");
      }
      GS_HIP(hipMemcpyAsync(e->clocks[m], &clk[m], sizeof(GsClock), hipMemcpyHostToDevice, s), "clock restore");
    }
    GS_HIP(hipStreamSynchronize(s), "sync");
    return GS_OK;
  };
  const int occ = e->occ;
  // resident grids of 1, 2, occ/2 and occ workgroups per CU, and (last) one wavefront per
  // tile, which the launch reaches by capping a larger grid at the tiles' need
  int per_cu[5] = {1, 2, occ > 1 ? occ / 2 : 1, occ, 0};
  const int64_t per_tile = (e->ntiles + 3) / 4;
  const int tpps[2] = {1, 2};
  hipEvent_t t0 = nullptr, t1 = nullptr;
  (void)hipEventCreate(&t0);
  (void)hipEventCreate(&t1);
  double best = 1e300;
  int best_b = e->blocks, best_c = e->tpp;
  for (int a = 0; a < 5 && rc == GS_OK; ++a) {
    bool dup = false;
    for (int b = 0; b < a; ++b) dup = dup || per_cu[b] == per_cu[a];
    if (a < 4 && (dup || per_cu[a] > occ)) continue;
    if (a == 4 && per_tile <= (int64_t)occ * e->cus) continue;   // already a resident grid
    for (int c = 0; c < 2 && rc == GS_OK; ++c) {
      if (a == 4 && c > 0) continue;
      e->blocks = a < 4 ? per_cu[a] * e->cus : (int)(per_tile < 0x7fffffff ? per_tile : 0x7fffffff);
      e->tpp = tpps[c];
      rc = restore();
      if (!rc) rc = ens_step(e, s);   // warm-up
      if (!rc && hipEventRecord(t0, s) != hipSuccess) rc = fail(GS_E_STATE, "event record failed");
      for (int i = 0; i < trial_steps && !rc; ++i) rc = ens_step(e, s);
      if (!rc && hipEventRecord(t1, s) != hipSuccess) rc = fail(GS_E_STATE, "event record failed");
      if (!rc && hipEventSynchronize(t1) != hipSuccess) rc = fail(GS_E_STATE, "event sync failed");
      float ms = 0.0f;
      if (!rc && hipEventElapsedTime(&ms, t0, t1) == hipSuccess && ms / trial_steps < best) {
        best = ms / trial_steps;
        best_b = e->blocks;
        best_c = e->tpp;
      }
    }
  }
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  e->blocks = best_b;
  e->tpp = best_c;
  const int rc2 = restore();
  cleanup();
  if (rc) return rc;
  if (rc2) return rc2;
  if (best_ms) *best_ms = best;
  return GS_OK;
}


}  // extern "C"

// ---------------------------------------------------------------------------
// the ensemble run with no allocation
// ---------------------------------------------------------------------------
// The device memory k_ens_step needs beyond the members' own buffers (the last-workgroup
// counter, the tile prefix and the member views) for gs_ens_run_noalloc, which allocates
// nothing: layout [counter, 256 B][tile prefix (n + 1) x int64, padded to 256 B][views].
// Every device holds its own copy at its own address (looked up per device); host code uses
// that address, never the array's name.
constexpr size_t kScratchBytes = 256 * 1024;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace gs {

constexpr int kScratchDevices = 64;
static std::mutex g_scratch_mu;                    // guards the table and the host staging copy
static void* g_scratch_dev[kScratchDevices];       // scratch's address on each device
alignas(256) static unsigned char g_scratch_host[kScratchBytes];   // what a call copies into scratch

// The address of `scratch` on the current device; the caller holds g_scratch_mu.
static int scratch_addr(void** out) {
  int dev = 0;
  GS_HIP(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= kScratchDevices) return fail(GS_E_RANGE, "device %d: the scratch table covers %d devices", dev,
                                                     kScratchDevices);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    GS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!p) return fail(GS_E_STATE, "hipGetSymbolAddress(scratch) returned NULL");
    g_scratch_dev[dev] = p;
  }
  *out = g_scratch_dev[dev];
  return GS_OK;
}

constexpr size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Bytes of scratch the no-allocation run needs for n members (0 if they do not fit).
template <typename T>
static size_t noalloc_bytes(int n) {
  const size_t need = align256(256 + 8 * ((size_t)n + 1)) + sizeof(FlatK<T>) * (size_t)n;
  return need <= kScratchBytes ? need : 0;
}

template <typename T>
static int ens_run_noalloc_t(const GsFlat* members, const GsPhys* phys, const GsForcing* forcings,
                             GsClock* const* clocks, int n, int64_t nsteps, hipStream_t s) {
  const size_t need = noalloc_bytes<T>(n);
  if (!need) {
    int cap = 0;
    while (noalloc_bytes<T>(cap + 1)) ++cap;
    return fail(GS_E_RANGE, "gs_ens_run_noalloc: %d members exceed the %zu-byte scratch (at most %d members of "
                "this dtype)", n, kScratchBytes, cap);
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  void* base = nullptr;
  int rc = scratch_addr(&base);
  if (rc) return rc;
  const size_t off_tile = 256, off_views = align256(256 + 8 * ((size_t)n + 1));
  memset(g_scratch_host, 0, off_views);            // the counter starts at zero
  int64_t* c0 = (int64_t*)(g_scratch_host + off_tile);
  FlatK<T>* views = (FlatK<T>*)(g_scratch_host + off_views);
  c0[0] = 0;
  for (int m = 0; m < n; ++m) {
    new (&views[m]) FlatK<T>();
    rc = ens_member_view<T>(members, phys, forcings, clocks, m, views[m]);
    if (rc) return rc;
    c0[m + 1] = c0[m] + views[m].n_tiles;
  }
  const int kind = ens_kind<T>(views, n);
  // steps of an earlier call on this stream may still read the scratch: let them finish
  // before it is overwritten (a synchronous copy, so the host staging is free on return)
  GS_HIP(hipStreamSynchronize(s), "hipStreamSynchronize");
  GS_HIP(hipMemcpy(base, g_scratch_host, need, hipMemcpyHostToDevice), "scratch h2d");
  int cus = 256, dev = 0;
  if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (cus < 1) cus = 256;
  EnsK<T> k;
  k.mem = (const FlatK<T>*)((unsigned char*)base + off_views);
  k.tile0 = (const int64_t*)((unsigned char*)base + off_tile);
  k.n_members = n;
  k.tpp = 1;
  k.ntiles = c0[n];
  k.done = (unsigned*)base;
  const int64_t blocks = ens_grid(0, ens_occupancy<T>(kind), cus, k.ntiles, 1);
  for (int64_t i = 0; i < nsteps; ++i) {
    ens_launch_kind<T>(kind, blocks, s, k);
    GS_HIP(hipGetLastError(), "k_ens_step launch");
  }
  return GS_OK;
}

}  // namespace gs

extern "C" {

int gs_ens_run_noalloc(const GsFlat* members, const GsPhys* phys, const GsForcing* forcings, GsClock* const* clocks,
                       int n_members, int64_t nsteps, void* stream) {
  using namespace gs;
  if (n_members < 1) return fail(GS_E_RANGE, "gs_ens_run_noalloc: n_members must be >= 1");
  if (!members || !phys || !forcings || !clocks) return fail(GS_E_ARG, "gs_ens_run_noalloc: member arrays are NULL");
  if (nsteps < 0) return fail(GS_E_ARG, "gs_ens_run_noalloc: nsteps < 0");
  const int dt = members[0].dtype;
  if (dt != GS_F32 && dt != GS_F64) return fail(GS_E_DTYPE, "unsupported dtype %d", dt);
  for (int m = 0; m < n_members; ++m) {
    if (members[m].dtype != dt) return fail(GS_E_DTYPE, "member %d dtype differs from member 0", m);
    if (!clocks[m]) return fail(GS_E_ARG, "clock of member %d is NULL", m);
  }
  hipStream_t s = (hipStream_t)stream;
  return dt == GS_F64 ? ens_run_noalloc_t<double>(members, phys, forcings, clocks, n_members, nsteps, s)
                      : ens_run_noalloc_t<float>(members, phys, forcings, clocks, n_members, nsteps, s);
}

}  // extern "C"
