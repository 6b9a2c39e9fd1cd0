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
};

template <typename T> __device__ __forceinline__ T fl_upper(T x) { return __shfl_down(x, 1, 64); }
template <typename T> __device__ __forceinline__ T fl_lower(T x) { return __shfl_up(x, 1, 64); }

template <typename T, int MODE>
__global__ __launch_bounds__(kFBlock) void k_flat_step(FlatK<T> a) {
  int par = 0;
  T dt = T(0), rain_u = T(0), eta = T(0);
  int ridx = 0;
  if (MODE == 1) {
    const GsClock* c = a.clock;
    if (c->status & GS_ST_DONE) return;
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
  __shared__ T s_ntab[32];
  if (MODE == 1 && threadIdx.x < (unsigned)a.n_classes) s_ntab[threadIdx.x] = a.ntab[threadIdx.x];
  __syncthreads();

  T* const* qin = MODE == 1 ? a.buf[par] : a.buf[a.which];
  T* const* qout = a.buf[par ^ 1];
  const Phys<T>& P = a.P;
  const T ix = P.inv_dx, iy = P.inv_dy, hdx = P.half_dx, hdy = P.half_dy, g = P.g, hm = P.h_min;
  const T* frame = (MODE == 1 && a.rain_mode == 2) ? a.rain_frames + (int64_t)ridx * a.frame_stride : nullptr;

  const int lane = threadIdx.x & 63;
  // Resident grid, each wavefront striding over chunks. Workgroups are dealt round-robin to
  // the 8 XCDs; the remap gives each XCD a contiguous range of wavefronts, hence of chunks
  // in every pass, so the rows i+-1 a chunk gathers are mostly in that XCD's L2.
  const unsigned nb = gridDim.x, b = blockIdx.x;
  const unsigned qd = nb / 8u, rm = nb % 8u, xcd = b % 8u, kk = b / 8u;
  const unsigned lb = (xcd < rm) ? xcd * (qd + 1u) + kk : rm * (qd + 1u) + (xcd - rm) * qd + kk;
  const int64_t nwaves = (int64_t)nb * (kFBlock / 64);
  uint64_t lam = 0ull;

  for (int64_t chunk = (int64_t)lb * (kFBlock / 64) + (threadIdx.x >> 6); chunk < a.nchunks; chunk += nwaves) {
    const int64_t k = chunk * kFOut + lane - 2;
    const int64_t kl = k < 0 ? 0 : (k >= a.n ? a.n - 1 : k);
    // this lane's cell (needed by every lane for the y-face shuffles)
    const T hC = qin[0][kl], huC = qin[1][kl], hvC = qin[2][kl];
    const T bC = a.bed[kl];
    T uC, vC;
    prims(hC, huC, hvC, hm, uC, vC);
    const T bl = fl_lower(bC), br = fl_upper(bC);
    const T gzy = (T(0.5) * (br - bl)) * iy;
    // upper y-face (k, k+1)
    const T hR = fl_upper(hC), uR = fl_upper(uC), vR = fl_upper(vC);
    const T gzyR = fl_upper(gzy);
    Face<T> fy;
    face<T>(bC, br, gzy, gzyR, hC, hR, vC, uC, vR, uR, hdy, g, hm, fy);
    const T yFhB = fl_lower(fy.Fh), yFNB = fl_lower(fy.FN), yFTB = fl_lower(fy.FT);
    const T ysRB = fl_lower(fy.srcR);
    const int ygB = fl_lower((int)fy.gate);

    // run of this cell
    int r = -1;
    if (lane >= 2 && lane < 2 + kFOut && k < a.n) {
      int rr = a.chunk_run[chunk];
      while (rr < a.n_runs && a.run_k0[rr] + a.run_len[rr] <= k) ++rr;
      if (rr < a.n_runs && a.run_k0[rr] <= k) r = rr;
    }
    if (r >= 0) {
      const int32_t* d = a.run_d + 4 * (int64_t)r;
      const int64_t kE = k + d[0], kW = k + d[1], kEE = k + d[2], kWW = k + d[3];
      const T bE = a.bed[kE], bW = a.bed[kW], bEE = a.bed[kEE], bWW = a.bed[kWW];
      const T hE = qin[0][kE], huE = qin[1][kE], hvE = qin[2][kE];
      const T hW = qin[0][kW], huW = qin[1][kW], hvW = qin[2][kW];
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
        a.rhs[0][k] = rh;
        a.rhs[1][k] = rhu;
        a.rhs[2][k] = rhv;
      } else {
        const uint8_t fg = a.flags ? a.flags[k] : (uint8_t)0;
        T rr0 = rh;
        if (a.rain_mode == 1) {
          rr0 = rr0 + rain_u;
        } else if (a.rain_mode == 2) {
          const int64_t ii = (int64_t)a.run_i[r] - a.ngh;
          const int64_t jj = (int64_t)a.run_j0[r] + (k - a.run_k0[r]) - a.ngh;
          rr0 = rr0 + frame[ii * a.rain_pitch + jj];
        }
        T n = T(0);
        if (P.friction) n = a.n_field ? a.n_field[k] : s_ntab[fg & GS_FLAG_CLASS_MASK];
        T h, hu, hv;
        update<T>(hC, huC, hvC, rr0, rhu, rhv, dt, n, P, h, hu, hv);
        if (fg & GS_FLAG_STAGE) {
          const T hs = eta - bC;
          h = hs > T(0) ? hs : T(0);
          hu = T(0);
          hv = T(0);
        }
        if (a.max_h && h > a.max_h[k]) a.max_h[k] = h;
        qout[0][k] = h;
        qout[1][k] = hu;
        qout[2][k] = hv;
        if (!(fg & GS_FLAG_CFL_SKIP)) {
          const uint64_t lb_ = lam_bits<T>(h, hu, hv, P);
          lam = lb_ > lam ? lb_ : lam;
        }
      }
    }
  }
  if (MODE == 1) grid_reduce_and_plan(lam, a.clock, a.plan, true);
}

// CFL wave speed of the current flat state over active cells, then plan dt.
template <typename T>
__global__ __launch_bounds__(kFBlock) void k_flat_cfl(FlatK<T> a) {
  const int par = (int)(a.clock->nsteps & 1);
  T* const* q = a.buf[par];
  uint64_t lam = 0ull;
  for (int64_t r = (int64_t)blockIdx.x * (kFBlock / 64) + (threadIdx.x >> 6); r < a.n_runs;
       r += (int64_t)gridDim.x * (kFBlock / 64)) {
    const int64_t k0 = a.run_k0[r];
    const int len = a.run_len[r];
    for (int t = threadIdx.x & 63; t < len; t += 64) {
      const int64_t k = k0 + t;
      if (a.flags && (a.flags[k] & GS_FLAG_CFL_SKIP)) continue;
      const uint64_t l = lam_bits<T>(q[0][k], q[1][k], q[2][k], a.P);
      lam = l > lam ? l : lam;
    }
  }
  grid_reduce_and_plan(lam, a.clock, a.plan, false);
}

template <typename T>
static int build_flat(const GsFlat* f, const GsPhys* p, const GsForcing& fo, GsClock* clk, FlatK<T>& a,
                      bool need_clock) {
  if (!f) return fail(GS_E_ARG, "flat descriptor is NULL");
  if (f->n_stored < 1 || f->n_stored > ((int64_t)1 << 40)) return fail(GS_E_RANGE, "n_stored out of range");
  if (f->n_runs < 0) return fail(GS_E_ARG, "n_runs < 0");
  if (f->n_runs > 0 && (!f->run_k0 || !f->run_len || !f->run_i || !f->run_j0 || !f->run_d || !f->chunk_run))
    return fail(GS_E_ARG, "run tables are NULL");
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
  a.run_k0 = f->run_k0;
  a.run_len = f->run_len;
  a.run_i = f->run_i;
  a.run_j0 = f->run_j0;
  a.run_d = f->run_d;
  a.chunk_run = f->chunk_run;
  a.P = make_phys<T>(*p, f->dx, f->dy);
  a.rain_mode = fo.n_rain > 0 ? (fo.rain_gridded ? 2 : 1) : 0;
  a.rain_frames = (const T*)fo.rain_frames;
  a.frame_stride = fo.rain_frame_stride;
  a.rain_pitch = fo.rain_row_pitch;
  a.clock = clk;
  a.plan = make_plan(*p, fo);
  return GS_OK;
}

// Workgroups of k_flat_step that are resident at once on the current device (cached per
// device and instantiation); the grid never exceeds the work.
template <typename T, int MODE>
static unsigned flat_grid(int64_t nchunks) {
  static int cache[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) dev = -1;
  int per = (dev >= 0 && dev < 64) ? cache[dev] : 0;
  if (per <= 0) {
    int occ = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_flat_step<T, MODE>, kFBlock, 0) != hipSuccess) occ = 4;
    if (dev < 0 || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) cus = 256;
    per = (occ > 0 ? occ : 1) * (cus > 0 ? cus : 1);
    if (dev >= 0 && dev < 64) cache[dev] = per;
  }
  const int64_t need = (nchunks + 3) / 4;
  return (unsigned)(need < per ? (need > 0 ? need : 1) : per);
}

template <typename T>
static int flat_step_t(const GsFlat* f, const GsPhys* p, GsForcing fo, GsClock* clk, hipStream_t s) {
  FlatK<T> a;
  int rc = build_flat<T>(f, p, fo, clk, a, true);
  if (rc) return rc;
  hipLaunchKernelGGL((k_flat_step<T, 1>), dim3(flat_grid<T, 1>(a.nchunks)), dim3(kFBlock), 0, s, a);
  GS_HIP(hipGetLastError(), "k_flat_step launch");
  return GS_OK;
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

}  // extern "C"
