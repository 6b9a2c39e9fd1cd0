// gs_adjoint.hip -- reverse step of the flat (compressed) fused step, for gradients of a
// scalar loss of the states with respect to the per-class Manning n and the per-cell bed.
//
// One reverse step maps the cotangent of q_{s+1} to that of q_s and accumulates the
// parameter cotangents, with the step's dt frozen. It runs in four passes over the stored
// cells, all gathers (no floating-point atomics on the state):
//   A  per active cell: reverse of the update (friction, wet/dry) from the stored
//      pre-friction state -> w = cotangent of (h1, hu1, hv1); residual cotangent is dt*w;
//      Manning n per class (workgroup sums, one atomic per class and workgroup);
//      stage cells pass -cot(h) to their bed where eta - b > 0.
//   B  per face, each face once: reverse of face() + prims() from the residual
//      cotangents of its two cells; 10 numbers into the face's slot, indexed by its lower
//      cell (x: (c, i+1), y: (c, j+1)). An active cell owns its upper faces and its lower
//      faces whose lower cell is not active.
//   C  per stored cell: cotangent of q_s = w + the four face slots touching the cell; bed
//      cotangent from the faces' direct terms; the cell's bed-gradient cotangents.
//   D  per stored cell: bed cotangent through the central bed gradients of its four
//      neighbours, gz_X = 0.5*(b_{X+1} - b_{X-1})*inv_d.
// The arithmetic is gs_adjoint.h's (checked against finite differences on the CPU).
#include "gs_common.h"
#include "gs_adjoint.h"

namespace gs {

constexpr int kABlock = 256;
constexpr int kAdjMaxBlocks = GS_ADJ_MAX_BLOCKS;

template <typename T>
struct AdjK {
  int64_t n;
  const uint8_t* active;
  const int32_t* nbr;     // [n][4]: i+1, i-1, j+1, j-1
  const T* bed;
  const uint8_t* flags;
  const T* ntab;
  int n_classes;
  const T* q[3];
  const T* h_next;
  const T* pre[3];
  T* lam[3];
  T* w[3];
  T* xf;                  // [10][n]
  T* yf;                  // [10][n]
  T* gz;                  // [2][n]
  const T* gz_prev;       // [2][n] of the previous reverse step, or NULL
  double* bbar;
  double* nbar;
  double* npart;          // [gridDim.x][32] per-workgroup class sums (deterministic order)
  Phys<T> P;
  T dt;
};

// Butterfly sum over the 64 lanes (fixed order, so every lane gets the same bits every run).
__device__ __forceinline__ double wave_sum_f64(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <typename T>
__global__ __launch_bounds__(kABlock) void k_adj_update(AdjK<T> a) {
  // per-class n cotangents: butterfly sums per wavefront, wavefronts combined in a fixed order
  // and added to this workgroup's own slot of npart (the grid is the same every reverse step),
  // so the gradient has the same bits on every run
  __shared__ double s_acc[kABlock / 64][32];
  const int wv = threadIdx.x >> 6;
  if (threadIdx.x < 32)
    for (int w = 0; w < kABlock / 64; ++w) s_acc[w][threadIdx.x] = 0.0;
  __syncthreads();
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < a.n; base += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = base + threadIdx.x;
    int cls = -1;
    T gn = T(0);
    if (c < a.n && a.active[c]) {
      const uint8_t fl = a.flags ? a.flags[c] : (uint8_t)0;
      const T lh = a.lam[0][c], lhu = a.lam[1][c], lhv = a.lam[2][c];
      if (fl & GS_FLAG_STAGE) {
        // h = max(eta - b, 0), hu = hv = 0: the step's own inputs get nothing
        if (a.h_next[c] > T(0)) a.bbar[c] -= (double)lh;
        a.w[0][c] = T(0);
        a.w[1][c] = T(0);
        a.w[2][c] = T(0);
      } else {
        cls = fl & GS_FLAG_CLASS_MASK;
        const T n = a.P.friction ? a.ntab[cls] : T(0);
        T g1, gu1, gv1;
        update_vjp<T>(a.pre[0][c], a.pre[1][c], a.pre[2][c], a.dt, n, a.P, lh, lhu, lhv, g1, gu1, gv1, gn);
        a.w[0][c] = g1;
        a.w[1][c] = gu1;
        a.w[2][c] = gv1;
      }
    }
    if (a.P.friction) {
      // one butterfly per class present in the wavefront (classes are spatially coherent:
      // usually one or two), each over all 64 lanes with zeros for the other lanes, so a
      // class's sum has the bits the loop over every class gave
      uint64_t pending = __ballot(cls >= 0);
      while (pending) {   // uniform: every lane shuffles
        const int k = __shfl(cls, __ffsll((unsigned long long)pending) - 1, 64);
        const bool mine = cls == k;
        const double v = wave_sum_f64(mine ? (double)gn : 0.0);
        if ((threadIdx.x & 63) == 0) s_acc[wv][k] += v;
        pending &= ~__ballot(mine);
      }
    }
  }
  __syncthreads();
  if (threadIdx.x < (unsigned)a.n_classes) {
    double v = 0.0;
    for (int w = 0; w < kABlock / 64; ++w) v += s_acc[w][threadIdx.x];
    a.npart[(int64_t)blockIdx.x * 32 + threadIdx.x] += v;
  }
}

// nbar[k] += sum over workgroup slots of npart[.][k], in slot order.
// One wavefront per class: lane l sums the slots of workgroups l, l+64, ... in order, then a
// butterfly over the lanes; the order is fixed, so the sum has the same bits on every run.
__global__ __launch_bounds__(64) void k_adj_nsum(const double* npart, int nblocks, int n_classes, double* nbar) {
  const int k = blockIdx.x;
  double v = 0.0;
  for (int b = (int)threadIdx.x; b < nblocks; b += 64) v += npart[(int64_t)b * 32 + k];
  v = wave_sum_f64(v);
  if (threadIdx.x == 0 && k < n_classes) nbar[k] += v;
}

// Reverse of one face (L, R) along x (axis 0) or y (axis 1) into slot L of `slots`.
template <typename T>
__device__ __forceinline__ void face_slot(const AdjK<T>& a, int axis, int64_t L, int64_t R, T* slots) {
  const T hm = a.P.h_min, g = a.P.g;
  const T inv = axis == 0 ? a.P.inv_dx : a.P.inv_dy;
  const T half = axis == 0 ? a.P.half_dx : a.P.half_dy;
  const int up = axis == 0 ? 0 : 2, dn = axis == 0 ? 1 : 3;
  const int64_t n = a.n;
  const T hL = a.q[0][L], huL = a.q[1][L], hvL = a.q[2][L];
  const T hR = a.q[0][R], huR = a.q[1][R], hvR = a.q[2][R];
  T* s = slots + L;
  if (hL < hm && hR < hm) {   // dry face: no contribution either way
    for (int k = 0; k < 10; ++k) s[k * n] = T(0);
    return;
  }
  const T bL = a.bed[L], bR = a.bed[R];
  // the cells either side of L and R along the axis are stored for every face of an
  // active cell (ring-2 set, checked when the table is built); clamp all the same
  const int32_t* nl = a.nbr + 4 * L;
  const int32_t* nr = a.nbr + 4 * R;
  const int64_t lu = nl[up] >= 0 ? nl[up] : L, ld = nl[dn] >= 0 ? nl[dn] : L;
  const int64_t ru = nr[up] >= 0 ? nr[up] : R, rd = nr[dn] >= 0 ? nr[dn] : R;
  const T gzL = (T(0.5) * (a.bed[lu] - a.bed[ld])) * inv;
  const T gzR = (T(0.5) * (a.bed[ru] - a.bed[rd])) * inv;
  T uL, vL, uR, vR;
  prims(hL, huL, hvL, hm, uL, vL);
  prims(hR, huR, hvR, hm, uR, vR);
  // residual cotangents dt*w of both cells (zero outside active non-stage cells)
  const T dt = a.dt;
  const T rhL = dt * a.w[0][L], ruL = dt * a.w[1][L], rvL = dt * a.w[2][L];
  const T rhR = dt * a.w[0][R], ruR = dt * a.w[1][R], rvR = dt * a.w[2][R];
  FaceGrad<T> G;
  if (axis == 0) {
    face_vjp<T>(bL, bR, gzL, gzR, hL, hR, uL, vL, uR, vR, half, g, hm, inv * (rhR - rhL), inv * (ruR - ruL),
                inv * (rvR - rvL), -(inv * ruL), inv * ruR, G);
  } else {
    face_vjp<T>(bL, bR, gzL, gzR, hL, hR, vL, uL, vR, uR, half, g, hm, inv * (rhR - rhL), inv * (rvR - rvL),
                inv * (ruR - ruL), -(inv * rvL), inv * rvR, G);
  }
  T ghL = G.hL, ghuL = T(0), ghvL = T(0), ghR = G.hR, ghuR = T(0), ghvR = T(0);
  if (axis == 0) {
    prims_vjp<T>(hL, huL, hvL, hm, G.uNL, G.uTL, ghL, ghuL, ghvL);
    prims_vjp<T>(hR, huR, hvR, hm, G.uNR, G.uTR, ghR, ghuR, ghvR);
  } else {
    prims_vjp<T>(hL, huL, hvL, hm, G.uTL, G.uNL, ghL, ghuL, ghvL);
    prims_vjp<T>(hR, huR, hvR, hm, G.uTR, G.uNR, ghR, ghuR, ghvR);
  }
  s[0] = ghL;
  s[n] = ghuL;
  s[2 * n] = ghvL;
  s[3 * n] = G.bL;
  s[4 * n] = G.gzL;
  s[5 * n] = ghR;
  s[6 * n] = ghuR;
  s[7 * n] = ghvR;
  s[8 * n] = G.bR;
  s[9 * n] = G.gzR;
}

// Every face with an active cell on either side, computed by its lower cell (the slots are
// indexed by it): each stored cell, active or ring, takes its upper face along each axis
// when it or the cell above is active. A wavefront then runs at most two face reverses (a
// scheme in which an active cell also took its lower face beside a ring cell ran up to
// four, the extra two for the few lanes on the active set's lower edges). One face per
// trip of a loop that is not unrolled, so the reverse of a face is compiled once.
template <typename T>
__global__ __launch_bounds__(kABlock) void k_adj_faces(AdjK<T> a) {
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < a.n; c += (int64_t)gridDim.x * blockDim.x) {
    const int32_t* nb = a.nbr + 4 * c;
    const bool act = a.active[c] != 0;
#pragma unroll 1
    for (int axis = 0; axis < 2; ++axis) {   // nb[0]: upper x, nb[2]: upper y
      const int64_t u = nb[2 * axis];
      if (u < 0 || !(act || a.active[u])) continue;
      face_slot<T>(a, axis, c, u, axis == 0 ? a.xf : a.yf);
    }
  }
}

template <typename T>
__global__ __launch_bounds__(kABlock) void k_adj_gather(AdjK<T> a) {
  const int64_t n = a.n;
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
    const int32_t* nb = a.nbr + 4 * c;
    const int64_t W = nb[1], S = nb[3];
    const T* x = a.xf;
    const T* y = a.yf;
    T gzx = x[4 * n + c], gzy = y[4 * n + c];
    T bb = x[3 * n + c] + y[3 * n + c];
    if (W >= 0) { gzx += x[9 * n + W]; bb += x[8 * n + W]; }
    if (S >= 0) { gzy += y[9 * n + S]; bb += y[8 * n + S]; }
    a.gz[c] = gzx;
    a.gz[n + c] = gzy;
    if (a.gz_prev) {   // the bed-gradient term of the previous reverse step (its gz is complete)
      const T hx = T(0.5) * a.P.inv_dx, hy = T(0.5) * a.P.inv_dy;
      const int64_t E = nb[0], Nn = nb[2];
      const T* gp_ = a.gz_prev;
      if (W >= 0) bb += hx * gp_[W];
      if (E >= 0) bb -= hx * gp_[E];
      if (S >= 0) bb += hy * gp_[n + S];
      if (Nn >= 0) bb -= hy * gp_[n + Nn];
    }
    a.bbar[c] += (double)bb;
    if (a.active[c]) {
      for (int k = 0; k < 3; ++k) {
        T l = a.w[k][c] + x[k * n + c] + y[k * n + c];
        if (W >= 0) l += x[(5 + k) * n + W];
        if (S >= 0) l += y[(5 + k) * n + S];
        a.lam[k][c] = l;
      }
    }
  }
}

template <typename T>
__global__ __launch_bounds__(kABlock) void k_adj_bedgrad(AdjK<T> a) {
  const int64_t n = a.n;
  const T hx = T(0.5) * a.P.inv_dx, hy = T(0.5) * a.P.inv_dy;
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
    const int32_t* nb = a.nbr + 4 * c;
    T v = T(0);
    if (nb[1] >= 0) v += hx * a.gz[nb[1]];
    if (nb[0] >= 0) v -= hx * a.gz[nb[0]];
    if (nb[3] >= 0) v += hy * a.gz[n + nb[3]];
    if (nb[2] >= 0) v -= hy * a.gz[n + nb[2]];
    a.bbar[c] += (double)v;
  }
}

template <typename T>
__global__ void k_adj_add(T* lam_h, const int64_t* ids, const T* coef, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) lam_h[ids[t]] += coef[t];
}

// The grid of every adjoint pass; fixed for a given n (npart has one slot row per workgroup).
static inline int64_t adj_blocks(int64_t n) {
  int64_t blocks = (n + kABlock - 1) / kABlock;
  if (blocks > kAdjMaxBlocks) blocks = kAdjMaxBlocks;
  return blocks < 1 ? 1 : blocks;
}

template <typename T>
static int adj_step_t(const GsAdjFlat* f, const GsPhys* p, double dt, hipStream_t s) {
  AdjK<T> a;
  memset(&a, 0, sizeof a);
  a.n = f->n_stored;
  a.active = f->active;
  a.nbr = f->nbr;
  a.bed = (const T*)f->bed;
  a.flags = f->flags;
  a.ntab = (const T*)f->ntab;
  a.n_classes = p->n_classes;
  for (int k = 0; k < 3; ++k) {
    a.q[k] = (const T*)f->q[k];
    a.pre[k] = (const T*)f->pre[k];
    a.lam[k] = (T*)f->lam[k];
    a.w[k] = (T*)f->w[k];
  }
  a.h_next = (const T*)f->h_next;
  a.xf = (T*)f->xf;
  a.yf = (T*)f->yf;
  a.gz = (T*)f->gz;
  a.gz_prev = (const T*)f->gz_prev;
  a.bbar = f->bbar;
  a.nbar = f->nbar;
  a.npart = f->npart;
  a.P = make_phys<T>(*p, f->dx, f->dy);
  a.dt = (T)dt;
  const int64_t blocks = adj_blocks(a.n);
  const dim3 gr((unsigned)blocks), bl(kABlock);
  hipLaunchKernelGGL(k_adj_update<T>, gr, bl, 0, s, a);
  hipLaunchKernelGGL(k_adj_faces<T>, gr, bl, 0, s, a);
  hipLaunchKernelGGL(k_adj_gather<T>, gr, bl, 0, s, a);
  // with gz_prev the caller folds this step's bed-gradient term into the next reverse step's
  // gather (double-buffered gz) and flushes the last one with gs_adj_flat_bedgrad
  if (!a.gz_prev) hipLaunchKernelGGL(k_adj_bedgrad<T>, gr, bl, 0, s, a);
  GS_HIP(hipGetLastError(), "adjoint step launch");
  return GS_OK;
}

template <typename T>
static int adj_bedgrad_t(const GsAdjFlat* f, const GsPhys* p, hipStream_t s) {
  AdjK<T> a;
  memset(&a, 0, sizeof a);
  a.n = f->n_stored;
  a.nbr = f->nbr;
  a.gz = (T*)f->gz;
  a.bbar = f->bbar;
  a.P = make_phys<T>(*p, f->dx, f->dy);
  const int64_t blocks = adj_blocks(a.n);
  hipLaunchKernelGGL(k_adj_bedgrad<T>, dim3((unsigned)blocks), dim3(kABlock), 0, s, a);
  GS_HIP(hipGetLastError(), "k_adj_bedgrad launch");
  return GS_OK;
}

}  // namespace gs

extern "C" {

int gs_adj_flat_step(const GsAdjFlat* f, const GsPhys* phys, double dt, void* stream) {
  using namespace gs;
  if (!f || !phys) return fail(GS_E_ARG, "NULL argument");
  if (f->n_stored < 1) return fail(GS_E_RANGE, "n_stored < 1");
  if (!f->active || !f->nbr || !f->bed || !f->h_next || !f->xf || !f->yf || !f->gz || !f->bbar || !f->nbar ||
      !f->npart)
    return fail(GS_E_ARG, "adjoint arrays are NULL");
  for (int k = 0; k < 3; ++k)
    if (!f->q[k] || !f->pre[k] || !f->lam[k] || !f->w[k]) return fail(GS_E_ARG, "state/cotangent arrays are NULL");
  int rc = check_phys(phys);
  if (rc) return rc;
  if (phys->friction && !f->ntab) return fail(GS_E_ARG, "friction needs ntab");
  if (!(dt > 0.0) || !isfinite(dt)) return fail(GS_E_ARG, "dt must be positive and finite");
  if (f->dtype == GS_F64) return adj_step_t<double>(f, phys, dt, (hipStream_t)stream);
  if (f->dtype == GS_F32) return adj_step_t<float>(f, phys, dt, (hipStream_t)stream);
  return fail(GS_E_DTYPE, "unsupported dtype %d", f->dtype);
}

int gs_adj_flat_nsum(const GsAdjFlat* f, void* stream) {
  using namespace gs;
  if (!f || !f->npart || !f->nbar) return fail(GS_E_ARG, "NULL argument");
  if (f->n_stored < 1) return fail(GS_E_RANGE, "n_stored < 1");
  if (f->n_classes < 1 || f->n_classes > 32) return fail(GS_E_RANGE, "n_classes must be in 1..32");
  hipLaunchKernelGGL(k_adj_nsum, dim3((unsigned)f->n_classes), dim3(64), 0, (hipStream_t)stream, f->npart,
                     (int)adj_blocks(f->n_stored), f->n_classes, f->nbar);
  GS_HIP(hipGetLastError(), "k_adj_nsum launch");
  return GS_OK;
}

int gs_adj_flat_bedgrad(const GsAdjFlat* f, const GsPhys* phys, void* stream) {
  using namespace gs;
  if (!f || !phys || !f->nbr || !f->gz || !f->bbar) return fail(GS_E_ARG, "NULL argument");
  if (f->n_stored < 1) return fail(GS_E_RANGE, "n_stored < 1");
  if (f->dtype == GS_F64) return adj_bedgrad_t<double>(f, phys, (hipStream_t)stream);
  if (f->dtype == GS_F32) return adj_bedgrad_t<float>(f, phys, (hipStream_t)stream);
  return fail(GS_E_DTYPE, "unsupported dtype %d", f->dtype);
}

int gs_adj_add(void* lam_h, int dtype, const int64_t* ids, const void* coef, int n, void* stream) {
  using namespace gs;
  if (n < 0) return fail(GS_E_ARG, "n < 0");
  if (n == 0) return GS_OK;
  if (!lam_h || !ids || !coef) return fail(GS_E_ARG, "NULL argument");
  const unsigned nb = (unsigned)((n + 255) / 256);
  if (dtype == GS_F64)
    hipLaunchKernelGGL(k_adj_add<double>, dim3(nb), dim3(256), 0, (hipStream_t)stream, (double*)lam_h, ids,
                       (const double*)coef, n);
  else if (dtype == GS_F32)
    hipLaunchKernelGGL(k_adj_add<float>, dim3(nb), dim3(256), 0, (hipStream_t)stream, (float*)lam_h, ids,
                       (const float*)coef, n);
  else
    return fail(GS_E_DTYPE, "unsupported dtype %d", dtype);
  GS_HIP(hipGetLastError(), "k_adj_add launch");
  return GS_OK;
}

int gs_flat_run_frozen(const GsFlat* f, const GsPhys* phys, GsForcing forc, GsClock* clock_dev, int n,
                       const double* dts, const int64_t* gauge_ids, int n_gauges, void* samples, int ckpt_every,
                       void* const* ckpt, void* stream) {
  using namespace gs;
  if (!f || !phys || !clock_dev || (n > 0 && !dts)) return fail(GS_E_ARG, "NULL argument");
  if (n < 0 || n_gauges < 0) return fail(GS_E_ARG, "negative count");
  if (n_gauges > 0 && (!gauge_ids || !samples)) return fail(GS_E_ARG, "gauge arrays are NULL");
  if (ckpt && ckpt_every < 1) return fail(GS_E_ARG, "ckpt_every must be >= 1");
  const size_t es = f->dtype == GS_F64 ? 8 : 4;
  const size_t bytes = (size_t)f->n_stored * es;
  hipStream_t s = (hipStream_t)stream;
  for (int j = 0; j < n; ++j) {
    if (ckpt && j % ckpt_every == 0) {   // the clock counts from 0: the state is q[j & 1]
      const int c = j / ckpt_every;
      for (int k = 0; k < 3; ++k)
        GS_HIP(hipMemcpyAsync(ckpt[3 * c + k], f->q[j & 1][k], bytes, hipMemcpyDeviceToDevice, s), "checkpoint");
    }
    int rc = gs_clock_set_dt(clock_dev, dts[j], phys, forc, stream);
    if (!rc) rc = gs_flat_step(f, phys, forc, clock_dev, stream);
    if (!rc && n_gauges > 0)
      rc = gs_flat_sample(f, clock_dev, gauge_ids, n_gauges, (char*)samples + (size_t)j * n_gauges * es, stream);
    if (rc) return rc;
  }
  return GS_OK;
}

int gs_flat_replay_pre(const GsFlat* f, const GsPhys* phys, GsForcing forc, GsClock* clock_dev, int n,
                       const double* dts, void* const* states, void* const* pre, void* stream) {
  using namespace gs;
  if (!f || !phys || !clock_dev || (n > 0 && (!dts || !states || !pre))) return fail(GS_E_ARG, "NULL argument");
  if (n < 0) return fail(GS_E_ARG, "n < 0");
  GsFlat d = *f;
  for (int j = 0; j < n; ++j) {
    const int par = j & 1;   // the clock counts the replayed steps from 0
    for (int k = 0; k < 3; ++k) {
      d.q[par][k] = states[3 * j + k];
      d.q[par ^ 1][k] = states[3 * (j + 1) + k];
    }
    int rc = gs_clock_set_dt(clock_dev, dts[j], phys, forc, stream);
    if (!rc) rc = gs_flat_step_pre(&d, phys, forc, clock_dev, pre[3 * j], pre[3 * j + 1], pre[3 * j + 2], stream);
    if (rc) return rc;
  }
  return GS_OK;
}

int gs_adj_flat_sweep(const GsAdjFlat* a, const GsPhys* phys, int n, const double* dts, const void* const* states,
                      const void* const* pre, const int64_t* gauge_ids, const void* const* coef, int n_gauges,
                      void* gz_pair, int64_t first_step, void* stream) {
  using namespace gs;
  if (!a || !phys || (n > 0 && (!dts || !states || !pre))) return fail(GS_E_ARG, "NULL argument");
  if (n < 0 || n_gauges < 0 || first_step < 0) return fail(GS_E_ARG, "negative count");
  GsAdjFlat d = *a;
  const size_t half = (size_t)2 * (size_t)a->n_stored * (a->dtype == GS_F64 ? 8 : 4);
  for (int j = n - 1; j >= 0; --j) {
    for (int k = 0; k < 3; ++k) {
      d.q[k] = states[3 * j + k];
      d.pre[k] = pre[3 * j + k];
    }
    d.h_next = states[3 * (j + 1)];
    if (gz_pair) {   // step s writes half s & 1 and folds in the half of step s + 1
      const int64_t st = first_step + j;
      d.gz = (char*)gz_pair + (size_t)(st & 1) * half;
      d.gz_prev = (char*)gz_pair + (size_t)((st + 1) & 1) * half;
    }
    int rc = gs_adj_flat_step(&d, phys, dts[j], stream);
    if (!rc && coef && coef[j] && n_gauges > 0) rc = gs_adj_add(d.lam[0], d.dtype, gauge_ids, coef[j], n_gauges, stream);
    if (rc) return rc;
  }
  return GS_OK;
}

}  // extern "C"
