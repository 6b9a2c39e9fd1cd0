// gs_util.hip -- device set-up and measurement kernels: constant fill, the seeded
// fractional-Brownian-motion bed generator used by the synthetic benchmark cases, and a
// plain streaming copy used to calibrate HBM bandwidth and traffic counters.
#include "gs_common.h"

namespace gs {

// Philox2x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as
// 1, 2, 3", SC'11): counter (c0, c1), key k; returns the first output word.
__device__ __forceinline__ uint32_t philox2x32(uint32_t c0, uint32_t c1, uint32_t k) {
  for (int r = 0; r < 10; ++r) {
    const uint64_t p = (uint64_t)0xD256D193u * (uint64_t)c0;
    const uint32_t hi = (uint32_t)(p >> 32), lo = (uint32_t)p;
    c0 = hi ^ k ^ c1;
    c1 = lo;
    k += 0x9E3779B9u;
  }
  return c0;
}

// Lattice value in [-1, 1) of octave `oct` at lattice point (ix, iy).
__device__ __forceinline__ double lattice(uint32_t seed, int oct, int64_t ix, int64_t iy) {
  const uint32_t c0 = (uint32_t)ix;
  const uint32_t c1 = (uint32_t)iy ^ ((uint32_t)oct << 24);
  const uint32_t r = philox2x32(c0, c1, seed + 0x1000193u * (uint32_t)oct);
  return (double)r * (2.0 / 4294967296.0) - 1.0;
}

// Value-noise fBm at cell (i, j): sum over octaves k of amp_k * bilinear-smoothstep
// interpolation of lattice values with spacing L0 / 2^k cells, amp_k = 2^(-H k).
// Every operation is an IEEE double rounding, mirrored by oracle/generators.py.
__device__ double fbm_at(int64_t i, int64_t j, uint32_t seed, int octaves, double L0, double gain) {
  double acc = 0.0, amp = 1.0, L = L0;
  for (int k = 0; k < octaves; ++k) {
    const double fx = ((double)i + 0.5) / L;
    const double fy = ((double)j + 0.5) / L;
    const double x0 = floor(fx), y0 = floor(fy);
    const double tx = fx - x0, ty = fy - y0;
    const double sx = (tx * tx) * (3.0 - 2.0 * tx);
    const double sy = (ty * ty) * (3.0 - 2.0 * ty);
    const int64_t ix = (int64_t)x0, iy = (int64_t)y0;
    const double v00 = lattice(seed, k, ix, iy), v10 = lattice(seed, k, ix + 1, iy);
    const double v01 = lattice(seed, k, ix, iy + 1), v11 = lattice(seed, k, ix + 1, iy + 1);
    const double a = v00 + sx * (v10 - v00);
    const double b = v01 + sx * (v11 - v01);
    acc = acc + amp * (a + sy * (b - a));
    amp = amp * gain;
    L = L * 0.5;
  }
  return acc;
}

template <typename T>
__global__ void k_gen_bed(T* out, int nx, int ny, int64_t pitch, int64_t i0, int64_t j0, int64_t stride,
                          uint32_t seed, int octaves, double L0, double gain, double mu, double scale, double slope_x,
                          double dx, double offset, int round_f32) {
  const int64_t n = (int64_t)nx * ny;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = k / ny, j = k - (k / ny) * ny;
    const int64_t gi = i0 + i * stride, gj = j0 + j * stride;
    const double v = (fbm_at(gi, gj, seed, octaves, L0, gain) - mu) * scale;
    const double b = slope_x * (((double)gi + 0.5) * dx) + v + offset;
    out[i * pitch + j] = round_f32 ? (T)(float)b : (T)b;
  }
}

template <typename T>
__global__ void k_fill(T* out, int64_t rows, int64_t cols, int64_t pitch, T value) {
  const int64_t n = rows * cols;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = k / cols, j = k - i * cols;
    out[i * pitch + j] = value;
  }
}

// Voronoi class field: each cell takes the class of the nearest site (squared distance in
// cell units, ties to the lowest site index), centres at (i0 + i + 0.5, j0 + j + 0.5).
__global__ __launch_bounds__(256) void k_gen_voronoi(uint8_t* out, int nx, int ny, int64_t pitch, int64_t i0,
                                                     int64_t j0, const float* __restrict__ sx,
                                                     const float* __restrict__ sy, const uint8_t* __restrict__ scls,
                                                     int nsites) {
  __shared__ float s_x[1024], s_y[1024];
  __shared__ uint8_t s_c[1024];
  for (int t = threadIdx.x; t < nsites; t += blockDim.x) {
    s_x[t] = sx[t];
    s_y[t] = sy[t];
    s_c[t] = scls[t];
  }
  __syncthreads();
  const int64_t n = (int64_t)nx * ny;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = k / ny, j = k - (k / ny) * ny;
    const float x = (float)(i0 + i) + 0.5f, y = (float)(j0 + j) + 0.5f;
    float best = INFINITY;
    uint8_t bc = 0;
    for (int t = 0; t < nsites; ++t) {
      const float dx = x - s_x[t], dy = y - s_y[t];
      const float d = dx * dx + dy * dy;
      if (d < best) { best = d; bc = s_c[t]; }
    }
    out[i * pitch + j] = bc;
  }
}

// Streaming copy, 16 B per lane per iteration.
__global__ __launch_bounds__(256) void k_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, int64_t n16) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n16; k += (int64_t)gridDim.x * blockDim.x)
    dst[k] = src[k];
}

static unsigned grid_for(int64_t n, int64_t cap = 65536) {
  int64_t b = (n + 255) / 256;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return (unsigned)b;
}

}  // namespace gs

using namespace gs;

// VALU issue-rate calibration: each lane runs 8 independent float32 chains of `iters` steps
// (MODE 0: a*m then +d, 16 VALU instructions per step; MODE 1: hardware square root then *m,
// hardware reciprocal then +d, 32 per step, half of them transcendental), 8 steps per loop
// trip so that loop control is negligible. The chains are stored so that nothing is dead;
// `out` holds gridDim.x * 256 floats. `iters` is rounded down to a multiple of 8.
template <int MODE>
__global__ __launch_bounds__(256) void k_valu(float* out, int iters, float m, float d) {
  const int gid = blockIdx.x * 256 + threadIdx.x;
  float a[8];
  for (int k = 0; k < 8; ++k) a[k] = (float)(gid + k) * 1.0e-3f;
  for (int i = 0; i < iters / 8; ++i) {
#pragma unroll
    for (int kk = 0; kk < 64; ++kk) {
      const int k = kk & 7;
      if (MODE == 0) {
        a[k] = a[k] * m;
        a[k] = a[k] + d;
      } else {
        a[k] = __builtin_amdgcn_sqrtf(a[k]) * m;
        a[k] = __builtin_amdgcn_rcpf(a[k]) + d;
      }
    }
  }
  float sum = 0.0f;
  for (int k = 0; k < 8; ++k) sum = sum + a[k];
  out[gid] = sum;
}

// The short correctly rounded division and square root of gs_crmath.h against the
// compiler's general IEEE expansions, over every float32 input (or 2^32 seeded pairs).
// CASE 0: 1/b for every b; 1: sqrt(x) for every x (the guarded form), and the short form for
// every x with |x| outside (0, 2^-96); 2: a/b for 2^32 pairs of seeded bit patterns (every exponent);
// 3: a/b for every a over four divisors with the hardest significands (all ones, 1 + ulp,
// 1 - ulp/2, 3). counts[2*k] = mismatches of check k, counts[2*k + 1] = 1 + the largest
// offending index (0: none).
__device__ __forceinline__ bool cr_same(float a, float b) {
  return (a != a && b != b) || __float_as_uint(a) == __float_as_uint(b);
}
__device__ __forceinline__ uint32_t cr_hash(uint64_t i, uint32_t k) {
  return philox2x32((uint32_t)i, (uint32_t)(i >> 32), k);
}

template <int CASE>
__global__ __launch_bounds__(256) void k_crmath_check(unsigned long long* counts) {
  const uint64_t n = CASE == 3 ? (4ull << 32) : (1ull << 32);
  unsigned long long bad0 = 0, bad1 = 0, last0 = 0, last1 = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
    if (CASE == 0) {
      const float b = __uint_as_float((uint32_t)i);
      if (!cr_same(div_cr(1.0f, b), 1.0f / b)) { bad0++; last0 = i + 1; }
    } else if (CASE == 1) {
      const float x = __uint_as_float((uint32_t)i);
      const float want = __builtin_sqrtf(x);
      if (!cr_same(sqrt_cr(x), want)) { bad0++; last0 = i + 1; }
      if (!(x != 0.0f && __builtin_fabsf(x) < 0x1p-96f) && !cr_same(sqrt_cr_big(x), want)) { bad1++; last1 = i + 1; }
    } else if (CASE == 2) {
      const float a = __uint_as_float(cr_hash(i, 0x2545F491u));
      const float b = __uint_as_float(cr_hash(i, 0x9E3779B9u));
      if (!cr_same(div_cr(a, b), a / b)) { bad0++; last0 = i + 1; }
    } else {
      const uint32_t dv[4] = {0x3FFFFFFFu, 0x3F800001u, 0x3F7FFFFFu, 0x40400000u};
      const float a = __uint_as_float((uint32_t)i);
      const float b = __uint_as_float(dv[i >> 32]);
      if (!cr_same(div_cr(a, b), a / b)) { bad0++; last0 = i + 1; }
    }
  }
  if (bad0) { atomicAdd(&counts[0], bad0); atomicMax(&counts[1], last0); }
  if (bad1) { atomicAdd(&counts[2], bad1); atomicMax(&counts[3], last1); }
}

extern "C" {

int gs_fill(void* dst, int dtype, int64_t rows, int64_t cols, int64_t pitch, double value, void* stream) {
  if (!dst || rows < 0 || cols < 0 || pitch < cols) return fail(GS_E_ARG, "gs_fill: bad arguments");
  if (rows * cols == 0) return GS_OK;
  const unsigned g = grid_for(rows * cols);
  if (dtype == GS_F32)
    hipLaunchKernelGGL(k_fill<float>, dim3(g), dim3(256), 0, (hipStream_t)stream, (float*)dst, rows, cols, pitch,
                       (float)value);
  else if (dtype == GS_F64)
    hipLaunchKernelGGL(k_fill<double>, dim3(g), dim3(256), 0, (hipStream_t)stream, (double*)dst, rows, cols, pitch,
                       value);
  else
    return fail(GS_E_DTYPE, "gs_fill: unsupported dtype %d", dtype);
  GS_HIP(hipGetLastError(), "k_fill launch");
  return GS_OK;
}

int gs_gen_bed(void* dst, int dtype, int nx, int ny, int64_t pitch, int64_t i0, int64_t j0, int64_t stride,
               uint32_t seed, int octaves, double L0, double hurst, double mu, double scale, double slope_x,
               double dx, double offset, int round_f32, void* stream) {
  if (!dst || nx < 1 || ny < 1 || pitch < ny || stride < 1 || octaves < 1 || octaves > 30 || !(L0 > 0.0))
    return fail(GS_E_ARG, "gs_gen_bed: bad arguments");
  const double gain = pow(2.0, -hurst);
  const unsigned g = grid_for((int64_t)nx * ny);
  if (dtype == GS_F32)
    hipLaunchKernelGGL(k_gen_bed<float>, dim3(g), dim3(256), 0, (hipStream_t)stream, (float*)dst, nx, ny, pitch, i0,
                       j0, stride, seed, octaves, L0, gain, mu, scale, slope_x, dx, offset, round_f32);
  else if (dtype == GS_F64)
    hipLaunchKernelGGL(k_gen_bed<double>, dim3(g), dim3(256), 0, (hipStream_t)stream, (double*)dst, nx, ny, pitch,
                       i0, j0, stride, seed, octaves, L0, gain, mu, scale, slope_x, dx, offset, round_f32);
  else
    return fail(GS_E_DTYPE, "gs_gen_bed: unsupported dtype %d", dtype);
  GS_HIP(hipGetLastError(), "k_gen_bed launch");
  return GS_OK;
}

int gs_gen_voronoi(uint8_t* dst, int nx, int ny, int64_t pitch, int64_t i0, int64_t j0, const float* site_x,
                   const float* site_y, const uint8_t* site_cls, int nsites, void* stream) {
  if (!dst || nx < 1 || ny < 1 || pitch < ny || !site_x || !site_y || !site_cls || nsites < 1 || nsites > 1024)
    return fail(GS_E_ARG, "gs_gen_voronoi: bad arguments (1..1024 sites)");
  hipLaunchKernelGGL(k_gen_voronoi, dim3(grid_for((int64_t)nx * ny)), dim3(256), 0, (hipStream_t)stream, dst, nx, ny,
                     pitch, i0, j0, site_x, site_y, site_cls, nsites);
  GS_HIP(hipGetLastError(), "k_gen_voronoi launch");
  return GS_OK;
}

int gs_bench_copy(void* dst, const void* src, size_t bytes, void* stream) {
  if (!dst || !src || (bytes % 16) != 0 || ((uintptr_t)dst % 16) || ((uintptr_t)src % 16))
    return fail(GS_E_ARG, "gs_bench_copy: 16-byte aligned pointers and a multiple of 16 bytes required");
  const int64_t n16 = (int64_t)(bytes / 16);
  hipLaunchKernelGGL(k_copy, dim3(grid_for(n16, 1 << 20)), dim3(256), 0, (hipStream_t)stream, (const uint4*)src,
                     (uint4*)dst, n16);
  GS_HIP(hipGetLastError(), "k_copy launch");
  return GS_OK;
}

int gs_bench_valu(float* out, int blocks, int iters, int mode, void* stream) {
  if (!out || blocks < 1 || iters < 1 || (mode != 0 && mode != 1))
    return fail(GS_E_ARG, "gs_bench_valu: out, blocks >= 1, iters >= 1, mode 0 or 1 required");
  if (mode == 0)
    hipLaunchKernelGGL(k_valu<0>, dim3(blocks), dim3(256), 0, (hipStream_t)stream, out, iters, 0.9999f, 1.0e-4f);
  else
    hipLaunchKernelGGL(k_valu<1>, dim3(blocks), dim3(256), 0, (hipStream_t)stream, out, iters, 0.9999f, 1.0e-4f);
  GS_HIP(hipGetLastError(), "k_valu launch");
  return GS_OK;
}

int gs_crmath_check(int which, unsigned long long* counts_dev, void* stream) {
  if (!counts_dev || which < 0 || which > 3) return fail(GS_E_ARG, "gs_crmath_check: counts and which in 0..3 required");
  hipStream_t s = (hipStream_t)stream;
  const dim3 g(8192), b(256);
  switch (which) {
    case 0: hipLaunchKernelGGL(k_crmath_check<0>, g, b, 0, s, counts_dev); break;
    case 1: hipLaunchKernelGGL(k_crmath_check<1>, g, b, 0, s, counts_dev); break;
    case 2: hipLaunchKernelGGL(k_crmath_check<2>, g, b, 0, s, counts_dev); break;
    default: hipLaunchKernelGGL(k_crmath_check<3>, g, b, 0, s, counts_dev); break;
  }
  GS_HIP(hipGetLastError(), "k_crmath_check launch");
  return GS_OK;
}

// Stream capture of step entry points into an executable HIP graph.
struct GsGraph {
  hipGraph_t graph;
  hipGraphExec_t exec;
  int64_t nodes;
};

int gs_graph_begin(void* stream) {
  if (!stream) return fail(GS_E_ARG, "gs_graph_begin: capture needs a created stream, not the null stream");
  GS_HIP(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
  return GS_OK;
}

int gs_graph_end(void* stream, GsGraph** out) {
  if (!stream || !out) return fail(GS_E_ARG, "gs_graph_end: stream and out are required");
  *out = nullptr;
  hipGraph_t g = nullptr;
  GS_HIP(hipStreamEndCapture((hipStream_t)stream, &g), "hipStreamEndCapture");
  if (!g) return fail(GS_E_STATE, "gs_graph_end: the capture recorded no graph");
  size_t nn = 0;
  hipError_t e = hipGraphGetNodes(g, nullptr, &nn);
  hipGraphExec_t x = nullptr;
  if (e == hipSuccess) e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(g);
    return hip_fail(e, "gs_graph_end: instantiate");
  }
  GsGraph* r = new GsGraph;
  r->graph = g;
  r->exec = x;
  r->nodes = (int64_t)nn;
  *out = r;
  return GS_OK;
}

int gs_graph_nodes(const GsGraph* g, int64_t* nodes) {
  if (!g || !nodes) return fail(GS_E_ARG, "gs_graph_nodes: graph and nodes are required");
  *nodes = g->nodes;
  return GS_OK;
}

int gs_graph_launch(const GsGraph* g, void* stream) {
  if (!g) return fail(GS_E_ARG, "gs_graph_launch: graph is NULL");
  GS_HIP(hipGraphLaunch(g->exec, (hipStream_t)stream), "hipGraphLaunch");
  return GS_OK;
}

int gs_graph_destroy(GsGraph* g) {
  if (!g) return GS_OK;
  hipError_t e1 = hipGraphExecDestroy(g->exec);
  hipError_t e2 = hipGraphDestroy(g->graph);
  delete g;
  if (e1 != hipSuccess) return hip_fail(e1, "hipGraphExecDestroy");
  if (e2 != hipSuccess) return hip_fail(e2, "hipGraphDestroy");
  return GS_OK;
}

}  // extern "C"
