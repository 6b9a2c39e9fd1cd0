// gs_mesh.hip -- host builder of the compressed ("flat") mesh: the stored set of an active
// mask on the padded grid, its row-major flat ids, the run table with its chunk index and
// the row-sweep tiles. Host code only. It produces, array for array and bit for bit, what
// the NumPy builder geoswe_amd/flatmesh.py produces (which follows the reference's
// CompressedMesh2D / CompressedSWE numbering, src/geoswe/compressed_mesh.py:108-187 and
// compressed_rhs.py:241-369), in one pass over the grid with five rolling rows of ids
// instead of full-grid temporaries.
#include <cstdlib>
#include <cstring>
#include <vector>

#include "gs_common.h"

namespace {

constexpr int kChunk = 60;                                   // flat ids written per chunk / window
constexpr uint64_t kWriteLanes = ((1ull << 62) - 1ull) & ~3ull;   // lanes 2..61

template <typename P>
static P* take(const std::vector<P>& v) {   // malloc'd copy handed to the caller
  P* p = (P*)std::malloc(v.empty() ? 1 : v.size() * sizeof(P));
  if (p && !v.empty()) std::memcpy(p, v.data(), v.size() * sizeof(P));
  return p;
}

}  // namespace

extern "C" {

int gs_flat_mesh_free(GsFlatMeshHost* m) {
  if (!m) return GS_OK;
  std::free(m->ij);
  std::free(m->active);
  std::free(m->run_k0);
  std::free(m->run_len);
  std::free(m->run_i);
  std::free(m->run_j0);
  std::free(m->run_d);
  std::free(m->chunk_run);
  std::free(m->tile_e0);
  std::free(m->tile_irc);
  std::free(m->ent);
  std::memset(m, 0, sizeof *m);
  return GS_OK;
}

int gs_flat_mesh_build(const uint8_t* inside, int nxp, int nyp, int ring, int sweep_rows, int tile_target,
                       int max_sweep_rows, GsFlatMeshHost* out) {
  using namespace gs;
  if (!inside || !out) return fail(GS_E_ARG, "gs_flat_mesh_build: NULL argument");
  std::memset(out, 0, sizeof *out);
  if (nxp < 1 || nyp < 1) return fail(GS_E_RANGE, "grid must be at least 1 x 1");
  if (ring < 2) return fail(GS_E_ARG, "ring must be >= 2 (the SRM stencil reaches two cells)");
  if ((int64_t)nyp >= (int64_t)(1u << 31) / 4) return fail(GS_E_RANGE, "padded rows too long");
  if (sweep_rows == 0 || sweep_rows < -1) return fail(GS_E_ARG, "sweep_rows must be >= 1 or -1 (automatic)");
  if (sweep_rows == -1 && (tile_target < 1 || max_sweep_rows < 1))
    return fail(GS_E_ARG, "tile_target and max_sweep_rows must be >= 1");
  const int64_t NX = nxp, NY = nyp, G = NX * NY;
  auto at = [&](int64_t i, int64_t j) { return i * NY + j; };
  for (int64_t i = 0; i < NX; ++i)
    for (int64_t j = 0; j < NY; ++j)
      if (inside[at(i, j)] && (i < 2 || i >= NX - 2 || j < 2 || j >= NY - 2))
        return fail(GS_E_ARG, "active cells must lie at least two cells inside the padded grid (ngh >= 2)");

  // stored set: `ring` dilations with the 4-neighbour cross
  std::vector<uint8_t> st(inside, inside + G), nx_(G);
  for (int it = 0; it < ring; ++it) {
    for (int64_t i = 0; i < NX; ++i)
      for (int64_t j = 0; j < NY; ++j) {
        const int64_t k = at(i, j);
        uint8_t v = st[k];
        if (i > 0) v |= st[k - NY];
        if (i < NX - 1) v |= st[k + NY];
        if (j > 0) v |= st[k - 1];
        if (j < NY - 1) v |= st[k + 1];
        nx_[k] = v ? 1 : 0;
      }
    st.swap(nx_);
  }
  std::vector<uint8_t>().swap(nx_);
  std::vector<int64_t> row_start(NX + 1, 0);
  for (int64_t i = 0; i < NX; ++i) {
    int64_t c = 0;
    for (int64_t j = 0; j < NY; ++j) c += st[at(i, j)];
    row_start[i + 1] = row_start[i] + c;
  }
  const int64_t n = row_start[NX];

  // flat ids of five rolling rows (i-2 .. i+2); -1 where not stored
  std::vector<std::vector<int64_t>> rid(5, std::vector<int64_t>(NY, -1));
  auto fill_row = [&](int64_t i, std::vector<int64_t>& r) {
    if (i < 0 || i >= NX) {
      std::fill(r.begin(), r.end(), (int64_t)-1);
      return;
    }
    int64_t k = row_start[i];
    for (int64_t j = 0; j < NY; ++j) r[j] = st[at(i, j)] ? k++ : -1;
  };
  for (int s = 0; s < 5; ++s) fill_row(s - 2, rid[s]);

  std::vector<int32_t> ij;
  std::vector<uint8_t> act;
  ij.reserve((size_t)(2 * n));
  act.reserve((size_t)n);
  std::vector<int64_t> run_k0;
  std::vector<int32_t> run_len, run_i, run_j0, run_d;
  int64_t n_active = 0;
  int64_t pk = -2, pi = -1, pj = -2;
  int64_t pd[4] = {0, 0, 0, 0};
  for (int64_t i = 0; i < NX; ++i) {
    if (i > 0) {   // roll: row i+2 enters
      std::vector<int64_t> tmp = std::move(rid[0]);
      for (int s = 0; s < 4; ++s) rid[s] = std::move(rid[s + 1]);
      rid[4] = std::move(tmp);
      fill_row(i + 2, rid[4]);
    }
    const std::vector<int64_t>& r0 = rid[2];
    for (int64_t j = 0; j < NY; ++j) {
      const int64_t k = r0[j];
      if (k < 0) continue;
      ij.push_back((int32_t)i);
      ij.push_back((int32_t)j);
      const bool a = inside[at(i, j)] != 0;
      act.push_back(a ? 1 : 0);
      if (!a) continue;
      ++n_active;
      const int64_t kE = rid[3][j], kW = rid[1][j], kEE = rid[4][j], kWW = rid[0][j];
      if (kE < 0 || kW < 0 || kEE < 0 || kWW < 0 || r0[j + 1] < 0 || r0[j - 1] < 0 || r0[j + 2] < 0 || r0[j - 2] < 0)
        return fail(GS_E_STATE, "ring does not cover the +-2 stencil of every active cell");
      if (r0[j + 1] != k + 1 || r0[j - 2] != k - 2)
        return fail(GS_E_STATE, "neighbours along j are not consecutive flat ids");
      const int64_t d[4] = {kE - k, kW - k, kEE - k, kWW - k};
      for (int q = 0; q < 4; ++q)
        if (d[q] >= ((int64_t)1 << 31) || d[q] <= -((int64_t)1 << 31)) return fail(GS_E_RANGE, "row offsets exceed int32");
      const bool brk = run_k0.empty() || k != pk + 1 || i != pi || j != pj + 1 || d[0] != pd[0] || d[1] != pd[1] ||
                       d[2] != pd[2] || d[3] != pd[3];
      if (brk) {
        run_k0.push_back(k);
        run_len.push_back(0);
        run_i.push_back((int32_t)i);
        run_j0.push_back((int32_t)j);
        for (int q = 0; q < 4; ++q) run_d.push_back((int32_t)d[q]);
      }
      ++run_len.back();
      pk = k;
      pi = i;
      pj = j;
      for (int q = 0; q < 4; ++q) pd[q] = d[q];
    }
  }
  const int64_t nruns = (int64_t)run_k0.size();
  for (int64_t r = 0; r < nruns; ++r) {
    int32_t lo = 0, hi = 0;
    for (int q = 0; q < 4; ++q) {
      lo = run_d[4 * r + q] < lo ? run_d[4 * r + q] : lo;
      hi = run_d[4 * r + q] > hi ? run_d[4 * r + q] : hi;
    }
    if (run_k0[r] + lo < 0 || run_k0[r] + run_len[r] - 1 + hi >= n)
      return fail(GS_E_STATE, "run offsets leave the stored range");
  }
  // first run whose end is past each chunk start (searchsorted(run_end, c*60, 'right'))
  const int64_t nchunks = (n + kChunk - 1) / kChunk;
  std::vector<int32_t> chunk_run((size_t)nchunks);
  {
    int64_t r = 0;
    for (int64_t c = 0; c < nchunks; ++c) {
      while (r < nruns && run_k0[r] + run_len[r] <= c * kChunk) ++r;
      chunk_run[c] = (int32_t)r;
    }
  }

  // tile height from the column grid 60w (windows with an active lane, over all rows)
  const int64_t nwin = (NY - 4) > 0 ? (NY - 4 + kChunk - 1) / kChunk : 1;
  const int64_t nwin1 = nwin < 1 ? 1 : nwin;
  int64_t busy = 0;
  for (int64_t i = 0; i < NX; ++i)
    for (int64_t w = 0; w < nwin1; ++w) {
      uint64_t am = 0;
      for (int l = 0; l < 64; ++l) {
        const int64_t j = kChunk * w + l;
        if (j >= NY) break;
        if (inside[at(i, j)]) am |= 1ull << l;
      }
      busy += (am & kWriteLanes) != 0;
    }
  int64_t rows = sweep_rows;
  if (rows == -1) {
    rows = 1;
    while (rows < max_sweep_rows && busy / (2 * rows) >= tile_target) rows *= 2;
  }
  // row-sweep tiles: in each band of `rows` rows, 64-lane windows placed from the left, each
  // starting two columns before the first active column the previous windows left uncovered
  // (lanes 2..61 are written), trimmed to the band's rows with an active lane
  const int64_t nb = (NX + rows - 1) / rows;
  std::vector<int64_t> tile_e0;
  std::vector<int32_t> tile_irc;
  std::vector<uint64_t> ent;
  std::vector<uint8_t> colany((size_t)NY);
  std::vector<int64_t> before;   // stored cells left of each column, rows r0-2 .. r1+1
  int64_t ne = 0;
  for (int64_t b = 0; b < nb; ++b) {
    const int64_t r0 = b * rows, r1 = (b + 1) * rows < NX ? (b + 1) * rows : NX;
    bool any = false;
    for (int64_t j = 0; j < NY; ++j) {
      uint8_t v = 0;
      for (int64_t i = r0; i < r1 && !v; ++i) v = inside[at(i, j)] != 0;
      colany[j] = v;
      any = any || v;
    }
    if (!any) continue;
    const int64_t lo = r0 - 2 > 0 ? r0 - 2 : 0, hi = r1 + 2 < NX ? r1 + 2 : NX;
    before.assign((size_t)((hi - lo) * (NY + 1)), 0);
    for (int64_t i = lo; i < hi; ++i) {
      int64_t* p = before.data() + (i - lo) * (NY + 1);
      for (int64_t j = 0; j < NY; ++j) p[j + 1] = p[j] + st[at(i, j)];
    }
    int64_t j = 0;
    while (true) {
      while (j < NY && !colany[j]) ++j;
      if (j >= NY) break;
      const int64_t c0 = j - 2;
      if (c0 < 0) return fail(GS_E_STATE, "a sweep window would start left of the padded grid");
      int64_t first = -1, last = -1;
      for (int64_t i = r0; i < r1; ++i) {
        bool live = false;
        for (int64_t jj = c0 + 2; jj < c0 + 62 && jj < NY && !live; ++jj) live = inside[at(i, jj)] != 0;
        if (live) {
          if (first < 0) first = i;
          last = i;
        }
      }
      if (first < 2 || last + 2 > NX - 1) return fail(GS_E_STATE, "a sweep tile's stencil rows leave the padded grid");
      tile_e0.push_back(ne);
      tile_irc.push_back((int32_t)first);
      tile_irc.push_back((int32_t)(last - first + 1));
      tile_irc.push_back((int32_t)c0);
      for (int64_t i = first - 2; i <= last + 2; ++i) {
        uint64_t sm = 0, am = 0;
        for (int l = 0; l < 64; ++l) {
          const int64_t jj = c0 + l;
          if (jj >= NY) break;
          if (st[at(i, jj)]) sm |= 1ull << l;
          if (inside[at(i, jj)]) am |= 1ull << l;
        }
        ent.push_back((uint64_t)(row_start[i] + before[(size_t)((i - lo) * (NY + 1) + c0)]));
        ent.push_back(sm);
        ent.push_back(am & kWriteLanes);
        ++ne;
      }
      j = c0 + 2 + kChunk;
    }
  }

  out->n_stored = n;
  out->n_active = n_active;
  out->n_runs = nruns;
  out->n_chunks = nchunks;
  out->n_tiles = (int64_t)tile_e0.size();
  out->n_entries = ne;
  out->sweep_rows = (int32_t)rows;
  out->ij = take(ij);
  out->active = take(act);
  out->run_k0 = take(run_k0);
  out->run_len = take(run_len);
  out->run_i = take(run_i);
  out->run_j0 = take(run_j0);
  out->run_d = take(run_d);
  out->chunk_run = take(chunk_run);
  out->tile_e0 = take(tile_e0);
  out->tile_irc = take(tile_irc);
  out->ent = take(ent);
  if (!out->ij || !out->active || !out->run_k0 || !out->run_len || !out->run_i || !out->run_j0 || !out->run_d ||
      !out->chunk_run || !out->tile_e0 || !out->tile_irc || !out->ent) {
    gs_flat_mesh_free(out);
    return fail(GS_E_RANGE, "host allocation failed");
  }
  return GS_OK;
}

}  // extern "C"
