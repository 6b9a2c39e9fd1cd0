/*
 * geoswe_abi.h -- C ABI of the MI355X (gfx950) fused finite-volume time step.
 *
 * The library replaces the reference's kernel-factory boundary: GeoSWE compiles CUDA C
 * strings at run time through backend.raw_kernel / cupy.RawKernel
 * (src/geoswe/backend.py:271-302) and calls them positionally from Solver2D and
 * CompressedSolver (src/geoswe/solver.py:2275-2279, compressed_solver.py:894-898).
 * Here the kernels are compiled ahead of time for gfx950 and reached through these
 * extern "C" entry points with plain pointers and sizes.
 *
 * Conventions
 *   - Every function returns 0 on success, a positive GS_E* code on a usage error, or
 *     a negative HIP error code (-hipError_t). gs_last_error() describes the last
 *     failure of the calling thread.
 *   - The caller owns every buffer. Step entry points never allocate and never
 *     synchronise the host: dt, t and the CFL wave speed stay in a device-resident
 *     GsClock, so K steps can be enqueued (or captured in a graph) back to back.
 *   - Streams are hipStream_t passed as void* (NULL = the null stream).
 *   - Arrays are row-major with the second index (j, "y") contiguous, as the
 *     reference's C-order (nx, ny) fields (src/geoswe/mesh.py:45-87).
 */
#ifndef GEOSWE_ABI_H
#define GEOSWE_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GS_ABI_VERSION 1

/* error codes (positive); negative values are -hipError_t */
#define GS_OK 0
#define GS_E_ARG 1        /* invalid argument (shape, pointer, enum) */
#define GS_E_DTYPE 2      /* unsupported dtype */
#define GS_E_RANGE 3      /* size out of the supported range */
#define GS_E_STATE 4      /* call order / state error */

/* boundary kinds: src/geoswe/bc.py:50-98 */
#define GS_BC_EXTRAPOLATE 0
#define GS_BC_PERIODIC 1
#define GS_BC_WALL 2
#define GS_BC_FALL 3

/* dtypes */
#define GS_F32 0
#define GS_F64 1
/* dense layout only: float32 planes (state, bed, Manning, rain frames, running max) stepped
 * with float64 arithmetic, every stored value rounded once to float32. No reference
 * counterpart (the reference computes a float32 state in float32). */
#define GS_F32X64 2

/* Manning classes: the flags byte holds ids 0..31; a uint8 class plane (the reference's
 * per-cell class array, solver.py:1524-1556) holds 0..255 and then takes precedence */
#define GS_MAX_CLASSES 256

/* per-cell flags byte (dense "flags" plane, flat "flags" array) */
#define GS_FLAG_CLASS_MASK 0x1F  /* Manning class id (0..31) into the ntab array */
#define GS_FLAG_CFL_SKIP 0x20    /* left out of the CFL wave-speed maximum (set_cfl_ghost_mask) */
#define GS_FLAG_INACTIVE 0x40    /* dense layout: outside the inside mask, zero residual, no rain */
#define GS_FLAG_FORCE 0x40       /* flat layout (which stores no inactive cell to update): a sponge
                                    or clamp cell, which reads GsFlat.sponge / GsFlat.clamp */
#define GS_FLAG_STAGE 0x80       /* stage-boundary cell: h = max(eta(t)-b, 0), hu = hv = 0 */

/* GsClock.status bits */
#define GS_ST_DONE 1      /* t reached t_end or nsteps reached max_steps: steps are no-ops */
#define GS_ST_NONFINITE 2 /* the wave speed of the current state is NaN or Inf */

/*
 * Device-resident schedule of a run (one per solver). Mirrors the host loop of
 * Solver2D.run / cfl_dt (solver.py:2569-2693, 2975-3027): the last workgroup of each
 * step turns the step's wave-speed maximum into the next dt on the device.
 */
typedef struct GsClock {
  double t;            /* simulated time of the current state */
  double dt;           /* step to take next (float64, as the reference's host dt) */
  double t_end;        /* run() target; +inf when stepping open-ended */
  double dt_max;       /* cap on dt, 0 = none */
  double eta64;        /* stage level at t + dt */
  int64_t nsteps;      /* steps taken so far; parity selects the input buffer */
  int64_t max_steps;   /* status gets GS_ST_DONE when nsteps reaches this */
  uint64_t lam_acc;    /* running maximum of the step in flight (float64 IEEE bits) */
  uint64_t lam_bits;   /* CFL wave speed of the current state (float64 IEEE bits) */
  uint32_t blocks_done;/* last-workgroup-done counter */
  int32_t status;      /* GS_ST_* */
  float dt32;          /* (float)dt */
  float rain32;        /* uniform rain rate of the next step, m/s, float */
  float eta32;         /* (float)eta64 */
  int32_t rain_idx;    /* rain segment / frame index of the next step */
  int32_t use_host_dt; /* 1: dt was set by the host for the next step */
  int32_t run_mode;    /* 1: Solver2D.run rules (rain-film cap, dt_max, clip to t_end) */
  int32_t pad_[6];
} GsClock;

/* Physics and scheme constants (Config, solver.py:629-860). */
typedef struct GsPhys {
  double g;            /* gravity */
  double h_min;        /* wet/dry threshold */
  double h_cfl;        /* dense: CFL depth below which a cell is skipped (h_min_cfl, floored at
                          1e-6 in float32); flat: the velocity divisor floor max(h, h_cfl) of
                          the wave speed, whose cells are skipped below h_min (h_min_cfl of the
                          flat stepper, compressed_solver.py:200-240, 599-604) */
  double cfl;          /* Courant number */
  double dmin;         /* min(dx, dy) */
  double lam_floor;    /* wave-speed floor computed by the host: (g*h_cfl)**0.5 dense
                          (solver.py:2663-2665), (g*h_min)**0.5 flat (compressed_solver.py:713-719) */
  double vcap;         /* friction velocity cap (m/s); inf disables it */
  int32_t friction;    /* 0 none, 1 Manning (point-implicit) */
  int32_t quadratic;   /* 1 quadratic-alpha friction root, 0 linearised */
  int32_t keep_h;      /* 1 keep the sub-threshold depth on dry cells, 0 zero it */
  int32_t cfl_l2;      /* CFL velocity norm: 0 max(|u|,|v|), 1 sqrt(u^2+v^2) */
  int32_t n_classes;   /* entries of the Manning table: 1..32 read through the flags byte,
                          up to GS_MAX_CLASSES with a class plane (GsDense.cls, GsFlat.cls) */
  int32_t pad_;
} GsPhys;

/* Forcing tables (device pointers, float64), shared by all layouts. */
typedef struct GsForcing {
  int32_t n_rain;            /* rain segments (0 = no rain) */
  int32_t rain_gridded;      /* 1: rain_frames holds one frame per segment */
  const double* rain_t;      /* [n_rain] segment start times, non-decreasing */
  const double* rain_rate;   /* [n_rain] uniform rate in m/s (rate_mm_h/3.6e6) */
  const double* rain_cap;    /* [n_rain] rain-film dt bound per segment (inf if dry) */
  const void* rain_frames;   /* gridded: [n_rain][frame_stride] rates in m/s (state dtype) */
  int64_t rain_frame_stride; /* elements between frames */
  int64_t rain_row_pitch;    /* elements between rows of a frame (dense: interior rows) */
  int32_t n_stage;           /* stage samples (0 = no stage boundary) */
  int32_t pad_;
  const double* stage_t;     /* [n_stage] */
  const double* stage_v;     /* [n_stage] */
} GsForcing;

/*
 * Dense layout. The state and bed hold interior cells only, nx rows of `pitch`
 * elements; ghost cells are not stored but synthesised inside the stencil loads
 * from the boundary kinds (bc.py:50-98, ghost bed solver.py:1466-1484).
 */
typedef struct GsDense {
  int32_t nx, ny;            /* interior cells */
  int64_t pitch;             /* elements between rows of every plane */
  int32_t bc_x, bc_y;        /* GS_BC_* */
  int32_t dtype;             /* GS_F32, GS_F64 or GS_F32X64 */
  int32_t tile_rows;         /* rows swept per wavefront (0 = library default) */
  double dx, dy;
  void* q[2][3];             /* double-buffered state planes h, hu, hv */
  const void* bed;           /* bed plane */
  const uint8_t* flags;      /* nullable per-cell GS_FLAG_* byte plane */
  const void* ntab;          /* Manning n per class, device array of n_classes (state dtype) */
  const void* n_field;       /* nullable per-cell Manning n plane (overrides ntab) */
  void* max_h;               /* nullable running-max depth plane */
  const uint8_t* cls;        /* nullable per-cell Manning class plane (ids 0..255 into ntab, in
                                place of the flags' class bits) */
} GsDense;

/*
 * Flat (compressed active-cell) layout: stored cells in row-major padded (i, j) order
 * (compressed_mesh.py:139-157). Active cells are listed by runs of consecutive flat
 * ids in one row whose +-1/+-2 row neighbours sit at constant id offsets; cells in no
 * run are the frozen ring (compressed_rhs.py:241-369). Within a row the +-j
 * neighbours of an active cell are k+-1, k+-2 (the ring-2 cross dilation).
 */
typedef struct GsFlat {
  int64_t n_stored;          /* stored cells N */
  int32_t n_runs;            /* runs of active cells */
  int32_t dtype;
  int32_t nyp;               /* padded row length (rain indexing) */
  int32_t ngh;               /* ghost width of the padded frame */
  double dx, dy;
  void* q[2][3];             /* double-buffered flat state */
  const void* bed;           /* flat bed */
  const uint8_t* flags;      /* nullable per-cell GS_FLAG_* byte (class, stage) */
  const void* ntab;          /* Manning n per class, device array (state dtype) */
  const void* n_field;       /* nullable per-cell Manning n */
  void* max_h;               /* nullable running max */
  const int64_t* run_k0;     /* [n_runs] first flat id of the run */
  const int32_t* run_len;    /* [n_runs] cells in the run */
  const int32_t* run_i;      /* [n_runs] padded row of the run */
  const int32_t* run_j0;     /* [n_runs] padded column of the first cell */
  const int32_t* run_d;      /* [n_runs][4] id offsets of rows i+1, i-1, i+2, i-2 */
  const int32_t* chunk_run;  /* [ceil(N/60)] first run whose end is past the chunk start */
  /* Row-sweep tiles (the step's traversal; n_tiles = 0 steps by chunks of the run table).
   * A tile's window covers padded columns c0 .. c0+63 (lanes 2..61 written; any c0 >= 0, as
   * long as every active cell is written by exactly one tile); entry = (flat id of the
   * row's first stored cell at column >= c0, stored-lane mask, active-lane mask) for one
   * padded row of the window. Tile t sweeps rows i0 .. i0+nr-1 of its window; its entries
   * are rows i0-2 .. i0+nr+1, from ent[3 * tile_e0[t]] on. */
  int64_t n_tiles;
  const int64_t* tile_e0;    /* [n_tiles] entry index of row i0 - 2 */
  const int32_t* tile_irc;   /* [n_tiles][3] i0, nr, padded column of lane 0 */
  const uint64_t* ent;       /* [entries][3] */
  /* Coastal forcings on GS_FLAG_FORCE cells, applied after the update, the stage and the
   * running max (compressed_solver.py:1320-1326 sponge_band, 424-440 stage_clamp_flat):
   * h <- keep*h + amb, hu, hv <- keep*hu, keep*hv; then where h > h_max:
   * r = h_max/h, h <- h_max, hu, hv <- r*hu, r*hv. */
  const void* sponge;        /* nullable [N][2] (keep, amb), state dtype */
  const void* clamp;         /* nullable [N] h_max, state dtype (+inf: no clamp) */
  /* Constant-rate infiltration (infiltrate_flat, compressed_solver.py:143-158) on every active
   * cell, after the running max and before the sponge: where h >= h_min,
   * h' = max(h - rate[class]*dt, 0), hu, hv scaled by h'/h. Karst drain (drain_cap, 269-286)
   * on GS_FLAG_FORCE cells, last (after the clamp): where h > h_tgt,
   * r = h > 1e-12 ? h_tgt/h : 0, h = h_tgt, hu, hv scaled by r. */
  const void* infil;         /* nullable [32] rate per Manning class (m/s), state dtype */
  const void* drain;         /* nullable [N] h_tgt, state dtype (+inf: no drain) */
  /* Green-Ampt infiltration and drain-tau (ga_drain_flat, compressed_solver.py:320-367; its ga
   * and drain modes, 369-421, are this with one part NULL) on every active cell with h > 0,
   * after the sponge and before the clamp: with K = ks[c] > 0, a = F + K dt,
   * disc = a*a + ((4 K dt)(psi[c] + h)) dth[c], F1 = (a + sqrt(max(disc, 0)))/2,
   * dF = min(max(F1 - F, 0), h), capped by the room fmax - F (0 when none), h -= dF, F += dF,
   * scale = h'/h; then with it = inv_tau > 0 and h > 0, decay = exp(-dt it) from IEEE
   * operations only (in float64, rounded to the state dtype), scale *= decay, h *= decay;
   * hu, hv *= scale. */
  const uint8_t* ga_cls;     /* [N] soil class, needed with ga_ks */
  const void* ga_ks;         /* nullable [classes] Ks (m/s), state dtype: no Green-Ampt when NULL */
  const void* ga_psi;        /* [classes] suction head (m) */
  const void* ga_dth;        /* [classes] moisture deficit */
  void* ga_F;                /* [N] cumulative infiltration (m), state dtype, updated in place */
  const void* ga_fmax;       /* nullable [N] storage cap of F */
  const void* ga_inv_tau;    /* nullable [N] drain rate 1/tau (1/s): no drain when NULL */
  /* Rain on a native grid (the run driver's bundle, compressed_solver.py:3157-3183, 2329):
   * with gridded rain, a cell reads column rain_lookup[k] of the step's row of the
   * (nt, npix) rate table (GsForcing.rain_frame_stride = npix) instead of its own cell of an
   * (nx, ny) frame. */
  const int32_t* rain_lookup; /* nullable [N] */
  /* Ring modes of the cached replay (compressed_solver.py:486-575, applied at 2134-2144):
   * at the top of each step, on the step's input state, before anything reads it, each
   * listed ring cell k with kind
   *   0 (extrapolate): h, hu, hv <- those of its active neighbour ring_mode_nb;
   *   1 (stage):       h <- max(eta - bed[k], 0), hu = hv = 0, with eta = ring_mode_eta;
   *   2 (stage, then velocity from the neighbour, ghost_uv 450-465): the stage, then
   *                    r = h_nb > h_min ? 1/h_nb : 0, hu <- h*(hu_nb*r), hv <- h*(hv_nb*r). */
  int64_t n_ring_mode;
  const int64_t* ring_mode_ids;  /* [n_ring_mode] */
  const int64_t* ring_mode_nb;   /* [n_ring_mode] neighbour id (kinds 0 and 2) */
  const uint8_t* ring_mode_kind; /* [n_ring_mode] */
  double ring_mode_eta;
  /* GS_FLAG_FORCE cells of the ring (stored, not active): the step writes
   * force(state) for them, as the reference's gathered kernels touch every stored cell
   * of the band; the residual and the CFL never see them. */
  int64_t n_force_ring;
  const int64_t* force_ring; /* [n_force_ring] flat ids */
  /* IDW gauge ring on ring cells (ring_bc_flat, compressed_solver.py:287-307, with the
   * per-step stage interpolation of 2338-2348), between the sponge and the clamp, at the
   * step's start time t: ti = clamp(#{t_common < t} - 1, 0, nt - 2),
   * wt = clamp((t - t_common[ti]) / (t_common[ti+1] - t_common[ti] + 1e-12), 0, 1) in
   * float64, s_g = S[g][ti]*(T)(1 - wt) + S[g][ti+1]*(T)wt, eta = sum over g ascending of
   * w[r][g]*s_g from 0, then h = max(0, eta - ring_bed[r]), hu = hv = 0. */
  const int32_t* force_ring_row; /* nullable [n_force_ring]: ring row r of each entry, -1 for none */
  int32_t ring_ng, ring_nt;  /* gauges, samples of the common time grid (>= 2) */
  const void* ring_bed;      /* [rows] state dtype */
  const void* ring_w;        /* [rows][ring_ng] state dtype */
  const double* ring_t;      /* [ring_nt] ascending */
  const void* ring_stage;    /* [ring_ng][ring_nt] state dtype */
  const uint8_t* cls;        /* nullable [N] Manning class ids 0..255 into ntab, in place of the
                                flags' class bits (also for the per-class infiltration rate) */
} GsFlat;

/* Host arrays of a flat mesh built by gs_flat_mesh_build (allocated by the library with
 * malloc; release them with gs_flat_mesh_free). Same meaning as the GsFlat tables. */
typedef struct GsFlatMeshHost {
  int64_t n_stored, n_active, n_runs, n_chunks, n_tiles, n_entries;
  int32_t sweep_rows;        /* tile height used */
  int32_t pad_;
  int32_t* ij;               /* [n_stored][2] padded (i, j) of each stored cell, row-major order */
  uint8_t* active;           /* [n_stored] 1 for an active cell, 0 for the frozen ring */
  int64_t* run_k0;           /* [n_runs] */
  int32_t* run_len;          /* [n_runs] */
  int32_t* run_i;            /* [n_runs] */
  int32_t* run_j0;           /* [n_runs] */
  int32_t* run_d;            /* [n_runs][4] */
  int32_t* chunk_run;        /* [n_chunks] */
  int64_t* tile_e0;          /* [n_tiles] */
  int32_t* tile_irc;         /* [n_tiles][3] */
  uint64_t* ent;             /* [n_entries][3] */
} GsFlatMeshHost;

/* ---- library ---- */
int gs_version(void);
const char* gs_last_error(void);
int gs_get_device_count(int* n);
int gs_set_device(int dev);

/* ---- memory and streams (the library is the allocator of the plain solver path) ---- */
int gs_malloc(void** ptr, size_t bytes);
int gs_free(void* ptr);
int gs_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream);
int gs_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream);
int gs_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream);
int gs_memcpy2d(void* dst, size_t dpitch, const void* src, size_t spitch, size_t width,
                size_t height, int kind, void* stream); /* kind: 1 h2d, 2 d2h, 3 d2d */
int gs_memset(void* dst, int value, size_t bytes, void* stream);
int gs_stream_create(void** stream);
int gs_stream_destroy(void* stream);
int gs_stream_synchronize(void* stream);
int gs_event_create(void** ev);
int gs_event_destroy(void* ev);
int gs_event_record(void* ev, void* stream);
int gs_event_elapsed_ms(float* ms, void* start, void* stop);

/* ---- clock ---- */
/* Reset a device clock to t0 (t_end = inf, no dt cap, nsteps 0). */
int gs_clock_reset(GsClock* clock_dev, double t0, void* stream);
/* Set the run target: t_end (inf = open), dt_max (0 = none), max_steps (<0 = none);
 * clears GS_ST_DONE and re-plans dt for the current state. */
int gs_clock_set_target(GsClock* clock_dev, double t_end, double dt_max, int64_t max_steps,
                        const GsPhys* phys, GsForcing forc, void* stream);
/* Use `dt` (float64) for the next step instead of the CFL step (Solver2D.step(dt)). */
int gs_clock_set_dt(GsClock* clock_dev, double dt, const GsPhys* phys, GsForcing forc,
                    void* stream);

/* ---- dense layout ---- */
/*
 * Replaces: the SRM-HLLC residual + update + friction + wet/dry + running-max fused
 * step fused_rhs_wb_srm_hllc_compact_fstepcfl_* (rhs_cuda.py:1746-1796, 2095-2149,
 * launched at solver.py:2271-2299), the ghost fill apply_bc_2d (bc.py:50-98), the
 * stage overwrite StageBoundary.apply (forcing.py:349-377) and the next step's
 * lammax_fp32_lean + cfl_dt host arithmetic (solver.py:152-205, 2569-2666).
 * Reads q[nsteps & 1], writes q[(nsteps+1) & 1], advances the clock.
 */
int gs_dense_step(const GsDense* d, const GsPhys* phys, GsForcing forc, GsClock* clock_dev,
                  void* stream);
/*
 * Replaces: fused_rhs_wb_srm_hllc_2d_* (rhs_cuda.py:900-1208) for parity checks:
 * residual of q[which] into rhs planes (interior cells; flags' inactive cells get 0).
 */
int gs_dense_residual(const GsDense* d, const GsPhys* phys, int which, void* rhs0, void* rhs1,
                      void* rhs2, void* stream);
/*
 * Replaces: lammax_fp32_lean (solver.py:152-205) + the dt rule of cfl_dt
 * (solver.py:2657-2666): wave speed of q[nsteps & 1] into clock->lam_bits and the
 * next dt planned on the device.
 */
int gs_dense_cfl(const GsDense* d, const GsPhys* phys, GsForcing forc, GsClock* clock_dev,
                 void* stream);

/* ---- flat layout ---- */
/*
 * Replaces: flat_srm_hllc_fstepcfl_* (compressed_rhs.py:673-865, 1000-1118) launched
 * from CompressedStepper.rhs_fused (compressed_solver.py:880-904) plus the dt rule of
 * dt_from_lam (compressed_solver.py:713-719).
 */
int gs_flat_step(const GsFlat* f, const GsPhys* phys, GsForcing forc, GsClock* clock_dev,
                 void* stream);
/* Replaces: flat_srm_hllc_* residual (compressed_rhs.py:129-158) for parity checks. */
int gs_flat_residual(const GsFlat* f, const GsPhys* phys, int which, void* rhs0, void* rhs1,
                     void* rhs2, void* stream);
/* Replaces: cfl_lammax_flat_blk + CompressedStepper.cfl_dt (compressed_solver.py:200-240, 721-748). */
int gs_flat_cfl(const GsFlat* f, const GsPhys* phys, GsForcing forc, GsClock* clock_dev,
                void* stream);

/* No reference counterpart (a launch query for tests and tuning): the kernel gs_flat_step
 * launches for this descriptor (0 chunk traversal, 1 plain row sweep, 2 extended row sweep),
 * its grid in workgroups and that kernel's resident grid on the current device (the grid
 * exceeds it when the step launches one wavefront per tile). No device work. */
int gs_flat_launch_info(const GsFlat* f, const GsPhys* phys, GsForcing forc, int* kernel, int64_t* blocks,
                        int64_t* resident);

/* Replaces: nothing in the reference (the adjoint recomputation needs it): gs_flat_step
 * that also stores the pre-friction state h + dt*rr0, hu + dt*r1, hv + dt*r2 of active
 * cells into pre0..2 (flat arrays of n_stored), the input of the update's reverse. */
int gs_flat_step_pre(const GsFlat* f, const GsPhys* phys, GsForcing forc, GsClock* clock_dev, void* pre0,
                     void* pre1, void* pre2, void* stream);
/* Depth of the current state (q[clock->nsteps & 1]) at n flat ids, into out (state dtype):
 * gauge sampling, the device side of reading s.depth() at gauge cells. */
int gs_flat_sample(const GsFlat* f, const GsClock* clock_dev, const int64_t* ids, int n, void* out,
                   void* stream);

/* ---- reverse (adjoint) step of the flat layout ---- */
/*
 * No reference counterpart (the reference has no adjoint). One reverse step of the flat
 * fused step with a frozen dt: the cotangent of q_{s+1} (lam, in) becomes that of q_s
 * (lam, out) and the cotangents of the per-class Manning n (nbar) and of the per-cell bed
 * (bbar) accumulate. Stored cells carry a full 4-neighbour table; w, xf, yf and gz are
 * scratch that must be zero-initialised once (w, xf, yf) before the first reverse step.
 */
#define GS_ADJ_MAX_BLOCKS 8192   /* workgroups of every adjoint pass (npart has this many rows at most) */
typedef struct GsAdjFlat {
  int64_t n_stored;
  int32_t dtype;
  int32_t n_classes;         /* Manning classes (1..32) */
  double dx, dy;
  const uint8_t* active;     /* [N] 1 for cells the step updates */
  const int32_t* nbr;        /* [N][4] flat ids of (i+1, j), (i-1, j), (i, j+1), (i, j-1); -1: not stored */
  const void* bed;           /* [N] */
  const uint8_t* flags;      /* [N] GS_FLAG_* (class, stage) */
  const void* ntab;          /* Manning n per class */
  const void* q[3];          /* state q_s */
  const void* h_next;        /* depth of q_{s+1} (stage cells) */
  const void* pre[3];        /* pre-friction state of the step s -> s+1 (gs_flat_step_pre) */
  void* lam[3];              /* cotangent of q_{s+1} in, of q_s out */
  void* w[3];                /* scratch [N] each */
  void* xf;                  /* scratch [10][N] */
  void* yf;                  /* scratch [10][N] */
  void* gz;                  /* scratch [2][N] */
  const void* gz_prev;       /* NULL, or the gz of the previous reverse step: its bed-gradient
                                term is then folded into this step's gather pass */
  double* bbar;              /* [N] accumulated d(loss)/d(bed) */
  double* nbar;              /* [n_classes] d(loss)/d(n), from gs_adj_flat_nsum */
  double* npart;             /* [min(ceil(N/256), GS_ADJ_MAX_BLOCKS)][32] per-workgroup class sums,
                                zero-initialised once; summed into nbar in a fixed order */
} GsAdjFlat;
int gs_adj_flat_step(const GsAdjFlat* a, const GsPhys* phys, double dt, void* stream);
/* nbar[k] += the per-workgroup class sums npart[.][k] in workgroup order (deterministic). */
int gs_adj_flat_nsum(const GsAdjFlat* a, void* stream);
/* The bed-gradient term of a->gz alone (the flush after the last step run with gz_prev). */
int gs_adj_flat_bedgrad(const GsAdjFlat* a, const GsPhys* phys, void* stream);
/* lam_h[ids[k]] += coef[k]: the loss's direct cotangent of a state (gauge misfit). */
int gs_adj_add(void* lam_h, int dtype, const int64_t* ids, const void* coef, int n, void* stream);
/* Forward pass of a calibration run: n frozen-dt steps (gs_clock_set_dt + gs_flat_step)
 * from q[0] with the clock at step 0, the depth at n_gauges flat ids after every step into
 * samples[n][n_gauges], and when ckpt is not NULL the state every ckpt_every steps
 * (before step j) into ckpt[3*(j/ckpt_every) .. +2]. */
int gs_flat_run_frozen(const GsFlat* f, const GsPhys* phys, GsForcing forc, GsClock* clock_dev, int n,
                       const double* dts, const int64_t* gauge_ids, int n_gauges, void* samples, int ckpt_every,
                       void* const* ckpt, void* stream);
/* Segment recomputation: n frozen-dt steps (gs_clock_set_dt + gs_flat_step_pre) where step j
 * reads the state states[3j..3j+2] and writes states[3j+3..3j+5] (device pointers; the
 * clock's step count starts at 0), keeping the pre-friction state in pre[3j..3j+2]. */
int gs_flat_replay_pre(const GsFlat* f, const GsPhys* phys, GsForcing forc, GsClock* clock_dev, int n,
                       const double* dts, void* const* states, void* const* pre, void* stream);
/* Reverse sweep of such a segment: for j = n-1 .. 0 the reverse step at dts[j] over
 * states[3j..], the depth states[3j+3] and pre[3j..], then, when coef[j] is not NULL,
 * gs_adj_add of coef[j] at the n_gauges gauge ids (the loss term of the state after step j).
 * With gz_pair (two [2][N] halves) step first_step + j writes half (first_step + j) & 1 and
 * folds in the other half's bed-gradient term; flush the last with gs_adj_flat_bedgrad. */
int gs_adj_flat_sweep(const GsAdjFlat* a, const GsPhys* phys, int n, const double* dts, const void* const* states,
                      const void* const* pre, const int64_t* gauge_ids, const void* const* coef, int n_gauges,
                      void* gz_pair, int64_t first_step, void* stream);

/* ---- host mesh builder ---- */
/*
 * Replaces: the CompressedMesh2D / CompressedSWE construction (compressed_mesh.py:108-187,
 * compressed_rhs.py:241-369) that CompressedSolver.from_dense (compressed_solver.py:2903-3124)
 * runs on the host. `inside` is the (nxp, nyp) active mask of the padded grid (row-major,
 * 0/1); the stored set is its `ring`-fold 4-neighbour dilation (ring >= 2), numbered
 * row-major. sweep_rows > 0 fixes the tile height; -1 picks the largest power of two <=
 * max_sweep_rows that still gives ~tile_target tiles. Host memory only, no device call.
 */
int gs_flat_mesh_build(const uint8_t* inside, int nxp, int nyp, int ring, int sweep_rows, int tile_target,
                       int max_sweep_rows, GsFlatMeshHost* out);
int gs_flat_mesh_free(GsFlatMeshHost* m);

/* ---- batched ensemble of flat domains ---- */
/*
 * Independent flat domains (members of uneven size, each with its own bed, Manning table,
 * rain, stage and clock) advanced together: one launch per step over all members' sweep
 * tiles (members need n_tiles > 0),
 * each member with its own CFL dt planned on the device, steps replayed from a hipGraph.
 * No reference counterpart: the reference runs members one after another through
 * CompressedSolver.run (compressed_solver.py:3359-3450). Member k's results are bitwise
 * those of gs_flat_step on the same member.
 */
typedef struct GsEns GsEns;
/* Build an ensemble over n_members flat descriptors (device buffers stay owned by the
 * caller and must outlive the ensemble). All members share one dtype. phys/forcings/clocks
 * are arrays of n_members. The clocks should already hold each member's first dt
 * (gs_flat_cfl + gs_clock_set_target), as for gs_flat_step. */
int gs_ens_create(const GsFlat* members, const GsPhys* phys, const GsForcing* forcings,
                  GsClock* const* clocks, int n_members, GsEns** out);
int gs_ens_destroy(GsEns* e);
/* Launch shape: resident workgroups (0 = one full occupancy wave of the device) and the
 * consecutive row-sweep tiles a wavefront takes per pass (1..64). */
int gs_ens_set_launch(GsEns* e, int blocks, int tiles_per_pass);
int gs_ens_get_launch(const GsEns* e, int* blocks, int* tiles_per_pass, int64_t* total_tiles);
/* One step of every member that is not done. */
int gs_ens_step(GsEns* e, void* stream);
/* nsteps steps; graph_batch > 0 replays a captured hipGraph of graph_batch steps. */
int gs_ens_run(GsEns* e, int64_t nsteps, int graph_batch, void* stream);
/* Time candidate launch shapes (trial_steps steps each, on a snapshot of the state that is
 * restored afterwards), keep the fastest; writes its mean ms per step. Allocates scratch
 * for the snapshot. */
int gs_ens_autotune(GsEns* e, int trial_steps, void* stream, double* best_ms);

/* No reference counterpart: for callers that cannot allocate device memory (tight serving
 * loops that step a batch of domains, harnesses that forbid allocation). nsteps steps of
 * every member, bitwise what gs_ens_create + gs_ens_run(nsteps, 0) + gs_ens_destroy give,
 * without any device allocation: the member table, the tile prefix and the last-workgroup
 * counter live in a static 256 KiB device array of the library (each GPU holds its own
 * until the process ends), so at most 409 float32 or 384 float64 members (GS_E_RANGE
 * beyond). Members use their own tile tables and one occupancy wave of workgroups. The call
 * waits for earlier work on `stream` (an earlier call's steps may still read the array),
 * copies the table synchronously, then enqueues the steps; calls that overlap on one GPU
 * (other streams or threads) race on the array. */
int gs_ens_run_noalloc(const GsFlat* members, const GsPhys* phys, const GsForcing* forcings,
                       GsClock* const* clocks, int n_members, int64_t nsteps, void* stream);

/* ---- graph capture of step entry points ---- */
/*
 * Replaces: the reference's K-step capture of its fixed-dt dense step in a CUDA graph
 * (tests/test_gpu_dense_no_host_sync.py:1-96 pins "K captured steps = K launches" byte for
 * byte). Begin capture on a created stream, call any step entry points (gs_dense_step,
 * gs_flat_step, ...) on that stream, end the capture into an executable graph, then launch
 * it on any stream as often as needed: each launch replays the recorded kernels with the
 * arguments they were recorded with. The step entry points keep dt, t and the wave speed in
 * the device clock and record no memset or host synchronisation, so a replay advances the
 * clock exactly as the same direct launches would. The buffers and tables a recorded call
 * names must outlive the graph.
 */
typedef struct GsGraph GsGraph;
int gs_graph_begin(void* stream);
int gs_graph_end(void* stream, GsGraph** out);
int gs_graph_nodes(const GsGraph* g, int64_t* nodes);
int gs_graph_launch(const GsGraph* g, void* stream);
int gs_graph_destroy(GsGraph* g);

/* ---- set-up and measurement utilities ---- */
/* Constant fill of a (rows, cols) plane with row pitch `pitch` (elements). No reference
 * counterpart beyond the host-side np.full/np.zeros of the solver constructors. */
int gs_fill(void* dst, int dtype, int64_t rows, int64_t cols, int64_t pitch, double value, void* stream);
/*
 * Seeded synthetic bed b(i, j) = slope_x*x_I + scale*(fbm(I, J) - mu) + offset on a (nx, ny)
 * plane, I = i0 + i*stride, J = j0 + j*stride: value-noise fBm over Philox2x32-10 lattice
 * values, `octaves` octaves of spacing L0/2^k cells and gain 2^-hurst, x_I = (I+0.5)*dx.
 * With round_f32 the float64 value is rounded to float32 first (a float64 run on the
 * float32 run's exact inputs). Replaces the host-side
 * synthetic-case set-up of the reference's scaling harness
 * (benchmark/scaling_640m/scaling_bench.py:59-83); oracle/generators.py mirrors it bit for bit.
 */
int gs_gen_bed(void* dst, int dtype, int nx, int ny, int64_t pitch, int64_t i0, int64_t j0, int64_t stride,
               uint32_t seed, int octaves, double L0, double hurst, double mu, double scale, double slope_x,
               double dx, double offset, int round_f32, void* stream);
/* Voronoi class field for synthetic Manning land cover: each cell of the (nx, ny) plane
 * takes the class of the nearest of `nsites` sites (float32 squared distance in cell units
 * from (i0+i+0.5, j0+j+0.5), ties to the lowest index). Replaces the host-side land-cover
 * preparation of the reference's cases (data_prep.py NLCD -> Manning), for synthetic
 * benchmark inputs; oracle/generators.py mirrors it bit for bit. */
int gs_gen_voronoi(uint8_t* dst, int nx, int ny, int64_t pitch, int64_t i0, int64_t j0, const float* site_x,
                   const float* site_y, const uint8_t* site_cls, int nsites, void* stream);
/* Plain streaming copy kernel (16 B per lane), used to calibrate HBM bandwidth and the
 * rocprofv3 FETCH_SIZE/WRITE_SIZE counters. No reference counterpart. */
int gs_bench_copy(void* dst, const void* src, size_t bytes, void* stream);
/* VALU issue-rate calibration: `blocks` workgroups of 256 lanes, each lane 8 independent
 * float32 chains of `iters` steps; mode 0: multiply + add (16 VALU instructions per step),
 * mode 1: hardware square root, multiply, hardware reciprocal, add (32 per step). `out`
 * holds blocks * 256 floats. No reference counterpart. */
int gs_bench_valu(float* out, int blocks, int iters, int mode, void* stream);
/* Check of the short correctly rounded float32 division and square root the kernels use
 * against the compiler's general IEEE expansions on this device. which 0: 1/b for every
 * float32 b; 1: sqrt(x) for every x (the guarded form; the short form where |x| is outside (0, 2^-96));
 * 2: a/b for 2^32 seeded bit-pattern pairs; 3: a/b for every a over four divisors of hard
 * significands. counts_dev (4 x uint64, zeroed by the caller) receives, per check, the
 * mismatch count and 1 + the largest offending index. No reference counterpart. */
int gs_crmath_check(int which, unsigned long long* counts_dev, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* GEOSWE_ABI_H */
