/* libsvimx -- MI355X-native narrow-band level-set erosion (C ABI).
 *
 * Every entry point takes the HIP stream as `void*` (a hipStream_t), raw
 * device pointers, int64 sizes and double scalars, and returns 0 or a
 * negative SVIMX_E* code; svimx_last_error() describes the last failure of
 * the calling thread.  No C++ exception crosses this boundary.
 *
 * Layout: dense fields are float32 C-order (x, y, z) grids of nx*ny*nz nodes
 * (flat strides ny*nz, nz, 1); a band is a sorted int32 list of flat node
 * indices; points are float32 index (cell) coordinates, 3 per point.
 *
 * Ownership: the caller allocates every buffer (the Python host uses the
 * PyTorch caching allocator) including one workspace per job of
 * svimx_workspace_bytes() bytes whose first SVIMX_R_COUNT doubles hold the
 * SVIMX_R_* results.  Plans (svimx_nr_plan_*) own their small device tables.
 *
 * Asynchrony: functions enqueue on `stream` and return, except those whose
 * comment says "synchronises" (they report a count the host needs).
 *
 * The reference interface each group replaces is cited as file:line of
 * jwcalder/SVIErosion.
 */
#ifndef SVIMX_H_
#define SVIMX_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  SVIMX_OK = 0,
  SVIMX_EBADARG = -1,   /* invalid argument (maps to ValueError) */
  SVIMX_EMARGIN = -2,   /* band reaches the grid boundary (maps to RuntimeError) */
  SVIMX_ECAPACITY = -3, /* an output buffer is too small; required size reported */
  SVIMX_EDEVICE = -4    /* HIP runtime error */
};

/* Result slots (doubles) at the start of the workspace. */
enum {
  SVIMX_R_VOLUME_SUM = 0, /* sum over the band of cell fractions (FP64) */
  SVIMX_R_HAS_NEG = 1,    /* band holds phi < 0 */
  SVIMX_R_HAS_POS = 2,    /* band holds phi > 0 */
  SVIMX_R_VMAX = 3,       /* max of V after the cutoff, over the band */
  SVIMX_R_QMIN = 4,       /* min / max of N_r (SVI) or H (curvature) over |phi| < h */
  SVIMX_R_QMAX = 5,
  SVIMX_R_DVMAX = 6,      /* max dV/dN over |phi| < h (SVI) */
  SVIMX_R_CAPPED = 7,     /* count of N > n_cap over |phi| < h */
  SVIMX_R_VSUM = 8,       /* sum of V before the cutoff over |phi| < h (FP64) */
  SVIMX_R_NNEAR = 9,      /* count of |phi| < h */
  SVIMX_R_DMAX = 10,      /* max dV/dk1 + dV/dk2 over |phi| < h (curvature) */
  SVIMX_R_QMIN_ALL = 11,  /* the same statistics over every evaluated band node, */
  SVIMX_R_QMAX_ALL = 12,  /* used when no node has |phi| < h                      */
  SVIMX_R_DVMAX_ALL = 13,
  SVIMX_R_CAPPED_ALL = 14,
  SVIMX_R_VSUM_ALL = 15,
  SVIMX_R_NALL = 16,
  SVIMX_R_DMAX_ALL = 17,
  SVIMX_R_TILES = 18,     /* N_r tiles evaluated by the last svimx_nr_exec */
  SVIMX_R_DT = 31,        /* stable dt: svimx_stable_dt, or the last tick of an engine job */
  SVIMX_R_COUNT = 32
};

/* Speed law parameters (svierosion/flows.py:29-213). */
typedef struct svimx_law {
  int32_t kind; /* 0 SVI power law, 1 mean curvature, 2 Gauss curvature */
  int32_t k;    /* SVI exponent */
  int32_t s;    /* SVI geometric switch (0/1) */
  int32_t reserved;
  double t;     /* SVI singularity */
  double eps;   /* additive constant */
  double n_cap; /* SVI cap on N (1 when s = 0) */
  double alpha; /* curvature exponent */
  double delta; /* curvature regularisation */
} svimx_law_t;

const char* svimx_last_error(void);
int svimx_version(void);
int64_t svimx_workspace_bytes(int nx, int ny, int nz);

/* The occupancy field f (1 inside, 0 outside, cell fraction on the band) is
 * stored padded: node (x, y, z) lives at ((x+16)*(ny+32) + (y+16))*pz + z+16
 * with pz = round_up(nz + 40, 4); the pad cells stay 0.  This is the element
 * count of that array (float32, 16-byte aligned, zero-initialised by the
 * caller). */
int64_t svimx_occupancy_elems(int nx, int ny, int nz);

/* ---------------------------------------------------------------- band
 * Clamp phi to +-width*h, band = nodes with |phi| < width*h*(1-1e-5) (float32
 * compare), write the sorted band, copy the clamped field into phi_copy1/2
 * (may be NULL) and reset `occupancy` to (phi < 0).  Synchronises.
 * info (host, 9 int64): band count, inside nodes outside the band, margin ok,
 * min i/j/k and max i/j/k over the band.
 * Replaces LevelSet.rebuild_band / _check_margin (svierosion/levelset.py:70-86);
 * the margin failure is reported through info[2] and SVIMX_EMARGIN. */
int svimx_band_rebuild(void* stream, float* phi, float* phi_copy1, float* phi_copy2, float* occupancy,
                       int nx, int ny, int nz, double h, double width, int margin, int32_t* band,
                       int64_t band_cap, void* ws, int64_t ws_bytes, int64_t* info);

/* ---------------------------------------------------------------- per-stage kernels
 * Cell fractions of the band written into `occupancy`, their FP64 sum
 * (SVIMX_R_VOLUME_SUM) and the interface flags.  When `nr_plan` is not NULL the
 * plan's tiles are reset and the tiles holding the tricubic stencils of the
 * closest points of band nodes with |phi| < cut_hi*h are marked for svimx_nr_exec.
 * Replaces LevelSet.cell_fractions / dense_fractions / volume / has_interface
 * (svierosion/levelset.py:93-126, svierosion/kernels.py:173-181). */
int svimx_fractions(void* stream, const float* phi, const int32_t* band, int64_t nb, int nx, int ny,
                    int nz, double h, float* occupancy, void* nr_plan, double cut_hi, void* ws,
                    int64_t ws_bytes);

/* Principal curvatures of the band scattered into a dense field of (k1, k2)
 * pairs (2 floats per grid node, interleaved so one 8-byte load fetches both
 * at a trilinear corner) that is zero elsewhere (svierosion/evolve.py:136-143,
 * svierosion/kernels.py:274-299). */
int svimx_curvatures(void* stream, const float* phi, const int32_t* band, int64_t nb, int nx, int ny,
                     int nz, double h, float* k12_dense);

/* Velocity of every band node: closest point, interpolated N_r (tricubic) or
 * k1/k2 (trilinear), speed law, smoothstep cutoff cut_lo*h -> cut_hi*h; with
 * stats != 0 the SVIMX_R_VMAX..SVIMX_R_DMAX_ALL results.  `field` is the dense
 * N_r for SVI laws; k12 the dense (k1, k2) pairs of svimx_curvatures otherwise.
 * Replaces ErosionSimulation._velocity (svierosion/evolve.py:108-155). */
int svimx_velocity(void* stream, const float* phi, const int32_t* band, int64_t nb, int nx, int ny,
                   int nz, double h, const svimx_law_t* law, const float* field, const float* k12,
                   double cut_lo, double cut_hi, int stats, float* V, void* ws, int64_t ws_bytes);
/* Velocity fused with the TVD-RK3 stage that consumes it (svierosion/evolve.py:
 * 108-160, :172-179): stage 0 computes the velocity with the dt statistics (as
 * svimx_velocity with stats) and stores L0 = V |grad phi_in| (band-indexed);
 * the update u0 + dt L0 is then svimx_rk_stage stage 3.  Stages 1 and 2 write
 * phi_out = 3/4 u0 + 1/4 (phi_in + dt L) and 1/3 u0 + 2/3 (phi_in + dt L) at
 * the band nodes, V never stored.  The same floats as svimx_velocity followed
 * by svimx_rk_stage (bit-identical fields and statistics).  A negative dt
 * (SVIMX_DT_DEVICE) makes stages 1 and 2 read the step svimx_stable_dt left in
 * the workspace's SVIMX_R_DT slot, so no host synchronisation sits between
 * stage 0 and the rest of the step. */
#define SVIMX_DT_DEVICE (-1.0)
int svimx_velocity_rk(void* stream, int stage, const float* phi_in, const float* u0, float* phi_out,
                      const int32_t* band, int64_t nb, int nx, int ny, int nz, double h, const svimx_law_t* law,
                      const float* field, const float* k12, double cut_lo, double cut_hi, double dt, float* L0,
                      void* ws, int64_t ws_bytes);
/* svimx_velocity_rk with each band entry's closest point given (points: nb x 3
 * float32, index coordinates, e.g. from svimx_closest_points_newton) instead of
 * the one-step projection of the node: the reference's velocity with
 * newton_steps > 1 (svierosion/evolve.py:118, svierosion/levelset.py:129-151). */
int svimx_velocity_rk_points(void* stream, int stage, const float* phi_in, const float* u0, float* phi_out,
                             const int32_t* band, int64_t nb, int nx, int ny, int nz, double h,
                             const svimx_law_t* law, const float* field, const float* k12, double cut_lo,
                             double cut_hi, double dt, float* L0, const float* points, void* ws, int64_t ws_bytes);
/* Closest points of n nodes (idx) on the zero set of phi, index coordinates
 * (n x 3 float32): the capped projection with the central gradient, then
 * newton_steps - 1 capped Newton steps with the tricubic value and gradient of
 * phi, each clamped to where tricubic stencils fit.  Replaces
 * LevelSet.closest_points (svierosion/levelset.py:129-151) for any newton_steps
 * in [1, 64]. */
int svimx_closest_points_newton(void* stream, const float* phi, const int32_t* idx, int64_t n, int nx, int ny,
                                int nz, double h, int newton_steps, float* out_p);
/* svimx_velocity without a workspace: no counterpart in the reference.  For
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The block partials of
 * the statistics live in a fixed 256 KiB device array per GPU (held until the
 * process ends; it covers every band size up to 2^31 - 1 nodes, larger nb is
 * refused); the statistics go to `results` (a caller device array of
 * SVIMX_R_COUNT doubles; slots SVIMX_R_VMAX .. SVIMX_R_DMAX_ALL are written,
 * the same values svimx_velocity writes into its workspace).  The stream must
 * belong to the current device.  Calls that overlap on one GPU share that array
 * and race: serialise them.  Allocates nothing. */
int svimx_velocity_static(void* stream, const float* phi, const int32_t* band, int64_t nb, int nx, int ny, int nz,
                          double h, const svimx_law_t* law, const float* field, const float* k12, double cut_lo,
                          double cut_hi, int stats, float* V, double* results);

/* One TVD-RK3 stage over the band (svierosion/evolve.py:157-160, :172-179):
 *   L = V * |grad phi_in|_Godunov-WENO5
 *   stage 0: out = u0 + dt L            (L saved into L0 when not NULL)
 *   stage 1: out = 3/4 u0 + 1/4 (in + dt L)
 *   stage 2: out = 1/3 u0 + 2/3 (in + dt L)
 *   stage 3: out = u0 + dt L0           (replays stage 0 with a new dt; no stencil)
 * u0 is the dense field at the start of the step; out may not alias in. */
int svimx_rk_stage(void* stream, int stage, const float* phi_in, const float* u0, float* phi_out,
                   const int32_t* band, int64_t nb, int nx, int ny, int nz, double h, double dt,
                   const float* V, float* L0);
/* svimx_rk_stage with dt read on the device from *dt_dev (FP64, rounded to
 * float as the host argument is; a step that is not finite and positive
 * applies 0).  No counterpart in the reference: with svimx_stable_dt it lets
 * a step's stages follow stage 0 without a host round trip. */
int svimx_rk_stage_dt(void* stream, int stage, const float* phi_in, const float* u0, float* phi_out,
                      const int32_t* band, int64_t nb, int nx, int ny, int nz, double h, const double* dt_dev,
                      const float* V, float* L0);
/* The stable step of ErosionSimulation.stable_dt (svierosion/evolve.py:162-170,
 * with the max_time clamp of :199-200 when clamp != 0: dt = min(dt, max_dt))
 * from the statistics a stage-0 velocity pass left in the workspace, written on
 * the device into the workspace's SVIMX_R_DT slot: FP64 in the host formula's
 * operation order, bit-identical to it.  radius is the SVI ball radius (unused
 * by curvature laws).  The stages then take SVIMX_DT_DEVICE (svimx_velocity_rk)
 * or the slot's address (svimx_rk_stage_dt). */
int svimx_stable_dt(void* stream, const svimx_law_t* law, double h, double cfl, double feedback_cfl,
                    double parabolic_cfl, double radius, int clamp, double max_dt, void* ws, int64_t ws_bytes);

/* ---------------------------------------------------------------- N_r (ball convolution)
 * A plan holds the FP32 ball weights (svierosion/svi.py:24-55, computed on the
 * host in FP64) as a per-column table, the tile marks and the tile list.
 * svimx_nr_exec evaluates N_r = 1 - h^3 (w * f)(x) / (4/3 pi r^3) by direct
 * summation at every node of every marked tile and writes it into nr_out
 * (other nodes untouched).  The weights must be mirror-symmetric in x, y and z
 * as FP32 values (a ball's are); others are refused with SVIMX_EBADARG.
 * Replaces BallConvolution (svierosion/svi.py:58-93). */
int svimx_nr_plan_create(const double* weights, int m, double h, double radius, int nx, int ny, int nz,
                         void** plan);
int svimx_nr_plan_destroy(void* plan);
int svimx_nr_reset(void* stream, void* plan);
int svimx_nr_mark_nodes(void* stream, void* plan, const int32_t* nodes, int64_t n);
/* Mark the tricubic stencils of n index-space points (n x 3 float32, device):
 * after select + sum, N_r is valid at every node an interpolation at those
 * points reads (axis-clamped).  Replaces the stencil-node lists the reference
 * evaluates at (svierosion/evolve.py:126-127, svierosion/api.py:220-223). */
int svimx_nr_mark_points(void* stream, void* plan, const float* points, int64_t n);
/* Evaluation method of a plan: 1 direct summation (at the marked nodes), 2
 * FFT over the whole grid (the reference's circular convolution,
 * svierosion/svi.py:84-93: forward 3-D FFT of f, product with the real
 * spectrum of the ball, and a forward FFT of the conjugate; mixed-radix
 * Stockham line transforms in LDS).  Plans of balls wider than 33 cells are
 * created with the FFT.  The FFT allocates 12 bytes per grid node. */
int svimx_nr_plan_set_method(void* plan, int method);
int svimx_nr_plan_method(void* plan);
/* Autotuner: times the run-mode summation over the shapes read through the
 * L1 (z-run length 4/8/16 x y-lines per unit 1/2/4; svimx_nr_tune_shape
 * names shape i of SVIMX_NR_TUNE_SHAPES, which also lists the LDS-staged
 * one-line kernels) on this job's field (fractions and selection recomputed
 * per candidate; these shapes sum in one order, so the N_r values at the
 * nodes the velocity reads are identical for every one) and keeps the
 * fastest.  ms_out (SVIMX_NR_TUNE_SHAPES doubles, NaN for the LDS shapes and
 * for shapes this plan cannot take; may be NULL) receives the per-launch
 * times.  Leaves occupancy, marks and unit list those of the chosen shape for
 * phi.  Synchronises.  No counterpart in the reference. */
#define SVIMX_NR_TUNE_SHAPES 11
int svimx_nr_tune_shape(int i, int* run_length, int* lines, int* lds);
int svimx_nr_plan_tune(void* stream, void* plan, const float* phi, const int32_t* band, int64_t nb, int nx, int ny,
                       int nz, double h, double cut_hi, float* occupancy, float* nr_out, void* ws, int64_t ws_bytes,
                       double* ms_out);
/* Run length of a run-mode plan (0: tile mode); the setter takes 4, 8 or 16
 * and invalidates the marks (reset and mark again before the next sum). */
int svimx_nr_plan_run_length(void* plan);
int svimx_nr_plan_set_run_length(void* plan, int rl);
/* y-lines per unit of a run-mode plan (1, 2 or 4): the lines of a unit share
 * every occupancy window they read, with each line's sums in the one-line
 * order (bit-identical).  Needs m + lines - 1 <= 16 and, for 4 lines, runs of
 * at most 8 nodes.  No counterpart in the reference. */
int svimx_nr_plan_set_lines(void* plan, int lines);
int svimx_nr_plan_lines(void* plan);
/* Run mode: 1 stages the occupancy through LDS (clusters of runs per wave,
 * one dx plane at a time, so the planes are summed in their plain order:
 * equal to the L1 shapes up to FP32 rounding), 0 reads each run's windows
 * through the L1.  No counterpart in the reference. */
int svimx_nr_plan_set_staging(void* plan, int lds);
int svimx_nr_plan_staging(void* plan);
int svimx_nr_exec(void* stream, void* plan, const float* occupancy, float* nr_out, void* ws,
                  int64_t ws_bytes);
/* The two halves of svimx_nr_exec: compact the marked tiles into the plan's
 * tile list, then run the summation kernel over that list. */
int svimx_nr_select(void* stream, void* plan);
int svimx_nr_sum(void* stream, void* plan, const float* occupancy, float* nr_out, void* ws, int64_t ws_bytes);

/* ---------------------------------------------------------------- reinitialisation
 * Min (2010) PDE reinitialisation on the band dilated by `dilation` cells minus
 * `margin` boundary layers: `iterations` TVD-RK2 pseudo-time iterations with
 * ENO2 + subcell fix, in place on `phi`, then the band rebuild of
 * svimx_band_rebuild restricted to the work set (the only nodes that changed):
 * clamp, copy into `scratch` and `phi_copy` (the other two stage fields),
 * reset `occupancy`, and write the new sorted band over `band` (capacity
 * nx*ny*nz; on entry it holds the nb nodes of the current band).
 * `inside_outside` is the count of inside nodes outside the current band.
 * work_idx/work_info hold work_cap nodes, geo 8*work_cap floats; mask_a/mask_b
 * are dense byte fields.  margin >= 2 (the sweeps read 2 nodes along each axis).
 * Synchronises once, at the end.  A work set above work_cap returns
 * SVIMX_ECAPACITY with info[9] = its size and leaves phi and the band as they
 * were.  info (host, 10 int64): new band count, inside nodes outside the new
 * band, margin ok, band min i/j/k, max i/j/k, work-set size.  Replaces
 * LevelSet.reinitialize (svierosion/levelset.py:154-183). */
int svimx_reinit(void* stream, float* phi, float* scratch, float* phi_copy, float* occupancy, int32_t* band,
                 int64_t nb, int64_t inside_outside, int nx, int ny, int nz, double h, double width,
                 int iterations, double cfl, int dilation, int margin, uint8_t* mask_a, uint8_t* mask_b,
                 int32_t* work_idx, int32_t* work_info, float* geo, int64_t work_cap, void* ws, int64_t ws_bytes,
                 int64_t* info);
/* svimx_reinit that runs only when the step whose results are in the
 * workspace (stage-0 statistics, the SVIMX_R_DT of svimx_stable_dt, the
 * volume of svimx_fractions) triggers it, decided on the device by the rule of
 * ErosionSimulation.run (svierosion/evolve.py:232-245) in the host's FP64
 * operation order: dt finite and positive, volume >= target (no secant retry),
 * volume > target_stop (not reached) and displacement + dt * max V >=
 * threshold.  So a caller can queue it before reading the step's results.
 * info holds 11 int64: those of svimx_reinit and info[10] = 1 when it ran; when
 * it did not, nothing changed (info[0] = nb, info[1] = inside_outside).  No
 * counterpart in the reference. */
typedef struct svimx_reinit_gate {
  double displacement;   /* front displacement accumulated before the step */
  double threshold;      /* reinit_distance * h */
  double inside_outside; /* inside nodes outside the band (volume = (this + sum) * h3) */
  double h3;             /* h ** 3 as the host computes it */
  double target;         /* target volume, when has_target */
  double target_stop;    /* target * (1 + tol) */
  int32_t has_target;
  int32_t reserved;
} svimx_reinit_gate_t;
int svimx_reinit_gated(void* stream, float* phi, float* scratch, float* phi_copy, float* occupancy, int32_t* band,
                       int64_t nb, int64_t inside_outside, int nx, int ny, int nz, double h, double width,
                       int iterations, double cfl, int dilation, int margin, uint8_t* mask_a, uint8_t* mask_b,
                       int32_t* work_idx, int32_t* work_info, float* geo, int64_t work_cap, void* ws,
                       int64_t ws_bytes, const svimx_reinit_gate_t* gate, int64_t* info);

/* ---------------------------------------------------------------- mesh -> signed distance
 * Dense phi = +-min(d, band)*h, negative inside, for a triangle mesh already in
 * index coordinates.  Inputs (device):
 *   tri_local  n_split x 9 float32: split triangles (edges <= 8 cells), vertices
 *              relative to tri_org (an integer node per triangle)
 *   tri_box    n_split x 6 int32: clipped node box (lo xyz, hi xyz) of the
 *              triangle dilated by exact_band cells
 *   tri_vid    n_split x 3 int32 vertex ids; star_first/star_deg (n_vertices)
 *              and star_tris (3 n_split) the vertex -> triangle CSR
 *   faces_idx  n_faces x 9 float64: original faces for the sign
 * sign_axes_mask: bit a casts lines along axis a (majority vote over the set);
 * signed_rule: winding number (1) or parity (0).  dist_out / tid_out (may be
 * NULL) receive the unclamped distance (cells) and closest split triangle.
 * Allocates its temporaries and synchronises.
 * Replaces mesh_to_sdf (svierosion/sdf.py:365-415) and its stages
 * _exact_near_field :100-145, _propagate :163-202, _refine_with_stars :211-265,
 * inside_mask :343-361. */
int svimx_sdf_build(void* stream, const float* tri_local, const int32_t* tri_org, const int32_t* tri_box,
                    int64_t n_split, const int32_t* tri_vid, const int32_t* star_first, const int32_t* star_deg,
                    const int32_t* star_tris, int64_t n_vertices, const double* faces_idx, int64_t n_faces,
                    int sign_axes_mask, int signed_rule, int nx, int ny, int nz, double h, double band,
                    double exact_band, float* phi_out, float* dist_out, int32_t* tid_out, void* ws,
                    int64_t ws_bytes);

/* ---------------------------------------------------------------- point-mesh distance
 * Unsigned distance (cells) from npts points (float32 index coordinates, 3
 * per point) to a triangle mesh given, as for svimx_sdf_build, by split
 * triangles in local frames (tri_local ntri x 9 float32 relative to tri_org
 * ntri x 3 int32), binned into buckets of bucket_cells cells over the
 * nx*ny*nz grid; the ring search stops on a lower bound, so the result is the
 * exact minimum.  Allocates stream-ordered temporaries; synchronises.
 * Replaces point_mesh_distance (svierosion/api.py:104-168) for the depth
 * annotation (:171-194). */
int svimx_point_mesh_distance(void* stream, const float* tri_local, const int32_t* tri_org, int64_t ntri, int nx,
                              int ny, int nz, double bucket_cells, const float* points, int64_t npts,
                              float* dist_out);

/* ---------------------------------------------------------------- surface
 * Zero set of phi (band cells) as a watertight, outward triangle mesh by
 * marching tetrahedra on the Kuhn split, vertices shared per edge and moved
 * by one capped Newton step (max_shift cells, 0 = none) onto the zero set of
 * the tricubic interpolant.  Outputs are device arrays the library allocates:
 * *xyz (3 float32 index coordinates per vertex) and *faces (3 int32 per
 * triangle); release them with svimx_free.  Synchronises.  Replaces
 * extract_surface (svierosion/surface.py:16-56). */
int svimx_extract_surface(void* stream, const float* phi, const int32_t* band, int64_t nb, int nx, int ny, int nz,
                          double max_shift, float** xyz, int64_t* nv, int32_t** faces, int64_t* nf);
/* Release an array of svimx_extract_surface, ordered on `stream`. */
int svimx_free(void* stream, void* device_ptr);
/* Copy `bytes` of device memory to host memory, ordered on `stream` (synchronises it). */
int svimx_copy_to_host(void* stream, void* dst, const void* device_src, int64_t bytes);
/* Timing events (no counterpart in the reference: measurement only).  Created
 * without the system-scope fence a default event performs when it is recorded
 * (hipEventDisableSystemFence), so bracketing a kernel with them does not
 * flush the caches between it and its neighbours.  _elapsed waits for `end`
 * and returns the milliseconds between the two records. */
int svimx_timing_event_create(void** event);
int svimx_timing_event_record(void* event, void* stream);
int svimx_timing_event_elapsed(void* start, void* end, float* ms);
int svimx_timing_event_destroy(void* event);

/* Per-point output fields of an eroded surface: p holds n index-coordinate
 * points (clamped to [1, dim - 3.001]).  SVI laws: q1 = N_r by tricubic
 * interpolation of `field`, speed = V(N).  Curvature laws: (k1, k2) by tricubic
 * interpolation of the dense pairs `k12` (2 floats per node, zero off the band),
 * q1 = H = (k1 + k2)/2, q2 = K = k1 k2 (may be NULL), speed = V(k1, k2).
 * Replaces the N / H / K / speed annotation of svierosion/api.py:171-194. */
int svimx_annotate_points(void* stream, const svimx_law_t* law, const float* field, const float* k12, int nx,
                          int ny, int nz, const float* p, int64_t n, float* q1, float* q2, float* speed);

/* ---------------------------------------------------------------- parameters (host)
 * Host-only: no device call, usable without a GPU. */
/* Smallest 2^a 3^b 5^c 7^d >= n.  Replaces next_fast_size (svierosion/grid.py:18-28). */
int svimx_smooth_size(int64_t n, int64_t* out);
/* Grid holding the box [lo, hi] centred with >= padding on every side:
 * shape = ceil((hi - lo + 2 padding) / spacing) + 1 per axis (rounded up to a
 * 7-smooth size when fast), origin = centre - (shape - 1) spacing / 2, in the
 * reference's FP64 operation order.  Replaces Grid.around (svierosion/grid.py:45-56). */
int svimx_grid_around(const double* lo, const double* hi, double spacing, double padding, int fast,
                      int64_t* shape, double* origin);
/* Validated speed-law record (kind 0 SVI, 1 mean curvature, 2 Gauss curvature);
 * for SVI laws with s = 1, n_cap is the largest double with
 * V(n_cap) < max_speed_ratio * V(1/2).  SVIMX_EBADARG with a message for invalid
 * parameters.  Replaces the law constructors of svierosion/flows.py:37-80,
 * :127-137. */
int svimx_law_init(int kind, double radius, double k, double s, double t, double eps, double max_speed_ratio,
                   double alpha, double delta, svimx_law_t* out);
/* *out = n^k / (t - n)^s + eps in FP64 (SVI laws).  Replaces SVIPowerLaw._raw
 * (svierosion/flows.py:82-87). */
int svimx_law_raw(const svimx_law_t* law, double n, double* out);

/* ---------------------------------------------------------------- mesh measures (host)
 * vertices: nv x 3 FP64, faces: nf x 3 int64 vertex indices (row-major).
 * Host-only: no device call, usable without a GPU. */
/* *volume = sum over faces of (p0 - c) . ((p1 - c) x (p2 - c)) / 6 with c the
 * vertex centroid: positive for an outward-oriented closed mesh.  Replaces
 * TriMesh volume (svierosion/meshio.py:80-85). */
int svimx_mesh_signed_volume(const double* vertices, int64_t nv, const int64_t* faces, int64_t nf, double* volume);
/* out[4] = {edges, boundary edges (1 face), non-manifold edges (> 2 faces),
 * misoriented edges (2 faces, same direction)}.  Replaces the edge counts of
 * svierosion/meshio.py:88-105. */
int svimx_mesh_edge_topology(const int64_t* faces, int64_t nf, int64_t nv, int64_t* out);

/* Host arrays of svimx_sdf_build for a mesh whose triangles all have edges
 * <= max_edge cells (the reference splits longer ones first,
 * svierosion/sdf.py:54-70): per face t the index coordinates tri_idx[9t..]
 * ((v - origin) / h, FP64), org = floor of corner 0, the corners relative to
 * org in FP32 (local), the node box [ceil(min - exact_band), floor(max +
 * exact_band)] clamped to the grid (box[6t..]: lo xyz, hi xyz), and the
 * vertex-star CSR (vid: an id per corner shared by bitwise-equal vertices;
 * star: the triangle of every corner grouped by id in corner order; deg,
 * first per id; *nvid ids; deg/first need nv entries).  If a longer edge
 * exists, only *nsplit (the number of such triangles) is written. */
int svimx_sdf_prepare(const double* vertices, int64_t nv, const int64_t* faces, int64_t nf, const double* origin,
                      double h, const int32_t* shape, double exact_band, double max_edge, double* tri_idx,
                      int32_t* org, float* local, int32_t* box, int32_t* vid, int32_t* star, int32_t* deg,
                      int32_t* first, int64_t* nvid, int64_t* nsplit);

/* ---------------------------------------------------------------- operators
 * Node-list forms of the reference's @kernel functions (svierosion/kernels.py),
 * sharing the device code of the fused stage kernels; used by LevelSet methods
 * and parity tests. */
int svimx_op_godunov(void* stream, const float* phi, const int32_t* idx, int64_t n, int nx, int ny,
                     int nz, double inv_h, float* out);                       /* :146-169 */
int svimx_op_cell_fractions(void* stream, const float* phi, const int32_t* idx, int64_t n, int nx,
                            int ny, int nz, float* out);                      /* :173-181 */
int svimx_op_closest_points(void* stream, const float* phi, const int32_t* idx, int64_t n, int nx,
                            int ny, int nz, double h, float* out_p);          /* :185-208 */
int svimx_op_tricubic(void* stream, const float* field, int nx, int ny, int nz, const float* p,
                      int64_t n, float* out);                                 /* :212-228 */
int svimx_op_tricubic_gradient(void* stream, const float* field, int nx, int ny, int nz, const float* p,
                               int64_t n, float* out, float* grad);                   /* :232-253 */
int svimx_op_trilinear(void* stream, const float* field, int nx, int ny, int nz, const float* p,
                       int64_t n, float* out);                                /* :257-270 */
int svimx_op_principal_curvatures(void* stream, const float* phi, const int32_t* idx, int64_t n,
                                  int nx, int ny, int nz, double h, float* k1,
                                  float* k2);                                 /* :274-299 */
int svimx_op_interface_mask(void* stream, const float* phi, const int32_t* idx, int64_t n, int nx,
                            int ny, int nz, uint8_t* out);                    /* :303-309 */
int svimx_op_reinit_geometry(void* stream, const float* phi, const int32_t* idx, int64_t n, int nx,
                             int ny, int nz, double h, double cfl, float* geo); /* :335-359 */
int svimx_op_point_triangle_distance(void* stream, const float* p, const float* tri, int64_t n,
                                     float* out);                             /* :382-427 */
/* Speed law at n values, with the velocity pass's device functions: SVI laws
 * read N from q1 and write V and dV/dN; curvature laws read k1, k2 from q1, q2
 * and write V and D = dV/dk1 + dV/dk2.  D may be NULL.  Replaces
 * SVIPowerLaw.speed / dspeed and speed_and_diffusivity (svierosion/flows.py:89-108,
 * :165-205). */
int svimx_op_law(void* stream, const svimx_law_t* law, const float* q1, const float* q2, int64_t n, float* V,
                 float* D);

/* ---------------------------------------------------------------- batch engine
 * Many erosion jobs of different grids and laws stepped together: every
 * engine tick advances each running job by one accepted TVD-RK3 step, with
 * one launch per kernel and RK stage for all jobs (a device job table maps
 * each virtual block to a job and its own block index), dt computed on the
 * device from each job's statistics, and one host synchronisation per tick
 * (plus the rare secant retries and reinitialisations, which run per job
 * set).  A job's arithmetic, block decomposition and reduction order are the
 * same as in the single-job entry points above, so its history and field do
 * not depend on which other jobs share the ticks.  The tick mirrors
 * ErosionSimulation.run (svierosion/evolve.py:182-250): stop tests, stable dt
 * (:162-170), RK3 (:172-179), regula falsi onto the target volume (:215-231),
 * reinitialisation every reinit_distance*h of displacement (:237-242).
 * The caller keeps a job's buffers alive until svimx_engine_remove; the
 * engine owns the band-sized velocity / L0 arrays and the reinit scratch. */
typedef struct svimx_job_desc {
  float* phi[3];         /* the three dense field buffers; phi[cur] holds phi */
  float* occupancy;      /* padded occupancy (svimx_occupancy_elems) */
  int32_t* band;         /* band list, capacity band_cap */
  int64_t band_cap;
  void* ws;              /* workspace (svimx_workspace_bytes) */
  int64_t ws_bytes;
  int32_t nx, ny, nz;
  int32_t cur;           /* buffer index of phi */
  int64_t nb;            /* band size */
  int64_t inside_outside;
  double h, h3;          /* spacing and its cube (volume factor) */
  double width;          /* band half-width in cells */
  int32_t margin;        /* band margin in cells */
  int32_t frac_valid;    /* buffer whose fractions / N_r marks are in occupancy, -1 none */
  svimx_law_t law;
  void* nr_plan;         /* SVI laws: the N_r plan */
  float* field;          /* SVI laws: dense N_r */
  float* k12;            /* curvature laws: dense (k1, k2) pairs, 2 floats per node */
  void* reserved_ptr;
  double radius;         /* SVI ball radius (feedback CFL) */
  double cfl, feedback_cfl, parabolic_cfl, reinit_distance;
  double cutoff_lo, cutoff_hi;  /* velocity cutoff in cells */
  int32_t reinit_iterations;
  int32_t reserved;
  double time, initial_volume, displacement;
  int64_t steps;
} svimx_job_desc_t;

typedef struct svimx_step_record {  /* StepRecord, svierosion/evolve.py:42-56 */
  int64_t step;
  double time, dt, volume, fraction, vmax, vmean, qmin, qmax;
  int64_t capped;
  int32_t reinit;
  int32_t reserved;
  double wall;
} svimx_step_record_t;

enum {
  SVIMX_STOP_NONE = 0,
  SVIMX_STOP_TARGET = 1,    /* "target_volume" */
  SVIMX_STOP_MAX_TIME = 2,  /* "max_time" */
  SVIMX_STOP_VANISHED = 3,  /* "vanished" */
  SVIMX_STOP_STALLED = 4,   /* "stalled" */
  SVIMX_STOP_MAX_STEPS = 5, /* "max_steps" */
  SVIMX_STOP_ERROR = 6      /* a kernel or reinitialisation failed: see error / message */
};

typedef struct svimx_job_status {
  int32_t running;       /* 1 while the job takes part in ticks */
  int32_t stop_reason;   /* SVIMX_STOP_* of the last run */
  int32_t error;         /* SVIMX_E* code when stop_reason is SVIMX_STOP_ERROR */
  int32_t cur, frac_valid;
  int32_t reserved;
  int64_t nb, inside_outside, steps, nrecords;
  double time, volume, displacement;
  char message[256];
} svimx_job_status_t;

/* stream NULL: the engine creates its own (needed for graph capture).  The
 * caller synchronises the streams that prepared a job before adding it;
 * every tick and run ends with the engine's stream synchronised. */
int svimx_engine_create(void* stream, int max_jobs, void** engine);
int svimx_engine_destroy(void* engine);
/* Register a prepared job (level set built, fractions state as in desc). */
int svimx_engine_add(void* engine, const svimx_job_desc_t* desc, int* job_id);
/* Start (or continue) job_id until its volume reaches target_fraction of
 * initial_volume (<= 0: no target), time reaches max_time (<= 0: none) or
 * max_steps more steps were taken.  Computes the current volume; synchronises. */
int svimx_engine_run(void* engine, int job_id, double target_fraction, double max_time, int64_t max_steps,
                     double tol);
/* One step of each running job that the schedule selects for this tick
 * (every running job under SVIMX_SCHED_LOCKSTEP); *n_running = jobs still
 * running after it, selected or not. */
int svimx_engine_tick(void* engine, int* n_running);
/* Which running jobs a tick steps.  The reference runs one erosion at a time
 * (svierosion/compare.py:36-85); this has no counterpart there.
 *   SVIMX_SCHED_LOCKSTEP: all of them (every job advances one step per tick).
 *   SVIMX_SCHED_FIFO:     jobs in the order they were added ...
 *   SVIMX_SCHED_LONGEST:  ... or by most estimated steps left first (from each
 *                         job's volume loss per step since its run started;
 *                         jobs with fewer than 4 steps come first),
 *   SVIMX_SCHED_SHORT:    in the order they were added, where behind the first
 *                         (oldest) job every job with an estimated work left
 *                         (steps left x band nodes) of at most short_work
 *                         node-steps, or fewer than 4 steps into its run,
 *                         joins, whatever its band size, and no other: long
 *                         jobs run one after another as on a one-at-a-time
 *                         queue while short ones pass them.
 * Except under SHORT, jobs are taken while the selected band nodes stay within
 * max(budget_nodes, (1 + slack) * band nodes of the first job): the first job
 * is always stepped, a job that does not fit is passed over for this tick.  A
 * job's results do not depend on the schedule. */
enum { SVIMX_SCHED_LOCKSTEP = 0, SVIMX_SCHED_FIFO = 1, SVIMX_SCHED_LONGEST = 2, SVIMX_SCHED_SHORT = 3 };
/* Up to max_ticks ticks in one call, ending early after the tick in which a
 * job stopped or when none is running; *n_ticks = ticks taken.  For request
 * loops that look at arrivals and stopped jobs every few ticks (no
 * counterpart in the reference). */
int svimx_engine_ticks(void* engine, int max_ticks, int* n_running, int* n_ticks);
int svimx_engine_schedule(void* engine, int policy, int64_t budget_nodes, double slack);
/* The work bound of SVIMX_SCHED_SHORT, in node-steps (default 2e10). */
int svimx_engine_short_work(void* engine, double node_steps);
int svimx_engine_status(void* engine, int job_id, svimx_job_status_t* out);
/* Copy step records [first, first + n) of job_id (records accumulate over runs). */
int svimx_engine_records(void* engine, int job_id, int64_t first, int64_t n, svimx_step_record_t* out);
/* Forget job_id (its buffers are the caller's again); frees the engine-owned arrays. */
int svimx_engine_remove(void* engine, int job_id);
/* Capture the main tick as a HIP graph and replay it (1) or launch eagerly (0).
 * Graphs are kept for the last few job sets a tick stepped. */
int svimx_engine_use_graphs(void* engine, int on);
/* Launch groups: a tick runs its jobs in groups whose band nodes add up to at
 * most `nodes` (a larger job alone; 0: one group), each group through every
 * stage before the next, so a large job's fields stay in the device's last-level
 * cache from one kernel to the next.  Results do not depend on it.  Default 2M
 * nodes.  No counterpart in the reference. */
int svimx_engine_group_nodes(void* engine, int64_t nodes);
/* Ids (up to cap) of jobs that stopped since the last call, oldest first. */
int svimx_engine_stopped(void* engine, int* ids, int cap, int* n);
/* Ticks whose graph capture was invalidated (by a device-wide operation in
 * another thread, e.g. a synchronous free) and that ran their launches eagerly
 * instead; results are unaffected. */
int64_t svimx_engine_capture_fallbacks(void* engine);
/* Milliseconds spent in ticks / secant retries / reinitialisations so far. */
int svimx_engine_timers(void* engine, double* out3);

#ifdef __cplusplus
}
#endif

#endif /* SVIMX_H_ */
