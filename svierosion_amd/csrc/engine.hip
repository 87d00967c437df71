// Batch engine: many erosion jobs of different grids and laws stepped in
// ticks (one accepted TVD-RK3 step for each job a tick selects).
//
// Device side.  Each job's pointers and scalars live in a JobDev record of a
// device table.  A launch covers a set of jobs through a segment table Seg:
// job s owns virtual blocks [start[s], start[s+1]), so virtual block vb maps
// to (job, local block lb, job grid G).  The kernels run a fixed persistent
// grid whose blocks walk the virtual blocks (vb = blockIdx.x, +gridDim.x ...)
// and call the same device bodies as the single-job kernels with the job's
// own (lb, G): a job's block decomposition, partial sums and finalisation are
// those of the single-job entry points, so its results are bit-identical
// alone, in any mix, and through the single-job API.  Per-job reductions end
// in one finalisation block per job, which also derives the job's stable dt
// on the device (svierosion/evolve.py:162-170), so a tick needs one host
// synchronisation: the gathered result slots of all jobs.
//
// Schedule.  Which running jobs a tick steps: all of them (lockstep), or a
// priority order (the order jobs were added, or most estimated steps left
// first) filled up to a band-node budget.  Lockstep holds every job to the
// tick rate of the whole set: a 4,000-step job sharing ticks with large grids
// advanced at ~0.7 ms per step instead of the ~0.16 ms of its own; the
// budget keeps ticks near the size that fills the GPU and lets the first
// job's steps pass at close to their own rate.  Graphs are captured per job
// set and the last few sets' graphs are kept, as the selection alternates.
//
// Host side.  The tick mirrors one iteration of ErosionSimulation.run
// (svierosion/evolve.py:182-250) for every selected job: stop tests, stage 0
// with statistics and device dt, stages 1-2, volume (each stage's velocity
// fused with the RK update that consumes it, as on the single-job path); then
// the regula falsi
// retries for the jobs that passed their target (in lockstep over that set),
// then reinitialisation per job where the displacement reached
// reinit_distance*h, followed by one batched volume pass.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "band_dev.h"
#include "compact.h"
#include "nr_dev.h"
#include "nr_plan.h"

namespace svimx {

constexpr int kMaxEngineJobs = 256;
constexpr int kEngGrid = 2048;  // persistent grid (multiple of the XCD count)
constexpr int kEngGridWaves = 2560;  // one-wave blocks of the LDS N_r family: 10 per CU
// A tick runs its jobs in groups whose band nodes add up to at most
// group_nodes (a job above it runs alone), each group through every stage
// before the next: the stages of a large job then reuse its fields in the
// 256 MB Infinity Cache from one kernel to the next, as on the single-job
// path, instead of evicting them with the other jobs' fields.  Small jobs
// still share launches.
constexpr int kMaxGroups = 8;
constexpr int64_t kDefaultGroupNodes = 2000000;

struct JobDev {
  float* buf[3];
  float* occ;
  const int32_t* band;
  int64_t nb;
  Dims d;
  int cur;
  int need_frac;  // the prepare of buffer `cur` must compute fractions
  float h, inv_h, frac_cut, cut_lo, cut_hi;
  Hi3 hi;
  LawF law;
  const float* field;
  float* k12;  // curvature laws: dense (k1, k2) pairs
  float* V;
  float* L0;
  double* results;
  double* partials;
  // N_r plan (run mode)
  int has_plan, run_mode, nzr, nty, ntz, ncols;
  uint8_t* marks;
  int64_t nmarks;
  uint8_t* dil;
  int32_t* list;
  int64_t* ccounts;
  int64_t nruns, nchunks;
  const NrColumn* cols;
  const int* col_start;
  int asplit, cfirst, csplit;  // the plan's N_r summation order
  const float* weights;
  const float* colrec;
  float scale, negbv;
  int m, lds;
  int ylines, nyu;  // run mode: y-lines per unit, units per x-plane column
  int64_t nunits;
  // dt control (FP64)
  double hd, cfl, feedback, parabolic, radius, time, max_time;
  int requires_svi, requires_curv, has_max_time, pad;
};

struct Seg {
  int32_t njobs, total;
  int32_t start[kMaxEngineJobs + 1];
  int32_t slot[kMaxEngineJobs];
};

// N_r tables: T_NR + 12 * variant + 3 * (MP / 4 - 1) + (0, 1, 2 for run lengths
// 4, 8, 16); variants 0: one line, 1: LDS staging, 2: two lines, 3: four lines
enum SegId { T_ALL, T_SVI, T_MARKS, T_RUNS, T_CHUNKS, T_JOBS, T_RUNJOBS, T_CURV, T_NR, T_COUNT = T_NR + 48 };
constexpr int kNrLdsVariant = 1;

struct Tables {
  Seg seg[T_COUNT];
};

// Virtual block -> (job slot, local block, job grid); segments are visited in
// increasing order by each persistent block.
struct SegCursor {
  const Seg* s;
  int seg;
  __device__ explicit SegCursor(const Seg* s_) : s(s_), seg(-1) {}
  __device__ bool at(int64_t vb, int& job, int64_t& lb, int64_t& G) {
    if (vb >= s->total) return false;
    if (seg < 0) {
      int lo = 0, hi = s->njobs - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (s->start[mid] <= vb) lo = mid; else hi = mid - 1;
      }
      seg = lo;
    }
    while (s->start[seg + 1] <= vb) ++seg;
    job = s->slot[seg];
    lb = vb - s->start[seg];
    G = s->start[seg + 1] - s->start[seg];
    return true;
  }
};

#define ENG_FOR(S)             \
  SegCursor _cur(S);           \
  int jslot;                   \
  int64_t lb, G;               \
  for (int64_t vb = blockIdx.x; _cur.at(vb, jslot, lb, G); vb += gridDim.x)

__device__ __forceinline__ float* jbuf(const JobDev& j, int rel) { return j.buf[(j.cur + rel) % 3]; }

__global__ void __launch_bounds__(256) k_eng_clear_marks(const JobDev* __restrict__ J, const Seg* S, int only_needed) {
  ENG_FOR(S) {
    const JobDev& j = J[jslot];
    if (only_needed && !j.need_frac) continue;
    // 16 bytes per thread (the plan's marks are hipMalloc'd), byte tail
    const int64_t n16 = j.nmarks >> 4;
    for (int64_t i = lb * 256 + threadIdx.x; i < n16; i += G * 256)
      reinterpret_cast<uint4*>(j.marks)[i] = make_uint4(0u, 0u, 0u, 0u);
    for (int64_t i = (n16 << 4) + lb * 256 + threadIdx.x; i < j.nmarks; i += G * 256) j.marks[i] = 0;
  }
}

__global__ void __launch_bounds__(kBandThreads) k_eng_fractions(const JobDev* __restrict__ J, const Seg* S, int rel,
                                                                int only_needed) {
  ENG_FOR(S) {
    const JobDev& j = J[jslot];
    if (only_needed && !j.need_frac) continue;
    fractions_body(jbuf(j, rel), j.band, j.nb, j.d, j.h, j.occ, j.has_plan ? j.marks : nullptr, j.run_mode, j.nty,
                   j.ntz, j.nzr, j.frac_cut, j.hi, j.partials, lb, G);
  }
}

// volume partials -> result slots 0..2 (svimx_fractions' finalisation)
__global__ void __launch_bounds__(kFinalizeThreads) k_eng_finalize_volume(const JobDev* __restrict__ J, const Seg* S) {
  ENG_FOR(S) {
    const JobDev& j = J[jslot];
    ReduceSpec spec;
    for (int s = 0; s < kPartialSlots; ++s) spec.op[s] = -1, spec.dst[s] = 0;
    spec.op[0] = 0; spec.dst[0] = SVIMX_R_VOLUME_SUM;
    spec.op[1] = 1; spec.dst[1] = SVIMX_R_HAS_NEG;
    spec.op[2] = 1; spec.dst[2] = SVIMX_R_HAS_POS;
    finalize_body(j.partials, reduce_grid(j.nb), j.results, spec);
  }
}

__global__ void __launch_bounds__(256) k_eng_unit_need(const JobDev* __restrict__ J, const Seg* S) {
  ENG_FOR(S) {
    const JobDev& j = J[jslot];
    unit_need_range(j.marks, j.dil, j.d.ny, j.nzr, j.nyu, j.ylines, j.nunits, lb * 256 + threadIdx.x, G * 256);
  }
}

__global__ void __launch_bounds__(kCompactThreads) k_eng_need_count(const JobDev* __restrict__ J, const Seg* S) {
  ENG_FOR(S) {
    const JobDev& j = J[jslot];
    compact_count_body(FlagPred{j.dil}, j.nunits, j.ccounts, lb);
  }
}

__global__ void __launch_bounds__(1024) k_eng_need_scan(const JobDev* __restrict__ J, const Seg* S) {
  ENG_FOR(S) {
    const JobDev& j = J[jslot];
    scan_counts_body(j.ccounts, j.nchunks);
  }
}

__global__ void __launch_bounds__(kCompactThreads) k_eng_need_scatter(const JobDev* __restrict__ J, const Seg* S) {
  ENG_FOR(S) {
    const JobDev& j = J[jslot];
    compact_scatter_body(FlagPred{j.dil}, j.nunits, j.ccounts, j.list, j.nunits,
                         IdentityMap(), lb);
  }
}

template <int MP, int RL>
__global__ void __launch_bounds__(256) k_eng_nr_runs(const JobDev* __restrict__ J, const Seg* S) {
  ENG_FOR(S) {
    const JobDev& j = J[jslot];
    const int64_t nrun = j.ccounts[j.nchunks];
    if (lb == 0 && threadIdx.x == 0) j.results[SVIMX_R_TILES] = (double)nrun;
    nr_runs_body<MP, RL>(j.occ, OccLayout::of(j.d.nx, j.d.ny, j.d.nz), j.nzr, j.list, nrun, j.cols, j.ncols, j.cfirst,
                         j.csplit, j.weights, j.scale, j.negbv, const_cast<float*>(j.field), lb, G);
  }
}

// Units of Y adjacent y-lines
template <int MP, int RL, int Y>
__global__ void __launch_bounds__(256) k_eng_nr_lines(const JobDev* __restrict__ J, const Seg* S) {
  __shared__ float xch[nr_lines_xch<RL, Y, kNrLineLanes>()];
  ENG_FOR(S) {
    const JobDev& j = J[jslot];
    const int64_t nunit = j.ccounts[j.nchunks];
    if (lb == 0 && threadIdx.x == 0) j.results[SVIMX_R_TILES] = (double)nunit;
    nr_lines_body<MP, RL, Y, kNrLineLanes, false>(j.occ, OccLayout::of(j.d.nx, j.d.ny, j.d.nz), j.nzr, j.list, nunit, j.cols,
                                           j.col_start, j.m, j.asplit, j.colrec, j.scale, j.negbv,
                                           const_cast<float*>(j.field), xch, lb, G);
  }
}

// LDS-staged variant: one wave per block (the job grid counts waves)
template <int MP, int RL>
__global__ void __launch_bounds__(64) k_eng_nr_lds(const JobDev* __restrict__ J, const Seg* S) {
  __shared__ float4 slab[kNrLdsCap4];
  __shared__ NrLdsClu clu[kNrLdsClusters];
  ENG_FOR(S) {
    const JobDev& j = J[jslot];
    const int64_t nrun = j.ccounts[j.nchunks];
    if (lb == 0 && threadIdx.x == 0) j.results[SVIMX_R_TILES] = (double)nrun;
    nr_lds_body<MP, RL>(j.occ, OccLayout::of(j.d.nx, j.d.ny, j.d.nz), j.nzr, j.list, nrun, j.cols, j.col_start, j.m,
                        j.weights, j.scale, j.negbv, const_cast<float*>(j.field), slab, clu, lb, G);
  }
}

__global__ void __launch_bounds__(kBandThreads) k_eng_curvatures(const JobDev* __restrict__ J, const Seg* S, int rel) {
  ENG_FOR(S) {
    const JobDev& j = J[jslot];
    curvatures_body(jbuf(j, rel), j.band, j.nb, j.d, j.h, j.k12, lb, G);
  }
}

// The velocity fused with the RK stage that consumes it (svimx_velocity_rk's
// body): MODE 1 stores L0 and the statistics of stage 0, MODE 2 and 3 write
// stages 1 and 2 into buffer rel_out with the job's device dt.
template <bool SVI, int MODE>
__global__ void __launch_bounds__(kBandThreads) k_eng_velocity_rk(const JobDev* __restrict__ J, const Seg* S,
                                                                  int rel_in, int rel_u0, int rel_out, int stats,
                                                                  const double* __restrict__ dts) {
  ENG_FOR(S) {
    const JobDev& j = J[jslot];
    RkFuse rk;
    rk.u0 = jbuf(j, rel_u0);
    rk.out = jbuf(j, rel_out);
    rk.L0 = j.L0;
    rk.inv_h = j.inv_h;
    rk.dt = MODE >= 2 ? (float)dts[jslot] : 0.0f;
    // the law by value: velocity_body takes it by reference, and a reference
    // into the job table had its fields reloaded by scalar loads at every node
    // of the inner loop (35% slower than the single-job kernel)
    const LawF law = j.law;
    velocity_body<SVI, MODE>(jbuf(j, rel_in), j.band, j.nb, j.d, j.h, law, j.field, j.k12, j.cut_lo, j.cut_hi,
                             j.hi, stats, j.V, j.partials, lb, G, rk);
  }
}

// Statistics finalisation (svimx_velocity's) and the stable dt of
// svierosion/evolve.py:162-170 with the max_time clamp of :199-200, in FP64
// with the operation order of the host formula.
__global__ void __launch_bounds__(kFinalizeThreads) k_eng_stats_dt(const JobDev* __restrict__ J, const Seg* S,
                                                      double* __restrict__ dts) {
#pragma clang fp contract(off)
  ENG_FOR(S) {
    const JobDev& j = J[jslot];
    ReduceSpec spec;
    const int8_t ops[15] = {1, 2, 1, 1, 0, 0, 0, 1, 2, 1, 1, 0, 0, 0, 1};
    for (int s = 0; s < kPartialSlots; ++s) spec.op[s] = -1, spec.dst[s] = 0;
    for (int s = 0; s < 15; ++s) spec.op[s] = ops[s], spec.dst[s] = (int8_t)(SVIMX_R_VMAX + s);
    finalize_body(j.partials, reduce_grid(j.nb), j.results, spec);
    if (threadIdx.x == 0) {
      const double dt = stable_dt_of(j.results, j.requires_svi, j.requires_curv, j.cfl, j.hd, j.feedback, j.radius,
                                     j.parabolic, j.has_max_time, j.has_max_time ? j.max_time - j.time : 0.0);
      dts[jslot] = dt;
      j.results[SVIMX_R_DT] = dt;
    }
  }
}

__global__ void __launch_bounds__(kBandThreads) k_eng_rk(const JobDev* __restrict__ J, const Seg* S, int stage,
                                                         int rel_in, int rel_u0, int rel_out,
                                                         const double* __restrict__ dts) {
  ENG_FOR(S) {
    const JobDev& j = J[jslot];
    const float dt = (float)dts[jslot];
    const float* in = jbuf(j, rel_in);
    const float* u0 = jbuf(j, rel_u0);
    float* out = jbuf(j, rel_out);
    for (int64_t q = lb * kBandThreads + threadIdx.x; q < j.nb; q += G * kBandThreads)
      rk_node(stage, in, u0, out, j.band, q, j.d, j.inv_h, dt, j.V, j.L0);
  }
}

__global__ void __launch_bounds__(256) k_eng_gather(const JobDev* __restrict__ J, const Seg* S,
                                                    double* __restrict__ out) {
  ENG_FOR(S) {
    const JobDev& j = J[jslot];
    if (threadIdx.x < SVIMX_R_COUNT) out[(int64_t)jslot * SVIMX_R_COUNT + threadIdx.x] = j.results[threadIdx.x];
  }
}

__global__ void k_eng_zero_at_band(float2* __restrict__ k12, const int32_t* __restrict__ band, int64_t nb) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nb) return;
  k12[band[q]] = make_float2(0.0f, 0.0f);
}

// ---------------------------------------------------------------- host side
template <class T>
static hipError_t grow(T*& p, int64_t& cap, int64_t need) {
  if (need <= cap && p) return hipSuccess;
  if (p) {
    hipError_t e = hipFree(p);
    if (e != hipSuccess) return e;
    p = nullptr;
  }
  cap = std::max<int64_t>(need + need / 4, 1024);
  return hipMalloc(&p, sizeof(T) * cap);
}

struct EngJob {
  bool used = false;
  svimx_job_desc_t desc{};
  NrPlan* plan = nullptr;
  float* V = nullptr;
  float* L0 = nullptr;
  int64_t vcap = 0, l0cap = 0;
  uint8_t* ma = nullptr;
  uint8_t* mb = nullptr;
  int64_t macap = 0, mbcap = 0;
  int32_t* widx = nullptr;
  int32_t* winfo = nullptr;
  float* geo = nullptr;
  int64_t wcap = 0, icap = 0, gcap = 0;
  // state
  int cur = 0, frac_valid = -1;
  int64_t nb = 0, inside_outside = 0, steps = 0;
  double time = 0, displacement = 0, V0 = 0, vol = 0;
  double res[SVIMX_R_COUNT] = {};
  int running = 0, stop = SVIMX_STOP_NONE, error = 0;
  std::string message;
  // run goal
  int64_t seq = 0;      // order of svimx_engine_add (the FIFO schedule's key)
  double vol_run0 = 0;  // volume when the current run started
  bool has_target = false, has_max_time = false;
  double target = 0, max_time = 0, tol = 1e-6;
  int64_t max_steps = 0, n_start = 0;
  std::vector<svimx_step_record_t> records;
  // tick scratch
  double dt = 0, vmax = 0, vnew = 0, tick_wall = 0;
  double lo_t = 0, lo_v = 0, hi_t = 0, hi_v = 0, dts = 0, v = 0;
  bool secant_live = false;
  double qmin = NAN, qmax = NAN, vmean = 0;
  int64_t capped = 0;

  bool svi() const { return desc.law.kind == 0; }
  bool run_plan() const { return plan && plan->run_mode && !plan->fft; }
};

struct Engine {
  hipStream_t st = nullptr;
  bool own_stream = false;
  int max_jobs = 0;
  std::vector<EngJob> jobs;
  JobDev* d_jobs = nullptr;
  Tables* d_tab = nullptr;
  double* d_dts = nullptr;
  double* d_res = nullptr;
  JobDev* h_jobs = nullptr;
  Tables* h_tab = nullptr;
  double* h_dts = nullptr;
  double* h_res = nullptr;
  double* res_map = nullptr;  // h_res as the device addresses it (the gather writes there)
  bool graphs = false;
  struct Captured {
    std::vector<int> set;     // job slots the graph was captured for
    std::vector<int> shape;   // their group sizes
    hipGraphExec_t exec;
    uint64_t last_use;
  };
  std::vector<Captured> captured;  // the graphs of the last few job sets
  uint64_t graph_clock = 0;
  int policy = SVIMX_SCHED_LOCKSTEP;
  int64_t budget_nodes = 0;
  double slack = 0.0;
  int64_t next_seq = 0;
  int64_t capture_fallbacks = 0;  // ticks run eagerly because their capture was invalidated
  bool graph_has_tile = false;
  double ms_tick = 0, ms_secant = 0, ms_reinit = 0;
  std::vector<int> set;      // job slots of the current phase
  std::vector<int> stopped;  // jobs that stopped since the last svimx_engine_stopped
  int64_t group_nodes = kDefaultGroupNodes;
  std::vector<int32_t> tab_sig;          // what the device tables hold (upload)
  double short_work = 2e10;              // SVIMX_SCHED_SHORT: riders' estimated node-steps left
  std::vector<std::vector<int>> groups;  // E.set split into launch groups (upload)
  int cur_group = 0;                     // the group whose tables the launches use
};

static double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static void fill_jobdev(const EngJob& e, JobDev& j) {
  const svimx_job_desc_t& D = e.desc;
  std::memset(&j, 0, sizeof(j));
  for (int b = 0; b < 3; ++b) j.buf[b] = D.phi[b];
  j.occ = D.occupancy;
  j.band = D.band;
  j.nb = e.nb;
  j.d = Dims{D.nx, D.ny, D.nz};
  j.cur = e.cur;
  j.need_frac = e.frac_valid != e.cur ? 1 : 0;
  j.h = (float)D.h;
  j.inv_h = (float)(1.0 / D.h);
  j.frac_cut = (float)(D.cutoff_hi * D.h);
  j.cut_lo = (float)(D.cutoff_lo * D.h);
  j.cut_hi = (float)(D.cutoff_hi * D.h);
  j.hi = hi_of(j.d);
  j.law = law_to_f(D.law);
  j.field = D.field;
  j.k12 = D.k12;
  j.V = e.V;
  j.L0 = e.L0;
  Workspace w;
  carve_workspace(D.ws, D.ws_bytes, j.d.size(), w);
  j.results = w.results;
  j.partials = w.partials;
  if (e.plan && !e.plan->fft) {
    const NrPlan* p = e.plan;
    j.has_plan = 1;
    j.run_mode = p->run_mode;
    j.nzr = p->nzr;
    j.nty = p->nty;
    j.ntz = p->ntz;
    j.ncols = p->ncols;
    j.marks = p->marks;
    j.nmarks = p->nmarks;
    j.dil = p->dil;
    j.list = p->list;
    j.ccounts = p->ccounts;
    j.nruns = p->nruns;
    j.ylines = p->ylines;
    j.nyu = p->nyu;
    j.nunits = p->nunits;
    j.nchunks = p->nchunks;
    j.cols = p->cols;
    j.col_start = p->col_start;
    j.asplit = p->asplit;
    j.cfirst = p->cfirst;
    j.csplit = p->csplit;
    j.m = p->m;
    j.lds = p->lds;
    j.weights = p->weights;
    j.colrec = p->colrec;
    j.scale = p->scale;
    j.negbv = p->negbv;
  }
  j.hd = D.h;
  j.cfl = D.cfl;
  j.feedback = D.feedback_cfl;
  j.parabolic = D.parabolic_cfl;
  j.radius = D.radius;
  j.time = e.time;
  j.max_time = e.max_time;
  j.has_max_time = e.has_max_time ? 1 : 0;
  j.requires_svi = D.law.kind == 0 ? 1 : 0;
  j.requires_curv = D.law.kind != 0 ? 1 : 0;
}

static int mp_of(const NrPlan* p) { return std::max(4, (p->m + 3) / 4 * 4); }

// virtual-block counts of a job in each table
static int64_t nr_grid(const EngJob& e) {
  // runs are not known on the host before the selection: size by the band
  const int64_t est = std::max<int64_t>(e.nb / 4 / e.plan->ylines, 256);
  const int threads = e.plan->lds ? 64 : 256;  // the LDS variant runs one wave per block
  int64_t g = (est + threads - 1) / threads;
  g = std::min<int64_t>(std::max<int64_t>(g, kXcds), e.plan->lds ? 2560 : 2048);
  return (g + kXcds - 1) / kXcds * kXcds;
}

static void seg_reset(Seg& s) {
  s.njobs = 0;
  s.total = 0;
  s.start[0] = 0;
}
// Append a job of `blocks` virtual blocks.  N_r tables add multiples of 8
// blocks only, so their segments start on multiples of 8 and a virtual block's
// XCD (vb mod 8 on a grid that is a multiple of 8) is its local block mod 8.
static void seg_add(Seg& s, int slot, int64_t blocks) {
  s.slot[s.njobs] = slot;
  s.njobs++;
  s.total = (int32_t)(s.total + blocks);
  s.start[s.njobs] = s.total;
}

static int top_slot(const Engine& E) {
  int top = 0;
  for (int slot : E.set) top = std::max(top, slot + 1);
  return top;
}

// The job records of E.set to the device (inside the captured graph on the
// tick path, so a replay needs no separate copy).
static int upload_jobs(Engine& E) {
  const int top = top_slot(E);
  if (top > 0)
    SVIMX_CHECK_HIP(hipMemcpyAsync(E.d_jobs, E.h_jobs, sizeof(JobDev) * top, hipMemcpyHostToDevice, E.st));
  return SVIMX_OK;
}

// Tables for the job slots in E.set, one set per launch group; uploads the job
// records (unless the caller's launches do) and the tables.
static int upload(Engine& E, bool jobs = true) {
  E.groups.clear();
  int64_t acc = 0;
  for (int slot : E.set) {
    const int64_t nb = std::max<int64_t>(E.jobs[slot].nb, 1);
    if (E.groups.empty() ||
        (E.group_nodes > 0 && acc + nb > E.group_nodes && (int)E.groups.size() < kMaxGroups)) {
      E.groups.emplace_back();
      acc = 0;
    }
    E.groups.back().push_back(slot);
    acc += nb;
  }
  for (size_t g = 0; g < E.groups.size(); ++g) {
  Tables& T = E.h_tab[g];
  for (int t = 0; t < T_COUNT; ++t) seg_reset(T.seg[t]);
  for (int slot : E.groups[g]) {
    EngJob& e = E.jobs[slot];
    fill_jobdev(e, E.h_jobs[slot]);
    seg_add(T.seg[T_ALL], slot, reduce_grid(e.nb));
    seg_add(T.seg[T_JOBS], slot, 1);
    if (e.svi()) seg_add(T.seg[T_SVI], slot, reduce_grid(e.nb));
    else seg_add(T.seg[T_CURV], slot, reduce_grid(e.nb));
    if (e.plan && !e.plan->fft) seg_add(T.seg[T_MARKS], slot, (e.plan->nmarks + 4095) / 4096);
    if (e.run_plan()) {
      seg_add(T.seg[T_RUNS], slot, (e.plan->nunits + 255) / 256);
      seg_add(T.seg[T_CHUNKS], slot, e.plan->nchunks);
      seg_add(T.seg[T_RUNJOBS], slot, 1);
      const int rl = e.plan->run_mode;
      const int variant = e.plan->lds ? kNrLdsVariant : (e.plan->ylines == 2 ? 2 : e.plan->ylines == 4 ? 3 : 0);
      const int t = T_NR + 12 * variant + 3 * (mp_of(e.plan) / 4 - 1) + (rl == 4 ? 0 : rl == 8 ? 1 : 2);
      seg_add(T.seg[t], slot, nr_grid(e));
    }
  }
  }
  E.cur_group = 0;
  if (jobs) {
    const int rc = upload_jobs(E);
    if (rc) return rc;
  }
  // the tables change only with the set, the groups or a band size: upload
  // them (ngroups x ~115 KB) only when their contents differ from the last
  std::vector<int32_t> sig;
  sig.push_back((int32_t)E.groups.size());
  for (size_t g = 0; g < E.groups.size(); ++g)
    for (int t = 0; t < T_COUNT; ++t) {
      const Seg& sg = E.h_tab[g].seg[t];
      sig.push_back(sg.njobs);
      sig.push_back(sg.total);
      sig.insert(sig.end(), sg.start + 1, sg.start + 1 + sg.njobs);
      sig.insert(sig.end(), sg.slot, sg.slot + sg.njobs);
    }
  if (sig != E.tab_sig) {
    SVIMX_CHECK_HIP(hipMemcpyAsync(E.d_tab, E.h_tab, sizeof(Tables) * std::max<size_t>(E.groups.size(), 1),
                                   hipMemcpyHostToDevice, E.st));
    E.tab_sig.swap(sig);
  }
  return SVIMX_OK;
}

// Run f once per launch group, with that group's tables current.
template <typename F>
static int each_group(Engine& E, F f) {
  int rc = SVIMX_OK;
  for (int g = 0; g < (int)E.groups.size() && !rc; ++g) {
    E.cur_group = g;
    rc = f();
  }
  E.cur_group = 0;
  return rc;
}

static unsigned grid_of(const Engine& E, int t) {
  const Seg& sg = E.h_tab[E.cur_group].seg[t];
  int64_t g;
  if (t == T_JOBS || t == T_RUNJOBS) {
    g = E.graphs ? E.max_jobs : sg.total;  // one block per job
  } else {
    // persistent grid: a captured graph keeps its grid while the band sizes
    // change, so it gets headroom over the totals at capture time
    g = E.graphs ? sg.total + sg.total / 2 + 64 : sg.total;
    g = std::min<int64_t>(g, t >= T_NR + 12 * kNrLdsVariant && t < T_NR + 12 * (kNrLdsVariant + 1) ? kEngGridWaves
                                                                                                     : kEngGrid);
  }
  g = std::max<int64_t>(g, 1);
  return (unsigned)((g + kXcds - 1) / kXcds * kXcds);
}

// Blocks of `kern` (with `threads` per block) that the device holds at once,
// rounded down to a multiple of the XCD count; cached per kernel.  A persistent
// grid above this leaves its last blocks to start after whole rows of virtual
// blocks of the first ones, so every job-table kernel runs with at most this
// many blocks (the virtual blocks, and so the results, do not change): with
// 6-wave occupancy 1536 blocks are resident, and 2048 blocks of four jobs'
// virtual blocks took 35% longer than the single-job kernels on the same jobs.
template <typename K>
static int64_t resident_blocks(K kern, int threads) {
  static std::mutex mu;
  static std::unordered_map<const void*, int64_t> cache;
  const void* key = reinterpret_cast<const void*>(kern);
  std::lock_guard<std::mutex> lock(mu);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  int dev = 0, per_cu = 0, cus = 0;
  int64_t r = kEngGrid;
  if (hipGetDevice(&dev) == hipSuccess &&
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, threads, 0) == hipSuccess && per_cu > 0 && cus > 0)
    r = std::max<int64_t>((int64_t)per_cu * cus / kXcds * kXcds, kXcds);
  cache.emplace(key, r);
  return r;
}

#define ENG_LAUNCH(kern, t, threads, ...)                                                                  \
  do {                                                                                                    \
    if (E.h_tab[E.cur_group].seg[t].njobs > 0 && E.h_tab[E.cur_group].seg[t].total > 0) {                 \
      const unsigned _g = (unsigned)std::min<int64_t>(grid_of(E, t), resident_blocks(kern, threads));     \
      hipLaunchKernelGGL(kern, dim3(_g), dim3(threads), 0, E.st, E.d_jobs, &E.d_tab[E.cur_group].seg[t],   \
                         ##__VA_ARGS__);                                                                  \
      SVIMX_CHECK_LAUNCH();                                                                               \
    }                                                                                                     \
  } while (0)

// fractions + N_r (SVI jobs) or curvatures of buffer cur+rel (the inputs of
// its velocity); curvature jobs compute fractions only for the end-of-step
// volume
static int inputs(Engine& E, int rel, int only_needed) {
  ENG_LAUNCH(k_eng_clear_marks, T_MARKS, 256, only_needed);
  ENG_LAUNCH(k_eng_fractions, T_SVI, kBandThreads, rel, only_needed);
  ENG_LAUNCH(k_eng_unit_need, T_RUNS, 256);
  ENG_LAUNCH(k_eng_need_count, T_CHUNKS, kCompactThreads);
  ENG_LAUNCH(k_eng_need_scan, T_RUNJOBS, 1024);
  ENG_LAUNCH(k_eng_need_scatter, T_CHUNKS, kCompactThreads);
#define ENG_NR(MPV, RLV, I) ENG_LAUNCH((k_eng_nr_runs<MPV, RLV>), T_NR + (I), 256)
  ENG_NR(4, 4, 0); ENG_NR(4, 8, 1); ENG_NR(4, 16, 2);
  ENG_NR(8, 4, 3); ENG_NR(8, 8, 4); ENG_NR(8, 16, 5);
  ENG_NR(12, 4, 6); ENG_NR(12, 8, 7); ENG_NR(12, 16, 8);
  ENG_NR(16, 4, 9); ENG_NR(16, 8, 10); ENG_NR(16, 16, 11);
#undef ENG_NR
#define ENG_NRL(MPV, RLV, I) ENG_LAUNCH((k_eng_nr_lds<MPV, RLV>), T_NR + 12 + (I), 64)
  ENG_NRL(4, 4, 0); ENG_NRL(4, 8, 1); ENG_NRL(4, 16, 2);
  ENG_NRL(8, 4, 3); ENG_NRL(8, 8, 4); ENG_NRL(8, 16, 5);
  ENG_NRL(12, 4, 6); ENG_NRL(12, 8, 7); ENG_NRL(12, 16, 8);
  ENG_NRL(16, 4, 9); ENG_NRL(16, 8, 10); ENG_NRL(16, 16, 11);
#undef ENG_NRL
#define ENG_NRY(MPV, RLV, YV, I) ENG_LAUNCH((k_eng_nr_lines<MPV, RLV, YV>), T_NR + 12 * (YV == 2 ? 2 : 3) + (I), 256)
  ENG_NRY(4, 4, 2, 0); ENG_NRY(4, 8, 2, 1); ENG_NRY(4, 16, 2, 2);
  ENG_NRY(8, 4, 2, 3); ENG_NRY(8, 8, 2, 4); ENG_NRY(8, 16, 2, 5);
  ENG_NRY(12, 4, 2, 6); ENG_NRY(12, 8, 2, 7); ENG_NRY(12, 16, 2, 8);
  ENG_NRY(16, 4, 2, 9); ENG_NRY(16, 8, 2, 10); ENG_NRY(16, 16, 2, 11);
  ENG_NRY(4, 4, 4, 0); ENG_NRY(4, 8, 4, 1);
  ENG_NRY(8, 4, 4, 3); ENG_NRY(8, 8, 4, 4);
  ENG_NRY(12, 4, 4, 6); ENG_NRY(12, 8, 4, 7);
  ENG_NRY(16, 4, 4, 9); ENG_NRY(16, 8, 4, 10);
#undef ENG_NRY
  // tile-mode plans (balls wider than 16 cells) run their own launches
  for (int slot : E.groups[E.cur_group]) {
    EngJob& e = E.jobs[slot];
    if (!e.plan || e.run_plan()) continue;
    int rc = svimx_nr_select(E.st, e.plan);
    if (!rc) rc = svimx_nr_sum(E.st, e.plan, e.desc.occupancy, e.desc.field, e.desc.ws, e.desc.ws_bytes);
    if (rc) return rc;
  }
  ENG_LAUNCH(k_eng_curvatures, T_CURV, kBandThreads, rel);
  return SVIMX_OK;
}

// Velocity of buffer cur+rel fused with the RK stage that consumes it (as the
// single-job path runs them): stage 0 stores L0 and the statistics and
// derives dt; stages 1 and 2 write buffer cur+rel_out from u0 = cur.
static int velocity_rk(Engine& E, int stage, int rel, int rel_out) {
  if (stage == 0) {
    ENG_LAUNCH((k_eng_velocity_rk<true, 1>), T_SVI, kBandThreads, rel, 0, rel, 1, E.d_dts);
    ENG_LAUNCH((k_eng_velocity_rk<false, 1>), T_CURV, kBandThreads, rel, 0, rel, 1, E.d_dts);
    ENG_LAUNCH(k_eng_stats_dt, T_JOBS, kFinalizeThreads, E.d_dts);
  } else if (stage == 1) {
    ENG_LAUNCH((k_eng_velocity_rk<true, 2>), T_SVI, kBandThreads, rel, 0, rel_out, 0, E.d_dts);
    ENG_LAUNCH((k_eng_velocity_rk<false, 2>), T_CURV, kBandThreads, rel, 0, rel_out, 0, E.d_dts);
  } else {
    ENG_LAUNCH((k_eng_velocity_rk<true, 3>), T_SVI, kBandThreads, rel, 0, rel_out, 0, E.d_dts);
    ENG_LAUNCH((k_eng_velocity_rk<false, 3>), T_CURV, kBandThreads, rel, 0, rel_out, 0, E.d_dts);
  }
  return SVIMX_OK;
}

static int volume(Engine& E, int rel) {
  ENG_LAUNCH(k_eng_clear_marks, T_MARKS, 256, 0);
  ENG_LAUNCH(k_eng_fractions, T_ALL, kBandThreads, rel, 0);
  ENG_LAUNCH(k_eng_finalize_volume, T_JOBS, kFinalizeThreads);
  ENG_LAUNCH(k_eng_gather, T_JOBS, 256, E.res_map);
  return SVIMX_OK;
}

static int rk(Engine& E, int stage, int rel_in, int rel_u0, int rel_out) {
  ENG_LAUNCH(k_eng_rk, T_ALL, kBandThreads, stage, rel_in, rel_u0, rel_out, E.d_dts);
  return SVIMX_OK;
}

// stages 1 and 2 after stage 0 (or its replay) wrote b = cur+1; then the
// volume of b
static int tail_and_volume(Engine& E) {
  int rc = inputs(E, 1, 0);
  if (!rc) rc = velocity_rk(E, 1, 1, 2);
  if (!rc) rc = inputs(E, 2, 0);
  if (!rc) rc = velocity_rk(E, 2, 2, 1);
  if (!rc) rc = volume(E, 1);
  return rc;
}

// stage 0 as velocity + L0 and statistics, then its update b = u0 + dt L0
// (the stage-3 replay of rk_node)
static int main_sequence(Engine& E) {
  return each_group(E, [&] {
    int rc = inputs(E, 0, 1);
    if (!rc) rc = velocity_rk(E, 0, 0, 0);
    if (!rc) rc = rk(E, 3, 0, 0, 1);
    if (!rc) rc = tail_and_volume(E);
    return rc;
  });
}

static std::vector<int> group_shape(const Engine& E) {
  std::vector<int> shape;
  for (const auto& g : E.groups) shape.push_back((int)g.size());
  return shape;
}

// The gather kernel wrote the results into host memory (res_map): waiting
// for the stream is the whole readback.
static int readback(Engine& E) {
  SVIMX_CHECK_HIP(hipStreamSynchronize(E.st));
  return SVIMX_OK;
}

static const double* res_of(const Engine& E, int slot) { return E.h_res + (int64_t)slot * SVIMX_R_COUNT; }

static bool tile_jobs_in(const Engine& E) {
  for (int slot : E.set)
    if (E.jobs[slot].plan && !E.jobs[slot].plan->run_mode) return true;
  return false;
}

constexpr int kGraphCache = 6;

// Drop the captured graphs (of every set, or of the sets holding slot).
static void drop_graph(Engine& E, int slot = -1) {
  auto holds = [&](const Engine::Captured& c) {
    return slot < 0 || std::find(c.set.begin(), c.set.end(), slot) != c.set.end();
  };
  for (auto& c : E.captured)
    if (holds(c)) hipGraphExecDestroy(c.exec);
  E.captured.erase(std::remove_if(E.captured.begin(), E.captured.end(), holds), E.captured.end());
}

// The main tick's launches for the current set (captured once per set when
// graphs are on: the tables it reads are refreshed before every replay).
static int jobs_and_main_sequence(Engine& E) {
  int rc = upload_jobs(E);
  if (!rc) rc = main_sequence(E);
  return rc;
}

static int launch_main(Engine& E) {
  if (!E.graphs || tile_jobs_in(E)) return jobs_and_main_sequence(E);
  Engine::Captured* hit = nullptr;
  const std::vector<int> shape = group_shape(E);
  for (auto& c : E.captured)
    if (c.set == E.set && c.shape == shape) hit = &c;
  if (!hit) {
    if ((int)E.captured.size() >= kGraphCache) {  // evict the least recently used
      auto lru = std::min_element(E.captured.begin(), E.captured.end(),
                                  [](const Engine::Captured& a, const Engine::Captured& b) {
                                    return a.last_use < b.last_use;
                                  });
      hipGraphExecDestroy(lru->exec);
      E.captured.erase(lru);
    }
    // Relaxed: other threads of the process (set-up and meshing workers) keep
    // allocating and freeing while this thread captures.
    SVIMX_CHECK_HIP(hipStreamBeginCapture(E.st, hipStreamCaptureModeRelaxed));
    int rc = jobs_and_main_sequence(E);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(E.st, &g);
    if (rc || e != hipSuccess || !g) {
      // The capture was invalidated (by a launch error, or by a device-wide
      // operation such as a synchronous free in another thread of the
      // process): none of the recorded work ran.  Clear the error and run
      // this tick eagerly; the next tick captures again.  A genuine launch
      // error fails the eager sequence as well and is returned from there.
      if (g) hipGraphDestroy(g);
      (void)hipGetLastError();
      hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
      SVIMX_CHECK_HIP(hipStreamIsCapturing(E.st, &cs));
      if (cs != hipStreamCaptureStatusNone) {
        // an invalidated capture whose end was refused: end it again and drop
        // what it recorded before running the tick eagerly
        hipGraph_t g2 = nullptr;
        (void)hipStreamEndCapture(E.st, &g2);
        if (g2) hipGraphDestroy(g2);
        (void)hipGetLastError();
        SVIMX_CHECK_HIP(hipStreamIsCapturing(E.st, &cs));
      }
      SVIMX_REQUIRE(cs == hipStreamCaptureStatusNone, SVIMX_EDEVICE, "engine stream left in capture state");
      ++E.capture_fallbacks;
      return jobs_and_main_sequence(E);
    }
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    SVIMX_CHECK_HIP(e);
    E.captured.push_back(Engine::Captured{E.set, shape, exec, 0});
    hit = &E.captured.back();
  }
  hit->last_use = ++E.graph_clock;
  SVIMX_CHECK_HIP(hipGraphLaunch(hit->exec, E.st));
  return SVIMX_OK;
}

static void stats_info(EngJob& e, const double* r) {
  const bool near = r[SVIMX_R_NNEAR] > 0;
  e.qmin = e.qmax = NAN;
  e.capped = 0;
  if (e.svi()) {
    if (r[SVIMX_R_NALL] > 0 || near) {
      e.qmin = r[near ? SVIMX_R_QMIN : SVIMX_R_QMIN_ALL];
      e.qmax = r[near ? SVIMX_R_QMAX : SVIMX_R_QMAX_ALL];
      e.capped = (int64_t)r[near ? SVIMX_R_CAPPED : SVIMX_R_CAPPED_ALL];
    }
  } else {
    e.qmin = r[near ? SVIMX_R_QMIN : SVIMX_R_QMIN_ALL];
    e.qmax = r[near ? SVIMX_R_QMAX : SVIMX_R_QMAX_ALL];
  }
  if (near)
    e.vmean = r[SVIMX_R_VSUM] / r[SVIMX_R_NNEAR];
  else if (r[SVIMX_R_NALL] > 0)
    e.vmean = r[SVIMX_R_VSUM_ALL] / r[SVIMX_R_NALL];
  else
    e.vmean = r[SVIMX_R_VMAX];
}

static double volume_of(const EngJob& e, const double* r) {
  return ((double)e.inside_outside + r[SVIMX_R_VOLUME_SUM]) * e.desc.h3;
}

static void halt(Engine& E, int slot, int reason);

static void fail(Engine& E, int slot, int rc) {
  EngJob& e = E.jobs[slot];
  if (e.stop == SVIMX_STOP_ERROR) return;
  const char* m = svimx_last_error();
  e.message = m ? m : "";
  e.error = rc;
  halt(E, slot, SVIMX_STOP_ERROR);
}

static int ensure_band_arrays(EngJob& e) {
  SVIMX_CHECK_HIP(grow(e.V, e.vcap, std::max<int64_t>(e.nb, 1)));
  SVIMX_CHECK_HIP(grow(e.L0, e.l0cap, std::max<int64_t>(e.nb, 1)));
  return SVIMX_OK;
}

// LevelSet.reinitialize (svierosion/levelset.py:154-183) of one job.
static int reinit_job(Engine& E, EngJob& e) {
  const svimx_job_desc_t& D = e.desc;
  const int64_t n = (int64_t)D.nx * D.ny * D.nz;
  if (!e.svi() && e.nb > 0) {
    // the dense curvature fields read 0 off the band (svierosion/evolve.py:138)
    hipLaunchKernelGGL(k_eng_zero_at_band, dim3(blocks_for(e.nb, 256)), dim3(256), 0, E.st, reinterpret_cast<float2*>(D.k12), D.band,
                       e.nb);
    SVIMX_CHECK_LAUNCH();
  }
  SVIMX_CHECK_HIP(grow(e.ma, e.macap, n));
  SVIMX_CHECK_HIP(grow(e.mb, e.mbcap, n));
  float* others[2];
  for (int b = 0, q = 0; b < 3; ++b)
    if (b != e.cur) others[q++] = D.phi[b];
  int64_t cap = std::min<int64_t>(n, 2 * e.nb + 4096);
  int rc = SVIMX_OK;
  int64_t info[10] = {};
  for (int attempt = 0; attempt < 2; ++attempt) {
    SVIMX_CHECK_HIP(grow(e.widx, e.wcap, cap));
    SVIMX_CHECK_HIP(grow(e.winfo, e.icap, cap));
    SVIMX_CHECK_HIP(grow(e.geo, e.gcap, 8 * cap));
    const int64_t have = std::min(std::min(e.wcap, e.icap), e.gcap / 8);
    rc = svimx_reinit(E.st, D.phi[e.cur], others[0], others[1], D.occupancy, D.band, e.nb, e.inside_outside, D.nx,
                      D.ny, D.nz, D.h, D.width, D.reinit_iterations, 0.45, 3, D.margin, e.ma, e.mb, e.widx, e.winfo,
                      e.geo, have, D.ws, D.ws_bytes, info);
    if (rc != SVIMX_ECAPACITY) break;
    cap = info[9];
  }
  if (rc == SVIMX_OK || rc == SVIMX_EMARGIN) {
    e.nb = info[0];
    e.inside_outside = info[1];
  }
  if (rc) return rc;
  return ensure_band_arrays(e);
}

static void push_record(EngJob& e, double vnew, bool did_reinit) {
  svimx_step_record_t r{};
  r.step = e.steps;
  r.time = e.time;
  r.dt = e.dt;
  r.volume = vnew;
  r.fraction = vnew / e.V0;
  r.vmax = e.vmax;
  r.vmean = e.vmean;
  r.qmin = e.qmin;
  r.qmax = e.qmax;
  r.capped = e.capped;
  r.reinit = did_reinit ? 1 : 0;
  r.wall = e.tick_wall;
  e.records.push_back(r);
}

static void halt(Engine& E, int slot, int reason) {
  EngJob& e = E.jobs[slot];
  e.running = 0;
  e.stop = reason;
  E.stopped.push_back(slot);
}

// stop tests at the top of the run loop (svierosion/evolve.py:187-197)
static int stop_test(const EngJob& e) {
#pragma clang fp contract(off)
  if (e.steps - e.n_start >= e.max_steps) return SVIMX_STOP_MAX_STEPS;
  if (e.has_target && e.vol <= e.target * (1 + e.tol)) return SVIMX_STOP_TARGET;
  if (e.has_max_time && e.time >= e.max_time * (1 - 1e-12)) return SVIMX_STOP_MAX_TIME;
  if (!(e.res[SVIMX_R_HAS_NEG] > 0 && e.res[SVIMX_R_HAS_POS] > 0)) return SVIMX_STOP_VANISHED;
  return SVIMX_STOP_NONE;
}

// Estimated steps left in the job's run: the volume still to go over the
// volume lost per step since the run started (infinite for the first steps
// of a run, so every job gets measured early); max_steps bounds it.
static double steps_left(const EngJob& e) {
  const double done = (double)(e.steps - e.n_start);
  const double cap = (double)e.max_steps - done;
  if (!e.has_target) return cap;
  if (done < 4) return INFINITY;
  const double rate = (e.vol_run0 - e.vol) / done;
  if (!(rate > 0)) return INFINITY;
  return std::min(cap, (e.vol - e.target) / rate);
}

// The jobs of this tick (E.set, increasing slot order) among the running
// candidates, by the engine's schedule.
static void select_jobs(Engine& E, const std::vector<int>& cand) {
  E.set.clear();
  if (E.policy == SVIMX_SCHED_LOCKSTEP || cand.size() <= 1) {
    E.set = cand;
    return;
  }
  std::vector<std::pair<double, int>> order;
  order.reserve(cand.size());
  for (int slot : cand) {
    const EngJob& e = E.jobs[slot];
    order.emplace_back(E.policy == SVIMX_SCHED_LONGEST ? -steps_left(e) : (double)e.seq, slot);
  }
  std::stable_sort(order.begin(), order.end(),
                   [](const std::pair<double, int>& a, const std::pair<double, int>& b) { return a.first < b.first; });
  const double cap = std::max((double)E.budget_nodes, (1.0 + E.slack) * (double)E.jobs[order[0].second].nb);
  double used = 0;
  for (const auto& o : order) {
    const EngJob& e = E.jobs[o.second];
    const double nb = (double)e.nb;
    // SHORT: behind the oldest job only jobs with little work left (or too few
    // steps in this run to tell) ride along, whatever their band size (they add
    // a few ticks' work each); long jobs wait for their turn
    bool joins = used + nb <= cap;
    if (E.policy == SVIMX_SCHED_SHORT)
      joins = e.steps - e.n_start < 4 || steps_left(e) * std::max(nb, 1.0) <= E.short_work;
    if (E.set.empty() || joins) {
      E.set.push_back(o.second);
      used += nb;
    }
  }
  std::sort(E.set.begin(), E.set.end());
}

static int tick(Engine& E, int* n_running) {
#pragma clang fp contract(off)
  const double t0 = now_ms();
  std::vector<int> cand;
  for (int slot = 0; slot < (int)E.jobs.size(); ++slot) {
    EngJob& e = E.jobs[slot];
    if (!e.used || !e.running) continue;
    const int why = stop_test(e);
    if (why) halt(E, slot, why);
    else cand.push_back(slot);
  }
  if (cand.empty()) {
    *n_running = 0;
    return SVIMX_OK;
  }
  select_jobs(E, cand);
  // one step of every job in the set (launch_main uploads the job records)
  int rc = upload(E, false);
  if (!rc) rc = launch_main(E);
  if (!rc) rc = readback(E);
  if (rc) {
    for (int slot : E.set) fail(E, slot, rc);
    return rc;
  }
  std::vector<int> secant, stepped;
  for (int slot : E.set) {
    EngJob& e = E.jobs[slot];
    const double* r = res_of(E, slot);
    e.dt = r[SVIMX_R_DT];
    e.vmax = r[SVIMX_R_VMAX];
    stats_info(e, r);
    if (!std::isfinite(e.dt) || e.dt <= 0) {
      halt(E, slot, SVIMX_STOP_STALLED);
      e.frac_valid = -1;  // the occupancy now holds another buffer's fractions
      continue;
    }
    e.vnew = volume_of(e, r);
    std::memcpy(e.res, r, sizeof(e.res));
    stepped.push_back(slot);
    if (e.has_target && e.vnew < e.target) {
      e.lo_t = 0.0;
      e.lo_v = e.vol;
      e.hi_t = e.dt;
      e.hi_v = e.vnew;
      e.dts = e.dt;
      e.v = e.vnew;
      e.secant_live = true;
      secant.push_back(slot);
    }
  }
  const double t1 = now_ms();
  // regula falsi onto the target volume (svierosion/evolve.py:215-231), in
  // lockstep over the jobs that passed their target
  for (int it = 0; it < 8 && !secant.empty(); ++it) {
    E.set.clear();
    for (int slot : secant) {
      EngJob& e = E.jobs[slot];
      if (!e.secant_live) continue;
      if (e.hi_v >= e.lo_v) {
        e.secant_live = false;
        continue;
      }
      e.dts = e.lo_t + (e.target - e.lo_v) * (e.hi_t - e.lo_t) / (e.hi_v - e.lo_v);
      E.h_dts[slot] = e.dts;
      E.set.push_back(slot);
    }
    if (E.set.empty()) break;
    const int top = top_slot(E);
    rc = upload(E);
    if (!rc) {
      hipError_t he = hipMemcpyAsync(E.d_dts, E.h_dts, sizeof(double) * top, hipMemcpyHostToDevice, E.st);
      if (he != hipSuccess) {
        set_error("dt upload failed: %s", hipGetErrorString(he));
        rc = SVIMX_EDEVICE;
      }
    }
    const bool saved = E.graphs;
    E.graphs = false;  // eager launches sized for the retry set
    if (!rc) rc = each_group(E, [&] {
      int r = rk(E, 3, 0, 0, 1);
      if (!r) r = tail_and_volume(E);
      return r;
    });
    E.graphs = saved;
    if (!rc) rc = readback(E);
    if (rc) {
      for (int slot : secant) fail(E, slot, rc);
      return rc;
    }
    for (int slot : E.set) {
      EngJob& e = E.jobs[slot];
      const double* r = res_of(E, slot);
      e.v = volume_of(e, r);
      std::memcpy(e.res, r, sizeof(e.res));
      if (std::fabs(e.v - e.target) <= e.tol * e.V0) {
        e.secant_live = false;
        continue;
      }
      if (e.v > e.target) {
        e.lo_t = e.dts;
        e.lo_v = e.v;
      } else {
        e.hi_t = e.dts;
        e.hi_v = e.v;
      }
    }
  }
  for (int slot : secant) {
    EngJob& e = E.jobs[slot];
    e.dt = e.dts;
    e.vnew = e.v;
  }
  const double t2 = now_ms();
  // commit the steps; reinitialise where the displacement reached the trigger
  std::vector<int> reinit;
  for (int slot : stepped) {
    EngJob& e = E.jobs[slot];
    if (e.stop == SVIMX_STOP_ERROR) continue;
    e.cur = (e.cur + 1) % 3;
    e.frac_valid = e.cur;
    e.time += e.dt;
    e.steps += 1;
    e.displacement += e.dt * e.vmax;
    const bool reached = e.has_target && e.vnew <= e.target * (1 + e.tol);
    if (e.displacement >= e.desc.reinit_distance * e.desc.h && !reached) reinit.push_back(slot);
  }
  E.set.clear();
  for (int slot : reinit) {
    EngJob& e = E.jobs[slot];
    int r2 = reinit_job(E, e);
    if (r2) {
      fail(E, slot, r2);
      continue;
    }
    e.frac_valid = -1;
    e.displacement = 0.0;
    E.set.push_back(slot);
  }
  if (!E.set.empty()) {
    rc = upload(E);
    const bool saved = E.graphs;
    E.graphs = false;
    if (!rc) rc = each_group(E, [&] { return volume(E, 0); });
    E.graphs = saved;
    if (!rc) rc = readback(E);
    if (rc) {
      for (int slot : E.set) fail(E, slot, rc);
      return rc;
    }
    for (int slot : E.set) {
      EngJob& e = E.jobs[slot];
      const double* r = res_of(E, slot);
      e.vnew = volume_of(e, r);
      std::memcpy(e.res, r, sizeof(e.res));
      e.frac_valid = e.cur;
    }
  }
  const double t3 = now_ms();
  for (int slot : stepped) {
    EngJob& e = E.jobs[slot];
    if (e.stop == SVIMX_STOP_ERROR) continue;
    e.tick_wall = (t3 - t0) * 1e-3;
    const bool did = std::find(reinit.begin(), reinit.end(), slot) != reinit.end();
    push_record(e, e.vnew, did);
    e.vol = e.vnew;
    // the next iteration's stop tests, so a finished job leaves at once
    const int why = stop_test(e);
    if (why) halt(E, slot, why);
  }
  E.ms_tick += t1 - t0;
  E.ms_secant += t2 - t1;
  E.ms_reinit += t3 - t2;
  int live = 0;
  for (auto& e : E.jobs) live += e.used && e.running;
  *n_running = live;
  return SVIMX_OK;
}

}  // namespace svimx

using namespace svimx;

extern "C" {

int svimx_engine_create(void* stream, int max_jobs, void** engine) {
  SVIMX_TRY_BEGIN
  SVIMX_REQUIRE(engine && max_jobs > 0 && max_jobs <= kMaxEngineJobs, SVIMX_EBADARG,
                "max_jobs must be in 1..%d", kMaxEngineJobs);
  Engine* E = new Engine();
  E->max_jobs = max_jobs;
  E->jobs.resize(max_jobs);
  hipError_t e = hipSuccess;
  if (stream) {
    E->st = as_stream(stream);
  } else {
    e = hipStreamCreateWithFlags(&E->st, hipStreamNonBlocking);
    E->own_stream = e == hipSuccess;
  }
  if (e == hipSuccess) e = hipMalloc(&E->d_jobs, sizeof(JobDev) * max_jobs);
  if (e == hipSuccess) e = hipMalloc(&E->d_tab, sizeof(Tables) * kMaxGroups);
  if (e == hipSuccess) e = hipMalloc(&E->d_dts, sizeof(double) * max_jobs);
  if (e == hipSuccess) e = hipMalloc(&E->d_res, sizeof(double) * SVIMX_R_COUNT * max_jobs);
  if (e == hipSuccess) e = hipHostMalloc(&E->h_jobs, sizeof(JobDev) * max_jobs);
  if (e == hipSuccess) e = hipHostMalloc(&E->h_tab, sizeof(Tables) * kMaxGroups);
  if (e == hipSuccess) e = hipHostMalloc(&E->h_dts, sizeof(double) * max_jobs);
  if (e == hipSuccess) e = hipHostMalloc(&E->h_res, sizeof(double) * SVIMX_R_COUNT * max_jobs, hipHostMallocMapped);
  if (e == hipSuccess) e = hipHostGetDevicePointer(reinterpret_cast<void**>(&E->res_map), E->h_res, 0);
  if (e != hipSuccess) {
    set_error("engine allocation failed: %s", hipGetErrorString(e));
    svimx_engine_destroy(E);
    return SVIMX_EDEVICE;
  }
  std::memset(E->h_tab, 0, sizeof(Tables) * kMaxGroups);
  *engine = E;
  return SVIMX_OK;
  SVIMX_TRY_END
}

int svimx_engine_destroy(void* engine) {
  Engine* E = static_cast<Engine*>(engine);
  if (!E) return SVIMX_OK;
  for (int id = 0; id < (int)E->jobs.size(); ++id)
    if (E->jobs[id].used) svimx_engine_remove(E, id);
  drop_graph(*E);
  hipFree(E->d_jobs);
  hipFree(E->d_tab);
  hipFree(E->d_dts);
  hipFree(E->d_res);
  hipHostFree(E->h_jobs);
  hipHostFree(E->h_tab);
  hipHostFree(E->h_dts);
  hipHostFree(E->h_res);
  if (E->own_stream) hipStreamDestroy(E->st);
  delete E;
  return SVIMX_OK;
}

int svimx_engine_add(void* engine, const svimx_job_desc_t* desc, int* job_id) {
  SVIMX_TRY_BEGIN
  Engine* E = static_cast<Engine*>(engine);
  SVIMX_REQUIRE(E && desc && job_id, SVIMX_EBADARG, "null argument");
  const svimx_job_desc_t& D = *desc;
  const int64_t n = (int64_t)D.nx * D.ny * D.nz;
  SVIMX_REQUIRE(D.phi[0] && D.phi[1] && D.phi[2] && D.occupancy && D.band && D.ws, SVIMX_EBADARG,
                "job buffers missing");
  SVIMX_REQUIRE(D.cur >= 0 && D.cur < 3 && D.nb >= 0 && D.nb <= D.band_cap && D.band_cap <= n, SVIMX_EBADARG,
                "bad job state");
  SVIMX_REQUIRE(workspace_bytes_for(n) <= D.ws_bytes, SVIMX_EBADARG, "job workspace too small");
  SVIMX_REQUIRE(D.law.kind >= 0 && D.law.kind <= 2, SVIMX_EBADARG, "bad law");
  SVIMX_REQUIRE(D.law.kind != 0 || (D.nr_plan && D.field), SVIMX_EBADARG, "SVI job needs an N_r plan and field");
  SVIMX_REQUIRE(D.law.kind == 0 || D.k12, SVIMX_EBADARG, "curvature job needs the (k1, k2) field");
  SVIMX_REQUIRE(D.h > 0 && D.initial_volume != 0, SVIMX_EBADARG, "bad spacing or initial volume");
  if (D.nr_plan) {
    const NrPlan* p = static_cast<const NrPlan*>(D.nr_plan);
    SVIMX_REQUIRE(p->d.nx == D.nx && p->d.ny == D.ny && p->d.nz == D.nz, SVIMX_EBADARG,
                  "N_r plan grid does not match");
  }
  int id = -1;
  for (int s = 0; s < (int)E->jobs.size(); ++s)
    if (!E->jobs[s].used) {
      id = s;
      break;
    }
  SVIMX_REQUIRE(id >= 0, SVIMX_ECAPACITY, "engine holds %d jobs already", E->max_jobs);
  EngJob& e = E->jobs[id];
  e = EngJob();
  e.used = true;
  e.desc = D;
  e.plan = D.law.kind == 0 ? static_cast<NrPlan*>(D.nr_plan) : nullptr;
  e.cur = D.cur;
  e.frac_valid = D.frac_valid;
  e.nb = D.nb;
  e.inside_outside = D.inside_outside;
  e.steps = D.steps;
  e.time = D.time;
  e.displacement = D.displacement;
  e.V0 = D.initial_volume;
  e.seq = E->next_seq++;
  int rc = ensure_band_arrays(e);
  if (rc) {
    svimx_engine_remove(E, id);
    return rc;
  }
  *job_id = id;
  return SVIMX_OK;
  SVIMX_TRY_END
}

int svimx_engine_run(void* engine, int id, double target_fraction, double max_time, int64_t max_steps, double tol) {
  SVIMX_TRY_BEGIN
  Engine* E = static_cast<Engine*>(engine);
  SVIMX_REQUIRE(E && id >= 0 && id < (int)E->jobs.size() && E->jobs[id].used, SVIMX_EBADARG, "bad job id");
  SVIMX_REQUIRE(max_steps >= 0, SVIMX_EBADARG, "max_steps must be >= 0");
  EngJob& e = E->jobs[id];
  e.has_target = target_fraction > 0;
  e.target = e.has_target ? target_fraction * e.V0 : 0.0;
  e.has_max_time = max_time > 0;
  e.max_time = e.has_max_time ? max_time : 0.0;
  e.max_steps = max_steps;
  e.n_start = e.steps;
  e.tol = tol;
  e.stop = SVIMX_STOP_NONE;
  e.error = 0;
  e.message.clear();
  // current volume (the fractions are recomputed, as _volume_of does)
  E->set.assign(1, id);
  e.frac_valid = -1;
  int rc = upload(*E);
  const bool saved = E->graphs;
  E->graphs = false;
  if (!rc) rc = volume(*E, 0);
  E->graphs = saved;
  if (!rc) rc = readback(*E);
  if (rc) {
    fail(*E, id, rc);
    return rc;
  }
  std::memcpy(e.res, res_of(*E, id), sizeof(e.res));
  e.vol = volume_of(e, e.res);
  e.vol_run0 = e.vol;
  e.frac_valid = e.cur;
  e.running = 1;
  return SVIMX_OK;
  SVIMX_TRY_END
}

int svimx_engine_tick(void* engine, int* n_running) {
  SVIMX_TRY_BEGIN
  Engine* E = static_cast<Engine*>(engine);
  SVIMX_REQUIRE(E && n_running, SVIMX_EBADARG, "null argument");
  return tick(*E, n_running);
  SVIMX_TRY_END
}

int svimx_engine_ticks(void* engine, int max_ticks, int* n_running, int* n_ticks) {
  SVIMX_TRY_BEGIN
  Engine* E = static_cast<Engine*>(engine);
  SVIMX_REQUIRE(E && n_running && n_ticks && max_ticks > 0, SVIMX_EBADARG, "null argument or max_ticks < 1");
  *n_ticks = 0;
  do {
    const int rc = tick(*E, n_running);
    if (rc) return rc;
    ++*n_ticks;
  } while (*n_ticks < max_ticks && *n_running > 0 && E->stopped.empty());
  return SVIMX_OK;
  SVIMX_TRY_END
}

int svimx_engine_status(void* engine, int id, svimx_job_status_t* out) {
  Engine* E = static_cast<Engine*>(engine);
  SVIMX_REQUIRE(E && out && id >= 0 && id < (int)E->jobs.size() && E->jobs[id].used, SVIMX_EBADARG, "bad job id");
  const EngJob& e = E->jobs[id];
  std::memset(out, 0, sizeof(*out));
  out->running = e.running;
  out->stop_reason = e.stop;
  out->error = e.error;
  out->cur = e.cur;
  out->frac_valid = e.frac_valid;
  out->nb = e.nb;
  out->inside_outside = e.inside_outside;
  out->steps = e.steps;
  out->nrecords = (int64_t)e.records.size();
  out->time = e.time;
  out->volume = e.vol;
  out->displacement = e.displacement;
  std::snprintf(out->message, sizeof(out->message), "%s", e.message.c_str());
  return SVIMX_OK;
}

int svimx_engine_records(void* engine, int id, int64_t first, int64_t n, svimx_step_record_t* out) {
  Engine* E = static_cast<Engine*>(engine);
  SVIMX_REQUIRE(E && id >= 0 && id < (int)E->jobs.size() && E->jobs[id].used, SVIMX_EBADARG, "bad job id");
  const EngJob& e = E->jobs[id];
  SVIMX_REQUIRE(first >= 0 && n >= 0 && first + n <= (int64_t)e.records.size() && (out || n == 0), SVIMX_EBADARG,
                "record range out of bounds");
  if (n) std::memcpy(out, e.records.data() + first, sizeof(svimx_step_record_t) * n);
  return SVIMX_OK;
}

int svimx_engine_remove(void* engine, int id) {
  Engine* E = static_cast<Engine*>(engine);
  SVIMX_REQUIRE(E && id >= 0 && id < (int)E->jobs.size() && E->jobs[id].used, SVIMX_EBADARG, "bad job id");
  // the job's kernels must be done before its arrays are freed
  SVIMX_CHECK_HIP(hipStreamSynchronize(E->st));
  EngJob& e = E->jobs[id];
  hipFree(e.V);
  hipFree(e.L0);
  hipFree(e.ma);
  hipFree(e.mb);
  hipFree(e.widx);
  hipFree(e.winfo);
  hipFree(e.geo);
  e = EngJob();
  E->stopped.erase(std::remove(E->stopped.begin(), E->stopped.end(), id), E->stopped.end());
  drop_graph(*E, id);
  return SVIMX_OK;
}

int svimx_engine_schedule(void* engine, int policy, int64_t budget_nodes, double slack) {
  Engine* E = static_cast<Engine*>(engine);
  SVIMX_REQUIRE(E, SVIMX_EBADARG, "null engine");
  SVIMX_REQUIRE(policy >= SVIMX_SCHED_LOCKSTEP && policy <= SVIMX_SCHED_SHORT, SVIMX_EBADARG,
                "unknown schedule %d", policy);
  SVIMX_REQUIRE(budget_nodes >= 0 && slack >= 0 && std::isfinite(slack), SVIMX_EBADARG,
                "budget_nodes and slack must be >= 0");
  E->policy = policy;
  E->budget_nodes = budget_nodes;
  E->slack = slack;
  return SVIMX_OK;
}

int svimx_engine_short_work(void* engine, double node_steps) {
  Engine* E = static_cast<Engine*>(engine);
  SVIMX_REQUIRE(E, SVIMX_EBADARG, "null engine");
  SVIMX_REQUIRE(node_steps >= 0 && std::isfinite(node_steps), SVIMX_EBADARG, "short_work must be finite and >= 0");
  E->short_work = node_steps;
  return SVIMX_OK;
}

int svimx_engine_group_nodes(void* engine, int64_t nodes) {
  Engine* E = static_cast<Engine*>(engine);
  SVIMX_REQUIRE(E, SVIMX_EBADARG, "null engine");
  SVIMX_REQUIRE(nodes >= 0, SVIMX_EBADARG, "group_nodes must be >= 0");
  E->group_nodes = nodes;
  return SVIMX_OK;
}

int svimx_engine_use_graphs(void* engine, int on) {
  Engine* E = static_cast<Engine*>(engine);
  SVIMX_REQUIRE(E, SVIMX_EBADARG, "null engine");
  if (!on) drop_graph(*E);
  E->graphs = on != 0;
  return SVIMX_OK;
}

int svimx_engine_stopped(void* engine, int* ids, int cap, int* n) {
  Engine* E = static_cast<Engine*>(engine);
  SVIMX_REQUIRE(E && n && (ids || cap == 0), SVIMX_EBADARG, "null argument");
  const int k = std::min<int>(cap, (int)E->stopped.size());
  for (int i = 0; i < k; ++i) ids[i] = E->stopped[i];
  E->stopped.erase(E->stopped.begin(), E->stopped.begin() + k);
  *n = k;
  return SVIMX_OK;
}

int svimx_engine_timers(void* engine, double* out3) {
  Engine* E = static_cast<Engine*>(engine);
  SVIMX_REQUIRE(E && out3, SVIMX_EBADARG, "null argument");
  out3[0] = E->ms_tick;
  out3[1] = E->ms_secant;
  out3[2] = E->ms_reinit;
  return SVIMX_OK;
}

int64_t svimx_engine_capture_fallbacks(void* engine) {
  const Engine* E = static_cast<const Engine*>(engine);
  return E ? E->capture_fallbacks : 0;
}

}  // extern "C"
