// Batched ensemble of CDKLM16 members of uneven size: one persistent launch
// per RK2 substep over all members, batched SOAR model error, and hipGraph
// capture of whole steps.  Built with -ffp-contract=off (noise kernels must
// match the per-member path bit for bit).
//
// Replaces the reference's sequential member loop
// (ensembles/OceanModelEnsemble.py:138-150), where each member is a separate
// CDKLM16 object with its own JIT module, stream and blocking dt read-back.

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "boundary_cells.h"
#include "cdklm_common.h"
#include "drift_cells.h"
#include "noise_cells.h"
#include "noise_tile.h"
#include "philox.h"

namespace gpo {
namespace {

struct NoiseMember {
  int nx, ny, per_ns, per_ew;
  int xi_rows, xi_cols, xi_pitch, c_pitch, pitch, hi_pitch, f_pitch;
  uint32_t k0, k1, stream_id;
  float W[25];
  float scale, dx, dy, g, f, beta, north_x, north_y, land;
  float* xi;
  float* coarse;
  float* eta;
  float* hu;
  float* hv;
  const float* Hi;
  const float* f_cells;
  // model error on a coarser grid: xi and SOAR over cnx x cny, then the
  // bicubic interpolation to this member's grid (coarse = 0: on the grid)
  int coarse_grid;
  BicubicGeo geo;
};

__global__ void __launch_bounds__(256) noise_fused(const NoiseMember* __restrict__ members,
                                                   const int2* __restrict__ tiles,
                                                   const uint32_t* __restrict__ draw) {
  const int2 t = tiles[blockIdx.x];
  const NoiseMember& m = members[t.x];
  NoiseTile a;
  a.nx = m.nx, a.ny = m.ny, a.per_ns = m.per_ns, a.per_ew = m.per_ew;
  a.xi_cols = m.xi_cols, a.pitch = m.pitch, a.hi_pitch = m.hi_pitch, a.f_pitch = m.f_pitch;
  a.k0 = m.k0, a.k1 = m.k1, a.stream_id = m.stream_id, a.draw = draw[t.x];
  a.W = m.W;
  a.scale = m.scale, a.dx = m.dx, a.dy = m.dy, a.g = m.g, a.f = m.f, a.beta = m.beta;
  a.north_x = m.north_x, a.north_y = m.north_y, a.land = m.land;
  a.eta = m.eta, a.hu = m.hu, a.hv = m.hv, a.Hi = m.Hi, a.f_cells = m.f_cells;
  a.pre_e = nullptr, a.d_out = nullptr, a.d_row = 0;
  noise_tile<false, false>(a, 2 + kTileH * (t.y >> 16), 2 + kTileW * (t.y & 0xFFFF));
}

// The model-error tile table serves both the fused kernel and the bicubic one.
static_assert(kTileW == kBiW && kTileH == kBiH, "model-error tiles of the two paths must agree");

// Coarse-grid model error in three launches, each over every member (block
// row y = member y; blocks past a member's extent and members left out of a
// selection return at once):
// 1. the member's normal field xi (rows x cols) of the current draw, one
//    Philox call per four consecutive elements (gpo_random_normal's values);
__global__ void __launch_bounds__(256) xi_batched(const NoiseMember* __restrict__ members,
                                                  const uint32_t* __restrict__ draw,
                                                  const uint8_t* __restrict__ active) {
  const NoiseMember& m = members[blockIdx.y];
  if (active && !active[blockIdx.y]) return;
  const uint32_t n = (uint32_t)m.xi_rows * (uint32_t)m.xi_cols;
  const uint32_t q = blockIdx.x * 256u + threadIdx.x;
  if (4u * q >= n) return;
  float v4[4];
  normals4(q, draw[blockIdx.y], m.stream_id, m.k0, m.k1, v4);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const uint32_t e = 4u * q + (uint32_t)i;
    if (e >= n) break;
    const uint32_t r = e / (uint32_t)m.xi_cols, c = e - r * (uint32_t)m.xi_cols;
    ((float*)((char*)m.xi + (size_t)r * m.xi_pitch))[c] = v4[i];
  }
}

// 2. the SOAR field over the coarse grid and its two ghost rings (gpo_soar);
__global__ void __launch_bounds__(256) soar_batched(const NoiseMember* __restrict__ members,
                                                    const uint8_t* __restrict__ active) {
  const NoiseMember& m = members[blockIdx.y];
  if (active && !active[blockIdx.y]) return;
  const int W = m.geo.cnx + 4, n = W * (m.geo.cny + 4);
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  const int tj = k / W, ti = k - tj * W;
  const float acc = soar_cell(m.geo.cnx, m.geo.cny, m.W, m.per_ns, m.per_ew, m.xi, m.xi_pitch, tj, ti);
  ((float*)((char*)m.coarse + (size_t)tj * m.c_pitch))[ti] = m.scale * acc;
}

// 3. the bicubic interpolation and its balance, one tile of the table per
//    block (gpo_bicubic_interpolation's tiles and arithmetic).
__global__ void __launch_bounds__(256) bicubic_batched(const NoiseMember* __restrict__ members,
                                                       const int2* __restrict__ tiles) {
  __shared__ BicubicLds lds;
  const int2 t = tiles[blockIdx.x];
  const NoiseMember& m = members[t.x];
  bicubic_tile(m.geo, m.g, m.f, m.beta, m.north_x, m.north_y, m.f == 0.0f ? m.f_cells : nullptr, m.f_pitch,
               m.coarse, m.c_pitch, m.eta, m.hu, m.hv, m.pitch, m.Hi, m.hi_pitch, m.land,
               2 + kBiW * (t.y & 0xFFFF), 2 + kBiH * (t.y >> 16), lds);
}

// Every member's draw index (or every selected one's) advances once the
// step's model error is applied.
__global__ void advance_draws(uint32_t* draw, int n, const uint8_t* active) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && (!active || active[i])) draw[i] = draw[i] + 1u;
}

// Flow-relaxation data of several times for one member: every level, the
// member's model time, and the two levels and weight of the current step
// (written by frs_clock at the start of each step).
struct FrsClock {
  double t;
  const double* times;
  const float* ns_levels;
  const float* ew_levels;
  long long ns_stride, ew_stride;  // floats per level
  const float *ns_cur, *ns_next, *ew_cur, *ew_next;
  float w;
  int n_times, member;
};

// BoundaryConditionsArakawaA.update_bc_values (utils/Common.py:1193-1308) at
// the member's time t, in FP64 as the host evaluates it: the levels around t
// (numpy.searchsorted's left insertion point k gives k-1 and k, clamped) and
// the weight clamp((t - t0) / max(1e-10, t1 - t0), 0, 1) rounded to FP32;
// then t advances by the member's dt, as CDKLM16.step accumulates it.
__device__ void clock_levels(FrsClock& c) {
  const double t = c.t;
  int k = 0;
  while (k < c.n_times && c.times[k] < t) ++k;
  const int i0 = k > 0 ? k - 1 : 0, i1 = k < c.n_times - 1 ? k : c.n_times - 1;
  const double elapsed = t - c.times[i0];
  const double interval = fmax(1.0e-10, c.times[i1] - c.times[i0]);
  c.w = (float)fmax(0.0, fmin(1.0, elapsed / interval));
  c.ns_cur = c.ns_levels ? c.ns_levels + i0 * c.ns_stride : nullptr;
  c.ns_next = c.ns_levels ? c.ns_levels + i1 * c.ns_stride : nullptr;
  c.ew_cur = c.ew_levels ? c.ew_levels + i0 * c.ew_stride : nullptr;
  c.ew_next = c.ew_levels ? c.ew_levels + i1 * c.ew_stride : nullptr;
}

__global__ void frs_clock(FrsClock* __restrict__ clocks, int n, const StepArgs* __restrict__ args,
                          const uint8_t* __restrict__ active) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  FrsClock& c = clocks[i];
  if (active && !active[c.member]) return;
  clock_levels(c);
  c.t = c.t + (double)args[2 * c.member].dt;
}

// Puts the members' model times t[member] into their clocks (NULL t: keep the
// times) and derives the levels and weight at those times, without advancing
// them, so the ghost-cell passes are valid before the next step's clock launch.
__global__ void clock_set_times(FrsClock* __restrict__ clocks, int n, const double* __restrict__ t) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  FrsClock& c = clocks[i];
  if (t) c.t = t[c.member];
  clock_levels(c);
}

// Ghost cells of one member, as gpo_boundary_conditions writes them, on its
// state (h0) and RK stage (h1).
struct BcMember {
  cells::BcTarget h0, h1;
  int frs;  // some side is flow relaxation: ghost cells after substep 0 too
  const FrsClock* clock;  // flow-relaxation levels and weight of this step, or NULL (fixed data)
};

// pass 0: north/south, pass 1: east/west; target 0: h0, 1: h1 (members with
// flow relaxation only).  Block row y is member y.
__global__ void __launch_bounds__(256) bc_batched(const BcMember* __restrict__ members, int pass, int target) {
  const BcMember& b = members[blockIdx.y];
  if (target == 1 && !b.frs) return;
  if (b.clock) {
    cells::BcTarget t = target ? b.h1 : b.h0;
    const FrsClock& c = *b.clock;
    t.ns_cur = c.ns_cur, t.ns_next = c.ns_next, t.ew_cur = c.ew_cur, t.ew_next = c.ew_next;
    t.frs_t = c.w;
    cells::bc_target_cell(t, pass, blockIdx.x * blockDim.x + threadIdx.x);
    return;
  }
  cells::bc_target_cell(target ? b.h1 : b.h0, pass, blockIdx.x * blockDim.x + threadIdx.x);
}

// Drifters of one member; the time step is the member's current one.
struct DriftMember {
  int nx, ny, per_ns, per_ew, rows, pos_pitch, pitch, bm_pitch, arg;
  float dx, dy, sensitivity;
  const float *eta, *hu, *hv, *Bm;
  float* pos;
};

__global__ void __launch_bounds__(64) drift_batched(const DriftMember* __restrict__ members,
                                                    const StepArgs* __restrict__ args) {
  const DriftMember& d = members[blockIdx.y];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d.rows) return;
  cells::drift_one(d.nx, d.ny, d.dx, d.dy, args[d.arg].dt, 2, 2, d.eta, d.hu, d.hv, d.pitch, d.Bm, d.bm_pitch,
                   d.per_ns, d.per_ew, cells::drift_row(d.pos, d.pos_pitch, i), d.sensitivity, nullptr, nullptr,
                   nullptr, nullptr, 0, 0, 0.0f, 0.0f);
}

int check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    char buf[256];
    snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
    set_error(buf);
    return (int)e;
  }
  return 0;
}

int bad(const char* msg) {
  set_error(msg);
  return GPO_ERR_BAD_ARGUMENT;
}

}  // namespace

struct Ensemble {
  int n = 0, n_work = 0, n_blocks = 0, exact = 1, has_noise = 0;
  int has_land = 0, f_array = 0, angle0 = 1;
  int any_stage = 0;  // some member has bottom friction: the general substep kernel
  int all_ok = 1;     // every member's fast divisions and square root verified
  int n_tiles = 0;  // model-error tiles over all members
  // model error on coarser grids: the largest normal field (in Philox quads)
  // and coarse buffer over the members, the x extents of the first two launches
  int coarse_noise = 0, max_xi_quads = 0, max_coarse_cells = 0;
  // flow-relaxation clocks (members whose boundary data has several times)
  std::vector<FrsClock> h_clock;
  FrsClock* d_clock = nullptr;
  StepArgs* d_args = nullptr;
  int2* d_tiles = nullptr;
  int2* d_work = nullptr;
  NoiseMember* d_noise = nullptr;
  uint32_t* d_draw = nullptr;
  BcMember* d_bc = nullptr;   // members whose ghost cells are written every step
  int n_bc = 0, bc_blocks = 0, any_frs = 0;
  // members with a periodic side whose ghost cells the steps do not write:
  // part 5 writes them once, after a run of steps
  BcMember* d_bc_end = nullptr;
  int n_bc_end = 0, bc_end_blocks = 0;
  double* d_times = nullptr;  // staging of gpo_ensemble_set_times (one per member)
  DriftMember* d_drift = nullptr;
  int n_drift = 0, drift_blocks = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_steps = 0;
  std::vector<StepArgs> h_args;  // host copy of the argument table (2 per member)
  // Host copies of the full tables, with the member of each BC and drifter
  // entry, from which a selection of members builds its own tables.
  std::vector<int2> h_work, h_tiles;
  std::vector<NoiseMember> h_noise;
  std::vector<BcMember> h_bc, h_bc_end;
  std::vector<DriftMember> h_drift;
  std::vector<int> bc_member, bc_end_member, drift_member;
  // Tables the launches use: the full ones, or those of the selected members.
  int selected = 0;
  const int2* cur_work = nullptr;
  const int2* cur_tiles = nullptr;
  const BcMember* cur_bc = nullptr;
  const BcMember* cur_bc_end = nullptr;
  const DriftMember* cur_drift = nullptr;
  int cur_n_work = 0, cur_n_tiles = 0, cur_n_bc = 0, cur_n_bc_end = 0, cur_n_drift = 0;
  int2* d_work_sel = nullptr;
  int2* d_tiles_sel = nullptr;
  BcMember* d_bc_sel = nullptr;
  BcMember* d_bc_end_sel = nullptr;
  DriftMember* d_drift_sel = nullptr;
  uint8_t* d_active = nullptr;
  const StepArgs* cur_args = nullptr;
  const uint8_t* cur_active = nullptr;  // per-member flags of the current selection (NULL: all)
  // Prepared selections (gpo_ensemble_prepare_sets): each set's tables, an
  // argument table with its time steps and its member flags, in one arena.
  struct Set {
    size_t work, tiles, bc, bc_end, drift, args, active;  // byte offsets in the arena
    int n_work, n_tiles, n_bc, n_bc_end, n_drift;
  };
  std::vector<Set> sets;
  char* d_arena = nullptr;
  size_t arena_bytes = 0;
  std::vector<char> h_arena;
};

static void use_full_tables(Ensemble* e) {
  e->selected = 0;
  e->cur_args = e->d_args;
  e->cur_active = nullptr;
  e->cur_work = e->d_work, e->cur_n_work = e->n_work;
  e->cur_tiles = e->d_tiles, e->cur_n_tiles = e->n_tiles;
  e->cur_bc = e->d_bc, e->cur_n_bc = e->n_bc;
  e->cur_bc_end = e->d_bc_end, e->cur_n_bc_end = e->n_bc_end;
  e->cur_drift = e->d_drift, e->cur_n_drift = e->n_drift;
}

static int ensemble_launch_substep(Ensemble* e, int rk, hipStream_t s) {
  const hipError_t err =
      e->exact ? launch_cdklm_batched_exact(e->cur_args, e->cur_work, e->cur_n_work, rk, e->n_blocks,
                                            e->has_land, e->f_array, e->angle0, e->any_stage, e->all_ok, s)
               : launch_cdklm_batched_fast(e->cur_args, e->cur_work, e->cur_n_work, rk, e->n_blocks, e->has_land,
                                           e->f_array, e->angle0, e->any_stage, e->all_ok, s);
  return err != hipSuccess ? check(err, "ensemble substep launch") : 0;
}

static int ensemble_launch_noise(Ensemble* e, hipStream_t s) {
  if (!e->has_noise) return 0;
  if (e->coarse_noise) {
    hipLaunchKernelGGL(xi_batched, dim3((e->max_xi_quads + 255) / 256, e->n), dim3(256), 0, s, e->d_noise, e->d_draw,
                       e->cur_active);
    hipLaunchKernelGGL(soar_batched, dim3((e->max_coarse_cells + 255) / 256, e->n), dim3(256), 0, s, e->d_noise,
                       e->cur_active);
    if (e->cur_n_tiles > 0)
      hipLaunchKernelGGL(bicubic_batched, dim3(e->cur_n_tiles), dim3(256), 0, s, e->d_noise, e->cur_tiles);
  } else if (e->cur_n_tiles > 0) {
    hipLaunchKernelGGL(noise_fused, dim3(e->cur_n_tiles), dim3(256), 0, s, e->d_noise, e->cur_tiles, e->d_draw);
  }
  hipLaunchKernelGGL(advance_draws, dim3((e->n + 255) / 256), dim3(256), 0, s, e->d_draw, e->n, e->cur_active);
  return check(hipGetLastError(), "ensemble noise launch");
}

static int ensemble_launch_bc(Ensemble* e, int target, hipStream_t s) {
  if (e->cur_n_bc == 0 || (target == 1 && !e->any_frs)) return 0;
  for (int pass = 0; pass < 2; ++pass)
    hipLaunchKernelGGL(bc_batched, dim3(e->bc_blocks, e->cur_n_bc), dim3(256), 0, s, e->cur_bc, pass, target);
  return check(hipGetLastError(), "ensemble boundary launch");
}

// Ghost cells of the state of members whose steps leave them unwritten
// (wall/periodic members without drifters), once after a run of steps.
static int ensemble_launch_bc_end(Ensemble* e, hipStream_t s) {
  if (e->cur_n_bc_end == 0) return 0;
  for (int pass = 0; pass < 2; ++pass)
    hipLaunchKernelGGL(bc_batched, dim3(e->bc_end_blocks, e->cur_n_bc_end), dim3(256), 0, s, e->cur_bc_end, pass, 0);
  return check(hipGetLastError(), "ensemble end-of-run boundary launch");
}

static int ensemble_launch_drift(Ensemble* e, hipStream_t s) {
  if (e->cur_n_drift == 0) return 0;
  hipLaunchKernelGGL(drift_batched, dim3(e->drift_blocks, e->cur_n_drift), dim3(64), 0, s, e->cur_drift, e->cur_args);
  return check(hipGetLastError(), "ensemble drift launch");
}

// Part 0: substep 0 (+ flow-relaxation ghost cells of the stage); part 1:
// substep 1; part 2: model error, ghost cells of the new state, drifters;
// part 3: part 2 without the model error; part 4: the model error alone;
// part 5: ghost cells of the members whose steps do not write them.
static int ensemble_launch_part(Ensemble* e, int part, hipStream_t s) {
  int rc = 0;
  if (part == 5) return ensemble_launch_bc_end(e, s);
  if (part == 0) {
    if (!e->h_clock.empty()) {
      const int nc = (int)e->h_clock.size();
      hipLaunchKernelGGL(frs_clock, dim3((nc + 63) / 64), dim3(64), 0, s, e->d_clock, nc, e->cur_args, e->cur_active);
      rc = check(hipGetLastError(), "ensemble boundary-data clock launch");
    }
    if (!rc) rc = ensemble_launch_substep(e, 0, s);
    if (!rc) rc = ensemble_launch_bc(e, 1, s);
  } else if (part == 1) {
    rc = ensemble_launch_substep(e, 1, s);
  } else if (part == 4) {
    rc = ensemble_launch_noise(e, s);
  } else {
    if (part == 2) rc = ensemble_launch_noise(e, s);
    if (!rc) rc = ensemble_launch_bc(e, 0, s);
    if (!rc) rc = ensemble_launch_drift(e, s);
  }
  return rc;
}

static int ensemble_launch_step(Ensemble* e, hipStream_t s) {
  int rc = 0;
  for (int part = 0; part < 3 && !rc; ++part) rc = ensemble_launch_part(e, part, s);
  return rc;
}

static void ensemble_free(Ensemble* e) {
  if (e->exec) (void)hipGraphExecDestroy(e->exec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  if (e->d_args) (void)hipFree(e->d_args);
  if (e->d_work) (void)hipFree(e->d_work);
  if (e->d_noise) (void)hipFree(e->d_noise);
  if (e->d_tiles) (void)hipFree(e->d_tiles);
  if (e->d_draw) (void)hipFree(e->d_draw);
  if (e->d_bc) (void)hipFree(e->d_bc);
  if (e->d_bc_end) (void)hipFree(e->d_bc_end);
  if (e->d_bc_end_sel) (void)hipFree(e->d_bc_end_sel);
  if (e->d_times) (void)hipFree(e->d_times);
  if (e->d_drift) (void)hipFree(e->d_drift);
  if (e->d_work_sel) (void)hipFree(e->d_work_sel);
  if (e->d_tiles_sel) (void)hipFree(e->d_tiles_sel);
  if (e->d_bc_sel) (void)hipFree(e->d_bc_sel);
  if (e->d_drift_sel) (void)hipFree(e->d_drift_sel);
  if (e->d_active) (void)hipFree(e->d_active);
  if (e->d_arena) (void)hipFree(e->d_arena);
  if (e->d_clock) (void)hipFree(e->d_clock);
  delete e;
}

}  // namespace gpo

using namespace gpo;

extern "C" {

int gpo_ensemble_create(const gpo_member* members, int32_t n, int32_t chunk_rows, int32_t waves_per_simd,
                        int32_t exact, gpo_ensemble* out, void* stream) {
  if (!members || n < 1 || n > 65535 || !out) return bad("gpo_ensemble_create: bad arguments (1 <= n <= 65535)");
  if (chunk_rows < 0 || chunk_rows > 4096 || waves_per_simd < 0 || waves_per_simd > 32)
    return bad("gpo_ensemble_create: chunk_rows or waves_per_simd out of range");
  hipStream_t s = (hipStream_t)stream;
  Ensemble* e = new (std::nothrow) Ensemble();
  if (!e) return bad("gpo_ensemble_create: out of host memory");
  e->n = n;
  e->exact = exact ? 1 : 0;
  e->has_land = members[0].params.has_land;
  e->f_array = members[0].params.f_is_array;
  e->angle0 = members[0].params.angle_is_zero;
  e->has_noise = members[0].noise ? 1 : 0;

  std::vector<StepArgs>& args = e->h_args;
  args.resize(2 * (size_t)n);
  std::vector<int2> work;
  std::vector<int2> tiles;
  std::vector<NoiseMember> noise;
  std::vector<BcMember> bcs;
  std::vector<DriftMember> drifts;
  long long max_bc_cells = 0, max_bc_end_cells = 0;
  int max_drift_rows = 0;
  std::vector<uint32_t> draws((size_t)n);
  for (int m = 0; m < n; ++m) {
    const gpo_member& M = members[m];
    gpo_cdklm16_params p = M.params;
    if (p.rk_order != 2) {
      ensemble_free(e);
      return bad("gpo_ensemble_create: members must use RK2");
    }
    if (p.has_land != e->has_land || p.f_is_array != e->f_array || p.angle_is_zero != e->angle0 ||
        (M.noise ? 1 : 0) != e->has_noise) {
      ensemble_free(e);
      return bad("gpo_ensemble_create: members differ in land, coriolis texture, angle or model error");
    }
    int frs = 0;
    for (int side = 0; side < 4; ++side) {
      const int code = (M.boundary_conditions >> (8 * side)) & 0xFF;
      if (code < 1 || code > 3) {
        ensemble_free(e);
        return bad("gpo_ensemble_create: only wall (1), periodic (2) and flow-relaxation (3) boundaries");
      }
      frs |= code == 3;
    }
    if (p.r > 0.0f) e->any_stage = 1;
    if (frs && !M.bc_each_step) {
      ensemble_free(e);
      return bad("gpo_ensemble_create: flow-relaxation members need bc_each_step");
    }
    if (chunk_rows > 0) p.chunk_rows = chunk_rows;
    int rc = make_step_args(&p, M.dt, 0, M.h0_eta, M.h0_hu, M.h0_hv, nullptr, nullptr, nullptr, M.h1_eta, M.h1_hu,
                            M.h1_hv, M.pitch, M.Hi, M.Hi_pitch, M.f_cells, nullptr, M.boundary_conditions,
                            &args[2 * m]);
    if (!rc)
      rc = make_step_args(&p, M.dt, 1, M.h1_eta, M.h1_hu, M.h1_hv, M.h0_eta, M.h0_hu, M.h0_hv, M.h0_eta, M.h0_hu,
                          M.h0_hv, M.pitch, M.Hi, M.Hi_pitch, M.f_cells, nullptr, M.boundary_conditions,
                          &args[2 * m + 1]);
    if (!rc && !(fast_arith_ok(args[2 * m]) && fast_arith_ok(args[2 * m + 1]))) e->all_ok = 0;
    if (rc) {
      ensemble_free(e);
      return rc;
    }
    const StepArgs& a = args[2 * m];
    if (a.n_strips > 0xFFFF || a.n_chunks > 0x7FFF) {
      ensemble_free(e);
      return bad("gpo_ensemble_create: member too large for the work table");
    }
    draws[m] = M.draw;
    if (M.bc_each_step) {
      BcMember b{};
      cells::BcTarget& t = b.h0;
      t.nx = p.nx;
      t.ny = p.ny;
      t.bn = (M.boundary_conditions >> 24) & 0xFF;
      t.bs = (M.boundary_conditions >> 16) & 0xFF;
      t.be = (M.boundary_conditions >> 8) & 0xFF;
      t.bw = M.boundary_conditions & 0xFF;
      t.sn = M.sponge[0];
      t.se = M.sponge[1];
      t.ss = M.sponge[2];
      t.sw = M.sponge[3];
      cells::bc_target_extent(&t);
      if ((t.bn == 3 || t.bs == 3) && t.bn != 2 && (!M.frs_ns_cur || !M.frs_ns_next || M.frs_ns_nx < 1)) {
        ensemble_free(e);
        return bad("gpo_ensemble_create: flow-relaxation north/south data missing");
      }
      if ((t.be == 3 || t.bw == 3) && t.be != 2 && (!M.frs_ew_cur || !M.frs_ew_next || M.frs_ew_ny < 1)) {
        ensemble_free(e);
        return bad("gpo_ensemble_create: flow-relaxation east/west data missing");
      }
      t.pitch = M.pitch;
      t.ns_cur = M.frs_ns_cur, t.ns_next = M.frs_ns_next, t.ns_nx = M.frs_ns_nx;
      t.ew_cur = M.frs_ew_cur, t.ew_next = M.frs_ew_next, t.ew_ny = M.frs_ew_ny;
      t.frs_t = M.frs_t;
      b.h1 = t;
      t.e = M.h0_eta, t.u = M.h0_hu, t.v = M.h0_hv;
      b.h1.e = M.h1_eta, b.h1.u = M.h1_hu, b.h1.v = M.h1_hv;
      b.frs = frs;
      b.clock = nullptr;
      if (frs && M.frs_n_times >= 1) {
        if (!M.frs_times || ((t.bn == 3 || t.bs == 3) && t.bn != 2 && !M.frs_ns_levels) ||
            ((t.be == 3 || t.bw == 3) && t.be != 2 && !M.frs_ew_levels)) {
          ensemble_free(e);
          return bad("gpo_ensemble_create: flow-relaxation time levels missing");
        }
        FrsClock c{};
        c.t = M.t;
        c.times = M.frs_times;
        c.n_times = M.frs_n_times;
        c.ns_levels = M.frs_ns_levels;
        c.ew_levels = M.frs_ew_levels;
        c.ns_stride = 2ll * M.frs_ns_nx * 3;
        c.ew_stride = 2ll * M.frs_ew_ny * 3;
        c.member = m;
        // the device address is known once the array exists: store the index + 1 for now
        b.clock = (const FrsClock*)(uintptr_t)(e->h_clock.size() + 1);
        e->h_clock.push_back(c);
      }
      bcs.push_back(b);
      e->bc_member.push_back(m);
      max_bc_cells = std::max(max_bc_cells, cells::bc_target_cells(t));
      e->any_frs |= frs;
    }
    if (!M.bc_each_step && (((M.boundary_conditions >> 24) & 0xFF) == 2 || ((M.boundary_conditions >> 8) & 0xFF) == 2)) {
      // periodic ghost cells, written once after a run of steps (part 5)
      BcMember b{};
      cells::BcTarget& t = b.h0;
      t.nx = p.nx;
      t.ny = p.ny;
      t.bn = (M.boundary_conditions >> 24) & 0xFF;
      t.bs = (M.boundary_conditions >> 16) & 0xFF;
      t.be = (M.boundary_conditions >> 8) & 0xFF;
      t.bw = M.boundary_conditions & 0xFF;
      cells::bc_target_extent(&t);
      t.pitch = M.pitch;
      t.e = M.h0_eta, t.u = M.h0_hu, t.v = M.h0_hv;
      b.h1 = t;
      e->h_bc_end.push_back(b);
      e->bc_end_member.push_back(m);
      max_bc_end_cells = std::max(max_bc_end_cells, cells::bc_target_cells(t));
    }
    if (M.drifters) {
      if (M.n_drifters < 0 || !M.positions || !M.Bm || M.pos_pitch < 8) {
        ensemble_free(e);
        return bad("gpo_ensemble_create: drifter buffers missing");
      }
      DriftMember d{};
      d.nx = p.nx;
      d.ny = p.ny;
      d.per_ns = ((M.boundary_conditions >> 24) & 0xFF) == 2 && ((M.boundary_conditions >> 16) & 0xFF) == 2;
      d.per_ew = ((M.boundary_conditions >> 8) & 0xFF) == 2 && (M.boundary_conditions & 0xFF) == 2;
      d.rows = M.n_drifters + 1;
      d.pos_pitch = M.pos_pitch;
      d.pitch = M.pitch;
      d.bm_pitch = M.Bm_pitch;
      d.arg = 2 * m;
      d.dx = p.dx;
      d.dy = p.dy;
      d.sensitivity = M.sensitivity;
      d.eta = M.h0_eta, d.hu = M.h0_hu, d.hv = M.h0_hv;
      d.Bm = M.Bm;
      d.pos = M.positions;
      drifts.push_back(d);
      e->drift_member.push_back(m);
      max_drift_rows = std::max(max_drift_rows, d.rows);
    }
    for (int sy = 0; sy < a.n_chunks; ++sy)
      for (int sx = 0; sx < a.n_strips; ++sx) work.push_back(make_int2(m, sx | (sy << 16)));
    if (e->has_noise) {
      NoiseMember nm{};
      nm.nx = p.nx;
      nm.ny = p.ny;
      nm.per_ns = ((M.boundary_conditions >> 24) & 0xFF) == 2 && ((M.boundary_conditions >> 16) & 0xFF) == 2;
      nm.per_ew = ((M.boundary_conditions >> 8) & 0xFF) == 2 && (M.boundary_conditions & 0xFF) == 2;
      nm.xi_rows = M.xi_rows;
      nm.xi_cols = M.xi_cols;
      nm.xi_pitch = M.xi_pitch;
      nm.c_pitch = M.coarse_pitch;
      nm.pitch = M.pitch;
      nm.hi_pitch = M.Hi_pitch;
      nm.f_pitch = M.pitch;
      nm.k0 = (uint32_t)(M.seed & 0xffffffffu);
      nm.k1 = (uint32_t)(M.seed >> 32);
      nm.stream_id = M.stream_id;
      for (int i = 0; i < 25; ++i) nm.W[i] = M.soar_weights[i];
      nm.scale = M.perturbation_scale;
      nm.dx = p.dx;
      nm.dy = p.dy;
      nm.g = M.g;
      nm.f = M.f;
      nm.beta = M.beta;
      nm.north_x = p.north_x;
      nm.north_y = p.north_y;
      nm.land = p.land_value;
      nm.xi = M.xi;
      nm.coarse = M.coarse;
      nm.eta = M.h0_eta;
      nm.hu = M.h0_hu;
      nm.hv = M.h0_hv;
      nm.Hi = M.Hi;
      nm.f_cells = M.f_cells;
      nm.coarse_grid = M.coarse_nx > 0 ? 1 : 0;
      const int cnx = nm.coarse_grid ? M.coarse_nx : p.nx, cny = nm.coarse_grid ? M.coarse_ny : p.ny;
      nm.geo = BicubicGeo{p.nx, p.ny, cnx, cny, M.coarse_offset_i, M.coarse_offset_j,
                          p.dx, p.dy, M.coarse_dx, M.coarse_dy};
      if (m == 0) e->coarse_noise = nm.coarse_grid;
      if (nm.coarse_grid != e->coarse_noise) {
        ensemble_free(e);
        return bad("gpo_ensemble_create: members differ in whether their model error lives on a coarser grid");
      }
      if (nm.coarse_grid &&
          (cny < 1 || !(M.coarse_dx > 0.0f) || !(M.coarse_dy > 0.0f) || M.coarse_offset_i < -cnx ||
           M.coarse_offset_i > cnx || M.coarse_offset_j < -cny || M.coarse_offset_j > cny)) {
        ensemble_free(e);
        return bad("gpo_ensemble_create: bad coarse model-error grid");
      }
      if (!nm.xi || !nm.coarse || nm.xi_rows != (nm.per_ns ? cny : cny + 8) ||
          nm.xi_cols != (nm.per_ew ? cnx : cnx + 8) || nm.xi_pitch < 4 * nm.xi_cols ||
          nm.c_pitch < 4 * (cnx + 4)) {
        ensemble_free(e);
        return bad("gpo_ensemble_create: model error buffers missing or mis-sized");
      }
      if (nm.coarse_grid) {
        const long long quads = ((long long)nm.xi_rows * nm.xi_cols + 3) / 4;
        const long long cells = (long long)(cnx + 4) * (cny + 4);
        if (quads > (1ll << 30) || cells > (1ll << 30)) {
          ensemble_free(e);
          return bad("gpo_ensemble_create: coarse model-error grid too large");
        }
        e->max_xi_quads = std::max(e->max_xi_quads, (int)quads);
        e->max_coarse_cells = std::max(e->max_coarse_cells, (int)cells);
      }
      noise.push_back(nm);
      const int tx_n = (p.nx + kTileW - 1) / kTileW, ty_n = (p.ny + kTileH - 1) / kTileH;
      if (tx_n > 0xFFFF || ty_n > 0x7FFF) {
        ensemble_free(e);
        return bad("gpo_ensemble_create: member too large for the model-error tile table");
      }
      for (int ty = 0; ty < ty_n; ++ty)
        for (int tx = 0; tx < tx_n; ++tx) tiles.push_back(make_int2((int)noise.size() - 1, tx | (ty << 16)));
    }
  }
  e->n_work = (int)work.size();
  int cus = 256;
  int dev = 0;
  if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  const int wps = waves_per_simd > 0 ? waves_per_simd : 4;
  const long long target_waves = (long long)cus * 4 * wps;
  const long long waves = target_waves < e->n_work ? target_waves : e->n_work;
  e->n_blocks = (int)((waves + kWavesPerBlock - 1) / kWavesPerBlock);

  int rc = check(hipMalloc(&e->d_args, sizeof(StepArgs) * args.size()), "gpo_ensemble_create hipMalloc");
  if (!rc) rc = check(hipMalloc(&e->d_work, sizeof(int2) * work.size()), "gpo_ensemble_create hipMalloc");
  if (!rc) rc = check(hipMalloc(&e->d_draw, sizeof(uint32_t) * n), "gpo_ensemble_create hipMalloc");
  if (!rc && e->has_noise)
    rc = check(hipMalloc(&e->d_noise, sizeof(NoiseMember) * noise.size()), "gpo_ensemble_create hipMalloc");
  if (!rc && e->has_noise)
    rc = check(hipMalloc(&e->d_tiles, sizeof(int2) * tiles.size()), "gpo_ensemble_create hipMalloc");
  if (!rc)
    rc = check(hipMemcpyAsync(e->d_args, args.data(), sizeof(StepArgs) * args.size(), hipMemcpyHostToDevice, s),
               "gpo_ensemble_create copy");
  if (!rc)
    rc = check(hipMemcpyAsync(e->d_work, work.data(), sizeof(int2) * work.size(), hipMemcpyHostToDevice, s),
               "gpo_ensemble_create copy");
  if (!rc)
    rc = check(hipMemcpyAsync(e->d_draw, draws.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, s),
               "gpo_ensemble_create copy");
  if (!rc && e->has_noise)
    rc = check(hipMemcpyAsync(e->d_noise, noise.data(), sizeof(NoiseMember) * noise.size(), hipMemcpyHostToDevice, s),
               "gpo_ensemble_create copy");
  if (!rc && e->has_noise)
    rc = check(hipMemcpyAsync(e->d_tiles, tiles.data(), sizeof(int2) * tiles.size(), hipMemcpyHostToDevice, s),
               "gpo_ensemble_create copy");
  e->n_tiles = (int)tiles.size();
  e->n_bc = (int)bcs.size();
  e->bc_blocks = (int)((max_bc_cells + 255) / 256);
  e->n_drift = (int)drifts.size();
  e->drift_blocks = (max_drift_rows + 63) / 64;
  if (!rc && !e->h_clock.empty()) {
    rc = check(hipMalloc(&e->d_clock, sizeof(FrsClock) * e->h_clock.size()), "gpo_ensemble_create hipMalloc");
    if (!rc)
      rc = check(hipMemcpyAsync(e->d_clock, e->h_clock.data(), sizeof(FrsClock) * e->h_clock.size(),
                                hipMemcpyHostToDevice, s),
                 "gpo_ensemble_create copy");
    // the members' clock indices (+1) become device addresses
    for (BcMember& b : bcs)
      if (b.clock) b.clock = e->d_clock + ((uintptr_t)b.clock - 1);
    // levels and weight at the members' current times, so that every clock
    // holds valid data pointers from the start
    if (!rc) {
      const int nc = (int)e->h_clock.size();
      hipLaunchKernelGGL(clock_set_times, dim3((nc + 63) / 64), dim3(64), 0, s, e->d_clock, nc,
                         (const double*)nullptr);
      rc = check(hipGetLastError(), "gpo_ensemble_create clock launch");
    }
    if (!rc) rc = check(hipMalloc(&e->d_times, sizeof(double) * n), "gpo_ensemble_create hipMalloc");
  }
  e->n_bc_end = (int)e->h_bc_end.size();
  e->bc_end_blocks = (int)((max_bc_end_cells + 255) / 256);
  if (!rc && e->n_bc_end > 0) {
    rc = check(hipMalloc(&e->d_bc_end, sizeof(BcMember) * e->h_bc_end.size()), "gpo_ensemble_create hipMalloc");
    if (!rc)
      rc = check(hipMemcpyAsync(e->d_bc_end, e->h_bc_end.data(), sizeof(BcMember) * e->h_bc_end.size(),
                                hipMemcpyHostToDevice, s),
                 "gpo_ensemble_create copy");
    if (!rc)
      rc = check(hipMalloc(&e->d_bc_end_sel, sizeof(BcMember) * e->h_bc_end.size()), "gpo_ensemble_create hipMalloc");
  }
  if (!rc && e->n_bc > 0 && e->bc_blocks > 0) {
    rc = check(hipMalloc(&e->d_bc, sizeof(BcMember) * bcs.size()), "gpo_ensemble_create hipMalloc");
    if (!rc)
      rc = check(hipMemcpyAsync(e->d_bc, bcs.data(), sizeof(BcMember) * bcs.size(), hipMemcpyHostToDevice, s),
                 "gpo_ensemble_create copy");
  } else {
    e->n_bc = 0;  // nothing to write (walls only)
  }
  if (!rc && e->n_drift > 0) {
    rc = check(hipMalloc(&e->d_drift, sizeof(DriftMember) * drifts.size()), "gpo_ensemble_create hipMalloc");
    if (!rc)
      rc = check(hipMemcpyAsync(e->d_drift, drifts.data(), sizeof(DriftMember) * drifts.size(),
                                hipMemcpyHostToDevice, s),
                 "gpo_ensemble_create copy");
  }
  // buffers for the tables of a selection of members (at most the full sizes)
  if (!rc) rc = check(hipMalloc(&e->d_work_sel, sizeof(int2) * work.size()), "gpo_ensemble_create hipMalloc");
  if (!rc) rc = check(hipMalloc(&e->d_active, (size_t)n), "gpo_ensemble_create hipMalloc");
  if (!rc && !tiles.empty())
    rc = check(hipMalloc(&e->d_tiles_sel, sizeof(int2) * tiles.size()), "gpo_ensemble_create hipMalloc");
  if (!rc && e->n_bc > 0)
    rc = check(hipMalloc(&e->d_bc_sel, sizeof(BcMember) * bcs.size()), "gpo_ensemble_create hipMalloc");
  if (!rc && e->n_drift > 0)
    rc = check(hipMalloc(&e->d_drift_sel, sizeof(DriftMember) * drifts.size()), "gpo_ensemble_create hipMalloc");
  if (!rc) rc = check(hipStreamSynchronize(s), "gpo_ensemble_create sync");
  if (rc) {
    ensemble_free(e);
    return rc;
  }
  e->h_work = std::move(work);
  e->h_tiles = std::move(tiles);
  e->h_noise = std::move(noise);
  e->h_bc = std::move(bcs);
  e->h_drift = std::move(drifts);
  if (e->n_bc == 0) e->h_bc.clear(), e->bc_member.clear();
  use_full_tables(e);
  *out = (gpo_ensemble)e;
  return 0;
}

int gpo_ensemble_info(gpo_ensemble h, int32_t* n_work, int32_t* n_blocks) {
  Ensemble* e = (Ensemble*)h;
  if (!e) return bad("gpo_ensemble_info: NULL ensemble");
  if (n_work) *n_work = e->n_work;
  if (n_blocks) *n_blocks = e->n_blocks;
  return 0;
}

int gpo_ensemble_step(gpo_ensemble h, int32_t n_steps, void* stream) {
  Ensemble* e = (Ensemble*)h;
  if (!e || n_steps < 0) return bad("gpo_ensemble_step: bad arguments");
  for (int i = 0; i < n_steps; ++i) {
    const int rc = ensemble_launch_step(e, (hipStream_t)stream);
    if (rc) return rc;
  }
  return 0;
}

int gpo_ensemble_part(gpo_ensemble h, int32_t part, void* stream) {
  Ensemble* e = (Ensemble*)h;
  if (!e || part < 0 || part > 5) return bad("gpo_ensemble_part: bad arguments");
  return ensemble_launch_part(e, part, (hipStream_t)stream);
}

int gpo_ensemble_capture(gpo_ensemble h, int32_t steps_per_graph, void* stream) {
  Ensemble* e = (Ensemble*)h;
  if (!e || steps_per_graph < 1) return bad("gpo_ensemble_capture: bad arguments");
  if (e->selected) return bad("gpo_ensemble_capture: a selection of members is active (select all first)");
  hipStream_t s = (hipStream_t)stream;
  if (e->exec) {
    (void)hipGraphExecDestroy(e->exec);
    e->exec = nullptr;
  }
  if (e->graph) {
    (void)hipGraphDestroy(e->graph);
    e->graph = nullptr;
  }
  int rc = check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "gpo_ensemble_capture begin");
  if (rc) return rc;
  int lrc = 0;
  for (int i = 0; i < steps_per_graph && !lrc; ++i) lrc = ensemble_launch_step(e, s);
  rc = check(hipStreamEndCapture(s, &e->graph), "gpo_ensemble_capture end");
  if (lrc) return lrc;
  if (rc) return rc;
  rc = check(hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0), "gpo_ensemble_capture instantiate");
  if (rc) return rc;
  e->graph_steps = steps_per_graph;
  return 0;
}

int gpo_ensemble_replay(gpo_ensemble h, int32_t n_replays, void* stream) {
  Ensemble* e = (Ensemble*)h;
  if (!e || !e->exec || n_replays < 0) return bad("gpo_ensemble_replay: no captured graph");
  for (int i = 0; i < n_replays; ++i) {
    const int rc = check(hipGraphLaunch(e->exec, (hipStream_t)stream), "gpo_ensemble_replay launch");
    if (rc) return rc;
  }
  return 0;
}

int gpo_ensemble_draws(gpo_ensemble h, uint32_t* draws, void* stream) {
  Ensemble* e = (Ensemble*)h;
  if (!e || !draws) return bad("gpo_ensemble_draws: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  int rc = check(hipMemcpyAsync(draws, e->d_draw, sizeof(uint32_t) * e->n, hipMemcpyDeviceToHost, s),
                 "gpo_ensemble_draws");
  if (!rc) rc = check(hipStreamSynchronize(s), "gpo_ensemble_draws sync");
  return rc;
}

int gpo_ensemble_update(gpo_ensemble h, const float* dts, const uint32_t* draws, void* stream) {
  Ensemble* e = (Ensemble*)h;
  if (!e) return bad("gpo_ensemble_update: NULL ensemble");
  hipStream_t s = (hipStream_t)stream;
  int rc = 0;
  if (dts) {
    for (int m = 0; m < e->n; ++m)
      if (!(dts[m] > 0.0f)) return bad("gpo_ensemble_update: time steps must be positive");
    for (int m = 0; m < e->n; ++m) {
      e->h_args[2 * m].dt = dts[m];
      e->h_args[2 * m + 1].dt = dts[m];
    }
    rc = check(hipMemcpyAsync(e->d_args, e->h_args.data(), sizeof(StepArgs) * e->h_args.size(),
                              hipMemcpyHostToDevice, s),
               "gpo_ensemble_update copy");
  }
  if (!rc && draws)
    rc = check(hipMemcpyAsync(e->d_draw, draws, sizeof(uint32_t) * e->n, hipMemcpyHostToDevice, s),
               "gpo_ensemble_update copy");
  // The host sources may change as soon as this returns.
  if (!rc) rc = check(hipStreamSynchronize(s), "gpo_ensemble_update sync");
  return rc;
}

int gpo_ensemble_set_times(gpo_ensemble h, const double* t, void* stream) {
  Ensemble* e = (Ensemble*)h;
  if (!e || !t) return bad("gpo_ensemble_set_times: NULL argument");
  if (e->h_clock.empty()) return 0;
  hipStream_t s = (hipStream_t)stream;
  // Only the times change: the kernel derives each clock's levels and weight
  // at its new time, so ghost-cell passes before the next step read valid data.
  int rc = check(hipMemcpyAsync(e->d_times, t, sizeof(double) * e->n, hipMemcpyHostToDevice, s),
                 "gpo_ensemble_set_times copy");
  if (!rc) {
    const int nc = (int)e->h_clock.size();
    hipLaunchKernelGGL(clock_set_times, dim3((nc + 63) / 64), dim3(64), 0, s, e->d_clock, nc,
                       (const double*)e->d_times);
    rc = check(hipGetLastError(), "gpo_ensemble_set_times launch");
  }
  // the host array may change as soon as this returns
  if (!rc) rc = check(hipStreamSynchronize(s), "gpo_ensemble_set_times sync");
  return rc;
}

int gpo_ensemble_select(gpo_ensemble h, const uint8_t* active, const float* noise_scale, void* stream) {
  Ensemble* e = (Ensemble*)h;
  if (!e) return bad("gpo_ensemble_select: NULL ensemble");
  hipStream_t s = (hipStream_t)stream;
  int rc = 0;
  if (noise_scale && e->has_noise) {
    for (int m = 0; m < e->n; ++m) e->h_noise[m].scale = noise_scale[m];
    rc = check(hipMemcpyAsync(e->d_noise, e->h_noise.data(), sizeof(NoiseMember) * e->h_noise.size(),
                              hipMemcpyHostToDevice, s),
               "gpo_ensemble_select copy");
  }
  if (!rc && !active) use_full_tables(e);
  if (!rc && active) {
    std::vector<int2> w, t;
    std::vector<BcMember> b;
    std::vector<DriftMember> d;
    for (const int2& x : e->h_work)
      if (active[x.x]) w.push_back(x);
    for (const int2& x : e->h_tiles)
      if (active[x.x]) t.push_back(x);
    std::vector<BcMember> be;
    for (size_t k = 0; k < e->h_bc.size(); ++k)
      if (active[e->bc_member[k]]) b.push_back(e->h_bc[k]);
    for (size_t k = 0; k < e->h_bc_end.size(); ++k)
      if (active[e->bc_end_member[k]]) be.push_back(e->h_bc_end[k]);
    for (size_t k = 0; k < e->h_drift.size(); ++k)
      if (active[e->drift_member[k]]) d.push_back(e->h_drift[k]);
    auto put = [&](void* dst, const void* src, size_t bytes) {
      if (!rc && bytes) rc = check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s), "gpo_ensemble_select copy");
    };
    put(e->d_work_sel, w.data(), sizeof(int2) * w.size());
    put(e->d_tiles_sel, t.data(), sizeof(int2) * t.size());
    put(e->d_bc_sel, b.data(), sizeof(BcMember) * b.size());
    put(e->d_bc_end_sel, be.data(), sizeof(BcMember) * be.size());
    put(e->d_drift_sel, d.data(), sizeof(DriftMember) * d.size());
    put(e->d_active, active, (size_t)e->n);
    // the host vectors are released when this returns: finish the copies first
    if (!rc) rc = check(hipStreamSynchronize(s), "gpo_ensemble_select sync");
    if (!rc) {
      e->selected = 1;
      e->cur_active = e->d_active;
      e->cur_work = e->d_work_sel, e->cur_n_work = (int)w.size();
      e->cur_tiles = e->d_tiles_sel, e->cur_n_tiles = (int)t.size();
      e->cur_bc = e->d_bc_sel, e->cur_n_bc = (int)b.size();
      e->cur_bc_end = e->d_bc_end_sel, e->cur_n_bc_end = (int)be.size();
      e->cur_drift = e->d_drift_sel, e->cur_n_drift = (int)d.size();
    }
  }
  if (!rc) rc = check(hipStreamSynchronize(s), "gpo_ensemble_select sync");
  return rc;
}

int gpo_ensemble_prepare_sets(gpo_ensemble h, int32_t n_sets, const uint8_t* active, const float* dts,
                              const float* noise_scale, void* stream) {
  Ensemble* e = (Ensemble*)h;
  if (!e || n_sets < 0 || (n_sets > 0 && !active)) return bad("gpo_ensemble_prepare_sets: bad arguments");
  if (dts)
    for (size_t k = 0; k < (size_t)n_sets * e->n; ++k)
      if (active[k] && !(dts[k] > 0.0f)) return bad("gpo_ensemble_prepare_sets: time steps must be positive");
  hipStream_t s = (hipStream_t)stream;
  std::vector<char>& A = e->h_arena;
  A.clear();
  e->sets.assign((size_t)n_sets, Ensemble::Set{});
  auto put = [&](const void* src, size_t bytes) {
    const size_t off = (A.size() + 255) & ~(size_t)255;
    A.resize(off + bytes);
    if (bytes) memcpy(A.data() + off, src, bytes);
    return off;
  };
  for (int k = 0; k < n_sets; ++k) {
    const uint8_t* act = active + (size_t)k * e->n;
    std::vector<int2> w, t;
    std::vector<BcMember> b, be;
    std::vector<DriftMember> d;
    for (const int2& x : e->h_work)
      if (act[x.x]) w.push_back(x);
    for (const int2& x : e->h_tiles)
      if (act[x.x]) t.push_back(x);
    for (size_t i = 0; i < e->h_bc.size(); ++i)
      if (act[e->bc_member[i]]) b.push_back(e->h_bc[i]);
    for (size_t i = 0; i < e->h_bc_end.size(); ++i)
      if (act[e->bc_end_member[i]]) be.push_back(e->h_bc_end[i]);
    for (size_t i = 0; i < e->h_drift.size(); ++i)
      if (act[e->drift_member[i]]) d.push_back(e->h_drift[i]);
    std::vector<StepArgs> args = e->h_args;
    if (dts)
      for (int m = 0; m < e->n; ++m)
        if (act[m]) args[2 * m].dt = args[2 * m + 1].dt = dts[(size_t)k * e->n + m];
    Ensemble::Set& S = e->sets[k];
    S.work = put(w.data(), sizeof(int2) * w.size()), S.n_work = (int)w.size();
    S.tiles = put(t.data(), sizeof(int2) * t.size()), S.n_tiles = (int)t.size();
    S.bc = put(b.data(), sizeof(BcMember) * b.size()), S.n_bc = (int)b.size();
    S.bc_end = put(be.data(), sizeof(BcMember) * be.size()), S.n_bc_end = (int)be.size();
    S.drift = put(d.data(), sizeof(DriftMember) * d.size()), S.n_drift = (int)d.size();
    S.args = put(args.data(), sizeof(StepArgs) * args.size());
    S.active = put(act, (size_t)e->n);
  }
  int rc = 0;
  // the arena may still be read by launches of earlier sets: finish them first
  if (A.size() > e->arena_bytes) {
    rc = check(hipStreamSynchronize(s), "gpo_ensemble_prepare_sets sync");
    if (!rc && e->d_arena) (void)hipFree(e->d_arena), e->d_arena = nullptr, e->arena_bytes = 0;
    if (!rc) rc = check(hipMalloc(&e->d_arena, A.size()), "gpo_ensemble_prepare_sets hipMalloc");
    if (!rc) e->arena_bytes = A.size();
  }
  if (!rc && !A.empty())
    rc = check(hipMemcpyAsync(e->d_arena, A.data(), A.size(), hipMemcpyHostToDevice, s),
               "gpo_ensemble_prepare_sets copy");
  if (!rc && noise_scale && e->has_noise) {
    for (int m = 0; m < e->n; ++m) e->h_noise[m].scale = noise_scale[m];
    rc = check(hipMemcpyAsync(e->d_noise, e->h_noise.data(), sizeof(NoiseMember) * e->h_noise.size(),
                              hipMemcpyHostToDevice, s),
               "gpo_ensemble_prepare_sets copy");
  }
  // the host sources are reused by the next call: let the copies finish
  if (!rc) rc = check(hipStreamSynchronize(s), "gpo_ensemble_prepare_sets sync");
  return rc;
}

int gpo_ensemble_run_set(gpo_ensemble h, int32_t set, int32_t n_steps, uint32_t parts, void* stream) {
  Ensemble* e = (Ensemble*)h;
  if (!e || set < 0 || set >= (int)e->sets.size() || n_steps < 0 || (parts & ~0x3Fu))
    return bad("gpo_ensemble_run_set: bad arguments");
  const Ensemble::Set& S = e->sets[set];
  e->selected = 1;
  e->cur_work = (const int2*)(e->d_arena + S.work), e->cur_n_work = S.n_work;
  e->cur_tiles = (const int2*)(e->d_arena + S.tiles), e->cur_n_tiles = S.n_tiles;
  e->cur_bc = (const BcMember*)(e->d_arena + S.bc), e->cur_n_bc = S.n_bc;
  e->cur_bc_end = (const BcMember*)(e->d_arena + S.bc_end), e->cur_n_bc_end = S.n_bc_end;
  e->cur_drift = (const DriftMember*)(e->d_arena + S.drift), e->cur_n_drift = S.n_drift;
  e->cur_args = (const StepArgs*)(e->d_arena + S.args);
  e->cur_active = (const uint8_t*)(e->d_arena + S.active);
  int rc = 0;
  for (int i = 0; i < n_steps && !rc; ++i)
    for (int part = 0; part < 6 && !rc; ++part)
      if (parts & (1u << part)) rc = ensemble_launch_part(e, part, (hipStream_t)stream);
  use_full_tables(e);
  return rc;
}

int gpo_ensemble_destroy(gpo_ensemble h) {
  if (h) ensemble_free((Ensemble*)h);
  return 0;
}

}  // extern "C"
