// Batched ensemble of CDKLM16 members of uneven size: one persistent launch
// per RK2 substep over all members, batched SOAR model error, and hipGraph
// capture of whole steps.  Built with -ffp-contract=off (noise kernels must
// match the per-member path bit for bit).
//
// Replaces the reference's sequential member loop
// (ensembles/OceanModelEnsemble.py:138-150), where each member is a separate
// CDKLM16 object with its own JIT module, stream and blocking dt read-back.

#include <cstdio>
#include <new>
#include <vector>

#include "cdklm_common.h"
#include "noise_cells.h"
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
};

__global__ void __launch_bounds__(256) noise_fill(const NoiseMember* __restrict__ members,
                                                  const uint32_t* __restrict__ draw) {
  const NoiseMember& m = members[blockIdx.y];
  const uint32_t d = draw[blockIdx.y];
  const int n = m.xi_rows * m.xi_cols;
  for (int idx = blockIdx.x * 256 + threadIdx.x; idx < n; idx += gridDim.x * 256) {
    const int j = idx / m.xi_cols, i = idx - j * m.xi_cols;
    mrow(m.xi, m.xi_pitch, j)[i] = normal_at((uint32_t)idx, d, m.stream_id, m.k0, m.k1);
  }
}

__global__ void __launch_bounds__(256) noise_soar(const NoiseMember* __restrict__ members) {
  const NoiseMember& m = members[blockIdx.y];
  const int w = m.nx + 4, n = (m.ny + 4) * w;
  for (int idx = blockIdx.x * 256 + threadIdx.x; idx < n; idx += gridDim.x * 256) {
    const int tj = idx / w, ti = idx - tj * w;
    const float acc = soar_cell(m.nx, m.ny, m.W, m.per_ns, m.per_ew, m.xi, m.xi_pitch, tj, ti);
    mrow(m.coarse, m.c_pitch, tj)[ti] = m.scale * acc;
  }
}

__global__ void __launch_bounds__(256) noise_balance(const NoiseMember* __restrict__ members, uint32_t* draw) {
  const NoiseMember& m = members[blockIdx.y];
  const int n = m.nx * m.ny;
  for (int idx = blockIdx.x * 256 + threadIdx.x; idx < n; idx += gridDim.x * 256) {
    const int j = idx / m.nx, i = idx - j * m.nx;
    balance_cell(m.dx, m.dy, m.g, m.f, m.beta, m.north_x, m.north_y, m.f_cells, m.f_pitch, m.coarse, m.c_pitch,
                 m.eta, m.hu, m.hv, m.pitch, m.Hi, m.hi_pitch, m.land, j + 2, i + 2);
  }
  // The fill kernel of this step has finished reading the member's draw index.
  if (blockIdx.x == 0 && threadIdx.x == 0) draw[blockIdx.y] = draw[blockIdx.y] + 1u;
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
  int noise_blocks = 64;
  StepArgs* d_args = nullptr;
  int2* d_work = nullptr;
  NoiseMember* d_noise = nullptr;
  uint32_t* d_draw = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_steps = 0;
  std::vector<StepArgs> h_args;  // host copy of the argument table (2 per member)
};

static int ensemble_launch_substep(Ensemble* e, int rk, hipStream_t s) {
  const hipError_t err =
      e->exact ? launch_cdklm_batched_exact(e->d_args, e->d_work, e->n_work, rk, e->n_blocks, e->has_land,
                                            e->f_array, e->angle0, s)
               : launch_cdklm_batched_fast(e->d_args, e->d_work, e->n_work, rk, e->n_blocks, e->has_land,
                                           e->f_array, e->angle0, s);
  return err != hipSuccess ? check(err, "ensemble substep launch") : 0;
}

static int ensemble_launch_noise(Ensemble* e, hipStream_t s) {
  if (!e->has_noise) return 0;
  const dim3 grid(e->noise_blocks, e->n);
  hipLaunchKernelGGL(noise_fill, grid, dim3(256), 0, s, e->d_noise, e->d_draw);
  hipLaunchKernelGGL(noise_soar, grid, dim3(256), 0, s, e->d_noise);
  hipLaunchKernelGGL(noise_balance, grid, dim3(256), 0, s, e->d_noise, e->d_draw);
  return check(hipGetLastError(), "ensemble noise launch");
}

static int ensemble_launch_step(Ensemble* e, hipStream_t s) {
  int rc = ensemble_launch_substep(e, 0, s);
  if (!rc) rc = ensemble_launch_substep(e, 1, s);
  if (!rc) rc = ensemble_launch_noise(e, s);
  return rc;
}

static void ensemble_free(Ensemble* e) {
  if (e->exec) (void)hipGraphExecDestroy(e->exec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  if (e->d_args) (void)hipFree(e->d_args);
  if (e->d_work) (void)hipFree(e->d_work);
  if (e->d_noise) (void)hipFree(e->d_noise);
  if (e->d_draw) (void)hipFree(e->d_draw);
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
  std::vector<NoiseMember> noise;
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
    for (int side = 0; side < 4; ++side) {
      const int code = (M.boundary_conditions >> (8 * side)) & 0xFF;
      if (code != 1 && code != 2) {
        ensemble_free(e);
        return bad("gpo_ensemble_create: only wall (1) and periodic (2) boundaries");
      }
    }
    if (chunk_rows > 0) p.chunk_rows = chunk_rows;
    int rc = make_step_args(&p, M.dt, 0, M.h0_eta, M.h0_hu, M.h0_hv, nullptr, nullptr, nullptr, M.h1_eta, M.h1_hu,
                            M.h1_hv, M.pitch, M.Hi, M.Hi_pitch, M.f_cells, nullptr, M.boundary_conditions,
                            &args[2 * m]);
    if (!rc)
      rc = make_step_args(&p, M.dt, 1, M.h1_eta, M.h1_hu, M.h1_hv, M.h0_eta, M.h0_hu, M.h0_hv, M.h0_eta, M.h0_hu,
                          M.h0_hv, M.pitch, M.Hi, M.Hi_pitch, M.f_cells, nullptr, M.boundary_conditions,
                          &args[2 * m + 1]);
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
      if (!nm.xi || !nm.coarse || nm.xi_rows != (nm.per_ns ? nm.ny : nm.ny + 8) ||
          nm.xi_cols != (nm.per_ew ? nm.nx : nm.nx + 8)) {
        ensemble_free(e);
        return bad("gpo_ensemble_create: model error buffers missing or mis-sized");
      }
      noise.push_back(nm);
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
  if (!rc) rc = check(hipStreamSynchronize(s), "gpo_ensemble_create sync");
  if (rc) {
    ensemble_free(e);
    return rc;
  }
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
  if (!e || part < 0 || part > 2) return bad("gpo_ensemble_part: bad arguments");
  return part < 2 ? ensemble_launch_substep(e, part, (hipStream_t)stream) : ensemble_launch_noise(e, (hipStream_t)stream);
}

int gpo_ensemble_capture(gpo_ensemble h, int32_t steps_per_graph, void* stream) {
  Ensemble* e = (Ensemble*)h;
  if (!e || steps_per_graph < 1) return bad("gpo_ensemble_capture: bad arguments");
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

int gpo_ensemble_destroy(gpo_ensemble h) {
  if (h) ensemble_free((Ensemble*)h);
  return 0;
}

}  // extern "C"
