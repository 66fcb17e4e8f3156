// Shared declarations of the CDKLM16 HIP library (internal, not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/gpuocean_amd.h"

namespace gpo {

// Values that must match between the step, bathymetry and CFL kernels
// (gpu_kernels/CDKLM16_kernel.cu:34-35, initHm_kernel.cu:27-28, max_dt.cu:30-35).
constexpr float kDryFlag = 1.0e-30f;
constexpr float kDryEps = 1.0e-3f;
constexpr float kMaxDtCap = 100000.0f;

// Output columns per 64-lane wavefront: lanes 0,1 and 62,63 are the halo.
constexpr int kStripOut = 60;
constexpr int kWave = 64;
constexpr int kWavesPerBlock = 4;
constexpr int kDefaultChunkRows = 6;
constexpr int kLargeChunkRows = 24;
constexpr int kDefaultWavesPerSimd = 5;

// Lane exchange through DPP wave shifts (verified on gfx950: wave_shr:1 gives
// lane i the value of lane i-1, wave_shl:1 that of lane i+1).
#ifndef GPO_USE_DPP
#define GPO_USE_DPP 1
#endif

struct StepArgs {
  // geometry
  int nx, ny, pitch, hi_pitch;  // pitches in elements
  int n_strips, n_chunks, chunk_rows;
  int grid_waves;  // wavefronts of the persistent substep grid (0: one per item)
  int bc_n, bc_e, bc_s, bc_w;
  int rk_order, rk_step, one_dim, has_forcing;
  int divc_ok;   // bit 0: Markstein division by dx verified, bit 1: by dy, bit 2: by 2g
  bool sqrt_ok;  // g*depth_cutoff >= 2^-96 and sqrt_rsq verified: the short square root is valid
  // constants as float32, derived on the host the same way the kernel would
  float dt, dx, dy, g, r, theta, desing_eps, flux_slope_eps, depth_cutoff, flux_balancer, land_value;
  float f, north_x, north_y, wind_stress_factor, rho_o;
  float g2, inv_g2, half_g, theta_g, Vc, Uc, inv_dx, inv_dy, one_m_fb, dxf, dyf;
  // buffers
  const float* in_e;
  const float* in_u;
  const float* in_v;
  const float* qn_e;
  const float* qn_u;
  const float* qn_v;
  float* out_e;
  float* out_u;
  float* out_v;
  const float* Hi;
  const float* fcell;
  gpo_forcing forcing;
};

// The exact build's fast divisions by dx, dy and 2g and its short square
// root are all verified for this argument block (the step kernel then needs
// no run-time flags for them).
inline bool fast_arith_ok(const StepArgs& a) { return a.divc_ok == 7 && a.sqrt_ok; }

// Square root from the hardware reciprocal square root y ~ 1/sqrt(x) and one
// correction: g = x y, then g + (x - g^2) y/2 (Markstein's step for the square
// root).  gpo_cdklm16_prepare checks on the device, for every float x in
// [2^-96, FLT_MAX], that it equals the IEEE square root bit for bit before
// the exact build uses it (bit 3 of gpo_cdklm16_params.divc_ok).
__device__ __forceinline__ float sqrt_rsq(float x) {
  const float y = __builtin_amdgcn_rsqf(x);
  const float g = x * y;
  const float h = 0.5f * y;
  const float r = __builtin_fmaf(-g, g, x);
  return __builtin_fmaf(r, h, g);
}

// Launchers of the two builds of the step kernel (fast / exact).
hipError_t launch_cdklm_fast(const StepArgs& a, int has_land, int f_array, int angle0, hipStream_t s);
hipError_t launch_cdklm_exact(const StepArgs& a, int has_land, int f_array, int angle0, hipStream_t s);

// any_stage: some member has RK3 or bottom friction (the general kernel);
// otherwise rk_step 0/1 select the kernels specialised to the RK2 stages.
// all_ok: fast_arith_ok holds for every member's argument blocks.
hipError_t launch_cdklm_batched_fast(const StepArgs* members, const int2* work, int n_work, int rk_step,
                                     int n_blocks, int has_land, int f_array, int angle0, int any_stage, int all_ok,
                                     hipStream_t s);
hipError_t launch_cdklm_batched_exact(const StepArgs* members, const int2* work, int n_work, int rk_step,
                                      int n_blocks, int has_land, int f_array, int angle0, int any_stage,
                                      int all_ok, hipStream_t s);

// Builds the argument block of one substep (aux_kernels.hip); returns 0 or an error code.
int make_step_args(const gpo_cdklm16_params* p, float dt, int32_t rk_step, const float* in_eta, const float* in_hu,
                   const float* in_hv, const float* qn_eta, const float* qn_hu, const float* qn_hv, float* out_eta,
                   float* out_hu, float* out_hv, int32_t pitch, const float* Hi, int32_t Hi_pitch,
                   const float* f_cells, const gpo_forcing* forcing, uint32_t boundary_conditions, StepArgs* out);

void set_error(const char* msg);

// Wavefronts of a persistent grid with `wps` waves per SIMD on the current device.
int persistent_waves(int wps);

}  // namespace gpo
