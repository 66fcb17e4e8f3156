// Per-cell bodies of the SOAR and geostrophic-balance kernels, shared by the
// per-member launches (noise_drift.hip) and the batched ensemble launches
// (ensemble.hip) so both compute identical values.  Include only from
// translation units built with -ffp-contract=off.
#pragma once

#include <hip/hip_runtime.h>

namespace gpo {

__device__ __forceinline__ const float* crow(const float* base, int pitch_bytes, int j) {
  return (const float*)((const char*)base + (size_t)pitch_bytes * j);
}
__device__ __forceinline__ float* mrow(float* base, int pitch_bytes, int j) {
  return (float*)((char*)base + (size_t)pitch_bytes * j);
}

// SOAR (gpu_kernels/ocean_noise.cu:141-241) at coarse cell (tj, ti): the 5x5
// weighted sum of xi around xi(tj+2, ti+2), or around ((tj-2) mod ny, (ti-2)
// mod nx) on a periodic axis, accumulated row by row with fused multiply-adds.
__device__ __forceinline__ float soar_cell(int nx, int ny, const float* W, int per_ns, int per_ew, const float* xi,
                                           int xi_pitch, int tj, int ti) {
  float acc = 0.0f;
#pragma unroll
  for (int k = -2; k <= 2; ++k) {
    const int yj = per_ns ? (tj - 2 + k + 2 * ny) % ny : tj + 2 + k;
    const float* row = crow(xi, xi_pitch, yj);
#pragma unroll
    for (int l = -2; l <= 2; ++l) {
      const int xx = per_ew ? (ti - 2 + l + 2 * nx) % nx : ti + 2 + l;
      acc = __builtin_fmaf(W[(k + 2) * 5 + (l + 2)], row[xx], acc);
    }
  }
  return acc;
}

// geostrophicBalance (gpu_kernels/ocean_noise.cu:256-390) at interior cell
// (tj, ti), coarse grid equal to the computational grid, given the coarse
// value d at the cell and at its four neighbours.
__device__ __forceinline__ void balance_apply(float dx, float dy, float g, float f, float beta, float north_x,
                                              float north_y, const float* f_cells, int f_pitch, float d, float d_xm,
                                              float d_xp, float d_ym, float d_yp, float* eta, float* hu, float* hv,
                                              int pitch, const float* Hi, int h_pitch, float land_value, int tj,
                                              int ti) {
  const float h00 = crow(Hi, h_pitch, tj)[ti], h01 = crow(Hi, h_pitch, tj)[ti + 1];
  const float h10 = crow(Hi, h_pitch, tj + 1)[ti], h11 = crow(Hi, h_pitch, tj + 1)[ti + 1];
  const bool dry = (h00 == land_value) || (h01 == land_value) || (h10 == land_value) || (h11 == land_value);
  const float H_mid = 0.25f * (h00 + h01 + h10 + h11);
  float coriolis = f + beta * (((float)ti + 0.5f) * dx * north_x + ((float)tj + 0.5f) * dy * north_y);
  if (f == 0.0f) coriolis = f_cells ? crow(f_cells, f_pitch, tj)[ti] : 0.0f;
  float* er = mrow(eta, pitch, tj);
  float* ur = mrow(hu, pitch, tj);
  float* vr = mrow(hv, pitch, tj);
  const float h_mid = d + H_mid + er[ti];
  const float ddx = (d_xp - d_xm) / (2.0f * dx);
  const float ddy = (d_yp - d_ym) / (2.0f * dy);
  const float gc = g / coriolis;
  const float d_hu = -gc * h_mid * ddy;
  const float d_hv = gc * h_mid * ddx;
  float e = er[ti] + d, u = ur[ti] + d_hu, v = vr[ti] + d_hv;
  if (dry) {
    e = 0.0f;
    u = 0.0f;
    v = 0.0f;
  }
  er[ti] = e;
  ur[ti] = u;
  vr[ti] = v;
}

// The same, reading the coarse field from a pitched array.
__device__ __forceinline__ void balance_cell(float dx, float dy, float g, float f, float beta, float north_x,
                                             float north_y, const float* f_cells, int f_pitch, const float* coarse,
                                             int c_pitch, float* eta, float* hu, float* hv, int pitch,
                                             const float* Hi, int h_pitch, float land_value, int tj, int ti) {
  const float* c0 = crow(coarse, c_pitch, tj);
  balance_apply(dx, dy, g, f, beta, north_x, north_y, f_cells, f_pitch, c0[ti], c0[ti - 1], c0[ti + 1],
                crow(coarse, c_pitch, tj - 1)[ti], crow(coarse, c_pitch, tj + 1)[ti], eta, hu, hv, pitch, Hi, h_pitch,
                land_value, tj, ti);
}

}  // namespace gpo
