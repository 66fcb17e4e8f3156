// Per-drifter body of the passive drifter step, shared by the per-member
// launch (noise_drift.hip) and the batched ensemble (ensemble.hip).  Callers
// compile with -ffp-contract=off.
#pragma once

#include <hip/hip_runtime.h>

namespace gpo {
namespace cells {

__device__ inline const float* drift_crow(const float* base, int pitch_bytes, int j) {
  return (const float*)((const char*)base + (size_t)pitch_bytes * j);
}
__device__ inline float* drift_row(float* base, int pitch_bytes, int j) {
  return (float*)((char*)base + (size_t)pitch_bytes * j);
}

// Normalised bilinear lookup of a 1-component texture (interpolation.cu:137-190).
__device__ inline float bilin1(const float* data, int dnx, int dny, float sx, float sy) {
  sx = fminf(fmaxf(sx - 0.5f / (float)dnx, 0.0f), 1.0f);
  sy = fminf(fmaxf(sy - 0.5f / (float)dny, 0.0f), 1.0f);
  const float x = sx * (float)dnx, y = sy * (float)dny;
  int x0 = (int)floorf(x), y0 = (int)floorf(y);
  const float fx = x - (float)x0, fy = y - (float)y0;
  x0 = min(x0, dnx - 1);
  y0 = min(y0, dny - 1);
  const int x1 = min(x0 + 1, dnx - 1), y1 = min(y0 + 1, dny - 1);
  const float d00 = data[y0 * dnx + x0], d01 = data[y0 * dnx + x1];
  const float d10 = data[y1 * dnx + x0], d11 = data[y1 * dnx + x1];
  return d00 + fx * (d01 - d00) + fy * (d10 - d00) + fx * fy * (d00 + d11 - d01 - d10);
}

__device__ inline float water_velocity(const float* eta, const float* mom, const float* Hm, int pitch, int hm_pitch,
                                       int cx, int cy) {
  const float h = drift_crow(Hm, hm_pitch, cy)[cx] + drift_crow(eta, pitch, cy)[cx];
  return drift_crow(mom, pitch, cy)[cx] / h;
}

// passiveDrifterKernel (gpu_kernels/driftKernels.cu:74-183) for the drifter at
// d = (x, y): bilinear velocity between cell centres, optional wind drag,
// forward Euler, periodic wrap.  A drifter whose four surrounding cell
// centres are not all inside the ghost-inclusive arrays (more than the ghost
// cells beyond an open side, or a non-finite position) is left where it is:
// the reference reads outside its arrays there.
__device__ inline void drift_one(int nx, int ny, float dx, float dy, float dt, int x0ref, int y0ref, const float* eta,
                                 const float* hu, const float* hv, int pitch, const float* Hm, int hm_pitch,
                                 int per_ns, int per_ew, float* d, float sensitivity, const float* wx_cur,
                                 const float* wx_next, const float* wy_cur, const float* wy_next, int w_nx, int w_ny,
                                 float wind_t, float wind_drift_factor) {
  float px = d[0], py = d[1];
  const float qx = px / dx, qy = py / dy;
  const float fcx = floorf(qx) + (float)x0ref, fcy = floorf(qy) + (float)y0ref;
  if (!(fcx >= 0.0f && fcx <= (float)(nx + 3) && fcy >= 0.0f && fcy <= (float)(ny + 3))) return;  // NaN too
  const int cx = (int)fcx;
  const int cy = (int)fcy;
  const float frx = qx - floorf(qx), fry = qy - floorf(qy);
  const int cx0 = frx < 0.5f ? cx - 1 : cx;
  const float xf = frx < 0.5f ? frx + 0.5f : frx - 0.5f;
  const int cy0 = fry < 0.5f ? cy - 1 : cy;
  const float yf = fry < 0.5f ? fry + 0.5f : fry - 0.5f;
  const int cx1 = cx0 + 1, cy1 = cy0 + 1;
  if (cx0 < 0 || cy0 < 0 || cx1 > nx + 3 || cy1 > ny + 3) return;
  const float u00 = water_velocity(eta, hu, Hm, pitch, hm_pitch, cx0, cy0);
  const float u10 = water_velocity(eta, hu, Hm, pitch, hm_pitch, cx1, cy0);
  const float u01 = water_velocity(eta, hu, Hm, pitch, hm_pitch, cx0, cy1);
  const float u11 = water_velocity(eta, hu, Hm, pitch, hm_pitch, cx1, cy1);
  const float v00 = water_velocity(eta, hv, Hm, pitch, hm_pitch, cx0, cy0);
  const float v10 = water_velocity(eta, hv, Hm, pitch, hm_pitch, cx1, cy0);
  const float v01 = water_velocity(eta, hv, Hm, pitch, hm_pitch, cx0, cy1);
  const float v11 = water_velocity(eta, hv, Hm, pitch, hm_pitch, cx1, cy1);
  const float uy0 = (1.0f - xf) * u00 + xf * u10;
  const float uy1 = (1.0f - xf) * u01 + xf * u11;
  const float vy0 = (1.0f - xf) * v00 + xf * v10;
  const float vy1 = (1.0f - xf) * v01 + xf * v11;
  float u = (1.0f - yf) * uy0 + yf * uy1;
  float v = (1.0f - yf) * vy0 + yf * vy1;
  if (wind_drift_factor != 0.0f && wx_cur) {
    const float Hmc = drift_crow(Hm, hm_pitch, cy)[cx];
    if (Hmc < 1e20f) {
      const float sx = px / ((float)nx * dx), sy = py / ((float)ny * dy);
      // The reference samples the y wind with the x wind's texture size
      // (driftKernels.cu:156); both textures share one shape here.
      const float wxc = bilin1(wx_cur, w_nx, w_ny, sx, sy), wxn = bilin1(wx_next, w_nx, w_ny, sx, sy);
      const float wyc = bilin1(wy_cur, w_nx, w_ny, sx, sy), wyn = bilin1(wy_next, w_nx, w_ny, sx, sy);
      const float wx = wind_t * wxn + (1.0f - wind_t) * wxc;
      const float wy = wind_t * wyn + (1.0f - wind_t) * wyc;
      u = u + (wx - u) * wind_drift_factor;
      v = v + (wy - v) * wind_drift_factor;
    }
  }
  px += sensitivity * u * dt;
  py += sensitivity * v * dt;
  const float Lx = (float)nx * dx, Ly = (float)ny * dy;
  if (per_ew && px < 0.0f) px += Lx;
  if (per_ew && px > Lx) px -= Lx;
  if (per_ns && py < 0.0f) py += Ly;
  if (per_ns && py > Ly) py -= Ly;
  d[0] = px;
  d[1] = py;
}

}  // namespace cells
}  // namespace gpo
