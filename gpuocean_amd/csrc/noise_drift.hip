// Stochastic model error (SOAR-correlated, geostrophically balanced) and
// Lagrangian drifter kernels of the CDKLM16 path, with their C ABI entry
// points.  Built with -ffp-contract=off: every float operation rounds as the
// host oracle's does.

#include <cstdio>

#include "cdklm_common.h"
#include "noise_cells.h"
#include "philox.h"

namespace gpo {
namespace {

inline int ceil_div(int a, int b) { return (a + b - 1) / b; }

__device__ __forceinline__ float* rowp(float* base, int pitch_bytes, int j) {
  return (float*)((char*)base + (size_t)pitch_bytes * j);
}
__device__ __forceinline__ const float* rowp(const float* base, int pitch_bytes, int j) {
  return (const float*)((const char*)base + (size_t)pitch_bytes * j);
}

__global__ void philox_batch(const uint32_t* ctr, const uint32_t* key, uint32_t* out, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const U4 r = philox4x32_10(U4{ctr[4 * i], ctr[4 * i + 1], ctr[4 * i + 2], ctr[4 * i + 3]}, key[2 * i],
                             key[2 * i + 1]);
  out[4 * i] = r.x;
  out[4 * i + 1] = r.y;
  out[4 * i + 2] = r.z;
  out[4 * i + 3] = r.w;
}

// xi(j, i) = normal number of element j*cols + i (replaces the XORWOW/LCG fill
// of utils/RandomNumbers.py:168-185).
__global__ void random_normal(float* out, int rows, int cols, int pitch, uint32_t k0, uint32_t k1, uint32_t stream,
                              uint32_t draw) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= cols || j >= rows) return;
  rowp(out, pitch, j)[i] = normal_at((uint32_t)j * (uint32_t)cols + (uint32_t)i, draw, stream, k0, k1);
}

// SOAR (gpu_kernels/ocean_noise.cu:141-241): coarse(tj, ti) over the (ny+4) x
// (nx+4) coarse grid = scale * sum over the 5x5 window of W * xi, the window
// centred at xi(tj+2, ti+2) (4 ghost cells of xi) or at xi((tj-2) mod ny, ...)
// on a periodic axis.  W (5x5, row-major from offset (-2,-2)) holds the SOAR
// covariance q0 (1 + d/L) exp(-d/L).
struct SoarW {
  float w[25];
};
__global__ void soar(int nx, int ny, SoarW W, float scale, int per_ns, int per_ew, const float* xi, int xi_pitch,
                     float* coarse, int c_pitch, int additive) {
  const int ti = blockIdx.x * blockDim.x + threadIdx.x;
  const int tj = blockIdx.y * blockDim.y + threadIdx.y;
  if (ti >= nx + 4 || tj >= ny + 4) return;
  const float acc = soar_cell(nx, ny, W.w, per_ns, per_ew, xi, xi_pitch, tj, ti);
  float* out = rowp(coarse, c_pitch, tj) + ti;
  if (additive)
    *out += scale * acc;
  else
    *out = scale * acc;
}

// geostrophicBalance (gpu_kernels/ocean_noise.cu:256-390) for a coarse grid
// equal to the computational grid.
__global__ void geo_balance(int nx, int ny, float dx, float dy, float g, float f, float beta, float north_x,
                            float north_y, const float* f_cells, int f_pitch, const float* coarse, int c_pitch,
                            float* eta, float* hu, float* hv, int pitch, const float* Hi, int h_pitch,
                            float land_value) {
  const int ti = blockIdx.x * blockDim.x + threadIdx.x + 2;
  const int tj = blockIdx.y * blockDim.y + threadIdx.y + 2;
  if (ti >= nx + 2 || tj >= ny + 2) return;
  balance_cell(dx, dy, g, f, beta, north_x, north_y, f_cells, f_pitch, coarse, c_pitch, eta, hu, hv, pitch, Hi,
               h_pitch, land_value, tj, ti);
}

__device__ float bilin1(const float* data, int dnx, int dny, float sx, float sy) {
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

__device__ __forceinline__ float water_velocity(const float* eta, const float* mom, const float* Hm, int pitch,
                                                int hm_pitch, int cx, int cy) {
  const float h = rowp(Hm, hm_pitch, cy)[cx] + rowp(eta, pitch, cy)[cx];
  return rowp(mom, pitch, cy)[cx] / h;
}

// passiveDrifterKernel (gpu_kernels/driftKernels.cu:74-183): bilinear velocity
// between cell centres, optional wind drag, forward Euler, periodic wrap.
// Positions are (n_drifters + 1) x 2 floats with row pitch `dpitch` bytes
// (the last row is the observation drifter).
__global__ void drift(int nx, int ny, float dx, float dy, float dt, int x0ref, int y0ref, const float* eta,
                      const float* hu, const float* hv, int pitch, const float* Hm, int hm_pitch, int per_ns,
                      int per_ew, int n_drifters, float* pos, int dpitch, float sensitivity, const float* wx_cur,
                      const float* wx_next, const float* wy_cur, const float* wy_next, int w_nx, int w_ny,
                      float wind_t, float wind_drift_factor) {
  const int ti = blockIdx.x * blockDim.x + threadIdx.x;
  if (ti >= n_drifters + 1) return;
  float* d = rowp(pos, dpitch, ti);
  float px = d[0], py = d[1];
  const float qx = px / dx, qy = py / dy;
  const int cx = (int)(floorf(qx) + (float)x0ref);
  const int cy = (int)(floorf(qy) + (float)y0ref);
  const float frx = qx - floorf(qx), fry = qy - floorf(qy);
  const int cx0 = frx < 0.5f ? cx - 1 : cx;
  const float xf = frx < 0.5f ? frx + 0.5f : frx - 0.5f;
  const int cy0 = fry < 0.5f ? cy - 1 : cy;
  const float yf = fry < 0.5f ? fry + 0.5f : fry - 0.5f;
  const int cx1 = cx0 + 1, cy1 = cy0 + 1;
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
    const float Hmc = rowp(Hm, hm_pitch, cy)[cx];
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
}  // namespace gpo

using namespace gpo;

extern "C" {

int gpo_philox4x32_10(const uint32_t* ctr, const uint32_t* key, uint32_t* out, int32_t n, void* stream) {
  if (n < 0 || (n > 0 && (!ctr || !key || !out))) return bad("gpo_philox4x32_10: bad arguments");
  if (n == 0) return 0;
  hipLaunchKernelGGL(philox_batch, dim3(ceil_div(n, 256)), dim3(256), 0, (hipStream_t)stream, ctr, key, out, n);
  return check(hipGetLastError(), "gpo_philox4x32_10 launch");
}

int gpo_random_normal(float* out, int32_t rows, int32_t cols, int32_t pitch, uint64_t seed, uint32_t stream_id,
                      uint32_t draw, void* stream) {
  if (!out || rows < 1 || cols < 1 || pitch < 4 * cols) return bad("gpo_random_normal: bad arguments");
  hipLaunchKernelGGL(random_normal, dim3(ceil_div(cols, 256), rows), dim3(256), 0, (hipStream_t)stream, out, rows,
                     cols, pitch, (uint32_t)(seed & 0xffffffffu), (uint32_t)(seed >> 32), stream_id, draw);
  return check(hipGetLastError(), "gpo_random_normal launch");
}

int gpo_soar(int32_t nx, int32_t ny, const float* weights25, float scale, int32_t periodic_ns,
             int32_t periodic_ew, const float* xi, int32_t xi_pitch, float* coarse, int32_t coarse_pitch,
             int32_t additive, void* stream) {
  if (nx < 1 || ny < 1 || !weights25 || !xi || !coarse) return bad("gpo_soar: bad arguments");
  SoarW W;
  for (int i = 0; i < 25; ++i) W.w[i] = weights25[i];
  hipLaunchKernelGGL(soar, dim3(ceil_div(nx + 4, 64), ceil_div(ny + 4, 4)), dim3(64, 4), 0, (hipStream_t)stream, nx,
                     ny, W, scale, periodic_ns, periodic_ew, xi, xi_pitch, coarse, coarse_pitch, additive);
  return check(hipGetLastError(), "gpo_soar launch");
}

int gpo_geostrophic_balance(int32_t nx, int32_t ny, float dx, float dy, float g, float f, float beta,
                            float north_x, float north_y, const float* f_cells, int32_t f_pitch,
                            const float* coarse, int32_t coarse_pitch, float* eta, float* hu, float* hv,
                            int32_t pitch, const float* Hi, int32_t Hi_pitch, float land_value, void* stream) {
  if (nx < 1 || ny < 1 || !coarse || !eta || !hu || !hv || !Hi) return bad("gpo_geostrophic_balance: bad arguments");
  hipLaunchKernelGGL(geo_balance, dim3(ceil_div(nx, 64), ceil_div(ny, 4)), dim3(64, 4), 0, (hipStream_t)stream, nx,
                     ny, dx, dy, g, f, beta, north_x, north_y, f_cells, f_pitch, coarse, coarse_pitch, eta, hu, hv,
                     pitch, Hi, Hi_pitch, land_value);
  return check(hipGetLastError(), "gpo_geostrophic_balance launch");
}

int gpo_drift(int32_t nx, int32_t ny, float dx, float dy, float dt, int32_t x0ref, int32_t y0ref, const float* eta,
              const float* hu, const float* hv, int32_t pitch, const float* Hm, int32_t Hm_pitch,
              int32_t periodic_ns, int32_t periodic_ew, int32_t n_drifters, float* positions, int32_t pos_pitch,
              float sensitivity, const float* wind_x_cur, const float* wind_x_next, const float* wind_y_cur,
              const float* wind_y_next, int32_t wind_nx, int32_t wind_ny, float wind_t, float wind_drift_factor,
              void* stream) {
  if (nx < 1 || ny < 1 || n_drifters < 0 || !positions || !eta || !hu || !hv || !Hm)
    return bad("gpo_drift: bad arguments");
  if (wind_drift_factor != 0.0f && wind_x_cur && (wind_nx < 1 || wind_ny < 1))
    return bad("gpo_drift: wind texture shape");
  hipLaunchKernelGGL(drift, dim3(ceil_div(n_drifters + 1, 64)), dim3(64), 0, (hipStream_t)stream, nx, ny, dx, dy,
                     dt, x0ref, y0ref, eta, hu, hv, pitch, Hm, Hm_pitch, periodic_ns, periodic_ew, n_drifters,
                     positions, pos_pitch, sensitivity, wind_x_cur, wind_x_next, wind_y_cur, wind_y_next, wind_nx,
                     wind_ny, wind_t, wind_drift_factor);
  return check(hipGetLastError(), "gpo_drift launch");
}

}  // extern "C"
