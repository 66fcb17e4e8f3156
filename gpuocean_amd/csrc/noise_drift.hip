// Stochastic model error (SOAR-correlated, geostrophically balanced) and
// Lagrangian drifter kernels of the CDKLM16 path, with their C ABI entry
// points.  Built with -ffp-contract=off: every float operation rounds as the
// host oracle's does.

#include <cstdio>

#include "cdklm_common.h"
#include "drift_cells.h"
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

// bicubicInterpolation (gpu_kernels/ocean_noise.cu:401-659).  A 64 x 16 tile
// of interior cells and its one-cell ring (the balance needs the neighbours'
// perturbation) covers a small block of coarse cells: their bicubic
// coefficients are formed once, in LDS, and every fine point of the tile
// evaluates the polynomial of its coarse cell from there, so the coefficient
// product is paid per coarse cell, not per fine cell.  Then the balance.
__global__ __launch_bounds__(256) void bicubic_balance(BicubicGeo q, float g, float f, float beta, float north_x,
                                                       float north_y, const float* f_cells, int f_pitch,
                                                       const float* coarse, int c_pitch, float* eta, float* hu,
                                                       float* hv, int pitch, const float* Hi, int h_pitch,
                                                       float land_value) {
  __shared__ BicubicLds lds;
  bicubic_tile(q, g, f, beta, north_x, north_y, f_cells, f_pitch, coarse, c_pitch, eta, hu, hv, pitch, Hi, h_pitch,
               land_value, blockIdx.x * kBiW + 2, blockIdx.y * kBiH + 2, lds);
}

__global__ void drift(int nx, int ny, float dx, float dy, float dt, int x0ref, int y0ref, const float* eta,
                      const float* hu, const float* hv, int pitch, const float* Hm, int hm_pitch, int per_ns,
                      int per_ew, int n_drifters, float* pos, int dpitch, float sensitivity, const float* wx_cur,
                      const float* wx_next, const float* wy_cur, const float* wy_next, int w_nx, int w_ny,
                      float wind_t, float wind_drift_factor) {
  const int ti = blockIdx.x * blockDim.x + threadIdx.x;
  if (ti >= n_drifters + 1) return;
  cells::drift_one(nx, ny, dx, dy, dt, x0ref, y0ref, eta, hu, hv, pitch, Hm, hm_pitch, per_ns, per_ew,
                   cells::drift_row(pos, dpitch, ti), sensitivity, wx_cur, wx_next, wy_cur, wy_next, w_nx, w_ny,
                   wind_t, wind_drift_factor);
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

int gpo_bicubic_interpolation(int32_t nx, int32_t ny, float dx, float dy, int32_t coarse_nx, int32_t coarse_ny,
                              float coarse_dx, float coarse_dy, int32_t offset_i, int32_t offset_j, float g, float f,
                              float beta, float north_x, float north_y, const float* f_cells, int32_t f_pitch,
                              const float* coarse, int32_t coarse_pitch, float* eta, float* hu, float* hv,
                              int32_t pitch, const float* Hi, int32_t Hi_pitch, float land_value, void* stream) {
  if (nx < 1 || ny < 1 || coarse_nx < 1 || coarse_ny < 1 || !coarse || !eta || !hu || !hv || !Hi ||
      !(coarse_dx > 0.0f) || !(coarse_dy > 0.0f))
    return bad("gpo_bicubic_interpolation: bad arguments");
  if (offset_i < -coarse_nx || offset_i > coarse_nx || offset_j < -coarse_ny || offset_j > coarse_ny)
    return bad("gpo_bicubic_interpolation: offset larger than the coarse grid");
  const BicubicGeo q{nx, ny, coarse_nx, coarse_ny, offset_i, offset_j, dx, dy, coarse_dx, coarse_dy};
  hipLaunchKernelGGL(bicubic_balance, dim3(ceil_div(nx, kBiW), ceil_div(ny, kBiH)), dim3(256), 0,
                     (hipStream_t)stream, q, g, f, beta, north_x, north_y, f_cells, f_pitch, coarse, coarse_pitch,
                     eta, hu, hv, pitch, Hi, Hi_pitch, land_value);
  return check(hipGetLastError(), "gpo_bicubic_interpolation launch");
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
