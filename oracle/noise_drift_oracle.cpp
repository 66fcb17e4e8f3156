// CPU oracle of the stochastic model error and drifter path -- TEST
// INFRASTRUCTURE ONLY (see cdklm_oracle.cpp).  Float arithmetic, no FMA
// contraction; the SOAR accumulation uses explicit fma as the GPU kernel
// specifies.  Random numbers are an input: the tests pass the generator's
// output so SOAR, balance and drift are checked on identical xi.

#include <cmath>
#include <cstdint>

#include "oracle.h"

extern "C" {

// SOAR (gpu_kernels/ocean_noise.cu:141-241) with precomputed 5x5 weights.
void gpo_oracle_soar(int nx, int ny, const float* W, float scale, int per_ns, int per_ew, const float* xi,
                     int xi_cols, float* coarse, int additive) {
  const int W4 = nx + 4;
  for (int tj = 0; tj < ny + 4; ++tj)
    for (int ti = 0; ti < nx + 4; ++ti) {
      float acc = 0.0f;
      for (int k = -2; k <= 2; ++k) {
        const int yj = per_ns ? (tj - 2 + k + 2 * ny) % ny : tj + 2 + k;
        for (int l = -2; l <= 2; ++l) {
          const int xx = per_ew ? (ti - 2 + l + 2 * nx) % nx : ti + 2 + l;
          acc = std::fmaf(W[(k + 2) * 5 + (l + 2)], xi[(size_t)yj * xi_cols + xx], acc);
        }
      }
      float* o = coarse + (size_t)tj * W4 + ti;
      *o = additive ? *o + scale * acc : scale * acc;
    }
}

// geostrophicBalance (gpu_kernels/ocean_noise.cu:256-390), interpolation factor 1.
void gpo_oracle_geostrophic_balance(int nx, int ny, float dx, float dy, float g, float f, float beta, float north_x,
                                    float north_y, const float* f_cells, const float* coarse, float* eta, float* hu,
                                    float* hv, const float* Hi, float land_value) {
  const int W = nx + 4, WI = nx + 5;
  for (int tj = 2; tj < ny + 2; ++tj)
    for (int ti = 2; ti < nx + 2; ++ti) {
      const float h00 = Hi[(size_t)tj * WI + ti], h01 = Hi[(size_t)tj * WI + ti + 1];
      const float h10 = Hi[(size_t)(tj + 1) * WI + ti], h11 = Hi[(size_t)(tj + 1) * WI + ti + 1];
      const bool dry = (h00 == land_value) || (h01 == land_value) || (h10 == land_value) || (h11 == land_value);
      const float H_mid = 0.25f * (h00 + h01 + h10 + h11);
      float coriolis = f + beta * (((float)ti + 0.5f) * dx * north_x + ((float)tj + 0.5f) * dy * north_y);
      if (f == 0.0f) coriolis = f_cells ? f_cells[(size_t)tj * W + ti] : 0.0f;
      const size_t k = (size_t)tj * W + ti;
      const float d = coarse[k];
      const float h_mid = d + H_mid + eta[k];
      const float ddx = (coarse[k + 1] - coarse[k - 1]) / (2.0f * dx);
      const float ddy = (coarse[k + W] - coarse[k - W]) / (2.0f * dy);
      const float gc = g / coriolis;
      const float d_hu = -gc * h_mid * ddy;
      const float d_hv = gc * h_mid * ddx;
      float e = eta[k] + d, u = hu[k] + d_hu, v = hv[k] + d_hv;
      if (dry) e = u = v = 0.0f;
      eta[k] = e;
      hu[k] = u;
      hv[k] = v;
    }
}

// passiveDrifterKernel (gpu_kernels/driftKernels.cu:74-183) without wind, or
// with a constant (1x1) wind texture pair.
void gpo_oracle_drift(int nx, int ny, float dx, float dy, float dt, int x0ref, int y0ref, const float* eta,
                      const float* hu, const float* hv, const float* Hm, int per_ns, int per_ew, int n_drifters,
                      float* pos, float sensitivity, float wind_x, float wind_y, float wind_drift_factor) {
  const int W = nx + 4;
  auto vel = [&](const float* mom, int cx, int cy) {
    const float h = Hm[(size_t)cy * W + cx] + eta[(size_t)cy * W + cx];
    return mom[(size_t)cy * W + cx] / h;
  };
  for (int i = 0; i < n_drifters + 1; ++i) {
    float px = pos[2 * i], py = pos[2 * i + 1];
    const float qx = px / dx, qy = py / dy;
    const int cx = (int)(std::floor(qx) + (float)x0ref);
    const int cy = (int)(std::floor(qy) + (float)y0ref);
    const float frx = qx - std::floor(qx), fry = qy - std::floor(qy);
    const int cx0 = frx < 0.5f ? cx - 1 : cx;
    const float xf = frx < 0.5f ? frx + 0.5f : frx - 0.5f;
    const int cy0 = fry < 0.5f ? cy - 1 : cy;
    const float yf = fry < 0.5f ? fry + 0.5f : fry - 0.5f;
    const float u00 = vel(hu, cx0, cy0), u10 = vel(hu, cx0 + 1, cy0);
    const float u01 = vel(hu, cx0, cy0 + 1), u11 = vel(hu, cx0 + 1, cy0 + 1);
    const float v00 = vel(hv, cx0, cy0), v10 = vel(hv, cx0 + 1, cy0);
    const float v01 = vel(hv, cx0, cy0 + 1), v11 = vel(hv, cx0 + 1, cy0 + 1);
    const float uy0 = (1.0f - xf) * u00 + xf * u10;
    const float uy1 = (1.0f - xf) * u01 + xf * u11;
    const float vy0 = (1.0f - xf) * v00 + xf * v10;
    const float vy1 = (1.0f - xf) * v01 + xf * v11;
    float u = (1.0f - yf) * uy0 + yf * uy1;
    float v = (1.0f - yf) * vy0 + yf * vy1;
    if (wind_drift_factor != 0.0f && Hm[(size_t)cy * W + cx] < 1e20f) {
      u = u + (wind_x - u) * wind_drift_factor;
      v = v + (wind_y - v) * wind_drift_factor;
    }
    px += sensitivity * u * dt;
    py += sensitivity * v * dt;
    const float Lx = (float)nx * dx, Ly = (float)ny * dy;
    if (per_ew && px < 0.0f) px += Lx;
    if (per_ew && px > Lx) px -= Lx;
    if (per_ns && py < 0.0f) py += Ly;
    if (per_ns && py > Ly) py -= Ly;
    pos[2 * i] = px;
    pos[2 * i + 1] = py;
  }
}

}  // extern "C"
