// CPU oracle of the stochastic model error and drifter path -- TEST
// INFRASTRUCTURE ONLY (see cdklm_oracle.cpp).  Float arithmetic, no FMA
// contraction; the SOAR accumulation uses explicit fma as the GPU kernel
// specifies.  Random numbers are an input: the tests pass the generator's
// output so SOAR, balance and drift are checked on identical xi.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

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

}  // extern "C"

namespace {

float dot4(const float* a, const float* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3]; }

// Bicubic value at fine cell (tj, ti) of the coarse buffer (cny+4) x (cnx+4):
// the interpolation loop of bicubicInterpolation (gpu_kernels/ocean_noise.cu:523-611),
// bicubic_interpolation_coefficients (interpolation.cu:39-117) and
// bicubic_evaluation (interpolation.cu:119-133).
float bicubic_value(int cnx, int cny, float dx, float dy, float cdx, float cdy, int oi, int oj, const float* coarse,
                    int tj, int ti) {
  const float x = ((float)(ti - 2) + 0.5f + (float)oi) * dx;
  const float y = ((float)(tj - 2) + 0.5f + (float)oj) * dy;
  int ci = (int)std::floor(x / cdx + 2.0f - 0.5f);
  int cj = (int)std::floor(y / cdy + 2.0f - 0.5f);
  if (ci > cnx + 1) ci -= 1;
  if (cj > cny + 1) cj -= 1;
  const float cx = ((float)(ci - 2) + 0.5f) * cdx;
  const float cy = ((float)(cj - 2) + 0.5f) * cdy;
  const int CW = cnx + 4;
  auto c = [&](int j, int i) {
    j = std::min(std::max(j, 0), cny + 3);
    i = std::min(std::max(i, 0), cnx + 3);
    return coarse[(size_t)j * CW + i];
  };
  // f matrix of corner values and central-difference derivatives (ocean_noise.cu:554-601)
  float f[4][4];
  f[0][0] = c(cj, ci);
  f[0][1] = c(cj + 1, ci);
  f[1][0] = c(cj, ci + 1);
  f[1][1] = c(cj + 1, ci + 1);
  f[2][0] = (c(cj, ci + 1) - c(cj, ci - 1)) / 2.0f;
  f[2][1] = (c(cj + 1, ci + 1) - c(cj + 1, ci - 1)) / 2.0f;
  f[3][0] = (c(cj, ci + 2) - c(cj, ci)) / 2.0f;
  f[3][1] = (c(cj + 1, ci + 2) - c(cj + 1, ci)) / 2.0f;
  f[0][2] = (c(cj + 1, ci) - c(cj - 1, ci)) / 2.0f;
  f[0][3] = (c(cj + 2, ci) - c(cj, ci)) / 2.0f;
  f[1][2] = (c(cj + 1, ci + 1) - c(cj - 1, ci + 1)) / 2.0f;
  f[1][3] = (c(cj + 2, ci + 1) - c(cj, ci + 1)) / 2.0f;
  const float fy_10 = (c(cj + 1, ci - 1) - c(cj - 1, ci - 1)) / 2.0f;
  const float fy_11 = (c(cj + 2, ci - 1) - c(cj, ci - 1)) / 2.0f;
  const float fy20 = (c(cj + 1, ci + 2) - c(cj - 1, ci + 2)) / 2.0f;
  const float fy21 = (c(cj + 2, ci + 2) - c(cj, ci + 2)) / 2.0f;
  f[2][2] = (f[1][2] - fy_10) / 2.0f;
  f[2][3] = (f[1][3] - fy_11) / 2.0f;
  f[3][2] = (fy20 - f[0][2]) / 2.0f;
  f[3][3] = (fy21 - f[0][3]) / 2.0f;
  static const float b[4][4] = {{1, 0, 0, 0}, {0, 0, 1, 0}, {-3, 3, -2, -1}, {2, -2, 1, 1}};
  float fbT[4][4], a[4][4];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) fbT[i][j] = dot4(f[j], b[i]);
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) a[i][j] = dot4(b[i], fbT[j]);
  const float rx = (x - cx) / cdx, ry = (y - cy) / cdy;
  const float xv[4] = {1.0f, rx, rx * rx, rx * rx * rx};
  const float yv[4] = {1.0f, ry, ry * ry, ry * ry * ry};
  float tmp[4];
  for (int i = 0; i < 4; ++i) tmp[i] = dot4(a[i], yv);
  return dot4(xv, tmp);
}

}  // namespace

extern "C" {

// bicubicInterpolation (gpu_kernels/ocean_noise.cu:401-659): the interpolated
// perturbation d_eta (interior and first ghost ring) and its balanced hu, hv
// from the still depth, added to the interior; cells with a land corner
// unchanged.  Dense arrays: state and f_cells rows of nx+4, Hi rows of nx+5.
void gpo_oracle_bicubic_interpolation(int nx, int ny, float dx, float dy, int cnx, int cny, float cdx, float cdy,
                                      int oi, int oj, float g, float f, float beta, float north_x, float north_y,
                                      const float* f_cells, const float* coarse, float* eta, float* hu, float* hv,
                                      const float* Hi, float land_value) {
  const int W = nx + 4, WI = nx + 5;
  std::vector<float> d((size_t)(ny + 4) * W, 0.0f);
  for (int tj = 1; tj < ny + 3; ++tj)
    for (int ti = 1; ti < nx + 3; ++ti)
      d[(size_t)tj * W + ti] = bicubic_value(cnx, cny, dx, dy, cdx, cdy, oi, oj, coarse, tj, ti);
  for (int tj = 2; tj < ny + 2; ++tj)
    for (int ti = 2; ti < nx + 2; ++ti) {
      const float h00 = Hi[(size_t)tj * WI + ti], h01 = Hi[(size_t)tj * WI + ti + 1];
      const float h10 = Hi[(size_t)(tj + 1) * WI + ti], h11 = Hi[(size_t)(tj + 1) * WI + ti + 1];
      const bool dry = (h00 == land_value) || (h01 == land_value) || (h10 == land_value) || (h11 == land_value);
      const float H_mid = 0.25f * (h00 + h01 + h10 + h11);
      float coriolis = f + beta * (((float)ti + 0.5f) * dx * north_x + ((float)tj + 0.5f) * dy * north_y);
      if (f == 0.0f) coriolis = f_cells ? f_cells[(size_t)tj * W + ti] : 0.0f;
      const size_t k = (size_t)tj * W + ti;
      const float eta_diff_x = (d[k + 1] - d[k - 1]) / (2.0f * dx);
      const float eta_diff_y = (d[k + W] - d[k - W]) / (2.0f * dy);
      float dd = d[k];
      float d_hu = -(g / coriolis) * (H_mid + dd) * eta_diff_y;
      float d_hv = (g / coriolis) * (H_mid + dd) * eta_diff_x;
      if (dry) dd = d_hu = d_hv = 0.0f;
      eta[k] += dd;
      hu[k] += d_hu;
      hv[k] += d_hv;
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
    // Cells outside the ghost-inclusive arrays (the reference reads them
    // unchecked, driftKernels.cu:118-141): the drifter stays where it is.
    const float fcx = std::floor(qx) + (float)x0ref, fcy = std::floor(qy) + (float)y0ref;
    if (!(fcx >= 0.0f && fcx <= (float)(nx + 3) && fcy >= 0.0f && fcy <= (float)(ny + 3))) continue;
    const int cx = (int)fcx;
    const int cy = (int)fcy;
    const float frx = qx - std::floor(qx), fry = qy - std::floor(qy);
    const int cx0 = frx < 0.5f ? cx - 1 : cx;
    const float xf = frx < 0.5f ? frx + 0.5f : frx - 0.5f;
    const int cy0 = fry < 0.5f ? cy - 1 : cy;
    const float yf = fry < 0.5f ? fry + 0.5f : fry - 0.5f;
    if (cx0 < 0 || cy0 < 0 || cx0 + 1 > nx + 3 || cy0 + 1 > ny + 3) continue;
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
