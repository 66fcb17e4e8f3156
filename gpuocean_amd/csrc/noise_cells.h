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

// What the balance of one cell reads from memory: its state, the four
// corner depths and (when f is a texture) its coriolis parameter.
struct BalanceIn {
  float e, u, v, h00, h01, h10, h11, fc;
};
__device__ __forceinline__ BalanceIn balance_load(const float* f_cells, int f_pitch, const float* eta,
                                                  const float* hu, const float* hv, int pitch, const float* Hi,
                                                  int h_pitch, int tj, int ti) {
  BalanceIn b;
  b.e = crow(eta, pitch, tj)[ti];
  b.u = crow(hu, pitch, tj)[ti];
  b.v = crow(hv, pitch, tj)[ti];
  b.h00 = crow(Hi, h_pitch, tj)[ti];
  b.h01 = crow(Hi, h_pitch, tj)[ti + 1];
  b.h10 = crow(Hi, h_pitch, tj + 1)[ti];
  b.h11 = crow(Hi, h_pitch, tj + 1)[ti + 1];
  b.fc = f_cells ? crow(f_cells, f_pitch, tj)[ti] : 0.0f;
  return b;
}

// geostrophicBalance (gpu_kernels/ocean_noise.cu:256-390) at interior cell
// (tj, ti), coarse grid equal to the computational grid, given the coarse
// value d at the cell and at its four neighbours and what balance_load read:
// the new (eta, hu, hv) of the cell.
__device__ __forceinline__ void balance_compute(float dx, float dy, float g, float f, float beta, float north_x,
                                                float north_y, float d, float d_xm, float d_xp, float d_ym,
                                                float d_yp, const BalanceIn& b, float land_value, int tj, int ti,
                                                float* e, float* u, float* v) {
  const bool dry = (b.h00 == land_value) || (b.h01 == land_value) || (b.h10 == land_value) || (b.h11 == land_value);
  const float H_mid = 0.25f * (b.h00 + b.h01 + b.h10 + b.h11);
  float coriolis = f + beta * (((float)ti + 0.5f) * dx * north_x + ((float)tj + 0.5f) * dy * north_y);
  if (f == 0.0f) coriolis = b.fc;
  const float h_mid = d + H_mid + b.e;
  const float ddx = (d_xp - d_xm) / (2.0f * dx);
  const float ddy = (d_yp - d_ym) / (2.0f * dy);
  const float gc = g / coriolis;
  const float d_hu = -gc * h_mid * ddy;
  const float d_hv = gc * h_mid * ddx;
  *e = dry ? 0.0f : b.e + d;
  *u = dry ? 0.0f : b.u + d_hu;
  *v = dry ? 0.0f : b.v + d_hv;
}

__device__ __forceinline__ void balance_store(float* eta, float* hu, float* hv, int pitch, int tj, int ti, float e,
                                              float u, float v) {
  mrow(eta, pitch, tj)[ti] = e;
  mrow(hu, pitch, tj)[ti] = u;
  mrow(hv, pitch, tj)[ti] = v;
}

__device__ __forceinline__ void balance_apply(float dx, float dy, float g, float f, float beta, float north_x,
                                              float north_y, const float* f_cells, int f_pitch, float d, float d_xm,
                                              float d_xp, float d_ym, float d_yp, float* eta, float* hu, float* hv,
                                              int pitch, const float* Hi, int h_pitch, float land_value, int tj,
                                              int ti) {
  const BalanceIn b = balance_load(f_cells, f_pitch, eta, hu, hv, pitch, Hi, h_pitch, tj, ti);
  float e, u, v;
  balance_compute(dx, dy, g, f, beta, north_x, north_y, d, d_xm, d_xp, d_ym, d_yp, b, land_value, tj, ti, &e, &u,
                  &v);
  balance_store(eta, hu, hv, pitch, tj, ti, e, u, v);
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

// ---- coarse-grid perturbations (interpolation factor > 1) ---------------------

__device__ __forceinline__ float dot4(float a0, float a1, float a2, float a3, float b0, float b1, float b2, float b3) {
  return a0 * b0 + a1 * b1 + a2 * b2 + a3 * b3;
}

// Where fine cell (tj, ti) of the ghosted fine grid lies on the coarse grid
// (2 ghost cells on both): the coarse cell (cj, ci) below-left of its centre
// and the centre's offset (ry, rx) in units of the coarse cell, as
// bicubicInterpolation finds them (gpu_kernels/ocean_noise.cu:523-545, 603-604).
struct CoarseSpot {
  int ci, cj;
  float rx, ry;
};
__device__ __forceinline__ CoarseSpot coarse_spot(int cnx, int cny, float dx, float dy, float cdx, float cdy, int oi,
                                                  int oj, int tj, int ti) {
  const float x = ((float)(ti - 2) + 0.5f + (float)oi) * dx;
  const float y = ((float)(tj - 2) + 0.5f + (float)oj) * dy;
  CoarseSpot p;
  p.ci = (int)floorf(x / cdx + 2.0f - 0.5f);
  p.cj = (int)floorf(y / cdy + 2.0f - 0.5f);
  if (p.ci > cnx + 1) p.ci -= 1;  // first ghost cell right/top of the domain: rel = 1
  if (p.cj > cny + 1) p.cj -= 1;
  const float cx = ((float)(p.ci - 2) + 0.5f) * cdx;
  const float cy = ((float)(p.cj - 2) + 0.5f) * cdy;
  p.rx = (x - cx) / cdx;
  p.ry = (y - cy) / cdy;
  return p;
}

// Bicubic coefficients B F B^T of coarse cell (cj, ci) (interpolation.cu:39-117):
// F holds the corner values and central-difference derivatives of the
// (cny+4) x (cnx+4) coarse buffer (ocean_noise.cu:554-601), B is the Hermite
// basis matrix.  Coarse indices are clamped to the buffer as the reference's
// shared-memory load does.
__device__ __forceinline__ void bicubic_coeffs(int cnx, int cny, const float* coarse, int c_pitch, int cj, int ci,
                                               float a[16]) {
  auto C = [&](int j, int i) {
    j = j < 0 ? 0 : (j > cny + 3 ? cny + 3 : j);
    i = i < 0 ? 0 : (i > cnx + 3 ? cnx + 3 : i);
    return crow(coarse, c_pitch, j)[i];
  };
  // F = [[f00, f01, fy00, fy01], [f10, f11, fy10, fy11], [fx00, fx01, fxy00, fxy01], [fx10, fx11, fxy10, fxy11]]
  const float f00 = C(cj, ci), f01 = C(cj + 1, ci), f10 = C(cj, ci + 1), f11 = C(cj + 1, ci + 1);
  const float fx00 = (C(cj, ci + 1) - C(cj, ci - 1)) / 2.0f;
  const float fx01 = (C(cj + 1, ci + 1) - C(cj + 1, ci - 1)) / 2.0f;
  const float fx10 = (C(cj, ci + 2) - C(cj, ci)) / 2.0f;
  const float fx11 = (C(cj + 1, ci + 2) - C(cj + 1, ci)) / 2.0f;
  const float fy00 = (C(cj + 1, ci) - C(cj - 1, ci)) / 2.0f;
  const float fy01 = (C(cj + 2, ci) - C(cj, ci)) / 2.0f;
  const float fy10 = (C(cj + 1, ci + 1) - C(cj - 1, ci + 1)) / 2.0f;
  const float fy11 = (C(cj + 2, ci + 1) - C(cj, ci + 1)) / 2.0f;
  const float fym0 = (C(cj + 1, ci - 1) - C(cj - 1, ci - 1)) / 2.0f;
  const float fym1 = (C(cj + 2, ci - 1) - C(cj, ci - 1)) / 2.0f;
  const float fy20 = (C(cj + 1, ci + 2) - C(cj - 1, ci + 2)) / 2.0f;
  const float fy21 = (C(cj + 2, ci + 2) - C(cj, ci + 2)) / 2.0f;
  const float F[4][4] = {{f00, f01, fy00, fy01},
                         {f10, f11, fy10, fy11},
                         {fx00, fx01, (fy10 - fym0) / 2.0f, (fy11 - fym1) / 2.0f},
                         {fx10, fx11, (fy20 - fy00) / 2.0f, (fy21 - fy01) / 2.0f}};
  const float B[4][4] = {{1.0f, 0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 1.0f, 0.0f},
                         {-3.0f, 3.0f, -2.0f, -1.0f}, {2.0f, -2.0f, 1.0f, 1.0f}};
  float T[4][4];  // (F B^T)^T: T[i][j] = F[j] . B[i]
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) T[i][j] = dot4(F[j][0], F[j][1], F[j][2], F[j][3], B[i][0], B[i][1], B[i][2], B[i][3]);
#pragma unroll
  for (int i = 0; i < 4; ++i)  // row i of B F B^T: a[4i + j] = B[i] . T[j]
#pragma unroll
    for (int j = 0; j < 4; ++j)
      a[4 * i + j] = dot4(B[i][0], B[i][1], B[i][2], B[i][3], T[j][0], T[j][1], T[j][2], T[j][3]);
}

// x^T A y with x = (1, rx, rx^2, rx^3), y likewise (interpolation.cu:119-133).
__device__ __forceinline__ float bicubic_eval(const float* a, float rx, float ry) {
  const float rx2 = rx * rx, ry2 = ry * ry;
  const float yv[4] = {1.0f, ry, ry2, ry2 * ry};
  float t[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) t[i] = dot4(a[4 * i], a[4 * i + 1], a[4 * i + 2], a[4 * i + 3], yv[0], yv[1], yv[2], yv[3]);
  return dot4(1.0f, rx, rx2, rx2 * rx, t[0], t[1], t[2], t[3]);
}

// The geostrophic step of bicubicInterpolation (gpu_kernels/ocean_noise.cu:616-657)
// at interior cell (tj, ti): d the interpolated perturbation there and at its
// four neighbours; the balanced momentum uses the still depth only, and a cell
// with a land corner is left unchanged.
__device__ __forceinline__ void bicubic_balance_apply(float dx, float dy, float g, float f, float beta,
                                                      float north_x, float north_y, const float* f_cells,
                                                      int f_pitch, float d, float d_xm, float d_xp, float d_ym,
                                                      float d_yp, float* eta, float* hu, float* hv, int pitch,
                                                      const float* Hi, int h_pitch, float land_value, int tj,
                                                      int ti) {
  const float h00 = crow(Hi, h_pitch, tj)[ti], h01 = crow(Hi, h_pitch, tj)[ti + 1];
  const float h10 = crow(Hi, h_pitch, tj + 1)[ti], h11 = crow(Hi, h_pitch, tj + 1)[ti + 1];
  const bool dry = (h00 == land_value) || (h01 == land_value) || (h10 == land_value) || (h11 == land_value);
  const float H_mid = 0.25f * (h00 + h01 + h10 + h11);
  float coriolis = f + beta * (((float)ti + 0.5f) * dx * north_x + ((float)tj + 0.5f) * dy * north_y);
  if (f == 0.0f) coriolis = f_cells ? crow(f_cells, f_pitch, tj)[ti] : 0.0f;
  const float ddx = (d_xp - d_xm) / (2.0f * dx);
  const float ddy = (d_yp - d_ym) / (2.0f * dy);
  const float gc = g / coriolis;
  float d_hu = -gc * (H_mid + d) * ddy;
  float d_hv = gc * (H_mid + d) * ddx;
  if (dry) d = d_hu = d_hv = 0.0f;
  mrow(eta, pitch, tj)[ti] += d;
  mrow(hu, pitch, tj)[ti] += d_hu;
  mrow(hv, pitch, tj)[ti] += d_hv;
}

constexpr int kBiW = 64, kBiH = 16;
constexpr int kBiMaxCoarse = 256;  // coarse cells per tile at factor >= 3: (66/3+2) x (18/3+2) = 192
struct BicubicGeo {
  int nx, ny, cnx, cny, oi, oj;
  float dx, dy, cdx, cdy;
};
// Coarse-grid model error of one tile of kBiH x kBiW interior cells whose
// first cell is (by, bx): bicubic interpolation of the coarse field to the
// tile and its one-cell ring (coefficients formed once per coarse cell, in
// LDS, when the tile spans few enough coarse cells) and its geostrophic
// balance added to the state.  One 256-thread block; coef and d_eta are its
// LDS arrays.
__device__ __forceinline__ void bicubic_tile(const BicubicGeo& q, float g, float f, float beta, float north_x,
                                             float north_y, const float* f_cells, int f_pitch, const float* coarse,
                                             int c_pitch, float* eta, float* hu, float* hv, int pitch,
                                             const float* Hi, int h_pitch, float land_value, const int bx,
                                             const int by, float (*coef)[17], float (*d_eta)[kBiW + 2]) {
  const int tid = threadIdx.x;
  // the coarse cells under the tile and its ring: the spot of a fine point
  // never decreases with its index, so the corners bound them
  const CoarseSpot lo = coarse_spot(q.cnx, q.cny, q.dx, q.dy, q.cdx, q.cdy, q.oi, q.oj, by - 1, bx - 1);
  const CoarseSpot hi = coarse_spot(q.cnx, q.cny, q.dx, q.dy, q.cdx, q.cdy, q.oi, q.oj, by + kBiH, bx + kBiW);
  const int ncw = hi.ci - lo.ci + 1, nch = hi.cj - lo.cj + 1;
  const bool shared = ncw * nch <= kBiMaxCoarse;  // false only for factors below 3
  if (shared) {
    for (int k = tid; k < ncw * nch; k += 256) {
      float a[16];
      bicubic_coeffs(q.cnx, q.cny, coarse, c_pitch, lo.cj + k / ncw, lo.ci + k % ncw, a);
#pragma unroll
      for (int i = 0; i < 16; ++i) coef[k][i] = a[i];
    }
  }
  __syncthreads();
  for (int k = tid; k < (kBiH + 2) * (kBiW + 2); k += 256) {
    const int j = k / (kBiW + 2), i = k % (kBiW + 2);
    const CoarseSpot p = coarse_spot(q.cnx, q.cny, q.dx, q.dy, q.cdx, q.cdy, q.oi, q.oj, by + j - 1, bx + i - 1);
    float a[16];
    if (shared) {
      const int c = (p.cj - lo.cj) * ncw + (p.ci - lo.ci);
#pragma unroll
      for (int t = 0; t < 16; ++t) a[t] = coef[c][t];
    } else {
      bicubic_coeffs(q.cnx, q.cny, coarse, c_pitch, p.cj, p.ci, a);
    }
    d_eta[j][i] = bicubic_eval(a, p.rx, p.ry);
  }
  __syncthreads();
  const int x = (tid & 63) + 1;
  for (int y = (tid >> 6) + 1; y <= kBiH; y += 4) {
    const int ti = bx + x - 1, tj = by + y - 1;
    if (ti >= q.nx + 2 || tj >= q.ny + 2) continue;
    bicubic_balance_apply(q.dx, q.dy, g, f, beta, north_x, north_y, f_cells, f_pitch, d_eta[y][x], d_eta[y][x - 1],
                          d_eta[y][x + 1], d_eta[y - 1][x], d_eta[y + 1][x], eta, hu, hv, pitch, Hi, h_pitch,
                          land_value, tj, ti);
  }
}


}  // namespace gpo
