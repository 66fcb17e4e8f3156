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
  // On an f-plane (beta = 0, f != 0) coriolis is f itself (f + 0 * x = f), so
  // g / coriolis is the uniform g / f, which callers looping over cells hoist.
  float gc;
  if (beta == 0.0f && f != 0.0f)
    gc = g / f;
  else
    gc = g / coriolis;
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
// One axis of coarse_spot: fine index t (ghost-inclusive) with offset o ->
// coarse index c and relative position r in [0, 1] within it.
__device__ __forceinline__ void coarse_axis(int cn, float d, float cd, int o, int t, int* c, float* r) {
  const float x = ((float)(t - 2) + 0.5f + (float)o) * d;
  int ci = (int)floorf(x / cd + 2.0f - 0.5f);
  if (ci > cn + 1) ci -= 1;  // first ghost cell right/top of the domain: rel = 1
  const float cx = ((float)(ci - 2) + 0.5f) * cd;
  *c = ci;
  *r = (x - cx) / cd;
}

__device__ __forceinline__ CoarseSpot coarse_spot(int cnx, int cny, float dx, float dy, float cdx, float cdy, int oi,
                                                  int oj, int tj, int ti) {
  CoarseSpot p;
  coarse_axis(cnx, dx, cdx, oi, ti, &p.ci, &p.rx);
  coarse_axis(cny, dy, cdy, oj, tj, &p.cj, &p.ry);
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
// A tile and its ring (66 x 18 fine cells) span at most 66/3 + 1 = 23 coarse
// columns and 18/3 + 1 = 7 coarse rows at factor 3 (fewer at 5 and 7).
constexpr int kBiMaxCoarse = 168;
constexpr int kBiMaxCw = 24;
struct BicubicGeo {
  int nx, ny, cnx, cny, oi, oj;
  float dx, dy, cdx, cdy;
};
// LDS of one bicubic tile: coefficients per coarse cell (dead once the row
// sums are formed, so the interpolated field reuses their space), the row
// sums sum_j a_ij ry^j per (tile row, coarse column), and the coarse index
// and relative position of every tile column and row.
struct BicubicLds {
  union {
    float coef[kBiMaxCoarse][17];  // +1 against bank conflicts
    float d_eta[kBiH + 2][kBiW + 2];
  };
  float tv[kBiH + 2][kBiMaxCw][4];
  float rx[kBiW + 2], ry[kBiH + 2];
  int ci[kBiW + 2], cj[kBiH + 2];
};
// Coarse-grid model error of one tile of kBiH x kBiW interior cells whose
// first cell is (by, bx): bicubic interpolation of the coarse field to the
// tile and its one-cell ring, and its geostrophic balance added to the state.
// A fine point's coarse column and x position depend on its column only, its
// coarse row and y position on its row only; so when the tile spans few
// enough coarse cells, the coefficients are formed once per coarse cell, the
// row sums t_i = sum_j a_ij ry^j once per (row, coarse column), and each point
// takes sum_i rx^i t_i: bicubic_eval's operations in bicubic_eval's order.
// One 256-thread block; L is its LDS.
__device__ __forceinline__ void bicubic_tile(const BicubicGeo& q, float g, float f, float beta, float north_x,
                                             float north_y, const float* f_cells, int f_pitch, const float* coarse,
                                             int c_pitch, float* eta, float* hu, float* hv, int pitch,
                                             const float* Hi, int h_pitch, float land_value, const int bx,
                                             const int by, BicubicLds& L) {
  const int tid = threadIdx.x;
  // the coarse cells under the tile and its ring: the spot of a fine point
  // never decreases with its index, so the corners bound them
  const CoarseSpot lo = coarse_spot(q.cnx, q.cny, q.dx, q.dy, q.cdx, q.cdy, q.oi, q.oj, by - 1, bx - 1);
  const CoarseSpot hi = coarse_spot(q.cnx, q.cny, q.dx, q.dy, q.cdx, q.cdy, q.oi, q.oj, by + kBiH, bx + kBiW);
  const int ncw = hi.ci - lo.ci + 1, nch = hi.cj - lo.cj + 1;
  const bool shared = ncw * nch <= kBiMaxCoarse && ncw <= kBiMaxCw;  // false only for factors below 3
  if (shared) {
    for (int k = tid; k < ncw * nch; k += 256) {
      float a[16];
      bicubic_coeffs(q.cnx, q.cny, coarse, c_pitch, lo.cj + k / ncw, lo.ci + k % ncw, a);
#pragma unroll
      for (int i = 0; i < 16; ++i) L.coef[k][i] = a[i];
    }
    for (int k = tid; k < (kBiW + 2) + (kBiH + 2); k += 256) {
      int c;
      float r;
      if (k < kBiW + 2) {
        coarse_axis(q.cnx, q.dx, q.cdx, q.oi, bx + k - 1, &c, &r);
        L.ci[k] = c - lo.ci;
        L.rx[k] = r;
      } else {
        coarse_axis(q.cny, q.dy, q.cdy, q.oj, by + (k - (kBiW + 2)) - 1, &c, &r);
        L.cj[k - (kBiW + 2)] = c - lo.cj;
        L.ry[k - (kBiW + 2)] = r;
      }
    }
    __syncthreads();
    for (int k = tid; k < (kBiH + 2) * ncw; k += 256) {
      const int j = k / ncw, c = k - j * ncw;
      const float* a = L.coef[L.cj[j] * ncw + c];
      const float ry = L.ry[j], ry2 = ry * ry;
      const float yv[4] = {1.0f, ry, ry2, ry2 * ry};
#pragma unroll
      for (int i = 0; i < 4; ++i)
        L.tv[j][c][i] = dot4(a[4 * i], a[4 * i + 1], a[4 * i + 2], a[4 * i + 3], yv[0], yv[1], yv[2], yv[3]);
    }
    __syncthreads();
    for (int k = tid; k < (kBiH + 2) * (kBiW + 2); k += 256) {
      const int j = k / (kBiW + 2), i = k % (kBiW + 2);
      const float rx = L.rx[i], rx2 = rx * rx;
      const float* t = L.tv[j][L.ci[i]];
      L.d_eta[j][i] = dot4(1.0f, rx, rx2, rx2 * rx, t[0], t[1], t[2], t[3]);
    }
  } else {
    for (int k = tid; k < (kBiH + 2) * (kBiW + 2); k += 256) {
      const int j = k / (kBiW + 2), i = k % (kBiW + 2);
      const CoarseSpot p = coarse_spot(q.cnx, q.cny, q.dx, q.dy, q.cdx, q.cdy, q.oi, q.oj, by + j - 1, bx + i - 1);
      float a[16];
      bicubic_coeffs(q.cnx, q.cny, coarse, c_pitch, p.cj, p.ci, a);
      L.d_eta[j][i] = bicubic_eval(a, p.rx, p.ry);
    }
  }
  __syncthreads();
  const int x = (tid & 63) + 1;
  for (int y = (tid >> 6) + 1; y <= kBiH; y += 4) {
    const int ti = bx + x - 1, tj = by + y - 1;
    if (ti >= q.nx + 2 || tj >= q.ny + 2) continue;
    bicubic_balance_apply(q.dx, q.dy, g, f, beta, north_x, north_y, f_cells, f_pitch, L.d_eta[y][x],
                          L.d_eta[y][x - 1], L.d_eta[y][x + 1], L.d_eta[y - 1][x], L.d_eta[y + 1][x], eta, hu, hv,
                          pitch, Hi, h_pitch, land_value, tj, ti);
  }
}

}  // namespace gpo
