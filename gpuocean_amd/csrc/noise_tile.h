// The fused model-error tile shared by the product's batched ensemble
// (ensemble.hip, noise_fused) and the differentiable batched ensemble
// (adjoint.hip): normal field, SOAR and geostrophic balance of one 64 x 16
// tile of interior cells in one block.  The differentiable path also keeps
// eta before the model error (KEEP_PRE: the reverse of the substep that
// produced it tests it for cells clamped dry) and the SOAR field itself
// (STORE_D), which the reverse of the balance reads.
#pragma once

#include <cstdint>

#include "noise_cells.h"
#include "philox.h"

namespace gpo {

// One member's model error: the arguments of a tile.
struct NoiseTile {
  int nx, ny, per_ns, per_ew;
  int xi_cols, pitch, hi_pitch, f_pitch;  // pitch, hi_pitch, f_pitch in bytes
  uint32_t k0, k1, stream_id, draw;
  const float* W;  // 25 SOAR weights
  float scale, dx, dy, g, f, beta, north_x, north_y, land;
  float *eta, *hu, *hv;
  const float* Hi;
  const float* f_cells;
  float* pre_e;                   // KEEP_PRE: eta before the model error (same pitch)
  float* d_out;                   // STORE_D: the SOAR field, dense rows of d_row elements
  int d_row;
};

// Model error of one output tile (kTileH rows x kTileW columns of interior
// cells of one member), fused: the normal field of the tile and its halo is
// generated into LDS (one Philox call per four consecutive elements, as the
// element index orders them), the SOAR field of the tile and a one-cell halo
// is formed from it in LDS, and the geostrophic balance is added to the
// state.  Every value equals the per-member path's (same element indices,
// same Box-Muller, same fused multiply-add order, same balance arithmetic).
constexpr int kTileW = 64, kTileH = 16;
constexpr int kXiW = kTileW + 6, kXiH = kTileH + 6;  // xi: coarse halo 1 + SOAR radius 2 (+ column offset)
constexpr int kCoW = kTileW + 2, kCoH = kTileH + 2;  // coarse: balance halo 1

// x mod n for n >= 1, without a division when x lies in [-n, 2n).
__device__ __forceinline__ int wrap(int x, int n) {
  if (x < 0) x += n;
  if (x >= n) x -= n;
  if (x < 0 || x >= n) x = ((x % n) + n) % n;
  return x;
}

template <bool STORE_D, bool KEEP_PRE>
__device__ __forceinline__ void noise_tile(const NoiseTile& m, int T0, int S0) {
  __shared__ float xs[kXiH][kXiW];
  __shared__ float cs[kCoH][kCoW];
  const uint32_t d = m.draw;
  // coarse cells needed: rows T0-1 .. cj_hi, columns S0-1 .. ci_hi
  const int cj_hi = min(T0 + kTileH, m.ny + 2), ci_hi = min(S0 + kTileW, m.nx + 2);
  // xi virtual rows V = T0-1+lr and columns U = S0-1+lc are needed up to cj_hi+4, ci_hi+4
  const int nrow = cj_hi + 4 - (T0 - 1) + 1, ncol = ci_hi + 4 - (S0 - 1) + 1;
  const int cols = m.xi_cols;

  // 0. the state and depths of this thread's four cells, loaded first so that
  // the HBM reads overlap the random numbers and SOAR below
  const int lx = threadIdx.x & 63, ly0 = threadIdx.x >> 6;
  constexpr int kRowsPerThread = kTileH / 4;
  BalanceIn in[kRowsPerThread];
#pragma unroll
  for (int k = 0; k < kRowsPerThread; ++k) {
    const int tj = T0 + ly0 + 4 * k, ti = S0 + lx;
    if (tj <= m.ny + 1 && ti <= m.nx + 1)
      in[k] = balance_load(m.f == 0.0f ? m.f_cells : nullptr, m.f_pitch, m.eta, m.hu, m.hv, m.pitch, m.Hi,
                           m.hi_pitch, tj, ti);
  }
  if (KEEP_PRE) {  // eta before the model error
#pragma unroll
    for (int k = 0; k < kRowsPerThread; ++k) {
      const int tj = T0 + ly0 + 4 * k, ti = S0 + lx;
      if (tj <= m.ny + 1 && ti <= m.nx + 1) mrow(m.pre_e, m.pitch, tj)[ti] = in[k].e;
    }
  }

  // 1. normal field into LDS
  const int a0 = m.per_ew ? wrap(S0 - 1 - 4, m.nx) : (S0 - 1);  // actual column of U = S0-1
  if (!m.per_ew || ncol <= m.nx) {
    // the needed actual columns form one run, or two when a periodic run wraps
    const int seg_len0 = m.per_ew ? min(ncol, m.nx - a0) : ncol;
    constexpr int kMaxQ = kXiW / 4 + 2;
    // work items (segment, row, quad): the second segment exists only where a
    // periodic window wraps, so most tiles keep every lane busy
    const int n_seg = seg_len0 < ncol ? 2 : 1;
    const int per_seg = nrow * kMaxQ;
    for (int u = threadIdx.x; u < n_seg * per_seg; u += 256) {
      const int seg = u >= per_seg ? 1 : 0, rem = u - seg * per_seg, lr = rem / kMaxQ, j = rem - lr * kMaxQ;
      const int V = T0 - 1 + lr;
      const int r = m.per_ns ? wrap(V - 4, m.ny) : V;
      const int ca = seg == 0 ? a0 : 0;
      const int len = seg == 0 ? seg_len0 : ncol - seg_len0;
      if (len <= 0) continue;
      const int cb = ca + len - 1;
      const uint32_t base = (uint32_t)r * (uint32_t)cols;
      const uint32_t q = ((base + (uint32_t)ca) >> 2) + (uint32_t)j;
      if (q > ((base + (uint32_t)cb) >> 2)) continue;
      float v4[4];
      normals4(q, d, m.stream_id, m.k0, m.k1, v4);
      // window column of actual column c: c - a0 in the first run, c - a0 + nx
      // in the wrapped second run (a0 = S0 - 1 without periodicity)
      const int c0 = (int)(4u * q - base), shift = seg == 0 ? -a0 : m.nx - a0;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int c = c0 + i;
        if (c < ca || c > cb) continue;
        xs[lr][c + shift] = v4[i];
      }
    }
  } else {
    // periodic member narrower than the tile window: element by element
    for (int u = threadIdx.x; u < kXiH * kXiW; u += 256) {
      const int lr = u / kXiW, lc = u - lr * kXiW;
      if (lr >= nrow || lc >= ncol) continue;
      const int V = T0 - 1 + lr, U = S0 - 1 + lc;
      const int r = m.per_ns ? (V - 4 + 2 * m.ny) % m.ny : V;
      const int c = (U - 4 + 2 * m.nx) % m.nx;
      xs[lr][lc] = normal_at((uint32_t)r * (uint32_t)cols + (uint32_t)c, d, m.stream_id, m.k0, m.k1);
    }
  }
  __syncthreads();

  // 2. SOAR field on the tile and its one-cell halo (soar_cell's order)
  for (int u = threadIdx.x; u < kCoH * kCoW; u += 256) {
    const int lr = u / kCoW, lc = u - lr * kCoW;
    if (T0 - 1 + lr > cj_hi || S0 - 1 + lc > ci_hi) continue;
    float acc = 0.0f;
#pragma unroll
    for (int k = 0; k < 5; ++k)
#pragma unroll
      for (int l = 0; l < 5; ++l) acc = __builtin_fmaf(m.W[k * 5 + l], xs[lr + k][lc + l], acc);
    cs[lr][lc] = m.scale * acc;
    // the SOAR field of the tile's cells, and of the halo cells on the ring
    // around the interior (rows 1 and ny+2, columns 1 and nx+2; not the
    // corners), which the balance reads at the grid edge
    if (STORE_D) {
      const int cj = T0 - 1 + lr, ci = S0 - 1 + lc;
      const bool own = lr >= 1 && lr <= kTileH && lc >= 1 && lc <= kTileW;
      const bool ring = (cj == 1 || cj == m.ny + 2) != (ci == 1 || ci == m.nx + 2);
      if (own || ring) m.d_out[(size_t)cj * m.d_row + ci] = cs[lr][lc];
    }
  }
  __syncthreads();

  // 3. geostrophic balance on the tile
#pragma unroll
  for (int k = 0; k < kRowsPerThread; ++k) {
    const int ly = ly0 + 4 * k;
    const int tj = T0 + ly, ti = S0 + lx;
    if (tj > m.ny + 1 || ti > m.nx + 1) continue;
    float e, u, v;
    balance_compute(m.dx, m.dy, m.g, m.f, m.beta, m.north_x, m.north_y, cs[ly + 1][lx + 1], cs[ly + 1][lx],
                    cs[ly + 1][lx + 2], cs[ly][lx + 1], cs[ly + 2][lx + 1], in[k], m.land, tj, ti, &e, &u, &v);
    balance_store(m.eta, m.hu, m.hv, m.pitch, tj, ti, e, u, v);
  }
}


}  // namespace gpo
