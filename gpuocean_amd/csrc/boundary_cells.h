// Per-cell bodies of the ghost-cell boundary conditions, shared by the
// per-member launches (aux_kernels.hip) and the batched ensemble
// (ensemble.hip).  Callers compile with -ffp-contract=off.
#pragma once

#include <hip/hip_runtime.h>

namespace gpo {
namespace cells {

__device__ inline float* bc_row(float* base, int pitch_bytes, int j) {
  return (float*)((char*)base + (size_t)pitch_bytes * j);
}

// periodicBoundary_NS (boundary_kernels.cu:36-72): ghost row g (0, 1, 2, 3 =
// rows 0, 1, ny+2, ny+3) copies the opposite interior row, at column ti
// (every column including ghost columns).
__device__ inline void periodic_ns_cell(int nx, int ny, float* h, float* u, float* v, int pitch, int ti, int g) {
  const int tj = (g < 2) ? g : ny + g;
  const int opp = (tj > ny + 1) ? tj - ny : tj + ny;
  bc_row(h, pitch, tj)[ti] = bc_row(h, pitch, opp)[ti];
  bc_row(u, pitch, tj)[ti] = bc_row(u, pitch, opp)[ti];
  bc_row(v, pitch, tj)[ti] = bc_row(v, pitch, opp)[ti];
}

// periodicBoundary_EW (boundary_kernels.cu:77-110): ghost column g of row tj.
__device__ inline void periodic_ew_cell(int nx, int ny, float* h, float* u, float* v, int pitch, int tj, int g) {
  const int ti = (g < 2) ? g : nx + g;
  const int opp = (ti > nx + 1) ? ti - nx : ti + nx;
  float* hr = bc_row(h, pitch, tj);
  float* ur = bc_row(u, pitch, tj);
  float* vr = bc_row(v, pitch, tj);
  hr[ti] = hr[opp];
  ur[ti] = ur[opp];
  vr[ti] = vr[opp];
}

// Normalised bilinear lookup of a 3-component texture (interpolation.cu:137-190).
__device__ inline float3 bilinear3(const float* data, int dnx, int dny, float nx_, float ny_) {
  nx_ = fminf(fmaxf(nx_ - 0.5f / (float)dnx, 0.0f), 1.0f);
  ny_ = fminf(fmaxf(ny_ - 0.5f / (float)dny, 0.0f), 1.0f);
  const float x = nx_ * (float)dnx, y = ny_ * (float)dny;
  int x0 = (int)floorf(x), y0 = (int)floorf(y);
  const float fx = x - (float)x0, fy = y - (float)y0;
  x0 = min(x0, dnx - 1);
  y0 = min(y0, dny - 1);
  const int x1 = min(x0 + 1, dnx - 1), y1 = min(y0 + 1, dny - 1);
  float res[3];
  for (int c = 0; c < 3; ++c) {
    const float d00 = data[(y0 * dnx + x0) * 3 + c], d01 = data[(y0 * dnx + x1) * 3 + c];
    const float d10 = data[(y1 * dnx + x0) * 3 + c], d11 = data[(y1 * dnx + x1) * 3 + c];
    res[c] = d00 + fx * (d01 - d00) + fy * (d10 - d00) + fx * fy * (d00 + d11 - d01 - d10);
  }
  return make_float3(res[0], res[1], res[2]);
}

// flowRelaxationScheme_NS (boundary_kernels.cu:406-472) at cell (ti, tj); cells
// outside the south/north sponges are left alone.
// linearInterpolation_NS (gpu_kernels/boundary_kernels.cu:265-325): on a
// code-4 side, the rows strictly between the outer ghost row and the row
// `sponge` cells in are interpolated linearly between those two rows, which
// are kept (two ghost cells).  Within one launch no cell reads a row another
// cell writes.
__device__ inline void lin_ns_cell(int bc_n, int bc_s, int nx, int ny, int sn, int ss, float* h, float* u, float* v,
                                   int pitch, int ti, int tj) {
  const int top = ny + 3;
  const bool in = ((bc_s == 4 && tj < ss && tj > 0) || (bc_n == 4 && tj > top - sn && tj < top)) && ti > 0 &&
                  ti < nx + 3;
  if (!in) return;
  int inner = ss, outer = 0;
  if (tj > ss) inner = top - sn, outer = top;
  const float ratio = (float)(tj - outer) / (float)(inner - outer);
  for (float* a : {h, u, v}) {
    const float vi = bc_row(a, pitch, inner)[ti], vo = bc_row(a, pitch, outer)[ti];
    bc_row(a, pitch, tj)[ti] = vo + ratio * (vi - vo);
  }
}

// linearInterpolation_EW (gpu_kernels/boundary_kernels.cu:330-384), by columns.
__device__ inline void lin_ew_cell(int bc_e, int bc_w, int nx, int ny, int se, int sw, float* h, float* u, float* v,
                                   int pitch, int ti, int tj) {
  const int right = nx + 3;
  const bool in = ((bc_w == 4 && ti < sw && ti > 0) || (bc_e == 4 && ti > right - se && ti < right)) && tj > 0 &&
                  tj < ny + 3;
  if (!in) return;
  int inner = sw, outer = 0;
  if (ti > sw) inner = right - se, outer = right;
  const float ratio = (float)(ti - outer) / (float)(inner - outer);
  for (float* a : {h, u, v}) {
    float* r = bc_row(a, pitch, tj);
    r[ti] = r[outer] + ratio * (r[inner] - r[outer]);
  }
}

__device__ inline void frs_ns_cell(int bc_n, int bc_s, int nx, int ny, int sn, int ss, float* h, float* u, float* v,
                                   int pitch, const float* cur, const float* next, int bc_nx, float t_, int ti,
                                   int tj) {
  const bool in_s = (bc_s == 3) && (tj < ss);
  const bool in_n = (bc_n == 3) && (tj > ny + 4 - 1 - sn);
  if (!(in_s || in_n)) return;
  const float j = (tj <= ss) ? (float)(tj - 1) : (float)((ny + 4 - 2) - tj);
  const float alpha = 1.0f - tanhf(fmaxf(0.0f, j) / 2.0f);
  const float s = (float)ti / (float)(nx + 4);
  const float t = ((double)tj < ny / 2.0) ? 0.25f : 0.75f;
  const float3 c = bilinear3(cur, bc_nx, 2, s, t);
  const float3 n = bilinear3(next, bc_nx, 2, s, t);
  const float eh = t_ * n.x + (1.0f - t_) * c.x;
  const float eu = t_ * n.y + (1.0f - t_) * c.y;
  const float ev = t_ * n.z + (1.0f - t_) * c.z;
  float* hr = bc_row(h, pitch, tj);
  float* ur = bc_row(u, pitch, tj);
  float* vr = bc_row(v, pitch, tj);
  hr[ti] = (1.0f - alpha) * hr[ti] + alpha * eh;
  ur[ti] = (1.0f - alpha) * ur[ti] + alpha * eu;
  vr[ti] = (1.0f - alpha) * vr[ti] + alpha * ev;
}

// flowRelaxationScheme_EW (boundary_kernels.cu:477-540) at cell (ti, tj).
__device__ inline void frs_ew_cell(int bc_e, int bc_w, int nx, int ny, int se, int sw, float* h, float* u, float* v,
                                   int pitch, const float* cur, const float* next, int bc_ny, float t_, int ti,
                                   int tj) {
  const bool in_w = (bc_w == 3) && (ti < sw);
  const bool in_e = (bc_e == 3) && (ti > nx + 4 - 1 - se);
  if (!(in_w || in_e)) return;
  const float j = (ti <= sw) ? (float)(ti - 1) : (float)((nx + 4 - 2) - ti);
  const float alpha = 1.0f - tanhf(fmaxf(0.0f, j) / 2.0f);
  const float s = ((double)ti < nx / 2.0) ? 0.25f : 0.75f;
  const float t = (float)tj / (float)(ny + 4);
  const float3 c = bilinear3(cur, 2, bc_ny, s, t);
  const float3 n = bilinear3(next, 2, bc_ny, s, t);
  const float eh = t_ * n.x + (1.0f - t_) * c.x;
  const float eu = t_ * n.y + (1.0f - t_) * c.y;
  const float ev = t_ * n.z + (1.0f - t_) * c.z;
  float* hr = bc_row(h, pitch, tj);
  float* ur = bc_row(u, pitch, tj);
  float* vr = bc_row(v, pitch, tj);
  hr[ti] = (1.0f - alpha) * hr[ti] + alpha * eh;
  ur[ti] = (1.0f - alpha) * ur[ti] + alpha * eu;
  vr[ti] = (1.0f - alpha) * vr[ti] + alpha * ev;
}

// The ghost-cell passes of gpo_boundary_conditions for one state of one
// member, as lists of cells: the north/south pass (periodic rows, or the
// flow-relaxation sponge rows: the first lo and the last hi rows, all
// columns) and the east/west pass (likewise columns).
struct BcTarget {
  int nx, ny, bn, be, bs, bw;
  int sn, se, ss, sw;
  int ns_rows_lo, ns_rows_hi, ew_cols_lo, ew_cols_hi;
  float *e, *u, *v;
  int pitch;
  const float *ns_cur, *ns_next, *ew_cur, *ew_next;
  int ns_nx, ew_ny;
  float frs_t;
};

// Rows and columns each pass writes: periodic when N (E) is 2, else flow
// relaxation when N or S (E or W) is 3, else none (walls).
inline void bc_target_extent(BcTarget* b) {
  b->ns_rows_lo = b->ns_rows_hi = b->ew_cols_lo = b->ew_cols_hi = 0;
  if (b->bn == 2) {
    b->ns_rows_lo = b->ns_rows_hi = 2;
  } else if (b->bn == 3 || b->bs == 3) {
    const int lo = b->bs == 3 ? (b->ss > 0 ? b->ss : 0) : 0, hi = b->bn == 3 ? (b->sn > 0 ? b->sn : 0) : 0;
    if (lo + hi >= b->ny + 4) b->ns_rows_lo = b->ny + 4;
    else b->ns_rows_lo = lo, b->ns_rows_hi = hi;
  }
  if (b->be == 2) {
    b->ew_cols_lo = b->ew_cols_hi = 2;
  } else if (b->be == 3 || b->bw == 3) {
    const int lo = b->bw == 3 ? (b->sw > 0 ? b->sw : 0) : 0, hi = b->be == 3 ? (b->se > 0 ? b->se : 0) : 0;
    if (lo + hi >= b->nx + 4) b->ew_cols_lo = b->nx + 4;
    else b->ew_cols_lo = lo, b->ew_cols_hi = hi;
  }
}
inline long long bc_target_cells(const BcTarget& b) {
  const long long ns = (long long)(b.ns_rows_lo + b.ns_rows_hi) * (b.nx + 4);
  const long long ew = (long long)(b.ew_cols_lo + b.ew_cols_hi) * (b.ny + 4);
  return ns > ew ? ns : ew;
}

// Cell k of pass `pass` (0: north/south, 1: east/west) of target b.
__device__ inline void bc_target_cell(const BcTarget& b, int pass, int k) {
  if (pass == 0) {
    const int w = b.nx + 4, rows = b.ns_rows_lo + b.ns_rows_hi;
    if (k >= rows * w) return;
    const int r = k / w, ti = k - r * w;
    const int tj = r < b.ns_rows_lo ? r : (b.ny + 4 - b.ns_rows_hi) + (r - b.ns_rows_lo);
    if (b.bn == 2)
      periodic_ns_cell(b.nx, b.ny, b.e, b.u, b.v, b.pitch, ti, tj < 2 ? tj : tj - b.ny);
    else
      frs_ns_cell(b.bn, b.bs, b.nx, b.ny, b.sn, b.ss, b.e, b.u, b.v, b.pitch, b.ns_cur, b.ns_next, b.ns_nx, b.frs_t, ti,
                  tj);
  } else {  // row by row, so that the lanes of a wavefront share rows
    const int hgt = b.ny + 4, cols = b.ew_cols_lo + b.ew_cols_hi;
    if (k >= cols * hgt) return;
    const int tj = k / cols, c = k - tj * cols;
    const int ti = c < b.ew_cols_lo ? c : (b.nx + 4 - b.ew_cols_hi) + (c - b.ew_cols_lo);
    if (b.be == 2)
      periodic_ew_cell(b.nx, b.ny, b.e, b.u, b.v, b.pitch, tj, ti < 2 ? ti : ti - b.nx);
    else
      frs_ew_cell(b.be, b.bw, b.nx, b.ny, b.se, b.sw, b.e, b.u, b.v, b.pitch, b.ew_cur, b.ew_next, b.ew_ny, b.frs_t, ti,
                  tj);
  }
}

}  // namespace cells
}  // namespace gpo
