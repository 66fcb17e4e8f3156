// Differentiable CDKLM16 substep on gfx950: kernels over the per-cell stages
// of adjoint_core.h (FP64 forward, tangent-linear model, FP32/FP64 adjoint),
// the reverse of the geostrophic-balance model error, the point-observation
// misfit and their C ABI.  Built with -ffp-contract=off.

#include <cmath>
#include <mutex>
#include <cstdint>
#include <algorithm>
#include <cstdio>
#include <new>
#include <vector>

#include "adjoint_core.h"
#include "adjoint_stages.h"
#include "boundary_cells.h"
#include "cdklm_common.h"
#include "noise_cells.h"
#include "noise_tile.h"
#include "philox.h"

namespace gpo {
namespace adj {

// ----------------------------------------------------------------------------- kernels

template <class S>
__global__ void f1_convert(Geo g, const S* in_e, const S* in_u, const S* in_v, const S* Hi, Fwd<S> w, int* flag) {
  f1_cell(g, in_e, in_u, in_v, Hi, w, flag, blockIdx.y * blockDim.y + threadIdx.y,
          blockIdx.x * blockDim.x + threadIdx.x);
}
template <class S>
__global__ void f2_slopes(Geo g, const S* Hi, const S* fc, Fwd<S> w) {
  f2_cell(g, Hi, fc, w, blockIdx.y * blockDim.y + threadIdx.y, blockIdx.x * blockDim.x + threadIdx.x);
}
template <class S>
__global__ void f3_faces(Geo g, const S* Hi, const S* fc, Fwd<S> w, int* flag) {
  f3_cell(g, Hi, fc, w, flag, blockIdx.y * blockDim.y + threadIdx.y, blockIdx.x * blockDim.x + threadIdx.x);
}
template <class S>
__global__ void f4_update(Geo g, S dt, int rk_step, const S* Hi, const S* fc, Fwd<S> w, const S* qn_e,
                          const S* qn_u, const S* qn_v, S* out_e, S* out_u, S* out_v, int* flag) {
  f4_cell(g, dt, rk_step, Hi, fc, w, qn_e, qn_u, qn_v, out_e, out_u, out_v, flag,
          blockIdx.y * blockDim.y + threadIdx.y + 2, blockIdx.x * blockDim.x + threadIdx.x + 2);
}
// The reverse-stage kernels with IEEE arithmetic (the FP64 build).
namespace exact {
#include "adjoint_stages.inc"
void launch_stages(const Geo& g, double dt, int rk, const double* ie, const double* iu, const double* iv,
                   const double* out_e, const double* le, const double* lu, const double* lv, double* lie, double* liu,
                   double* liv, double* lqe, double* lqu, double* lqv, int accumulate, int lq_mode, const double* Hi,
                   const double* fc, void* ws, int* flag, hipStream_t s) {
  stages_t<double>(g, dt, rk, ie, iu, iv, out_e, le, lu, lv, lie, liu, liv, lqe, lqu, lqv, accumulate, lq_mode, Hi,
                   fc, ws, flag, s);
}
}  // namespace exact

// ---- flow relaxation (boundary_kernels.cu:406-540): q <- (1 - alpha) q + alpha q_ext
// on the sponge rows (north/south pass) and then columns (east/west pass),
// alpha = 1 - tanh(max(0, j) / 2) in float as the forward computes it.
struct Sponge {
  int sn, se, ss, sw;
};
__device__ inline float frs_alpha_ns(const Geo& g, const Sponge& sp, int tj, bool* in) {
  *in = (g.bc_s == 3 && tj < sp.ss) || (g.bc_n == 3 && tj > g.ny + 4 - 1 - sp.sn);
  const float j = (tj <= sp.ss) ? (float)(tj - 1) : (float)((g.ny + 4 - 2) - tj);
  return 1.0f - tanhf(fmaxf(0.0f, j) / 2.0f);
}
__device__ inline float frs_alpha_ew(const Geo& g, const Sponge& sp, int ti, bool* in) {
  *in = (g.bc_w == 3 && ti < sp.sw) || (g.bc_e == 3 && ti > g.nx + 4 - 1 - sp.se);
  const float j = (ti <= sp.sw) ? (float)(ti - 1) : (float)((g.nx + 4 - 2) - ti);
  return 1.0f - tanhf(fmaxf(0.0f, j) / 2.0f);
}

// Transpose (and tangent) of the relaxation: a <- (1 - alpha) a per pass, the
// east/west pass first (it was applied last).
template <class T>
__global__ void frs_scale(Geo g, Sponge sp, T* a, T* b, T* c) {
  const int ti = blockIdx.x * blockDim.x + threadIdx.x, tj = blockIdx.y * blockDim.y + threadIdx.y;
  if (ti >= g.nx + 4 || tj >= g.ny + 4) return;
  const size_t k = (size_t)tj * g.sp + ti;
  bool in_ew, in_ns;
  const float aew = frs_alpha_ew(g, sp, ti, &in_ew), ans = frs_alpha_ns(g, sp, tj, &in_ns);
  if (in_ew) {
    const T f = T(1.0f - aew);
    a[k] *= f, b[k] *= f, c[k] *= f;
  }
  if (in_ns) {
    const T f = T(1.0f - ans);
    a[k] *= f, b[k] *= f, c[k] *= f;
  }
}

__device__ inline float3 frs_bilinear3(const float* data, int dnx, int dny, float nx_, float ny_) {
  nx_ = fminf(fmaxf(nx_ - 0.5f / (float)dnx, 0.0f), 1.0f);
  ny_ = fminf(fmaxf(ny_ - 0.5f / (float)dny, 0.0f), 1.0f);
  const float x = nx_ * (float)dnx, y = ny_ * (float)dny;
  int x0 = (int)floorf(x), y0 = (int)floorf(y);
  const float fx = x - (float)x0, fy = y - (float)y0;
  x0 = min(x0, dnx - 1);
  y0 = min(y0, dny - 1);
  const int x1 = min(x0 + 1, dnx - 1), y1 = min(y0 + 1, dny - 1);
  float r[3];
  for (int q = 0; q < 3; ++q) {
    const float d00 = data[(y0 * dnx + x0) * 3 + q], d01 = data[(y0 * dnx + x1) * 3 + q];
    const float d10 = data[(y1 * dnx + x0) * 3 + q], d11 = data[(y1 * dnx + x1) * 3 + q];
    r[q] = d00 + fx * (d01 - d00) + fy * (d10 - d00) + fx * fy * (d00 + d11 - d01 - d10);
  }
  return make_float3(r[0], r[1], r[2]);
}

// The relaxation itself on an FP64 state (exterior values from the float
// boundary data, as the FP32 kernels form them).
__global__ void frs_blend_f64(Geo g, Sponge sp, const float* ns_cur, const float* ns_next, int ns_nx,
                              const float* ew_cur, const float* ew_next, int ew_ny, float t_, double* e, double* u,
                              double* v) {
  const int ti = blockIdx.x * blockDim.x + threadIdx.x, tj = blockIdx.y * blockDim.y + threadIdx.y;
  if (ti >= g.nx + 4 || tj >= g.ny + 4) return;
  const size_t k = (size_t)tj * g.sp + ti;
  bool in_ns, in_ew;
  const float ans = frs_alpha_ns(g, sp, tj, &in_ns), aew = frs_alpha_ew(g, sp, ti, &in_ew);
  if (in_ns) {
    const float s = (float)ti / (float)(g.nx + 4), t = ((double)tj < g.ny / 2.0) ? 0.25f : 0.75f;
    const float3 c = frs_bilinear3(ns_cur, ns_nx, 2, s, t), n = frs_bilinear3(ns_next, ns_nx, 2, s, t);
    const double a = ans;
    e[k] = (1.0 - a) * e[k] + a * (double)(t_ * n.x + (1.0f - t_) * c.x);
    u[k] = (1.0 - a) * u[k] + a * (double)(t_ * n.y + (1.0f - t_) * c.y);
    v[k] = (1.0 - a) * v[k] + a * (double)(t_ * n.z + (1.0f - t_) * c.z);
  }
  if (in_ew) {
    const float s = ((double)ti < g.nx / 2.0) ? 0.25f : 0.75f, t = (float)tj / (float)(g.ny + 4);
    const float3 c = frs_bilinear3(ew_cur, 2, ew_ny, s, t), n = frs_bilinear3(ew_next, 2, ew_ny, s, t);
    const double a = aew;
    e[k] = (1.0 - a) * e[k] + a * (double)(t_ * n.x + (1.0f - t_) * c.x);
    u[k] = (1.0 - a) * u[k] + a * (double)(t_ * n.y + (1.0f - t_) * c.y);
    v[k] = (1.0 - a) * v[k] + a * (double)(t_ * n.z + (1.0f - t_) * c.z);
  }
}

template <class T>
__global__ void frs_scale_b(const AdjItem<T>* __restrict__ tab) {
  const AdjItem<T>& it = tab[blockIdx.z];
  if (!it.frs) return;
  const int ti = blockIdx.x * blockDim.x + threadIdx.x, tj = blockIdx.y * blockDim.y + threadIdx.y;
  if (ti >= it.g.nx + 4 || tj >= it.g.ny + 4) return;
  const Sponge sp{it.sn, it.se, it.ss, it.sw};
  const size_t k = (size_t)tj * it.g.sp + ti;
  T* a = (T*)it.le;
  T* b = (T*)it.lu;
  T* c = (T*)it.lv;
  bool in_ew, in_ns;
  const float aew = frs_alpha_ew(it.g, sp, ti, &in_ew), ans = frs_alpha_ns(it.g, sp, tj, &in_ns);
  if (in_ew) {
    const T f = T(1.0f - aew);
    a[k] *= f, b[k] *= f, c[k] *= f;
  }
  if (in_ns) {
    const T f = T(1.0f - ans);
    a[k] *= f, b[k] *= f, c[k] *= f;
  }
}

// The transpose of the sponge blend over the sponge cells only (frs_scale_b
// over the whole grid, in its order: east/west factor first (pass 0), then
// north/south (pass 1)); block row y = member y.
template <class T>
__global__ void frs_scale_sponge_b(const AdjItem<T>* __restrict__ tab, int pass) {
  const AdjItem<T>& it = tab[blockIdx.y];
  if (!it.frs) return;
  const Geo& g = it.g;
  const int W = g.nx + 4, H = g.ny + 4;
  int lo, hi;  // sponge columns (pass 0) or rows (pass 1) at each end
  if (pass == 0) {
    lo = g.bc_w == 3 ? min(it.sw, W) : 0, hi = g.bc_e == 3 ? min(it.se, W) : 0;
    if (lo + hi > W) lo = W, hi = 0;
  } else {
    lo = g.bc_s == 3 ? min(it.ss, H) : 0, hi = g.bc_n == 3 ? min(it.sn, H) : 0;
    if (lo + hi > H) lo = H, hi = 0;
  }
  const int along = pass == 0 ? H : W;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (lo + hi) * along) return;
  // lanes follow rows in both passes (row by row over the sponge columns)
  const int c = pass == 0 ? k % (lo + hi) : k / along, t = pass == 0 ? k / (lo + hi) : k - c * along;
  const int across = c < lo ? c : (pass == 0 ? W : H) - hi + (c - lo);
  const int ti = pass == 0 ? across : t, tj = pass == 0 ? t : across;
  const Sponge sp{it.sn, it.se, it.ss, it.sw};
  bool in;
  const float a = pass == 0 ? frs_alpha_ew(g, sp, ti, &in) : frs_alpha_ns(g, sp, tj, &in);
  if (!in) return;
  const size_t kk = (size_t)tj * g.sp + ti;
  const T f = T(1.0f - a);
  T* x = (T*)it.le;
  T* y = (T*)it.lu;
  T* z = (T*)it.lv;
  x[kk] *= f, y[kk] *= f, z[kk] *= f;
}

// Ghost-cell passes of flow-relaxation members on one state each (block row
// y = item y; members without flow relaxation have no item).
__global__ void __launch_bounds__(256) bc_targets_b(const cells::BcTarget* __restrict__ tab, int pass) {
  cells::bc_target_cell(tab[blockIdx.y], pass, blockIdx.x * blockDim.x + threadIdx.x);
}

// Interior copy of one state per member: dst[m] = src[m] (3 planes each).
struct CopyItem {
  const GPO_GLOBAL float* src;
  GPO_GLOBAL float* dst;
  int nx, ny, sp;  // sp: row pitch in elements
  size_t plane;    // elements between the eta, hu, hv planes
};
__global__ void copy_states_b(const CopyItem* __restrict__ tab) {
  const CopyItem& it = tab[blockIdx.z];
  const int c = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y * blockDim.y + threadIdx.y;
  if (c >= it.nx + 4 || r >= it.ny + 4) return;
  const size_t k = (size_t)r * it.sp + c;
  for (int q = 0; q < 3; ++q) it.dst[q * it.plane + k] = it.src[q * it.plane + k];
}

// Dual <-> split value/tangent arrays for the tangent-linear entry point.
__global__ void pack_dual(Geo g, const double* v, const double* d, Dual* out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y * blockDim.y + threadIdx.y;
  if (c >= g.nx + 4 || r >= g.ny + 4) return;
  const size_t k = (size_t)r * g.sp + c;
  out[(size_t)r * g.W + c] = Dual(v[k], d ? d[k] : 0.0);
}
// Hi (the caller's, land unflagged; geometry gc) into a dense Dual copy with
// land corners flagged (rows of gh.W).
__global__ void pack_dual_hi(Geo gh, Geo gc, const double* Hi, Dual* out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y * blockDim.y + threadIdx.y;
  if (c >= gh.nx + 4 || r >= gh.ny + 4) return;
  out[(size_t)r * gh.W + c] = Dual(corner_land(gc, Hi, r, c), 0.0);
}
__global__ void unpack_dual(Geo g, const Dual* in, double* v, double* d) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x + 2, r = blockIdx.y * blockDim.y + threadIdx.y + 2;
  if (c > g.nx + 1 || r > g.ny + 1) return;
  const size_t k = (size_t)r * g.W + c, o = (size_t)r * g.sp + c;
  v[o] = in[k].v;
  d[o] = in[k].d;
}

// ----------------------------------------------------------------------------- model error and misfit

// Reverse of the state dependence of geostrophicBalance
// (gpu_kernels/ocean_noise.cu:256-390): hu += -g/f (d + H + eta) d_y d,
// hv += g/f (d + H + eta) d_x d, eta += d, so only eta feeds back:
// lambda_eta += -g/f d_y d lambda_hu + g/f d_x d lambda_hv.
template <class T>
__device__ void balance_rev_cell(const Geo& g, double beta, const float* f_cells, int f_pitch, const float* coarse,
                                 int c_pitch, const T* Hi, float land, T* l_e, T* l_u, T* l_v, int ti, int tj) {
  if (ti > g.nx + 1 || tj > g.ny + 1) return;
  const T h00 = corner(g, Hi, tj, ti), h01 = corner(g, Hi, tj, ti + 1);
  const T h10 = corner(g, Hi, tj + 1, ti), h11 = corner(g, Hi, tj + 1, ti + 1);
  if (h00 == T(land) || h01 == T(land) || h10 == T(land) || h11 == T(land)) {  // dry: state set to 0
    const size_t k = (size_t)tj * g.sp + ti;
    l_e[k] = T(0);
    l_u[k] = T(0);
    l_v[k] = T(0);
    return;
  }
  float cor = (float)g.f + (float)beta * (((float)ti + 0.5f) * (float)g.dx * (float)g.north_x +
                                          ((float)tj + 0.5f) * (float)g.dy * (float)g.north_y);
  if ((float)g.f == 0.0f) cor = f_cells ? ((const float*)((const char*)f_cells + (size_t)f_pitch * tj))[ti] : 0.0f;
  auto crow = [&](int j) { return (const float*)((const char*)coarse + (size_t)c_pitch * j); };
  const T ddx = T(crow(tj)[ti + 1] - crow(tj)[ti - 1]) / (T(2) * T(g.dx));
  const T ddy = T(crow(tj + 1)[ti] - crow(tj - 1)[ti]) / (T(2) * T(g.dy));
  const T gc = T(g.g) / T(cor);
  const size_t k = (size_t)tj * g.sp + ti;
  l_e[k] = l_e[k] - gc * ddy * l_u[k] + gc * ddx * l_v[k];
}
template <class T>
__global__ void balance_rev(Geo g, double beta, const float* f_cells, int f_pitch, const float* coarse, int c_pitch,
                            const T* Hi, float land, T* l_e, T* l_u, T* l_v) {
  balance_rev_cell<T>(g, beta, f_cells, f_pitch, coarse, c_pitch, Hi, land, l_e, l_u, l_v,
                      blockIdx.x * blockDim.x + threadIdx.x + 2, blockIdx.y * blockDim.y + threadIdx.y + 2);
}

// geostrophicBalance (gpu_kernels/ocean_noise.cu:256-390) on a state of type
// T, the verification counterpart of the product's FP32 balance, with its
// tangent (d* may be NULL): eta += d, hu -= g/f (d + H + eta) d_y d,
// hv += g/f (d + H + eta) d_x d with the eta before the update; a cell with a
// land corner is set to 0.  The coriolis value is formed in FP32 as the
// product forms it.
template <class T>
__global__ void balance_fwd(Geo g, double beta, const float* f_cells, int f_pitch, const float* coarse, int c_pitch,
                            const T* Hi, float land, T* e, T* u, T* v, T* de, T* du, T* dv) {
  const int ti = blockIdx.x * blockDim.x + threadIdx.x + 2, tj = blockIdx.y * blockDim.y + threadIdx.y + 2;
  if (ti > g.nx + 1 || tj > g.ny + 1) return;
  const size_t k = (size_t)tj * g.sp + ti;
  const T h00 = corner(g, Hi, tj, ti), h01 = corner(g, Hi, tj, ti + 1);
  const T h10 = corner(g, Hi, tj + 1, ti), h11 = corner(g, Hi, tj + 1, ti + 1);
  if (h00 == T(land) || h01 == T(land) || h10 == T(land) || h11 == T(land)) {
    e[k] = T(0), u[k] = T(0), v[k] = T(0);
    if (de) de[k] = T(0), du[k] = T(0), dv[k] = T(0);
    return;
  }
  float cor = (float)g.f + (float)beta * (((float)ti + 0.5f) * (float)g.dx * (float)g.north_x +
                                          ((float)tj + 0.5f) * (float)g.dy * (float)g.north_y);
  if ((float)g.f == 0.0f) cor = f_cells ? ((const float*)((const char*)f_cells + (size_t)f_pitch * tj))[ti] : 0.0f;
  auto crow = [&](int j) { return (const float*)((const char*)coarse + (size_t)c_pitch * j); };
  const T d = T(crow(tj)[ti]);
  const T ddx = T(crow(tj)[ti + 1] - crow(tj)[ti - 1]) / (T(2) * T(g.dx));
  const T ddy = T(crow(tj + 1)[ti] - crow(tj - 1)[ti]) / (T(2) * T(g.dy));
  const T gc = T(g.g) / T(cor);
  const T H_mid = T(0.25) * (h00 + h01 + h10 + h11);
  const T h_mid = d + H_mid + e[k];
  u[k] = u[k] - gc * h_mid * ddy;
  v[k] = v[k] + gc * h_mid * ddx;
  e[k] = e[k] + d;
  if (de) {
    du[k] = du[k] - gc * ddy * de[k];
    dv[k] = dv[k] + gc * ddx * de[k];
  }
}

// ----------------------------------------------------------------------------- drifters

// Where a drifter reads the state (passiveDrifterKernel, gpu_kernels/
// driftKernels.cu:74-183, reference cell 2): the lower-left of the four cell
// centres around it and its position between them, or !ok when the four
// cells are not all inside the ghost-inclusive arrays (it then stays put).
template <class T>
struct DriftAt {
  bool ok;
  int cx0, cy0;
  T xf, yf;
};
template <class T>
__device__ DriftAt<T> drift_at(int nx, int ny, T dx, T dy, T px, T py) {
  DriftAt<T> a{false, 0, 0, T(0), T(0)};
  const T qx = px / dx, qy = py / dy;
  const T fx = floor(qx), fy = floor(qy);
  const T fcx = fx + T(2), fcy = fy + T(2);
  if (!(fcx >= T(0) && fcx <= T(nx + 3) && fcy >= T(0) && fcy <= T(ny + 3))) return a;
  const int cx = (int)fcx, cy = (int)fcy;
  const T frx = qx - fx, fry = qy - fy;
  a.cx0 = frx < T(0.5) ? cx - 1 : cx;
  a.xf = frx < T(0.5) ? frx + T(0.5) : frx - T(0.5);
  a.cy0 = fry < T(0.5) ? cy - 1 : cy;
  a.yf = fry < T(0.5) ? fry + T(0.5) : fry - T(0.5);
  a.ok = !(a.cx0 < 0 || a.cy0 < 0 || a.cx0 + 1 > nx + 3 || a.cy0 + 1 > ny + 3);
  return a;
}

struct DriftGeo {
  int nx, ny, sp, bp;  // state and Bm pitches in elements
  int per_ns, per_ew;
  int n;               // drifters (rows of the position array, 2 values each)
  double dx, dy, dt, sens;
};

// Cell whose state a stencil cell reads: periodic ghosts hold copies of the
// interior cells they wrap (the ghost-cell pass runs before the drift), so
// they map there; other ghost cells keep stored constants (mapped = -1 row).
__device__ inline int wrap_axis(int c, int n, int periodic) {
  if (c >= 2 && c <= n + 1) return c;
  if (!periodic) return -1;
  return c < 2 ? c + n : c - n;
}

// Forward Euler drift of the positions in T with its tangent: d_pos and the
// state tangent d_* (NULL: none).  The FP64 counterpart of the product drift.
template <class T>
__global__ void drift_fwd(DriftGeo q, const T* e, const T* hu, const T* hv, const T* Bm, T* pos, T* dpos,
                          const T* de, const T* du, const T* dv) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= q.n) return;
  const T px = pos[2 * d], py = pos[2 * d + 1];
  const DriftAt<T> a = drift_at<T>(q.nx, q.ny, T(q.dx), T(q.dy), px, py);
  if (!a.ok) return;
  T uu[4], vv[4], hh[4], w[4], du_[4], dv_[4];
  const T wq[4] = {(T(1) - a.xf) * (T(1) - a.yf), a.xf * (T(1) - a.yf), (T(1) - a.xf) * a.yf, a.xf * a.yf};
  T u = T(0), v = T(0), tu = T(0), tv = T(0);
  for (int c = 0; c < 4; ++c) {
    const int cc = a.cx0 + (c & 1), rr = a.cy0 + (c >> 1);
    const int mc = wrap_axis(cc, q.nx, q.per_ew), mr = wrap_axis(rr, q.ny, q.per_ns);
    const bool dyn = mc >= 0 && mr >= 0;
    const size_t ks = dyn ? (size_t)mr * q.sp + mc : (size_t)rr * q.sp + cc;
    const T h = Bm[(size_t)rr * q.bp + cc] + e[ks];
    uu[c] = hu[ks] / h;
    vv[c] = hv[ks] / h;
    hh[c] = h;
    w[c] = wq[c];
    u = u + w[c] * uu[c];
    v = v + w[c] * vv[c];
    if (dpos && dyn && de) {
      const T dh = de[ks];
      du_[c] = (du[ks] - uu[c] * dh) / h;
      dv_[c] = (dv[ks] - vv[c] * dh) / h;
    } else {
      du_[c] = dv_[c] = T(0);
    }
  }
  const T s_dt = T(q.sens) * T(q.dt);
  T nx_ = px + T(q.sens) * u * T(q.dt), ny_ = py + T(q.sens) * v * T(q.dt);
  if (dpos) {
    const T dudx = ((T(1) - a.yf) * (uu[1] - uu[0]) + a.yf * (uu[3] - uu[2])) / T(q.dx);
    const T dudy = ((T(1) - a.xf) * (uu[2] - uu[0]) + a.xf * (uu[3] - uu[1])) / T(q.dy);
    const T dvdx = ((T(1) - a.yf) * (vv[1] - vv[0]) + a.yf * (vv[3] - vv[2])) / T(q.dx);
    const T dvdy = ((T(1) - a.xf) * (vv[2] - vv[0]) + a.xf * (vv[3] - vv[1])) / T(q.dy);
    T su = T(0), sv = T(0);
    for (int c = 0; c < 4; ++c) su = su + w[c] * du_[c], sv = sv + w[c] * dv_[c];
    const T dx0 = dpos[2 * d], dy0 = dpos[2 * d + 1];
    dpos[2 * d] = dx0 + s_dt * (dudx * dx0 + dudy * dy0 + su);
    dpos[2 * d + 1] = dy0 + s_dt * (dvdx * dx0 + dvdy * dy0 + sv);
  }
  const T Lx = T(q.nx) * T(q.dx), Ly = T(q.ny) * T(q.dy);
  if (q.per_ew && nx_ < T(0)) nx_ = nx_ + Lx;
  if (q.per_ew && nx_ > Lx) nx_ = nx_ - Lx;
  if (q.per_ns && ny_ < T(0)) ny_ = ny_ + Ly;
  if (q.per_ns && ny_ > Ly) ny_ = ny_ - Ly;
  pos[2 * d] = nx_;
  pos[2 * d + 1] = ny_;
  (void)hh;
}

// In-order accumulation of one chunk of contributions by a block of 256
// threads: item i (< kOrdItems) adds sv[i] = (u, v, eta) parts to cell
// sk[i] >> 10 of (l_u, l_v, l_e); sk[i] = (cell << 10) | i, or kNoKey for an
// item without a contribution, thread t owning items 4t..4t+3.  The items
// are sorted by (cell, item) (bitonic: partners within a thread in
// registers, within a wavefront by lane shuffles, across wavefronts through
// shared memory), and the first item of each cell adds that cell's items in
// item order: the sums do not depend on the order the hardware runs the
// items in, as float atomics would.  Chunks added one after another keep the
// order across chunks.
constexpr int kOrdItems = 1024;
constexpr uint64_t kNoKey = ~0ull;
__device__ inline uint64_t shfl_xor_u64(uint64_t v, int m) {
  const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, m), hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), m);
  return ((uint64_t)hi << 32) | lo;
}
template <class T>
__device__ void ordered_add(uint64_t* sk, const T (*sv)[3], T* l_u, T* l_v, T* l_e) {
  const int t = threadIdx.x;
  uint64_t key[4];
  for (int r = 0; r < 4; ++r) key[r] = sk[4 * t + r];
  for (int k = 2; k <= kOrdItems; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      uint64_t other[4];
      if (j >= 256) {  // partner in another wavefront (items 256w..256w+255 belong to wavefront w)
        __syncthreads();
        for (int r = 0; r < 4; ++r) sk[4 * t + r] = key[r];
        __syncthreads();
        for (int r = 0; r < 4; ++r) other[r] = sk[(4 * t + r) ^ j];
      } else if (j >= 4) {  // partner in lane t ^ (j / 4) of the same wavefront
        for (int r = 0; r < 4; ++r) other[r] = shfl_xor_u64(key[r], j >> 2);
      } else {
        for (int r = 0; r < 4; ++r) other[r] = key[r ^ j];
      }
      for (int r = 0; r < 4; ++r) {  // the lower index of a pair keeps the minimum in an ascending run
        const int x = 4 * t + r;
        const bool keep_min = ((x & j) == 0) == ((x & k) == 0);
        key[r] = keep_min ? (key[r] < other[r] ? key[r] : other[r]) : (key[r] < other[r] ? other[r] : key[r]);
      }
    }
  __syncthreads();
  for (int r = 0; r < 4; ++r) sk[4 * t + r] = key[r];
  __syncthreads();
  for (int i = threadIdx.x; i < kOrdItems; i += 256) {
    const uint64_t cell = sk[i] >> 10;
    if (sk[i] == kNoKey || (i > 0 && (sk[i - 1] >> 10) == cell)) continue;  // empty, or not the cell's first item
    T u = l_u[cell], v = l_v[cell], e = l_e[cell];
    for (int j = i; j < kOrdItems && sk[j] != kNoKey && (sk[j] >> 10) == cell; ++j) {
      const int it = (int)(sk[j] & (kOrdItems - 1));
      u = u + sv[it][0];
      v = v + sv[it][1];
      e = e + sv[it][2];
    }
    l_u[cell] = u;
    l_v[cell] = v;
    l_e[cell] = e;
  }
  __syncthreads();
}

// Reverse of one drift step at positions pos (before the step) over the state
// it read, by one block of 256 threads for all q.n drifters (chunks of 256):
// lam_pos (in: at the new positions; out: at the old ones) and the state
// sensitivities l_* (accumulated in drifter order; cells whose values are
// stored constants receive nothing).
template <class T>
__device__ void drift_rev_block(const DriftGeo& q, const T* e, const T* hu, const T* hv, const T* Bm, const T* pos,
                                T* lpos, T* l_e, T* l_u, T* l_v) {
  __shared__ uint64_t sk[kOrdItems];
  __shared__ T sv[kOrdItems][3];
  const int t = threadIdx.x;
  for (int base = 0; base < q.n; base += 256) {
    const int d = base + t;
    for (int c = 0; c < 4; ++c) sk[4 * t + c] = kNoKey;
    const DriftAt<T> a = d < q.n ? drift_at<T>(q.nx, q.ny, T(q.dx), T(q.dy), pos[2 * d], pos[2 * d + 1])
                                 : DriftAt<T>{false, 0, 0, T(0), T(0)};
    if (a.ok) {  // otherwise the drifter stayed put: lam_pos passes through unchanged
      const T s_dt = T(q.sens) * T(q.dt);
      const T lx = lpos[2 * d], ly = lpos[2 * d + 1];
      const T ga = s_dt * lx, gb = s_dt * ly;  // sensitivities of u and v
      const T wq[4] = {(T(1) - a.xf) * (T(1) - a.yf), a.xf * (T(1) - a.yf), (T(1) - a.xf) * a.yf, a.xf * a.yf};
      T uu[4], vv[4];
      for (int c = 0; c < 4; ++c) {
        const int cc = a.cx0 + (c & 1), rr = a.cy0 + (c >> 1);
        const int mc = wrap_axis(cc, q.nx, q.per_ew), mr = wrap_axis(rr, q.ny, q.per_ns);
        const bool dyn = mc >= 0 && mr >= 0;
        const size_t ks = dyn ? (size_t)mr * q.sp + mc : (size_t)rr * q.sp + cc;
        const T h = Bm[(size_t)rr * q.bp + cc] + e[ks];
        uu[c] = hu[ks] / h;
        vv[c] = hv[ks] / h;
        if (dyn) {
          const T a_u = ga * wq[c] / h, a_v = gb * wq[c] / h;
          sk[4 * t + c] = ((uint64_t)ks << 10) | (uint64_t)(4 * t + c);
          sv[4 * t + c][0] = a_u;
          sv[4 * t + c][1] = a_v;
          sv[4 * t + c][2] = -(a_u * uu[c] + a_v * vv[c]);
        }
      }
      const T dudx = ((T(1) - a.yf) * (uu[1] - uu[0]) + a.yf * (uu[3] - uu[2])) / T(q.dx);
      const T dudy = ((T(1) - a.xf) * (uu[2] - uu[0]) + a.xf * (uu[3] - uu[1])) / T(q.dy);
      const T dvdx = ((T(1) - a.yf) * (vv[1] - vv[0]) + a.yf * (vv[3] - vv[2])) / T(q.dx);
      const T dvdy = ((T(1) - a.xf) * (vv[2] - vv[0]) + a.xf * (vv[3] - vv[1])) / T(q.dy);
      lpos[2 * d] = lx + ga * dudx + gb * dvdx;
      lpos[2 * d + 1] = ly + ga * dudy + gb * dvdy;
    }
    __syncthreads();
    ordered_add<T>(sk, sv, l_u, l_v, l_e);
  }
}
template <class T>
__global__ void __launch_bounds__(256) drift_rev(DriftGeo q, const T* e, const T* hu, const T* hv, const T* Bm,
                                                 const T* pos, T* lpos, T* l_e, T* l_u, T* l_v) {
  drift_rev_block<T>(q, e, hu, hv, Bm, pos, lpos, l_e, l_u, l_v);
}

// ---- batched model error and drifters of the differentiable ensemble

// One member's model error (fixed normals) and drifters, with the buffers the
// batched forward, trajectory and sweep address by step.
struct NoiseDriftItem {
  Geo g;                 // state pitch sp, Hi pitch hp (elements), dx, dy, g, f, north
  int per_ns, per_ew;
  int noise;
  int xi_rows, xi_cols;
  uint32_t k0, k1, stream_id, draw0;
  float W[25];
  float beta, land;
  float *xi, *coarse;    // dense rows: xi_cols, nx + 4
  float* coarse_seg;     // k SOAR fields of the segment being recomputed (NULL: regenerated in the sweep)
  const float* Hi;
  float *state, *qs, *pre, *lam;  // forward state; segment states; pre-noise eta planes; lambda buffers
  size_t plane, st;      // elements per plane and per state
  int n_drifters;
  const float* Bm;
  int bm_pitch;          // elements
  float *pos, *pos_ckpt, *pos_seg, *lam_pos;
  float dt;
};

// SOAR field buffer: the member's own (slot < 0) or the segment's slot.
__device__ inline float* nd_coarse(const NoiseDriftItem& m, int slot) {
  return (slot < 0 || !m.coarse_seg) ? m.coarse : m.coarse_seg + (size_t)slot * (m.g.nx + 4) * (m.g.ny + 4);
}

// The normal field of draw draw0 + step (gpo_random_normal's values).
__global__ void __launch_bounds__(256) nd_xi_b(const NoiseDriftItem* __restrict__ tab, int step) {
  const NoiseDriftItem& m = tab[blockIdx.y];
  const uint32_t n = (uint32_t)m.xi_rows * (uint32_t)m.xi_cols;
  const uint32_t e = blockIdx.x * 256u + threadIdx.x;
  if (!m.noise || e >= n) return;
  m.xi[e] = normal_at(e, m.draw0 + (uint32_t)step, m.stream_id, m.k0, m.k1);
}

// Its SOAR field over the grid and two ghost rings (gpo_soar, scale 1).
__global__ void __launch_bounds__(256) nd_soar_b(const NoiseDriftItem* __restrict__ tab, int slot) {
  const NoiseDriftItem& m = tab[blockIdx.y];
  const int W = m.g.nx + 4, n = W * (m.g.ny + 4);
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (!m.noise || k >= n) return;
  const int tj = k / W, ti = k - tj * W;
  nd_coarse(m, slot)[k] = 1.0f * soar_cell(m.g.nx, m.g.ny, m.W, m.per_ns, m.per_ew, m.xi, 4 * m.xi_cols, tj, ti);
}

// The model error of step `step` (draw draw0 + step) in one launch over the
// tiles of every member with model error (noise_tile: normal field, SOAR and
// geostrophic balance of a 64 x 16 tile, the values of gpo_random_normal,
// gpo_soar with scale 1 and gpo_geostrophic_balance): on the forward state
// (TRAJ false), or on segment state `which` keeping its pre-noise eta in
// `keep` and the SOAR field in the segment's slot (TRAJ true).
template <bool TRAJ>
__global__ void __launch_bounds__(256) nd_noise_b(const NoiseDriftItem* __restrict__ tab,
                                                  const int2* __restrict__ tiles, int step, int which, int keep,
                                                  int slot) {
  const int2 t = tiles[blockIdx.x];
  const NoiseDriftItem& m = tab[t.x];
  NoiseTile a;
  a.nx = m.g.nx, a.ny = m.g.ny, a.per_ns = m.per_ns, a.per_ew = m.per_ew;
  a.xi_cols = m.xi_cols, a.pitch = 4 * m.g.sp, a.hi_pitch = 4 * m.g.hp, a.f_pitch = 0;
  a.k0 = m.k0, a.k1 = m.k1, a.stream_id = m.stream_id, a.draw = m.draw0 + (uint32_t)step;
  a.W = m.W;
  a.scale = 1.0f, a.dx = (float)m.g.dx, a.dy = (float)m.g.dy, a.g = (float)m.g.g, a.f = (float)m.g.f;
  a.beta = m.beta, a.north_x = (float)m.g.north_x, a.north_y = (float)m.g.north_y, a.land = m.land;
  float* base = TRAJ ? m.qs + (size_t)which * m.st : m.state;
  a.eta = base, a.hu = base + m.plane, a.hv = base + 2 * m.plane;
  a.Hi = m.Hi, a.f_cells = nullptr;
  a.pre_e = nullptr, a.d_out = nullptr, a.d_row = m.g.nx + 4;
  if (TRAJ) {  // the sweep reads the eta plane of the pre-noise copy (its dry-cell test)
    a.pre_e = m.pre + (size_t)keep * m.plane;
    a.d_out = nd_coarse(m, slot);
  }
  noise_tile<TRAJ, TRAJ>(a, 2 + kTileH * (t.y >> 16), 2 + kTileW * (t.y & 0xFFFF));
}

// Drift of the forward positions over the forward state (which < 0), or of
// segment positions which-1 into `which` over segment state `which`, with
// gpo_drift's arithmetic; periodic ghost cells are read as the interior cells
// they copy (the values the ghost-cell pass would write).
__global__ void __launch_bounds__(64) nd_drift_b(const NoiseDriftItem* __restrict__ tab, int which) {
  const NoiseDriftItem& m = tab[blockIdx.y];
  const int d = blockIdx.x * 64 + threadIdx.x;
  if (d >= m.n_drifters) return;
  const float* base = which < 0 ? m.state : m.qs + (size_t)which * m.st;
  const float *eta = base, *hu = base + m.plane, *hv = base + 2 * m.plane;
  float* out = which < 0 ? m.pos : m.pos_seg + (size_t)which * 2 * m.n_drifters;
  const float* in = which < 0 ? m.pos : m.pos_seg + (size_t)(which - 1) * 2 * m.n_drifters;
  float px = in[2 * d], py = in[2 * d + 1];
  const int nx = m.g.nx, ny = m.g.ny;
  const float dx = (float)m.g.dx, dy = (float)m.g.dy;
  const float qx = px / dx, qy = py / dy;
  const float fcx = floorf(qx) + 2.0f, fcy = floorf(qy) + 2.0f;
  bool ok = fcx >= 0.0f && fcx <= (float)(nx + 3) && fcy >= 0.0f && fcy <= (float)(ny + 3);
  if (ok) {
    const int cx = (int)fcx, cy = (int)fcy;
    const float frx = qx - floorf(qx), fry = qy - floorf(qy);
    const int cx0 = frx < 0.5f ? cx - 1 : cx;
    const float xf = frx < 0.5f ? frx + 0.5f : frx - 0.5f;
    const int cy0 = fry < 0.5f ? cy - 1 : cy;
    const float yf = fry < 0.5f ? fry + 0.5f : fry - 0.5f;
    ok = !(cx0 < 0 || cy0 < 0 || cx0 + 1 > nx + 3 || cy0 + 1 > ny + 3);
    if (ok) {
      auto vel = [&](const float* mom, int c, int r) {
        const int mc = wrap_axis(c, nx, m.per_ew), mr = wrap_axis(r, ny, m.per_ns);
        const size_t ks = (mc >= 0 && mr >= 0) ? (size_t)mr * m.g.sp + mc : (size_t)r * m.g.sp + c;
        const float h = m.Bm[(size_t)r * m.bm_pitch + c] + eta[ks];
        return mom[ks] / h;
      };
      const float u00 = vel(hu, cx0, cy0), u10 = vel(hu, cx0 + 1, cy0);
      const float u01 = vel(hu, cx0, cy0 + 1), u11 = vel(hu, cx0 + 1, cy0 + 1);
      const float v00 = vel(hv, cx0, cy0), v10 = vel(hv, cx0 + 1, cy0);
      const float v01 = vel(hv, cx0, cy0 + 1), v11 = vel(hv, cx0 + 1, cy0 + 1);
      const float uy0 = (1.0f - xf) * u00 + xf * u10;
      const float uy1 = (1.0f - xf) * u01 + xf * u11;
      const float vy0 = (1.0f - xf) * v00 + xf * v10;
      const float vy1 = (1.0f - xf) * v01 + xf * v11;
      const float u = (1.0f - yf) * uy0 + yf * uy1;
      const float v = (1.0f - yf) * vy0 + yf * vy1;
      px += 1.0f * u * m.dt;
      py += 1.0f * v * m.dt;
      const float Lx = (float)nx * dx, Ly = (float)ny * dy;
      if (m.per_ew && px < 0.0f) px += Lx;
      if (m.per_ew && px > Lx) px -= Lx;
      if (m.per_ns && py < 0.0f) py += Ly;
      if (m.per_ns && py > Ly) py -= Ly;
    }
  }
  out[2 * d] = px;
  out[2 * d + 1] = py;
}

// Reverse of the drift of step j of the swept segment: positions pos_seg[j],
// state qs[j+1], lambda buffer (j+1) & 1; one block per member.
__global__ void __launch_bounds__(256) nd_drift_rev_b(const NoiseDriftItem* __restrict__ tab, int j) {
  const NoiseDriftItem& m = tab[blockIdx.y];
  if (m.n_drifters == 0) return;
  const DriftGeo q{m.g.nx, m.g.ny, m.g.sp, m.bm_pitch, m.per_ns, m.per_ew, m.n_drifters, m.g.dx, m.g.dy,
                   (double)m.dt, 1.0};
  const float* s = m.qs + (size_t)(j + 1) * m.st;
  float* l = m.lam + (size_t)((j + 1) & 1) * m.st;
  drift_rev_block<float>(q, s, s + m.plane, s + 2 * m.plane, m.Bm, m.pos_seg + (size_t)j * 2 * m.n_drifters,
                         m.lam_pos, l, l + m.plane, l + 2 * m.plane);
}

// Reverse of the model error of step j on lambda buffer (j+1) & 1 (the
// member's coarse field holds that step's SOAR field), over the 64 x 16
// tiles of the members with model error: four rows per thread.
__global__ void __launch_bounds__(256) nd_balance_rev_b(const NoiseDriftItem* __restrict__ tab,
                                                        const int2* __restrict__ tiles, int j, int slot) {
  const int2 t = tiles[blockIdx.x];
  const NoiseDriftItem& m = tab[t.x];
  float* l = m.lam + (size_t)((j + 1) & 1) * m.st;
  const Geo g = m.g;
  const float* coarse = nd_coarse(m, slot);
  const int T0 = 2 + kTileH * (t.y >> 16), ti = 2 + kTileW * (t.y & 0xFFFF) + (threadIdx.x & 63);
#pragma unroll
  for (int k = 0; k < kTileH / 4; ++k)
    balance_rev_cell<float>(g, (double)m.beta, nullptr, 0, coarse, 4 * (g.nx + 4), m.Hi, m.land, l, l + m.plane,
                            l + 2 * m.plane, ti, T0 + (threadIdx.x >> 6) + 4 * k);
}

// Drifter positions: mode 0 forward -> checkpoint seg, mode 1 checkpoint seg ->
// segment start, mode 2 checkpoint seg -> forward, mode 3 forward -> segment
// start.
__global__ void __launch_bounds__(64) nd_pos_copy_b(const NoiseDriftItem* __restrict__ tab, int mode, int seg) {
  const NoiseDriftItem& m = tab[blockIdx.y];
  const int i = blockIdx.x * 64 + threadIdx.x;
  if (i >= 2 * m.n_drifters) return;
  float* ck = m.pos_ckpt + (size_t)seg * 2 * m.n_drifters;
  if (mode == 0) ck[i] = m.pos[i];
  else if (mode == 1) m.pos_seg[i] = ck[i];
  else if (mode == 2) m.pos[i] = ck[i];
  else m.pos_seg[i] = m.pos[i];
}

// Observation of the cell-centre velocity (u, v) = (hu, hv)/(eta + Hm),
// bilinear between the four cell centres around each position (x, y).
// J += sum (u - u_obs)^2 + (v - v_obs)^2; lambda += dJ/d(eta, hu, hv).  One
// block of 256 threads takes the points in chunks of 256; the sensitivities
// are added in point order (ordered_add) and J in a fixed order, so both are
// the same run to run.
template <class T>
__global__ void __launch_bounds__(256) misfit(Geo g, int n, const float* pos, const float* obs, const T* e,
                                              const T* hu, const T* hv, const T* Hi, double* J, T* l_e, T* l_u,
                                              T* l_v) {
  __shared__ uint64_t sk[kOrdItems];
  __shared__ T sv[kOrdItems][3];
  __shared__ double sj[256];
  const int t = threadIdx.x;
  double jsum = 0.0;
  for (int base = 0; base < n; base += 256) {
    const int d = base + t;
    for (int q = 0; q < 4; ++q) sk[4 * t + q] = kNoKey;
    if (d < n) {
      const T sx = T(pos[2 * d]) / T(g.dx) - T(0.5), sy = T(pos[2 * d + 1]) / T(g.dy) - T(0.5);
      int i0 = (int)floor((double)sx), j0 = (int)floor((double)sy);
      i0 = min(max(i0, 0), g.nx - 2);
      j0 = min(max(j0, 0), g.ny - 2);
      const T ax = sx - T(i0), ay = sy - T(j0);
      const T wgt[4] = {(T(1) - ax) * (T(1) - ay), ax * (T(1) - ay), (T(1) - ax) * ay, ax * ay};
      const int rr[4] = {j0 + 2, j0 + 2, j0 + 3, j0 + 3}, cc[4] = {i0 + 2, i0 + 3, i0 + 2, i0 + 3};
      T u = 0, v = 0, hh[4];
      for (int q = 0; q < 4; ++q) {
        const size_t k = (size_t)rr[q] * g.sp + cc[q];
        hh[q] = e[k] + cell_depth_land(g, Hi, rr[q], cc[q]);
        u += wgt[q] * hu[k] / hh[q];
        v += wgt[q] * hv[k] / hh[q];
      }
      const T du = u - T(obs[2 * d]), dv = v - T(obs[2 * d + 1]);
      jsum += (double)(du * du + dv * dv);
      if (l_e) {
        for (int q = 0; q < 4; ++q) {
          const size_t k = (size_t)rr[q] * g.sp + cc[q];
          const T a = T(2) * du * wgt[q] / hh[q], b = T(2) * dv * wgt[q] / hh[q];
          sk[4 * t + q] = ((uint64_t)k << 10) | (uint64_t)(4 * t + q);
          sv[4 * t + q][0] = a;
          sv[4 * t + q][1] = b;
          sv[4 * t + q][2] = -(a * hu[k] + b * hv[k]) / hh[q];
        }
      }
    }
    __syncthreads();
    if (l_e) ordered_add<T>(sk, sv, l_u, l_v, l_e);
  }
  // J: each thread's points in order, then a fixed tree over the threads
  sj[t] = jsum;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) sj[t] += sj[t + w];
    __syncthreads();
  }
  if (t == 0) *J = *J + sj[0];
}

// Copy of the interior cells of three state arrays (pitch sp).
template <class T>
__global__ void state_copy(Geo g, const T* a, const T* b, const T* c, T* x, T* y, T* z) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x + 2, r = blockIdx.y * blockDim.y + threadIdx.y + 2;
  if (col > g.nx + 1 || r > g.ny + 1) return;
  const size_t k = (size_t)r * g.sp + col;
  x[k] = a[k];
  y[k] = b[k];
  z[k] = c[k];
}

}  // namespace adj
}  // namespace gpo

// ============================================================================= C ABI

using namespace gpo;
using namespace gpo::adj;

namespace {

int bad_arg(const char* msg) {
  set_error(msg);
  return GPO_ERR_BAD_ARGUMENT;
}
int hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    char buf[256];
    snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
    set_error(buf);
    return (int)e;
  }
  return 0;
}

int make_geo(const gpo_cdklm16_params* p, uint32_t bc, int32_t pitch_bytes, int32_t hi_pitch_bytes,
             int32_t f_pitch_bytes, int elem, Geo* g) {
  if (!p || p->nx < 2 || p->ny < 2) return bad_arg("differentiable substep: bad parameters");
  if (p->rk_order != 2) return bad_arg("differentiable substep: RK2 only");
  if (p->r != 0.0f) return bad_arg("differentiable substep: friction is not supported");

  if (p->one_dimensional) return bad_arg("differentiable substep: one-dimensional mode is not supported");
  g->nx = p->nx;
  g->ny = p->ny;
  g->W = p->nx + 4;
  g->bc_n = (bc >> 24) & 0xFF;
  g->bc_s = (bc >> 16) & 0xFF;
  g->bc_e = (bc >> 8) & 0xFF;
  g->bc_w = bc & 0xFF;
  for (int code : {g->bc_n, g->bc_s, g->bc_e, g->bc_w})
    if (code < 1 || code > 3)
      return bad_arg("differentiable substep: wall (1), periodic (2) and flow-relaxation (3) boundaries only");
  if (pitch_bytes % elem || hi_pitch_bytes % elem || f_pitch_bytes % elem || pitch_bytes < elem * (p->nx + 4) ||
      hi_pitch_bytes < elem * (p->nx + 5))
    return bad_arg("differentiable substep: pitches must be multiples of the element size and cover a row");
  g->sp = pitch_bytes / elem;
  g->hp = hi_pitch_bytes / elem;
  g->fp = f_pitch_bytes / elem;
  g->dx = p->dx;
  g->dy = p->dy;
  g->g = p->g;
  g->f = p->f;
  g->theta = p->theta;
  g->desing_eps = p->desing_eps;
  g->depth_cutoff = p->depth_cutoff;
  g->flux_slope_eps = p->flux_slope_eps;
  g->flux_balancer = p->flux_balancer;
  g->has_land = p->has_land ? 1 : 0;
  g->land = p->land_value;
  g->north_x = p->north_x;
  g->north_y = p->north_y;
  return 0;
}

dim3 cell_grid(const Geo& g) { return dim3((g.nx + 4 + 63) / 64, (g.ny + 4 + 3) / 4); }
dim3 inner_grid(const Geo& g) { return dim3((g.nx + 63) / 64, (g.ny + 3) / 4); }
const dim3 kBlock(64, 4);

// Land-flagged copy of Hi (land corners -> kDry) in the workspace tail, dense
// rows of nx+5; the stages then read corners without testing them.
template <class T>
__global__ void flag_hi(Geo g, const T* Hi, T* out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y * blockDim.y + threadIdx.y;
  if (c >= g.nx + 5 || r >= g.ny + 5) return;
  out[(size_t)r * (g.nx + 5) + c] = corner_land(g, Hi, r, c);
}
template <class T>
const T* land_hi(Geo* g, const T* Hi, void* ws, hipStream_t s) {
  if (!g->has_land) return Hi;
  T* out = (T*)ws + (size_t)kAdjWorkArrays * (g->nx + 4) * (g->ny + 4);
  hipLaunchKernelGGL(flag_hi<T>, dim3((g->nx + 5 + 63) / 64, (g->ny + 5 + 3) / 4), kBlock, 0, s, *g, Hi, out);
  g->hp = g->nx + 5;
  return out;
}

template <class S>
int forward_stages(const Geo& g, S dt, int rk, const S* ie, const S* iu, const S* iv, const S* qe, const S* qu,
                   const S* qv, S* oe, S* ou, S* ov, const S* Hi, const S* fc, void* ws, int* flag, hipStream_t s) {
  Fwd<S> w = carve<S>(ws, g.nx, g.ny);
  hipLaunchKernelGGL(f1_convert<S>, cell_grid(g), kBlock, 0, s, g, ie, iu, iv, Hi, w, flag);
  hipLaunchKernelGGL(f2_slopes<S>, cell_grid(g), kBlock, 0, s, g, Hi, fc, w);
  hipLaunchKernelGGL(f3_faces<S>, cell_grid(g), kBlock, 0, s, g, Hi, fc, w, flag);
  hipLaunchKernelGGL(f4_update<S>, inner_grid(g), kBlock, 0, s, g, dt, rk, Hi, fc, w, qe, qu, qv, oe, ou, ov, flag);
  return hip_check(hipGetLastError(), "differentiable forward launch");
}

// Reverse of one substep.  lambda_in = (or +=, with accumulate) the
// sensitivity of the stage input; for rk == 1 the direct Q^n term is added to
// lambda_qn, or written to it with lq_mode 1.  The FP32 reverse runs the
// stage kernels built with fast arithmetic (adjoint_rev32.hip), the FP64 one
// those built here.
template <class T>
int adjoint_stages(const Geo& g, T dt, int rk, const T* ie, const T* iu, const T* iv, const T* out_e, const T* le,
                   const T* lu, const T* lv, T* lie, T* liu, T* liv, T* lqe, T* lqu, T* lqv, int accumulate,
                   int lq_mode, const T* Hi, const T* fc, void* ws, int* flag, hipStream_t s);
template <>
int adjoint_stages<float>(const Geo& g, float dt, int rk, const float* ie, const float* iu, const float* iv,
                          const float* out_e, const float* le, const float* lu, const float* lv, float* lie,
                          float* liu, float* liv, float* lqe, float* lqu, float* lqv, int accumulate, int lq_mode,
                          const float* Hi, const float* fc, void* ws, int* flag, hipStream_t s) {
  fast32::launch_stages(g, dt, rk, ie, iu, iv, out_e, le, lu, lv, lie, liu, liv, lqe, lqu, lqv, accumulate, lq_mode, Hi,
                        fc, ws, flag, s);
  return hip_check(hipGetLastError(), "adjoint launch");
}
template <>
int adjoint_stages<double>(const Geo& g, double dt, int rk, const double* ie, const double* iu, const double* iv,
                           const double* out_e, const double* le, const double* lu, const double* lv, double* lie,
                           double* liu, double* liv, double* lqe, double* lqu, double* lqv, int accumulate,
                           int lq_mode, const double* Hi, const double* fc, void* ws, int* flag, hipStream_t s) {
  exact::launch_stages(g, dt, rk, ie, iu, iv, out_e, le, lu, lv, lie, liu, liv, lqe, lqu, lqv, accumulate, lq_mode, Hi,
                       fc, ws, flag, s);
  return hip_check(hipGetLastError(), "adjoint launch");
}

// ---- batched differentiable ensemble

// Rows per work item of the fused reverse march.
constexpr int kDefaultMarchRows = 24;

struct AdjBatch {
  int n = 0, k = 0, n_seg = 0;
  int max_w = 0, max_h = 0;  // largest (nx+4), (ny+4) over the members
  int n_work = 0, n_blocks = 0;
  int has_land = 0, f_array = 0, angle0 = 1;  // shared by all members (kernel variant)
  int any_fc = 0;                             // some member has a coriolis array
  // reverse substeps as one strip march per work item (adjoint_march.inc)
  // rather than the staged kernels; the work table for chunks of march_rows
  int use_march = 1, march_rows = 0, n_march = 0, max_fold = 0;
  int4* d_march = nullptr;
  std::vector<int> nxs, nys;
  // model-error tiles (member, tile row << 16 | tile column) of members with model error
  int2* d_ntiles = nullptr;
  int n_ntiles = 0;
  int all_ok = 1;                             // every member's fast divisions and square root verified
  StepArgs* d_fwd = nullptr;         // 2n: forward state -> stage -> state
  StepArgs* d_traj = nullptr;        // k x 2n: qs[j] -> q1s[j] -> qs[j+1]
  int2* d_work = nullptr;            // (member, strip | chunk << 16)
  AdjItem<float>* d_items = nullptr;  // (2j + rk) x n
  // save (n_seg x n), restore (n_seg x n), lambda parity (n), load (n_seg x n),
  // forward state -> segment start (n); n_seg is the number of checkpoint slots
  CopyItem* d_copy = nullptr;
  // flow-relaxation members (n_frs of them): ghost-cell passes on the forward
  // stage, the forward state, then per step j the stage and the new state
  cells::BcTarget* d_bc = nullptr;   // (2 + 2k) x n_frs
  int n_frs = 0, bc_blocks = 0;
  // model error and drifters (per member, n items)
  NoiseDriftItem* d_nd = nullptr;
  int any_noise = 0, any_drift = 0, max_xi = 0, max_cells = 0, max_drifters = 0;
  int sponge_cells = 0;  // most sponge cells of one pass over the members
  int coarse_saved = 0;  // the segment's SOAR fields are kept for the sweep
  int step = 0;       // step index of the forward run (the draw offset of its model error)
  int seg_start = 0;  // first step of the segment being recomputed and swept
};

void adj_batch_free(AdjBatch* b) {
  for (void* p : {(void*)b->d_fwd, (void*)b->d_traj, (void*)b->d_work, (void*)b->d_items, (void*)b->d_copy,
                  (void*)b->d_bc, (void*)b->d_nd, (void*)b->d_march, (void*)b->d_ntiles})
    if (p) (void)hipFree(p);
  delete b;
}

// The model error of step `step` added to the forward state (which < 0) or to
// segment state `which` (its pre-noise copy kept in `keep`).
int adj_batch_noise(AdjBatch* b, int step, int which, int keep, hipStream_t s) {
  if (!b->any_noise) return 0;
  if (which < 0)
    hipLaunchKernelGGL(nd_noise_b<false>, dim3(b->n_ntiles), dim3(256), 0, s, b->d_nd, b->d_ntiles, step, which, keep,
                       keep);
  else  // the segment's slot of this step's SOAR field is `keep`
    hipLaunchKernelGGL(nd_noise_b<true>, dim3(b->n_ntiles), dim3(256), 0, s, b->d_nd, b->d_ntiles, step, which, keep,
                       keep);
  return hip_check(hipGetLastError(), "batched model error launch");
}

int adj_batch_drift(AdjBatch* b, int which, hipStream_t s) {
  if (!b->any_drift) return 0;
  hipLaunchKernelGGL(nd_drift_b, dim3((b->max_drifters + 63) / 64, b->n), dim3(64), 0, s, b->d_nd, which);
  return hip_check(hipGetLastError(), "batched drifter launch");
}

int adj_batch_copy(AdjBatch* b, const CopyItem* tab, hipStream_t s) {
  hipLaunchKernelGGL(copy_states_b, dim3((b->max_w + 63) / 64, (b->max_h + 3) / 4, b->n), kBlock, 0, s, tab);
  return hip_check(hipGetLastError(), "batched state copy");
}

int adj_batch_steps(AdjBatch* b, const StepArgs* tab, int rk, hipStream_t s) {
  const hipError_t e =
      launch_cdklm_batched_exact(tab, b->d_work, b->n_work, rk, b->n_blocks, b->has_land, b->f_array, b->angle0,
                                 /*any_stage=*/0, b->all_ok, s);  // RK2 without friction
  return e != hipSuccess ? hip_check(e, "batched substep launch") : 0;
}

// Ghost-cell passes (north/south, then east/west) of the flow-relaxation
// members on table slot `slot` (0: forward stage, 1: forward state, 2 + 2j:
// stage of step j, 3 + 2j: state after step j).
int adj_batch_bc(AdjBatch* b, int slot, hipStream_t s) {
  if (b->n_frs == 0) return 0;
  const cells::BcTarget* tab = b->d_bc + (size_t)slot * b->n_frs;
  for (int pass = 0; pass < 2; ++pass)
    hipLaunchKernelGGL(bc_targets_b, dim3(b->bc_blocks, b->n_frs), dim3(256), 0, s, tab, pass);
  return hip_check(hipGetLastError(), "batched boundary launch");
}

void adj_batch_sponge(AdjBatch* b, const AdjItem<float>* tab, hipStream_t s) {
  for (int pass = 0; pass < 2; ++pass)
    hipLaunchKernelGGL(frs_scale_sponge_b<float>, dim3((b->sponge_cells + 255) / 256, b->n), dim3(256), 0, s, tab,
                       pass);
}

int adj_batch_stage(AdjBatch* b, const AdjItem<float>* tab, int* flag, hipStream_t s, bool scale_first = true) {
  if (b->n_frs && scale_first) adj_batch_sponge(b, tab, s);
  if (b->use_march)
    fast32::launch_batch_march(tab, b->d_march, b->n_march, b->n, b->max_fold, b->any_fc != 0, s);
  else
    fast32::launch_batch_stage(tab, b->max_w, b->max_h, b->n, b->any_fc != 0, flag, s);
  return hip_check(hipGetLastError(), "batched adjoint launch");
}

// (Re)build the march work table for chunks of `rows` rows (waits for the
// stream: the old table may be in use).
int adj_batch_march_table(AdjBatch* b, int rows, hipStream_t s) {
  if (rows == b->march_rows && b->d_march) return 0;
  std::vector<int4> work;
  fast32::march_work_items(b->nxs.data(), b->nys.data(), b->n, rows, work);
  int rc = hip_check(hipStreamSynchronize(s), "march table sync");
  if (!rc && b->d_march) rc = hip_check(hipFree(b->d_march), "march table free");
  b->d_march = nullptr;
  if (!rc) rc = hip_check(hipMalloc(&b->d_march, sizeof(int4) * work.size()), "march table hipMalloc");
  if (!rc)
    rc = hip_check(hipMemcpyAsync(b->d_march, work.data(), sizeof(int4) * work.size(), hipMemcpyHostToDevice, s),
                   "march table copy");
  if (!rc) rc = hip_check(hipStreamSynchronize(s), "march table sync");
  if (rc) return rc;
  b->n_march = (int)work.size();
  b->march_rows = rows;
  b->max_fold = 0;
  for (int m = 0; m < b->n; ++m) b->max_fold = std::max(b->max_fold, fast32::march_fold_cells(b->nxs[m], b->nys[m]));
  return 0;
}

// One member's march has one wavefront per (strip, chunk): below this many the
// GPU is mostly idle and the staged kernels, which spread a grid over many
// more wavefronts, are faster (512 x 512: 198 march wavefronts, 2.8x slower).
constexpr int kMarchMinItems = 2048;
bool member_march(const Geo& g) {
  const int strips = (g.nx + 4 + fast32::kMarchColsHost - 1) / fast32::kMarchColsHost;
  return (long long)strips * ((g.ny + 4 + kDefaultMarchRows - 1) / kDefaultMarchRows) >= kMarchMinItems;
}

// One member's reverse substep through the fused march (fast32::launch_march):
// lambda_in (li) must not alias lambda_out (lo) or lambda_qn (lq).
int march_stage(const Geo& g, float dt, int rk, const float* const* in, const float* out_e, float* const* lo,
                float* const* li, float* const* lq, int accumulate, int lq_mode, const float* Hi, const float* fc,
                void* ws, hipStream_t s) {
  AdjItem<float> it{};
  it.g = g;
  it.dt = dt;
  it.rk = rk;
  it.accumulate = accumulate;
  it.lq_mode = lq_mode;
  it.ie = gptr(in[0]), it.iu = gptr(in[1]), it.iv = gptr(in[2]);
  it.out_e = gptr(out_e);
  it.le = gptr((const float*)lo[0]), it.lu = gptr((const float*)lo[1]), it.lv = gptr((const float*)lo[2]);
  it.lie = gptr(li[0]), it.liu = gptr(li[1]), it.liv = gptr(li[2]);
  if (lq) it.lqe = gptr(lq[0]), it.lqu = gptr(lq[1]), it.lqv = gptr(lq[2]);
  it.Hi = gptr(Hi);
  it.fc = gptr(fc);
  it.w = carve<float>(ws, g.nx, g.ny);
  it.rv = carve_rev<float>(ws, g.nx, g.ny);
  fast32::launch_march(it, kDefaultMarchRows, s);
  return hip_check(hipGetLastError(), "adjoint march launch");
}

// The reverse sweep of gpo_cdklm16_adjoint_steps_f32 over caller-given (or
// library scratch) lambda arrays and stage workspace.
int adjoint_steps_impl(const gpo_cdklm16_params* p, int32_t n_steps, float dt, const float* const* qs,
                       const float* const* q1s, float* const* lam, float* const* work, int32_t pitch, const float* Hi,
                       int32_t Hi_pitch, const float* f_cells, uint32_t boundary_conditions, const gpo_frs* frs,
                       void* workspace, int32_t* flag, void* stream) {
  Geo g;
  int rc = make_geo(p, boundary_conditions, pitch, Hi_pitch, f_cells ? pitch : 4, 4, &g);
  if (rc) return rc;
  if (n_steps < 0 || !qs || !q1s || !lam || !work || !workspace)
    return bad_arg("gpo_cdklm16_adjoint_steps_f32: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  const Sponge sp = frs ? Sponge{frs->sponge[0], frs->sponge[1], frs->sponge[2], frs->sponge[3]} : Sponge{0, 0, 0, 0};
  Hi = land_hi(&g, Hi, workspace, s);
  float* cur[3] = {lam[0], lam[1], lam[2]};     // lambda at Q^{j+1}, then at Q^j
  float* nxt[3] = {work[0], work[1], work[2]};  // lambda being formed at Q^j
  float* l1[3] = {work[3], work[4], work[5]};   // lambda at the RK stage Q^(1)
  const bool march = member_march(g);
  for (int j = n_steps - 1; j >= 0; --j) {
    const float* const* a = qs + 3 * j;
    const float* const* b = q1s + 3 * j;
    const float* const* c = qs + 3 * (j + 1);
    // reverse of the second RK stage (writes its direct Q^n term to nxt), then of
    // the first (adds its sensitivity of Q^n to nxt); with flow relaxation,
    // the transpose of each sponge blend before the stage that produced it
    if (frs) hipLaunchKernelGGL(frs_scale<float>, cell_grid(g), kBlock, 0, s, g, sp, cur[0], cur[1], cur[2]);
    if (march)
      rc = march_stage(g, dt, 1, b, c[0], cur, l1, nxt, 0, 1, Hi, f_cells, workspace, s);
    else
      rc = adjoint_stages<float>(g, dt, 1, b[0], b[1], b[2], c[0], cur[0], cur[1], cur[2], l1[0], l1[1], l1[2],
                                 nxt[0], nxt[1], nxt[2], 0, 1, Hi, f_cells, workspace, flag, s);
    if (!rc && frs) hipLaunchKernelGGL(frs_scale<float>, cell_grid(g), kBlock, 0, s, g, sp, l1[0], l1[1], l1[2]);
    if (!rc && march) rc = march_stage(g, dt, 0, a, b[0], l1, nxt, nullptr, 1, 0, Hi, f_cells, workspace, s);
    if (!rc && !march)
      rc = adjoint_stages<float>(g, dt, 0, a[0], a[1], a[2], b[0], l1[0], l1[1], l1[2], nxt[0], nxt[1], nxt[2],
                                 nullptr, nullptr, nullptr, 1, 0, Hi, f_cells, workspace, flag, s);
    if (rc) return rc;
    for (int k = 0; k < 3; ++k) {
      float* t = cur[k];
      cur[k] = nxt[k];
      nxt[k] = t;
    }
  }
  if (cur[0] != lam[0])
    hipLaunchKernelGGL(state_copy<float>, inner_grid(g), kBlock, 0, s, g, cur[0], cur[1], cur[2], lam[0], lam[1],
                       lam[2]);
  return hip_check(hipGetLastError(), "gpo_cdklm16_adjoint_steps_f32 launch");
}

// Scratch of gpo_cdklm16_adjoint_steps_f32_noalloc: one array per device,
// held until the process ends.  Host code never passes the symbol itself (its
// host-side name is no device address): the device address is looked up once
// per device, with that device current, and kept.
constexpr size_t kAdjScratchBytes = (size_t)48 << 20;
__device__ __align__(256) unsigned char scratch[kAdjScratchBytes];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
char* g_scratch_ptr[kMaxDevices] = {};

int scratch_base(char** out) {
  int dev = 0;
  int rc = hip_check(hipGetDevice(&dev), "scratch: hipGetDevice");
  if (rc) return rc;
  if (dev < 0 || dev >= kMaxDevices) return bad_arg("scratch: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    rc = hip_check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "scratch: hipGetSymbolAddress");
    if (rc) return rc;
    if (!ptr) return bad_arg("scratch: no device address");
    g_scratch_ptr[dev] = (char*)ptr;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

__global__ void fill_zero_words(uint32_t* p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0u;
}

}  // namespace

extern "C" {

size_t gpo_adjoint_workspace_bytes(int32_t nx, int32_t ny, int32_t elem_bytes) {
  if (nx < 1 || ny < 1 || (elem_bytes != 4 && elem_bytes != 8)) return 0;
  return ((size_t)kAdjWorkArrays * (nx + 4) * (ny + 4) + (size_t)(nx + 5) * (ny + 5)) * elem_bytes;
}

int gpo_cdklm16_substep_f64(const gpo_cdklm16_params* p, double dt, int32_t rk_step, const double* in_eta,
                            const double* in_hu, const double* in_hv, const double* qn_eta, const double* qn_hu,
                            const double* qn_hv, double* out_eta, double* out_hu, double* out_hv, int32_t pitch,
                            const double* Hi, int32_t Hi_pitch, const double* f_cells, uint32_t boundary_conditions,
                            void* workspace, int32_t* flag, void* stream) {
  Geo g;
  int rc = make_geo(p, boundary_conditions, pitch, Hi_pitch, f_cells ? pitch : 8, 8, &g);
  if (rc) return rc;
  if (!in_eta || !out_eta || !Hi || !workspace || (rk_step == 1 && !qn_eta) || rk_step < 0 || rk_step > 1)
    return bad_arg("gpo_cdklm16_substep_f64: bad arguments");
  const double* hi = land_hi(&g, Hi, workspace, (hipStream_t)stream);
  return forward_stages<double>(g, dt, rk_step, in_eta, in_hu, in_hv, qn_eta, qn_hu, qn_hv, out_eta, out_hu, out_hv,
                                hi, f_cells, workspace, flag, (hipStream_t)stream);
}

int gpo_cdklm16_substep_tlm_f64(const gpo_cdklm16_params* p, double dt, int32_t rk_step, const double* in_eta,
                                const double* in_hu, const double* in_hv, const double* d_in_eta,
                                const double* d_in_hu, const double* d_in_hv, const double* qn_eta,
                                const double* qn_hu, const double* qn_hv, const double* d_qn_eta,
                                const double* d_qn_hu, const double* d_qn_hv, double* out_eta, double* out_hu,
                                double* out_hv, double* d_out_eta, double* d_out_hu, double* d_out_hv,
                                int32_t pitch, const double* Hi, int32_t Hi_pitch, const double* f_cells,
                                uint32_t boundary_conditions, void* workspace, int32_t* flag, void* stream) {
  Geo g;
  int rc = make_geo(p, boundary_conditions, pitch, Hi_pitch, f_cells ? pitch : 8, 8, &g);
  if (rc) return rc;
  if (!in_eta || !d_in_eta || !out_eta || !d_out_eta || !Hi || !workspace || rk_step < 0 || rk_step > 1 ||
      (rk_step == 1 && (!qn_eta || !d_qn_eta)))
    return bad_arg("gpo_cdklm16_substep_tlm_f64: bad arguments");
  // Dual workspace: the 15 stage arrays, then the state, Q^n, output (3 each), Hi and f
  const size_t n = (size_t)(g.nx + 4) * (g.ny + 4), nh = (size_t)(g.ny + 5) * (g.nx + 5);
  Dual* base = (Dual*)workspace;
  Dual* st = base + 15 * n;
  Dual *ie = st, *iu = st + n, *iv = st + 2 * n, *qe = st + 3 * n, *qu = st + 4 * n, *qv = st + 5 * n;
  Dual *oe = st + 6 * n, *ou = st + 7 * n, *ov = st + 8 * n;
  Dual* hi = st + 9 * n;
  Dual* fcd = hi + nh;
  hipStream_t s = (hipStream_t)stream;
  Geo gd = g;  // dense Dual copies use the workspace row length
  gd.sp = g.W;
  gd.hp = g.nx + 5;
  gd.fp = g.W;
  hipLaunchKernelGGL(pack_dual, cell_grid(g), kBlock, 0, s, g, in_eta, d_in_eta, ie);
  hipLaunchKernelGGL(pack_dual, cell_grid(g), kBlock, 0, s, g, in_hu, d_in_hu, iu);
  hipLaunchKernelGGL(pack_dual, cell_grid(g), kBlock, 0, s, g, in_hv, d_in_hv, iv);
  if (rk_step == 1) {
    hipLaunchKernelGGL(pack_dual, cell_grid(g), kBlock, 0, s, g, qn_eta, d_qn_eta, qe);
    hipLaunchKernelGGL(pack_dual, cell_grid(g), kBlock, 0, s, g, qn_hu, d_qn_hu, qu);
    hipLaunchKernelGGL(pack_dual, cell_grid(g), kBlock, 0, s, g, qn_hv, d_qn_hv, qv);
  }
  // Hi (pitch Hi_pitch) -> dense Dual (row length nx+5); f likewise (row length nx+4)
  {
    Geo gh = g;
    gh.sp = g.hp;
    gh.W = g.nx + 5;
    gh.nx = g.nx + 1;
    gh.ny = g.ny + 1;
    hipLaunchKernelGGL(pack_dual_hi, cell_grid(gh), kBlock, 0, s, gh, g, Hi, hi);
    if (f_cells) hipLaunchKernelGGL(pack_dual, cell_grid(g), kBlock, 0, s, g, f_cells, (const double*)nullptr, fcd);
  }
  rc = forward_stages<Dual>(gd, Dual(dt), rk_step, ie, iu, iv, qe, qu, qv, oe, ou, ov, hi, f_cells ? fcd : nullptr,
                            base, flag, s);
  if (rc) return rc;
  // dense Dual rows (length W) -> the caller's arrays (pitch sp)
  hipLaunchKernelGGL(unpack_dual, inner_grid(g), kBlock, 0, s, g, oe, out_eta, d_out_eta);
  hipLaunchKernelGGL(unpack_dual, inner_grid(g), kBlock, 0, s, g, ou, out_hu, d_out_hu);
  hipLaunchKernelGGL(unpack_dual, inner_grid(g), kBlock, 0, s, g, ov, out_hv, d_out_hv);
  return hip_check(hipGetLastError(), "tangent-linear launch");
}

size_t gpo_tlm_workspace_bytes(int32_t nx, int32_t ny) {
  if (nx < 1 || ny < 1) return 0;
  const size_t n = (size_t)(nx + 4) * (ny + 4), nh = (size_t)(ny + 5) * (nx + 5);
  return (25 * n + nh) * sizeof(Dual);
}

int gpo_cdklm16_trajectory_f32(const gpo_cdklm16_params* p, int32_t n_steps, float dt, float* const* qs,
                               float* const* q1s, int32_t pitch, const float* Hi, int32_t Hi_pitch,
                               const float* f_cells, uint32_t boundary_conditions, const gpo_frs* frs,
                               void* stream) {
  if (!p || n_steps < 0 || !qs || !q1s || !Hi) return bad_arg("gpo_cdklm16_trajectory_f32: bad arguments");
  if (p->rk_order != 2) return bad_arg("gpo_cdklm16_trajectory_f32: RK2 only");
  hipStream_t s = (hipStream_t)stream;
  for (int j = 0; j < n_steps; ++j) {
    float* const* a = qs + 3 * j;
    float* const* b = q1s + 3 * j;
    float* const* c = qs + 3 * (j + 1);
    StepArgs a0, a1;
    int rc = make_step_args(p, dt, 0, a[0], a[1], a[2], nullptr, nullptr, nullptr, b[0], b[1], b[2], pitch, Hi,
                            Hi_pitch, f_cells, nullptr, boundary_conditions, &a0);
    if (!rc)
      rc = make_step_args(p, dt, 1, b[0], b[1], b[2], a[0], a[1], a[2], c[0], c[1], c[2], pitch, Hi, Hi_pitch,
                          f_cells, nullptr, boundary_conditions, &a1);
    if (rc) return rc;
    hipError_t e = launch_cdklm_exact(a0, p->has_land, p->f_is_array, p->angle_is_zero, s);
    if (e != hipSuccess) return hip_check(e, "gpo_cdklm16_trajectory_f32 launch");
    if (frs) {  // ghost cells and sponge of the stage, as the step does between substeps
      rc = gpo_boundary_conditions(p->nx, p->ny, boundary_conditions, frs->sponge, b[0], b[1], b[2], pitch,
                                   frs->ns_cur, frs->ns_next, frs->ns_nx, frs->ew_cur, frs->ew_next, frs->ew_ny,
                                   frs->t, stream);
      if (rc) return rc;
    }
    e = launch_cdklm_exact(a1, p->has_land, p->f_is_array, p->angle_is_zero, s);
    if (e != hipSuccess) return hip_check(e, "gpo_cdklm16_trajectory_f32 launch");
    if (frs) {
      rc = gpo_boundary_conditions(p->nx, p->ny, boundary_conditions, frs->sponge, c[0], c[1], c[2], pitch,
                                   frs->ns_cur, frs->ns_next, frs->ns_nx, frs->ew_cur, frs->ew_next, frs->ew_ny,
                                   frs->t, stream);
      if (rc) return rc;
    }
  }
  return 0;
}

int gpo_cdklm16_adjoint_steps_f32(const gpo_cdklm16_params* p, int32_t n_steps, float dt, const float* const* qs,
                                  const float* const* q1s, float* const* lam, float* const* work, int32_t pitch,
                                  const float* Hi, int32_t Hi_pitch, const float* f_cells,
                                  uint32_t boundary_conditions, const gpo_frs* frs, void* workspace, int32_t* flag,
                                  void* stream) {
  return adjoint_steps_impl(p, n_steps, dt, qs, q1s, lam, work, pitch, Hi, Hi_pitch, f_cells, boundary_conditions,
                            frs, workspace, flag, stream);
}

int gpo_cdklm16_adjoint_steps_f32_noalloc(const gpo_cdklm16_params* p, int32_t n_steps, float dt,
                                          const float* const* qs, const float* const* q1s, float* const* lam,
                                          int32_t pitch, const float* Hi, int32_t Hi_pitch, const float* f_cells,
                                          uint32_t boundary_conditions, const gpo_frs* frs, int32_t* flag,
                                          void* stream) {
  if (!p || p->nx < 1 || p->ny < 1 || pitch < 4 * (p->nx + 4) || pitch % 4)
    return bad_arg("gpo_cdklm16_adjoint_steps_f32_noalloc: bad parameters");
  // scratch layout: six lambda arrays of the state's shape, then the stage workspace
  const size_t plane = (((size_t)pitch * (p->ny + 4)) + 255) & ~(size_t)255;
  const size_t need = 6 * plane + gpo_adjoint_workspace_bytes(p->nx, p->ny, 4);
  if (need > kAdjScratchBytes) {
    char buf[240];
    snprintf(buf, sizeof buf,
             "gpo_cdklm16_adjoint_steps_f32_noalloc: this grid needs %zu bytes of scratch, more than the library's "
             "%zu (use gpo_cdklm16_adjoint_steps_f32 with caller buffers)",
             need, kAdjScratchBytes);
    return bad_arg(buf);
  }
  char* base = nullptr;
  int rc = scratch_base(&base);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  // The lambda arrays' ghost cells are read as zeros and the previous call's
  // values remain: clear all six, with a fill kernel (no memset node, so the
  // call can be captured in a graph).
  const size_t words = 6 * plane / 4;
  hipLaunchKernelGGL(fill_zero_words, dim3((unsigned)std::min<size_t>((words + 255) / 256, 65535)), dim3(256), 0, s,
                     (uint32_t*)base, words);
  rc = hip_check(hipGetLastError(), "gpo_cdklm16_adjoint_steps_f32_noalloc clear");
  if (rc) return rc;
  float* work[6];
  for (int q = 0; q < 6; ++q) work[q] = (float*)(base + q * plane);
  return adjoint_steps_impl(p, n_steps, dt, qs, q1s, lam, work, pitch, Hi, Hi_pitch, f_cells, boundary_conditions,
                            frs, base + 6 * plane, flag, stream);
}

int gpo_cdklm16_substep_adj_f32(const gpo_cdklm16_params* p, float dt, int32_t rk_step, const float* in_eta,
                                const float* in_hu, const float* in_hv, const float* out_eta,
                                const float* lam_out_eta, const float* lam_out_hu, const float* lam_out_hv,
                                float* lam_in_eta, float* lam_in_hu, float* lam_in_hv, float* lam_qn_eta,
                                float* lam_qn_hu, float* lam_qn_hv, int32_t accumulate, int32_t pitch,
                                const float* Hi, int32_t Hi_pitch, const float* f_cells,
                                uint32_t boundary_conditions, void* workspace, int32_t* flag, void* stream) {
  Geo g;
  int rc = make_geo(p, boundary_conditions, pitch, Hi_pitch, f_cells ? pitch : 4, 4, &g);
  if (rc) return rc;
  if (!in_eta || !out_eta || !lam_out_eta || !lam_in_eta || !Hi || !workspace || rk_step < 0 || rk_step > 1)
    return bad_arg("gpo_cdklm16_substep_adj_f32: bad arguments");
  const float* hi = land_hi(&g, Hi, workspace, (hipStream_t)stream);
  return adjoint_stages<float>(g, dt, rk_step, in_eta, in_hu, in_hv, out_eta, lam_out_eta, lam_out_hu, lam_out_hv,
                               lam_in_eta, lam_in_hu, lam_in_hv, lam_qn_eta, lam_qn_hu, lam_qn_hv, accumulate, 0, hi,
                               f_cells, workspace, flag, (hipStream_t)stream);
}

int gpo_cdklm16_substep_adj_f64(const gpo_cdklm16_params* p, double dt, int32_t rk_step, const double* in_eta,
                                const double* in_hu, const double* in_hv, const double* out_eta,
                                const double* lam_out_eta, const double* lam_out_hu, const double* lam_out_hv,
                                double* lam_in_eta, double* lam_in_hu, double* lam_in_hv, double* lam_qn_eta,
                                double* lam_qn_hu, double* lam_qn_hv, int32_t accumulate, int32_t pitch,
                                const double* Hi, int32_t Hi_pitch, const double* f_cells,
                                uint32_t boundary_conditions, void* workspace, int32_t* flag, void* stream) {
  Geo g;
  int rc = make_geo(p, boundary_conditions, pitch, Hi_pitch, f_cells ? pitch : 8, 8, &g);
  if (rc) return rc;
  if (!in_eta || !out_eta || !lam_out_eta || !lam_in_eta || !Hi || !workspace || rk_step < 0 || rk_step > 1)
    return bad_arg("gpo_cdklm16_substep_adj_f64: bad arguments");
  const double* hi = land_hi(&g, Hi, workspace, (hipStream_t)stream);
  return adjoint_stages<double>(g, dt, rk_step, in_eta, in_hu, in_hv, out_eta, lam_out_eta, lam_out_hu, lam_out_hv,
                                lam_in_eta, lam_in_hu, lam_in_hv, lam_qn_eta, lam_qn_hu, lam_qn_hv, accumulate, 0, hi,
                                f_cells, workspace, flag, (hipStream_t)stream);
}

int gpo_balance_adj_f32(const gpo_cdklm16_params* p, float g_, float f, float beta, const float* f_cells,
                        int32_t f_pitch, const float* coarse, int32_t coarse_pitch, float* lam_eta,
                        float* lam_hu, float* lam_hv, int32_t pitch, const float* Hi, int32_t Hi_pitch,
                        void* stream) {
  if (!p || !coarse || !lam_eta || !lam_hu || !lam_hv || !Hi) return bad_arg("gpo_balance_adj_f32: bad arguments");
  Geo g{};
  g.nx = p->nx;
  g.ny = p->ny;
  g.W = p->nx + 4;
  g.sp = pitch / 4;
  g.hp = Hi_pitch / 4;
  g.dx = p->dx;
  g.dy = p->dy;
  g.g = g_;
  g.f = f;
  g.north_x = p->north_x;
  g.north_y = p->north_y;
  hipLaunchKernelGGL(balance_rev<float>, inner_grid(g), kBlock, 0, (hipStream_t)stream, g, (double)beta, f_cells,
                     f_pitch, coarse, coarse_pitch, Hi, p->land_value, lam_eta, lam_hu, lam_hv);
  return hip_check(hipGetLastError(), "gpo_balance_adj_f32 launch");
}

int gpo_balance_adj_f64(const gpo_cdklm16_params* p, double g_, double f, double beta, const float* f_cells,
                        int32_t f_pitch, const float* coarse, int32_t coarse_pitch, double* lam_eta,
                        double* lam_hu, double* lam_hv, int32_t pitch, const double* Hi, int32_t Hi_pitch,
                        void* stream) {
  if (!p || !coarse || !lam_eta || !lam_hu || !lam_hv || !Hi) return bad_arg("gpo_balance_adj_f64: bad arguments");
  Geo g{};
  g.nx = p->nx, g.ny = p->ny, g.W = p->nx + 4, g.sp = pitch / 8, g.hp = Hi_pitch / 8;
  g.dx = p->dx, g.dy = p->dy, g.g = g_, g.f = f, g.north_x = p->north_x, g.north_y = p->north_y;
  hipLaunchKernelGGL(balance_rev<double>, inner_grid(g), kBlock, 0, (hipStream_t)stream, g, beta, f_cells, f_pitch,
                     coarse, coarse_pitch, Hi, p->land_value, lam_eta, lam_hu, lam_hv);
  return hip_check(hipGetLastError(), "gpo_balance_adj_f64 launch");
}

int gpo_balance_f64(const gpo_cdklm16_params* p, double g_, double f, double beta, const float* f_cells,
                    int32_t f_pitch, const float* coarse, int32_t coarse_pitch, double* eta, double* hu, double* hv,
                    double* d_eta, double* d_hu, double* d_hv, int32_t pitch, const double* Hi, int32_t Hi_pitch,
                    void* stream) {
  if (!p || !coarse || !eta || !hu || !hv || !Hi || (d_eta && (!d_hu || !d_hv)))
    return bad_arg("gpo_balance_f64: bad arguments");
  Geo g{};
  g.nx = p->nx, g.ny = p->ny, g.W = p->nx + 4, g.sp = pitch / 8, g.hp = Hi_pitch / 8;
  g.dx = p->dx, g.dy = p->dy, g.g = g_, g.f = f, g.north_x = p->north_x, g.north_y = p->north_y;
  hipLaunchKernelGGL(balance_fwd<double>, inner_grid(g), kBlock, 0, (hipStream_t)stream, g, beta, f_cells, f_pitch,
                     coarse, coarse_pitch, Hi, p->land_value, eta, hu, hv, d_eta, d_hu, d_hv);
  return hip_check(hipGetLastError(), "gpo_balance_f64 launch");
}

}  // extern "C"

namespace {
template <class T>
int drift_geo(int32_t nx, int32_t ny, double dx, double dy, double dt, int32_t pitch, int32_t Bm_pitch,
              uint32_t bc, int32_t n, double sens, DriftGeo* q) {
  if (nx < 1 || ny < 1 || n < 0 || !(dx > 0) || !(dy > 0) || pitch % (int)sizeof(T) || Bm_pitch % (int)sizeof(T))
    return bad_arg("drifter derivative: bad arguments");
  q->nx = nx, q->ny = ny, q->sp = pitch / (int)sizeof(T), q->bp = Bm_pitch / (int)sizeof(T);
  q->per_ns = ((bc >> 24) & 0xFF) == 2 && ((bc >> 16) & 0xFF) == 2;
  q->per_ew = ((bc >> 8) & 0xFF) == 2 && (bc & 0xFF) == 2;
  q->n = n, q->dx = dx, q->dy = dy, q->dt = dt, q->sens = sens;
  return 0;
}
}  // namespace

extern "C" {

int gpo_drift_f64(int32_t nx, int32_t ny, double dx, double dy, double dt, const double* eta, const double* hu,
                  const double* hv, int32_t pitch, const double* Bm, int32_t Bm_pitch, uint32_t boundary_conditions,
                  int32_t n, double* positions, double* d_positions, const double* d_eta, const double* d_hu,
                  const double* d_hv, double sensitivity, void* stream) {
  DriftGeo q;
  int rc = drift_geo<double>(nx, ny, dx, dy, dt, pitch, Bm_pitch, boundary_conditions, n, sensitivity, &q);
  if (rc) return rc;
  if (!eta || !hu || !hv || !Bm || !positions) return bad_arg("gpo_drift_f64: bad arguments");
  if (n == 0) return 0;
  hipLaunchKernelGGL(drift_fwd<double>, dim3((n + 63) / 64), dim3(64), 0, (hipStream_t)stream, q, eta, hu, hv, Bm,
                     positions, d_positions, d_eta, d_hu, d_hv);
  return hip_check(hipGetLastError(), "gpo_drift_f64 launch");
}

int gpo_drift_adj_f32(int32_t nx, int32_t ny, float dx, float dy, float dt, const float* eta, const float* hu,
                      const float* hv, int32_t pitch, const float* Bm, int32_t Bm_pitch, uint32_t boundary_conditions,
                      int32_t n, const float* positions, float* lam_positions, float* lam_eta, float* lam_hu,
                      float* lam_hv, float sensitivity, void* stream) {
  DriftGeo q;
  int rc = drift_geo<float>(nx, ny, dx, dy, dt, pitch, Bm_pitch, boundary_conditions, n, sensitivity, &q);
  if (rc) return rc;
  if (!eta || !hu || !hv || !Bm || !positions || !lam_positions || !lam_eta || !lam_hu || !lam_hv)
    return bad_arg("gpo_drift_adj_f32: bad arguments");
  if (n == 0) return 0;
  hipLaunchKernelGGL(drift_rev<float>, dim3(1), dim3(256), 0, (hipStream_t)stream, q, eta, hu, hv, Bm,
                     positions, lam_positions, lam_eta, lam_hu, lam_hv);
  return hip_check(hipGetLastError(), "gpo_drift_adj_f32 launch");
}

int gpo_drift_adj_f64(int32_t nx, int32_t ny, double dx, double dy, double dt, const double* eta, const double* hu,
                      const double* hv, int32_t pitch, const double* Bm, int32_t Bm_pitch,
                      uint32_t boundary_conditions, int32_t n, const double* positions, double* lam_positions,
                      double* lam_eta, double* lam_hu, double* lam_hv, double sensitivity, void* stream) {
  DriftGeo q;
  int rc = drift_geo<double>(nx, ny, dx, dy, dt, pitch, Bm_pitch, boundary_conditions, n, sensitivity, &q);
  if (rc) return rc;
  if (!eta || !hu || !hv || !Bm || !positions || !lam_positions || !lam_eta || !lam_hu || !lam_hv)
    return bad_arg("gpo_drift_adj_f64: bad arguments");
  if (n == 0) return 0;
  hipLaunchKernelGGL(drift_rev<double>, dim3(1), dim3(256), 0, (hipStream_t)stream, q, eta, hu, hv, Bm,
                     positions, lam_positions, lam_eta, lam_hu, lam_hv);
  return hip_check(hipGetLastError(), "gpo_drift_adj_f64 launch");
}

int gpo_point_misfit_f32(int32_t nx, int32_t ny, float dx, float dy, const float* eta, const float* hu,
                         const float* hv, int32_t pitch, const float* Hi, int32_t Hi_pitch, int32_t n,
                         const float* positions, const float* observed, double* J, float* lam_eta, float* lam_hu,
                         float* lam_hv, void* stream) {
  if (nx < 2 || ny < 2 || n < 0 || !eta || !hu || !hv || !Hi || !positions || !observed || !J)
    return bad_arg("gpo_point_misfit_f32: bad arguments");
  Geo g{};
  g.nx = nx;
  g.ny = ny;
  g.sp = pitch / 4;
  g.hp = Hi_pitch / 4;
  g.dx = dx;
  g.dy = dy;
  if (n == 0) return 0;
  hipLaunchKernelGGL(misfit<float>, dim3(1), dim3(256), 0, (hipStream_t)stream, g, n, positions,
                     observed, eta, hu, hv, Hi, J, lam_eta, lam_hu, lam_hv);
  return hip_check(hipGetLastError(), "gpo_point_misfit_f32 launch");
}

int gpo_point_misfit_f64(int32_t nx, int32_t ny, double dx, double dy, const double* eta, const double* hu,
                         const double* hv, int32_t pitch, const double* Hi, int32_t Hi_pitch, int32_t n,
                         const float* positions, const float* observed, double* J, double* lam_eta, double* lam_hu,
                         double* lam_hv, void* stream) {
  if (nx < 2 || ny < 2 || n < 0 || !eta || !hu || !hv || !Hi || !positions || !observed || !J)
    return bad_arg("gpo_point_misfit_f64: bad arguments");
  Geo g{};
  g.nx = nx;
  g.ny = ny;
  g.sp = pitch / 8;
  g.hp = Hi_pitch / 8;
  g.dx = dx;
  g.dy = dy;
  if (n == 0) return 0;
  hipLaunchKernelGGL(misfit<double>, dim3(1), dim3(256), 0, (hipStream_t)stream, g, n, positions,
                     observed, eta, hu, hv, Hi, J, lam_eta, lam_hu, lam_hv);
  return hip_check(hipGetLastError(), "gpo_point_misfit_f64 launch");
}

int gpo_frs_adj_f32(const gpo_cdklm16_params* p, uint32_t boundary_conditions, const int32_t sponge[4],
                    float* lam_eta, float* lam_hu, float* lam_hv, int32_t pitch, void* stream) {
  Geo g;
  int rc = make_geo(p, boundary_conditions, pitch, 4 * (p ? p->nx + 5 : 0), 4, 4, &g);
  if (rc) return rc;
  if (!sponge || !lam_eta || !lam_hu || !lam_hv) return bad_arg("gpo_frs_adj_f32: bad arguments");
  hipLaunchKernelGGL(frs_scale<float>, cell_grid(g), kBlock, 0, (hipStream_t)stream, g,
                     Sponge{sponge[0], sponge[1], sponge[2], sponge[3]}, lam_eta, lam_hu, lam_hv);
  return hip_check(hipGetLastError(), "gpo_frs_adj_f32 launch");
}

int gpo_frs_adj_f64(const gpo_cdklm16_params* p, uint32_t boundary_conditions, const int32_t sponge[4],
                    double* lam_eta, double* lam_hu, double* lam_hv, int32_t pitch, void* stream) {
  Geo g;
  int rc = make_geo(p, boundary_conditions, pitch, 8 * (p ? p->nx + 5 : 0), 8, 8, &g);
  if (rc) return rc;
  if (!sponge || !lam_eta || !lam_hu || !lam_hv) return bad_arg("gpo_frs_adj_f64: bad arguments");
  hipLaunchKernelGGL(frs_scale<double>, cell_grid(g), kBlock, 0, (hipStream_t)stream, g,
                     Sponge{sponge[0], sponge[1], sponge[2], sponge[3]}, lam_eta, lam_hu, lam_hv);
  return hip_check(hipGetLastError(), "gpo_frs_adj_f64 launch");
}

int gpo_frs_apply_f64(const gpo_cdklm16_params* p, uint32_t boundary_conditions, const gpo_frs* frs, double* eta,
                      double* hu, double* hv, int32_t pitch, void* stream) {
  Geo g;
  int rc = make_geo(p, boundary_conditions, pitch, 8 * (p ? p->nx + 5 : 0), 8, 8, &g);
  if (rc) return rc;
  if (!frs || !eta || !hu || !hv) return bad_arg("gpo_frs_apply_f64: bad arguments");
  const bool ns = g.bc_n == 3 || g.bc_s == 3, ew = g.bc_e == 3 || g.bc_w == 3;
  if ((ns && (!frs->ns_cur || !frs->ns_next || frs->ns_nx < 1)) || (ew && (!frs->ew_cur || !frs->ew_next || frs->ew_ny < 1)))
    return bad_arg("gpo_frs_apply_f64: boundary data missing");
  hipLaunchKernelGGL(frs_blend_f64, cell_grid(g), kBlock, 0, (hipStream_t)stream, g,
                     Sponge{frs->sponge[0], frs->sponge[1], frs->sponge[2], frs->sponge[3]}, frs->ns_cur,
                     frs->ns_next, frs->ns_nx, frs->ew_cur, frs->ew_next, frs->ew_ny, frs->t, eta, hu, hv);
  return hip_check(hipGetLastError(), "gpo_frs_apply_f64 launch");
}

// ---- batched differentiable ensemble

int gpo_adj_batch_create(const gpo_adj_member* members, int32_t n, int32_t k, int32_t n_segments,
                         gpo_adj_batch* out, void* stream) {
  if (!members || n < 1 || n > 65535 || k < 1 || n_segments < 1 || !out)
    return bad_arg("gpo_adj_batch_create: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  AdjBatch* b = new (std::nothrow) AdjBatch();
  if (!b) return bad_arg("gpo_adj_batch_create: out of host memory");
  b->n = n, b->k = k, b->n_seg = n_segments;
  std::vector<StepArgs> fwd(2 * (size_t)n), traj(2 * (size_t)n * k);
  std::vector<AdjItem<float>> items(2 * (size_t)n * k);
  std::vector<CopyItem> copies((3 * (size_t)n_segments + 2) * n);
  std::vector<int2> work;
  std::vector<NoiseDriftItem> nds;
  std::vector<int2> ntiles;
  std::vector<std::vector<cells::BcTarget>> bcs(2 + 2 * (size_t)k);  // per slot, one per flow-relaxation member
  long long max_bc_cells = 0;
  for (int m = 0; m < n; ++m) {
    const gpo_adj_member& M = members[m];
    gpo_cdklm16_params p = M.params;
    Geo g;
    int rc = make_geo(&p, M.boundary_conditions, M.pitch, M.Hi_pitch, M.f_cells ? M.pitch : 4, 4, &g);
    if (!rc && (!M.state || !M.stage || !M.ckpt || !M.qs || !M.q1s || !M.lam || !M.workspace || !M.Hi ||
                M.plane < (int64_t)M.pitch * (p.ny + 4) || !(M.dt > 0.0f)))
      rc = bad_arg("gpo_adj_batch_create: member buffers, plane or dt");
    const bool frs = g.bc_n == 3 || g.bc_s == 3 || g.bc_e == 3 || g.bc_w == 3;
    if (!rc && frs && !M.has_frs) rc = bad_arg("gpo_adj_batch_create: flow-relaxation member without frs data");
    if (m == 0) b->has_land = p.has_land ? 1 : 0, b->f_array = p.f_is_array, b->angle0 = p.angle_is_zero;
    if (!rc && ((p.has_land ? 1 : 0) != b->has_land || p.f_is_array != b->f_array ||
                p.angle_is_zero != b->angle0 || (p.f_is_array && !M.f_cells)))
      rc = bad_arg("gpo_adj_batch_create: members differ in land, coriolis texture or north angle");
    if (rc) {
      adj_batch_free(b);
      return rc;
    }
    if (p.chunk_rows == 0) p.chunk_rows = 16;
    Geo ga = g;  // the reverse stages read the land-flagged copy of Hi
    const float* hi_adj = land_hi(&ga, M.Hi, M.workspace, s);
    const size_t pl = (size_t)M.plane / 4, st = 3 * pl;  // elements per plane / per state
    auto planes = [&](float* base, size_t i, float** e, float** u, float** v) {
      *e = base + i * st, *u = base + i * st + pl, *v = base + i * st + 2 * pl;
    };
    float *se, *su, *sv, *ge, *gu, *gv;
    planes(M.state, 0, &se, &su, &sv);
    planes(M.stage, 0, &ge, &gu, &gv);
    cells::BcTarget bt{};
    if (frs) {
      bt.nx = p.nx, bt.ny = p.ny, bt.bn = g.bc_n, bt.be = g.bc_e, bt.bs = g.bc_s, bt.bw = g.bc_w;
      bt.sn = M.frs.sponge[0], bt.se = M.frs.sponge[1], bt.ss = M.frs.sponge[2], bt.sw = M.frs.sponge[3];
      cells::bc_target_extent(&bt);
      bt.pitch = M.pitch;
      bt.ns_cur = M.frs.ns_cur, bt.ns_next = M.frs.ns_next, bt.ns_nx = M.frs.ns_nx;
      bt.ew_cur = M.frs.ew_cur, bt.ew_next = M.frs.ew_next, bt.ew_ny = M.frs.ew_ny;
      bt.frs_t = M.frs.t;
      if (((bt.bn == 3 || bt.bs == 3) && bt.bn != 2 && (!bt.ns_cur || !bt.ns_next || bt.ns_nx < 1)) ||
          ((bt.be == 3 || bt.bw == 3) && bt.be != 2 && (!bt.ew_cur || !bt.ew_next || bt.ew_ny < 1))) {
        adj_batch_free(b);
        return bad_arg("gpo_adj_batch_create: flow-relaxation data missing");
      }
      max_bc_cells = std::max(max_bc_cells, cells::bc_target_cells(bt));
      auto put = [&](size_t slot, float* e_, float* u_, float* v_) {
        cells::BcTarget t = bt;
        t.e = e_, t.u = u_, t.v = v_;
        bcs[slot].push_back(t);
      };
      put(0, ge, gu, gv);
      put(1, se, su, sv);
      for (int j = 0; j < k; ++j) {
        float *be_, *bu_, *bv_, *ce_, *cu_, *cv_;
        planes(M.q1s, j, &be_, &bu_, &bv_);
        planes(M.qs, j + 1, &ce_, &cu_, &cv_);
        put(2 + 2 * (size_t)j, be_, bu_, bv_);
        put(3 + 2 * (size_t)j, ce_, cu_, cv_);
      }
    }
    rc = make_step_args(&p, M.dt, 0, se, su, sv, nullptr, nullptr, nullptr, ge, gu, gv, M.pitch, M.Hi, M.Hi_pitch,
                        M.f_cells, nullptr, M.boundary_conditions, &fwd[2 * m]);
    if (!rc)
      rc = make_step_args(&p, M.dt, 1, ge, gu, gv, se, su, sv, se, su, sv, M.pitch, M.Hi, M.Hi_pitch, M.f_cells,
                          nullptr, M.boundary_conditions, &fwd[2 * m + 1]);
    if (!rc && !(fast_arith_ok(fwd[2 * m]) && fast_arith_ok(fwd[2 * m + 1]))) b->all_ok = 0;
    for (int j = 0; j < k && !rc; ++j) {
      float *ae, *au, *av, *be, *bu, *bv, *ce, *cu, *cv;
      planes(M.qs, j, &ae, &au, &av);
      planes(M.q1s, j, &be, &bu, &bv);
      planes(M.qs, j + 1, &ce, &cu, &cv);
      StepArgs* t = &traj[(size_t)j * 2 * n];
      rc = make_step_args(&p, M.dt, 0, ae, au, av, nullptr, nullptr, nullptr, be, bu, bv, M.pitch, M.Hi, M.Hi_pitch,
                          M.f_cells, nullptr, M.boundary_conditions, &t[2 * m]);
      if (!rc)
        rc = make_step_args(&p, M.dt, 1, be, bu, bv, ae, au, av, ce, cu, cv, M.pitch, M.Hi, M.Hi_pitch, M.f_cells,
                            nullptr, M.boundary_conditions, &t[2 * m + 1]);
      // lambda buffers: buf[0] = lam, buf[1] = lam + 1 state, stage lambda = lam + 2 states
      float* buf[2] = {M.lam, M.lam + st};
      float* l1 = M.lam + 2 * st;
      const Fwd<float> w = carve<float>(M.workspace, p.nx, p.ny);
      const Rev<float> rv = carve_rev<float>(M.workspace, p.nx, p.ny);
      for (int rk = 0; rk < 2; ++rk) {
        AdjItem<float>& it = items[((size_t)2 * j + rk) * n + m];
        it.g = ga;
        it.dt = M.dt;
        it.rk = rk;
        it.Hi = gptr(hi_adj);
        it.fc = gptr(M.f_cells);
        if (M.f_cells) b->any_fc = 1;
        it.w = w;
        it.rv = rv;
        it.frs = frs ? 1 : 0;
        if (frs) it.sn = M.frs.sponge[0], it.se = M.frs.sponge[1], it.ss = M.frs.sponge[2], it.sw = M.frs.sponge[3];
        float* lo = rk ? buf[(j + 1) & 1] : l1;
        float* li = rk ? l1 : buf[j & 1];
        it.le = gptr(lo), it.lu = gptr(lo + pl), it.lv = gptr(lo + 2 * pl);
        it.lie = gptr(li), it.liu = gptr(li + pl), it.liv = gptr(li + 2 * pl);
        if (rk) {
          // the second substep's own output: before the model error when there is one
          it.ie = gptr(be), it.iu = gptr(bu), it.iv = gptr(bv), it.out_e = gptr(M.has_noise ? M.pre + j * pl : ce);
          float* lq = buf[j & 1];
          it.lqe = gptr(lq), it.lqu = gptr(lq + pl), it.lqv = gptr(lq + 2 * pl);
          it.accumulate = 0, it.lq_mode = 1;
        } else {
          it.ie = gptr(ae), it.iu = gptr(au), it.iv = gptr(av), it.out_e = gptr(be);
          it.lqe = it.lqu = it.lqv = nullptr;
          it.accumulate = 1, it.lq_mode = 0;
        }
      }
    }
    if (rc) {
      adj_batch_free(b);
      return rc;
    }
    for (int sg = 0; sg < n_segments; ++sg) {
      copies[(size_t)sg * n + m] = CopyItem{gptr(M.state), gptr(M.ckpt + sg * st), p.nx, p.ny, g.sp, pl};
      copies[((size_t)n_segments + sg) * n + m] = CopyItem{gptr(M.ckpt + sg * st), gptr(M.qs), p.nx, p.ny, g.sp, pl};
    }
    copies[(size_t)2 * n_segments * n + m] = CopyItem{gptr(M.lam), gptr(M.lam + st), p.nx, p.ny, g.sp, pl};
    for (int sg = 0; sg < n_segments; ++sg)
      copies[((size_t)2 * n_segments + 1 + sg) * n + m] =
          CopyItem{gptr(M.ckpt + sg * st), gptr(M.state), p.nx, p.ny, g.sp, pl};
    copies[((size_t)3 * n_segments + 1) * n + m] = CopyItem{gptr(M.state), gptr(M.qs), p.nx, p.ny, g.sp, pl};
    const StepArgs& a = fwd[2 * m];
    if (a.n_strips > 0xFFFF || a.n_chunks > 0x7FFF) {
      adj_batch_free(b);
      return bad_arg("gpo_adj_batch_create: member too large for the work table");
    }
    if (frs) {
      const int W = p.nx + 4, H = p.ny + 4;
      const int cols = std::min(W, (g.bc_w == 3 ? M.frs.sponge[3] : 0) + (g.bc_e == 3 ? M.frs.sponge[1] : 0));
      const int rows = std::min(H, (g.bc_s == 3 ? M.frs.sponge[2] : 0) + (g.bc_n == 3 ? M.frs.sponge[0] : 0));
      b->sponge_cells = std::max(b->sponge_cells, std::max(cols * H, rows * W));
    }
    for (int sy = 0; sy < a.n_chunks; ++sy)
      for (int sx = 0; sx < a.n_strips; ++sx) work.push_back(make_int2(m, sx | (sy << 16)));
    b->max_w = std::max(b->max_w, p.nx + 4);
    b->max_h = std::max(b->max_h, p.ny + 4);
    // model error and drifters
    NoiseDriftItem nd{};
    nd.g = g;
    nd.per_ns = g.bc_n == 2 && g.bc_s == 2;
    nd.per_ew = g.bc_e == 2 && g.bc_w == 2;
    nd.Hi = M.Hi;
    nd.state = M.state, nd.qs = M.qs, nd.pre = M.pre, nd.lam = M.lam;
    nd.plane = pl, nd.st = st;
    nd.dt = M.dt;
    nd.land = p.land_value;
    if (M.has_noise) {
      if (!M.xi || !M.coarse || !M.pre || M.xi_rows != (nd.per_ns ? p.ny : p.ny + 8) ||
          M.xi_cols != (nd.per_ew ? p.nx : p.nx + 8) || p.f_is_array || M.f_cells) {
        adj_batch_free(b);
        return bad_arg("gpo_adj_batch_create: model-error buffers (xi, coarse, pre) missing or mis-sized, or f not "
                       "scalar");
      }
      nd.noise = 1;
      nd.xi_rows = M.xi_rows, nd.xi_cols = M.xi_cols;
      nd.k0 = (uint32_t)(M.seed & 0xffffffffu), nd.k1 = (uint32_t)(M.seed >> 32);
      nd.stream_id = M.stream_id, nd.draw0 = M.draw0;
      for (int i = 0; i < 25; ++i) nd.W[i] = M.soar_weights[i];
      nd.g.g = M.g, nd.g.f = M.f, nd.beta = M.beta;
      nd.xi = M.xi, nd.coarse = M.coarse;
      nd.coarse_seg = M.coarse_seg;
      if (m == 0) b->coarse_saved = M.coarse_seg ? 1 : 0;
      if ((M.coarse_seg ? 1 : 0) != b->coarse_saved) {
        adj_batch_free(b);
        return bad_arg("gpo_adj_batch_create: members differ in whether their segment SOAR fields are kept");
      }
      b->any_noise = 1;
      for (int ty = 0; ty < (p.ny + kTileH - 1) / kTileH; ++ty)
        for (int tx = 0; tx < (p.nx + kTileW - 1) / kTileW; ++tx) ntiles.push_back(make_int2(m, (ty << 16) | tx));
      b->max_xi = std::max(b->max_xi, M.xi_rows * M.xi_cols);
      b->max_cells = std::max(b->max_cells, (p.nx + 4) * (p.ny + 4));
    }
    if (M.n_drifters > 0) {
      if (!M.Bm || !M.pos || !M.pos_ckpt || !M.pos_seg || !M.lam_pos || M.Bm_pitch % 4) {
        adj_batch_free(b);
        return bad_arg("gpo_adj_batch_create: drifter buffers missing");
      }
      nd.n_drifters = M.n_drifters;
      nd.Bm = M.Bm, nd.bm_pitch = M.Bm_pitch / 4;
      nd.pos = M.pos, nd.pos_ckpt = M.pos_ckpt, nd.pos_seg = M.pos_seg, nd.lam_pos = M.lam_pos;
      b->any_drift = 1;
      b->max_drifters = std::max(b->max_drifters, M.n_drifters);
    }
    nds.push_back(nd);
  }
  for (int m = 0; m < n; ++m) b->nxs.push_back(members[m].params.nx), b->nys.push_back(members[m].params.ny);
  b->n_work = (int)work.size();
  const long long waves = std::min<long long>((long long)persistent_waves(8), b->n_work);
  b->n_blocks = (int)((waves + kWavesPerBlock - 1) / kWavesPerBlock);
  int rc = hip_check(hipMalloc(&b->d_fwd, sizeof(StepArgs) * fwd.size()), "gpo_adj_batch_create hipMalloc");
  if (!rc) rc = hip_check(hipMalloc(&b->d_traj, sizeof(StepArgs) * traj.size()), "gpo_adj_batch_create hipMalloc");
  if (!rc) rc = hip_check(hipMalloc(&b->d_work, sizeof(int2) * work.size()), "gpo_adj_batch_create hipMalloc");
  if (!rc)
    rc = hip_check(hipMalloc(&b->d_items, sizeof(AdjItem<float>) * items.size()), "gpo_adj_batch_create hipMalloc");
  if (!rc) rc = hip_check(hipMalloc(&b->d_copy, sizeof(CopyItem) * copies.size()), "gpo_adj_batch_create hipMalloc");
  if (!rc)
    rc = hip_check(hipMemcpyAsync(b->d_fwd, fwd.data(), sizeof(StepArgs) * fwd.size(), hipMemcpyHostToDevice, s),
                   "gpo_adj_batch_create copy");
  if (!rc)
    rc = hip_check(hipMemcpyAsync(b->d_traj, traj.data(), sizeof(StepArgs) * traj.size(), hipMemcpyHostToDevice, s),
                   "gpo_adj_batch_create copy");
  if (!rc)
    rc = hip_check(hipMemcpyAsync(b->d_work, work.data(), sizeof(int2) * work.size(), hipMemcpyHostToDevice, s),
                   "gpo_adj_batch_create copy");
  if (!rc)
    rc = hip_check(hipMemcpyAsync(b->d_items, items.data(), sizeof(AdjItem<float>) * items.size(),
                                  hipMemcpyHostToDevice, s),
                   "gpo_adj_batch_create copy");
  if (!rc)
    rc = hip_check(hipMemcpyAsync(b->d_copy, copies.data(), sizeof(CopyItem) * copies.size(), hipMemcpyHostToDevice,
                                  s),
                   "gpo_adj_batch_create copy");
  b->n_frs = (int)bcs[0].size();
  if (!rc && b->n_frs > 0) {
    std::vector<cells::BcTarget> flat;
    for (const auto& v : bcs) flat.insert(flat.end(), v.begin(), v.end());
    b->bc_blocks = (int)std::max<long long>(1, (max_bc_cells + 255) / 256);
    rc = hip_check(hipMalloc(&b->d_bc, sizeof(cells::BcTarget) * flat.size()), "gpo_adj_batch_create hipMalloc");
    if (!rc)
      rc = hip_check(hipMemcpyAsync(b->d_bc, flat.data(), sizeof(cells::BcTarget) * flat.size(),
                                    hipMemcpyHostToDevice, s),
                     "gpo_adj_batch_create copy");
  }
  b->n_ntiles = (int)ntiles.size();
  if (!rc && b->n_ntiles > 0) {
    rc = hip_check(hipMalloc(&b->d_ntiles, sizeof(int2) * ntiles.size()), "gpo_adj_batch_create hipMalloc");
    if (!rc)
      rc = hip_check(hipMemcpyAsync(b->d_ntiles, ntiles.data(), sizeof(int2) * ntiles.size(), hipMemcpyHostToDevice,
                                    s),
                     "gpo_adj_batch_create copy");
  }
  if (!rc && (b->any_noise || b->any_drift)) {
    rc = hip_check(hipMalloc(&b->d_nd, sizeof(NoiseDriftItem) * nds.size()), "gpo_adj_batch_create hipMalloc");
    if (!rc)
      rc = hip_check(hipMemcpyAsync(b->d_nd, nds.data(), sizeof(NoiseDriftItem) * nds.size(), hipMemcpyHostToDevice,
                                    s),
                     "gpo_adj_batch_create copy");
  }
  if (!rc) rc = hip_check(hipStreamSynchronize(s), "gpo_adj_batch_create sync");
  if (!rc) rc = adj_batch_march_table(b, kDefaultMarchRows, s);
  if (rc) {
    adj_batch_free(b);
    return rc;
  }
  *out = (gpo_adj_batch)b;
  return 0;
}

int gpo_adj_batch_set_reverse(gpo_adj_batch h, int32_t fused, int32_t rows, void* stream) {
  AdjBatch* b = (AdjBatch*)h;
  if (!b || rows < 0 || rows > 4096) return bad_arg("gpo_adj_batch_set_reverse: bad arguments");
  int rc = hip_check(hipStreamSynchronize((hipStream_t)stream), "gpo_adj_batch_set_reverse sync");
  if (!rc && rows > 0) rc = adj_batch_march_table(b, rows, (hipStream_t)stream);
  if (!rc) b->use_march = fused ? 1 : 0;
  return rc;
}

int gpo_adj_batch_forward(gpo_adj_batch h, int32_t n_steps, void* stream) {
  AdjBatch* b = (AdjBatch*)h;
  if (!b || n_steps < 0) return bad_arg("gpo_adj_batch_forward: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  for (int i = 0; i < n_steps; ++i) {
    int rc = adj_batch_steps(b, b->d_fwd, 0, s);
    if (!rc) rc = adj_batch_bc(b, 0, s);
    if (!rc) rc = adj_batch_steps(b, b->d_fwd, 1, s);
    if (!rc) rc = adj_batch_noise(b, b->step, -1, -1, s);
    if (!rc) rc = adj_batch_bc(b, 1, s);
    if (!rc) rc = adj_batch_drift(b, -1, s);
    if (rc) return rc;
    ++b->step;
  }
  return 0;
}

int gpo_adj_batch_slot(gpo_adj_batch h, int32_t slot, int32_t mode, int32_t segment, void* stream) {
  AdjBatch* b = (AdjBatch*)h;
  if (!b || slot < 0 || slot >= b->n_seg || mode < 0 || mode > 3 || segment < 0)
    return bad_arg("gpo_adj_batch_slot: bad slot, mode or segment");
  size_t base = 0;
  if (mode == 0) base = (size_t)slot * b->n;                            // forward state -> slot
  if (mode == 1) base = ((size_t)b->n_seg + slot) * b->n;               // slot -> segment start
  if (mode == 2) base = ((size_t)2 * b->n_seg + 1 + slot) * b->n;       // slot -> forward state
  if (mode == 3) base = ((size_t)3 * b->n_seg + 1) * b->n;              // forward state -> segment start
  hipStream_t s = (hipStream_t)stream;
  if (mode == 1 || mode == 3)
    b->seg_start = segment * b->k;
  else
    b->step = segment * b->k;
  if (b->any_drift) {
    hipLaunchKernelGGL(nd_pos_copy_b, dim3((2 * b->max_drifters + 63) / 64, b->n), dim3(64), 0, s, b->d_nd, mode,
                       slot);
    const int rc = hip_check(hipGetLastError(), "batched drifter checkpoint");
    if (rc) return rc;
  }
  return adj_batch_copy(b, b->d_copy + base, s);
}

int gpo_adj_batch_checkpoint(gpo_adj_batch h, int32_t segment, int32_t restore, void* stream) {
  AdjBatch* b = (AdjBatch*)h;
  if (!b || segment < 0 || segment >= b->n_seg) return bad_arg("gpo_adj_batch_checkpoint: bad segment");
  return gpo_adj_batch_slot(h, segment, restore ? 1 : 0, segment, stream);
}

int gpo_adj_batch_trajectory(gpo_adj_batch h, int32_t m_steps, void* stream) {
  AdjBatch* b = (AdjBatch*)h;
  if (!b || m_steps < 0 || m_steps > b->k) return bad_arg("gpo_adj_batch_trajectory: 0 <= steps <= k");
  hipStream_t s = (hipStream_t)stream;
  for (int j = 0; j < m_steps; ++j) {
    const StepArgs* t = b->d_traj + (size_t)j * 2 * b->n;
    int rc = adj_batch_steps(b, t, 0, s);
    if (!rc) rc = adj_batch_bc(b, 2 + 2 * j, s);
    if (!rc) rc = adj_batch_steps(b, t, 1, s);
    if (!rc) rc = adj_batch_noise(b, b->seg_start + j, j + 1, j, s);
    if (!rc) rc = adj_batch_bc(b, 3 + 2 * j, s);
    if (!rc) rc = adj_batch_drift(b, j + 1, s);
    if (rc) return rc;
  }
  return 0;
}

int gpo_adj_batch_sweep(gpo_adj_batch h, int32_t m_steps, int32_t* flag, void* stream) {
  AdjBatch* b = (AdjBatch*)h;
  if (!b || m_steps < 0 || m_steps > b->k || !flag) return bad_arg("gpo_adj_batch_sweep: 0 <= steps <= k, flag");
  hipStream_t s = (hipStream_t)stream;
  // lambda at Q^{j+1} lives in buffer (j+1) & 1: start from buffer m & 1
  int rc = (m_steps & 1) ? adj_batch_copy(b, b->d_copy + (size_t)2 * b->n_seg * b->n, s) : 0;
  for (int j = m_steps - 1; j >= 0 && !rc; --j) {
    const AdjItem<float>* t1 = b->d_items + ((size_t)2 * j + 1) * b->n;
    if (b->any_drift) {  // the drift read the state after the step
      hipLaunchKernelGGL(nd_drift_rev_b, dim3(1, b->n), dim3(256), 0, s, b->d_nd, j);
      rc = hip_check(hipGetLastError(), "batched drifter reverse launch");
    }
    if (!rc && b->any_noise) {  // sponge blend, then the model error, in reverse
      if (b->n_frs) adj_batch_sponge(b, t1, s);
      if (!b->coarse_saved) {   // regenerate the step's SOAR fields
        hipLaunchKernelGGL(nd_xi_b, dim3((b->max_xi + 255) / 256, b->n), dim3(256), 0, s, b->d_nd, b->seg_start + j);
        hipLaunchKernelGGL(nd_soar_b, dim3((b->max_cells + 255) / 256, b->n), dim3(256), 0, s, b->d_nd, -1);
      }
      hipLaunchKernelGGL(nd_balance_rev_b, dim3(b->n_ntiles), dim3(256), 0, s, b->d_nd, b->d_ntiles, j,
                         b->coarse_saved ? j : -1);
      rc = hip_check(hipGetLastError(), "batched model-error reverse launch");
    }
    if (!rc) rc = adj_batch_stage(b, t1, flag, s, !b->any_noise);
    if (!rc) rc = adj_batch_stage(b, b->d_items + ((size_t)2 * j) * b->n, flag, s);
  }
  return rc;
}

int gpo_adj_batch_destroy(gpo_adj_batch h) {
  if (h) adj_batch_free((AdjBatch*)h);
  return 0;
}

}  // extern "C"
