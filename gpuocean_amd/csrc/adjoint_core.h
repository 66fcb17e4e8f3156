// Per-cell stages of the differentiable CDKLM16 substep, shared by the gfx950
// kernels (adjoint.hip) and a host build of the same code used by the CPU
// tests.  Include from translation units built with -ffp-contract=off.
//
//  * Forward stages templated on the arithmetic type: double (the FP64
//    verification build), float (recomputation of intermediates for the FP32
//    adjoint) and Dual (value + one tangent, the tangent-linear model).
//    F1 conversion of every cell (eta, u, v) with wall mirror / periodic wrap,
//    F2 x- and y-slopes (minmod, well-balanced potential, adjustSlopes),
//    F3 face fluxes (central-upwind), F4 sources + RK update.
//  * The hand-written reverse of the same stages: A1 face fluxes, A2 per-cell
//    gather (update + sources, face parts, slope subgradients, conversion),
//    A3 fold of ghost-cell sensitivities onto the interior cells they mirror
//    or wrap.
//
// Scope of the derivative: RK2 substeps of wet cells (desingularised below
// eps, dried to a constant state at the depth cutoff) and land cells
// (constants, as the step kernel treats them), wall and periodic boundaries (flow-relaxation sides read their ghost
// cells as stored constants; the sponge blend between substeps is handled by
// the callers), no friction, wind or pressure forcing.  Branches outside that scope set bits in the
// caller's flag word.  Minmod, max and min take the subgradient of the first
// minimal/maximal argument, the same choice the tangent-linear model makes.
//
// The arithmetic follows gpu_kernels/CDKLM16_kernel.cu:475-1090 (forward).
#pragma once

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <type_traits>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define GPO_HD __host__ __device__
#endif

// Workspace pointers are global memory: in the device pass they carry the
// global address space, so every access through them is a global (not flat)
// instruction.  The host pass and the host build see plain pointers of the
// same size.
#if defined(__HIP_DEVICE_COMPILE__)
#define GPO_GLOBAL __attribute__((address_space(1)))
#else
#define GPO_GLOBAL
#endif

// A pointer converted to the global address space of the device pass.
template <class T> inline GPO_GLOBAL T* gptr(T* p) { return (GPO_GLOBAL T*)p; }

#if defined(__HIPCC__)
#define GPO_FLAG(flag, bits) atomicOr((flag), (bits))
#else
#define GPO_HD
#define GPO_FLAG(flag, bits) (*(flag) |= (bits))
#endif

namespace gpo {
namespace adj {

using std::copysign;
using std::fabs;
using std::floor;
using std::sqrt;

// ----------------------------------------------------------------------------- number types

struct Dual {
  double v, d;
  Dual() = default;
  GPO_HD constexpr Dual(double a) : v(a), d(0.0) {}
  GPO_HD constexpr Dual(double a, double b) : v(a), d(b) {}
};
GPO_HD inline Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
GPO_HD inline Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
GPO_HD inline Dual operator-(Dual a) { return {-a.v, -a.d}; }
GPO_HD inline Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
GPO_HD inline Dual operator/(Dual a, Dual b) {
  const double q = a.v / b.v;
  return {q, (a.d - q * b.d) / b.v};
}
GPO_HD inline Dual tsqrt(Dual a) {
  const double s = sqrt(a.v);
  return {s, s > 0.0 ? a.d / (2.0 * s) : 0.0};
}
GPO_HD inline float tsqrt(float a) {
#if defined(GPO_ADJ_FAST_F32) && defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_sqrtf(a);
#else
  return std::sqrt(a);
#endif
}
GPO_HD inline double tsqrt(double a) { return sqrt(a); }

// Division and square root of the stages.  IEEE for double and Dual, and for
// float by default; a translation unit that defines GPO_ADJ_FAST_F32 before
// including this header (the product's FP32 reverse) gets the hardware
// reciprocal (1 ulp) times the numerator and the hardware square root for
// float on the device.  Written as builtins, these are the same instructions
// wherever the code is inlined, so two kernels that evaluate the same
// expression on the same values agree bit for bit.
template <class T> GPO_HD inline T qdiv(T a, T b) {
#if defined(GPO_ADJ_FAST_F32) && defined(__HIP_DEVICE_COMPILE__)
  if constexpr (std::is_same<T, float>::value) return a * __builtin_amdgcn_rcpf(b);
#endif
  return a / b;
}
GPO_HD inline double val(Dual a) { return a.v; }
GPO_HD inline float val(float a) { return a; }
GPO_HD inline double val(double a) { return a; }

// sign as copysign(1, x): a constant of the derivative
template <class T> GPO_HD inline T sgn1(T x) { return T(copysign(1.0, (double)val(x))); }
template <> GPO_HD inline float sgn1<float>(float x) { return std::copysign(1.0f, x); }
template <class T> GPO_HD inline T tabs(T x) { return x * sgn1(x); }
// std::min / std::max: the first argument on ties
template <class T> GPO_HD inline T tmin(T a, T b) { return (val(b) < val(a)) ? b : a; }
template <class T> GPO_HD inline T tmax(T a, T b) { return (val(a) < val(b)) ? b : a; }

// ----------------------------------------------------------------------------- geometry

struct Geo {
  int nx, ny;
  int W;                    // dense workspace row length (nx + 4)
  int sp, hp, fp;           // state, Hi and f_cells pitches in elements
  int bc_n, bc_e, bc_s, bc_w;
  double dx, dy, g, f, theta, desing_eps, depth_cutoff, flux_slope_eps, flux_balancer, north_x, north_y;
  int has_land;  // corner depths equal to `land` are land
  double land;
};

// Land sentinels of the step kernel (cdklm_common.h): a land corner depth reads
// as kDry, a land cell's eta is kDry (CDKLM16_kernel.cu:552-655).
constexpr float kDry = 1.0e-30f;
constexpr double kDryTol = 1.0e-3;

// Cell of the input array that supplies stencil cell (r, c): wall ghosts
// mirror (normal momentum negated), periodic ghosts wrap; only for ghost
// columns of interior rows and ghost rows of interior columns
// (CDKLM16_kernel.cu:419-470, boundary_kernels.cu:36-110).
struct Src {
  int r, c;
  float shu, shv;
};
GPO_HD inline Src map_cell(const Geo& g, int r, int c) {
  Src s{r, c, 1.0f, 1.0f};
  const bool row_in = r >= 2 && r <= g.ny + 1, col_in = c >= 2 && c <= g.nx + 1;
  if (row_in && c < 2) {
    if (g.bc_w == 1) { s.c = 3 - c; s.shu = -1.0f; }
    else if (g.bc_w == 2) s.c = c + g.nx;
  } else if (row_in && c > g.nx + 1) {
    if (g.bc_e == 1) { s.c = 2 * g.nx + 3 - c; s.shu = -1.0f; }
    else if (g.bc_e == 2) s.c = c - g.nx;
  } else if (col_in && r < 2) {
    if (g.bc_s == 1) { s.r = 3 - r; s.shv = -1.0f; }
    else if (g.bc_s == 2) s.r = r + g.ny;
  } else if (col_in && r > g.ny + 1) {
    if (g.bc_n == 1) { s.r = 2 * g.ny + 3 - r; s.shv = -1.0f; }
    else if (g.bc_n == 2) s.r = r - g.ny;
  }
  return s;
}

// Dense workspace arrays ((ny+4) x (nx+4) each, element type S).
template <class S> struct Fwd {
  GPO_GLOBAL S *ce, *cu, *cv;    // converted cells
  GPO_GLOBAL S *qx0, *qx1, *kx;  // x-slopes of u, v and the potential K
  GPO_GLOBAL S *qy0, *qy1, *ly;  // y-slopes of u, v and the potential L
  GPO_GLOBAL S *fx, *fy, *fz;    // x-face fluxes (face c between c and c+1)
  GPO_GLOBAL S *gx, *gy, *gz;    // y-face fluxes (face r between r and r+1)
};

template <class S> inline Fwd<S> carve(void* ws, int nx, int ny) {
  const size_t n = (size_t)(nx + 4) * (ny + 4);
  GPO_GLOBAL S* p = (GPO_GLOBAL S*)ws;
  Fwd<S> f;
  GPO_GLOBAL S** slots[] = {&f.ce, &f.cu, &f.cv, &f.qx0, &f.qx1, &f.kx, &f.qy0, &f.qy1,
                            &f.ly, &f.fx, &f.fy, &f.fz, &f.gx, &f.gy, &f.gz};
  for (int i = 0; i < 15; ++i) *slots[i] = p + n * i;
  return f;
}

// Corner depth.  With land, the stages read a copy of Hi whose land corners
// already hold kDry (made once per call by the entry points); corner_land
// reads the caller's Hi and flags on the fly.
template <class T> GPO_HD inline T corner(const Geo& g, const T* Hi, int r, int c) {
  return Hi[(size_t)r * g.hp + c];
}
template <class T> GPO_HD inline T corner_land(const Geo& g, const T* Hi, int r, int c) {
  const T h = Hi[(size_t)r * g.hp + c];
  if (g.has_land && fabs((double)val(h) - g.land) < kDryTol) return T(kDry);
  return h;
}
// Mean of the four corner depths (land corners as kDry): Hm of the update.
template <class T> GPO_HD inline T cell_depth(const Geo& g, const T* Hi, int r, int c) {
  const T a = corner(g, Hi, r, c), b = corner(g, Hi, r + 1, c);
  const T c_ = corner(g, Hi, r, c + 1), d = corner(g, Hi, r + 1, c + 1);
  return T(0.25) * (a + b + c_ + d);
}
template <class T> GPO_HD inline T cell_depth_land(const Geo& g, const T* Hi, int r, int c) {
  const T a = corner_land(g, Hi, r, c), b = corner_land(g, Hi, r + 1, c);
  const T c_ = corner_land(g, Hi, r, c + 1), d = corner_land(g, Hi, r + 1, c + 1);
  return T(0.25) * (a + b + c_ + d);
}
// Depth of the conversion (initHm_kernel.cu:81-100) from the four corner
// depths (r,c), (r+1,c), (r,c+1), (r+1,c+1): the mean of the wet corners,
// `land` when all four are land.
template <class T> GPO_HD inline T bm_of(const Geo& g, T a, T b, T c_, T d) {
  const T S = a + b + c_ + d;
  if (!g.has_land) return T(0.25) * S;
  const int wet = 4 - (val(a) == (double)kDry) - (val(b) == (double)kDry) - (val(c_) == (double)kDry) -
                  (val(d) == (double)kDry);
  return wet == 0 ? T(g.land) : (wet == 4 ? T(0.25) * S : S / T(wet));
}
template <class T> GPO_HD inline T cell_bm(const Geo& g, const T* Hi, int r, int c) {
  return bm_of(g, corner(g, Hi, r, c), corner(g, Hi, r + 1, c), corner(g, Hi, r, c + 1), corner(g, Hi, r + 1, c + 1));
}
template <class T> GPO_HD inline bool land_cell(const Geo& g, T bm) {
  return g.has_land && fabs((double)val(bm) - g.land) <= kDryTol;
}
template <class T> GPO_HD inline T fcor(const Geo& g, const T* fc, int r, int c) {
  return fc ? fc[(size_t)r * g.fp + c] : T(g.f);
}

// Converted cells, coriolis values and corner depths of a slope stencil:
// cells 0, 1, 2 are s-1, s, s+1 along the slope's axis; the corners are
// (r,s), (r+1,s), (r,s+1), (r+1,s+1) for an x-slope at (r, s) and
// (s,c), (s,c+1), (s+1,c), (s+1,c+1) for a y-slope at (s, c).
template <class T> struct Sten {
  T e0, e1, e2, u0, u1, u2, v0, v1, v2;
  T F0, F1, F2;
  T Ha, Hb, Hc, Hd;
};
template <class T>
GPO_HD inline Sten<T> xsten(const Geo& g, const Fwd<T>& w, const T* Hi, const T* fc, int r, int s) {
  const size_t o = (size_t)r * g.W + s;
  return {w.ce[o - 1], w.ce[o], w.ce[o + 1], w.cu[o - 1], w.cu[o], w.cu[o + 1], w.cv[o - 1], w.cv[o], w.cv[o + 1],
          fcor(g, fc, r, s - 1), fcor(g, fc, r, s), fcor(g, fc, r, s + 1), corner(g, Hi, r, s),
          corner(g, Hi, r + 1, s), corner(g, Hi, r, s + 1), corner(g, Hi, r + 1, s + 1)};
}
template <class T>
GPO_HD inline Sten<T> ysten(const Geo& g, const Fwd<T>& w, const T* Hi, const T* fc, int s, int c) {
  const size_t o = (size_t)s * g.W + c, up = g.W;
  return {w.ce[o - up], w.ce[o], w.ce[o + up], w.cu[o - up], w.cu[o], w.cu[o + up], w.cv[o - up], w.cv[o],
          w.cv[o + up], fcor(g, fc, s - 1, c), fcor(g, fc, s, c), fcor(g, fc, s + 1, c), corner(g, Hi, s, c),
          corner(g, Hi, s, c + 1), corner(g, Hi, s + 1, c), corner(g, Hi, s + 1, c + 1)};
}

// ----------------------------------------------------------------------------- forward stages

// desingularize (common.cu:664-666), the double formula
template <class S> GPO_HD inline S desing(S h, S hu, double eps) {
  const S a = qdiv(h * h, S(2.0 * eps)) + S(0.5 * eps);
  return qdiv(hu, tmax(tmin(a, S(eps)), tabs(h)));
}

// The conversion of one cell (CDKLM16_kernel.cu:614-655): loaded eta and
// momenta (wall mirror signs applied), the cell's conversion depth Hm ->
// (eta, u, v).
template <class S>
GPO_HD inline void conv_v(const Geo& g, S e, S hu, S hv, S Hm, S& oe, S& ou, S& ov) {
  const S h = e + Hm;
  if (land_cell(g, Hm)) {  // land: constants
    oe = S(kDry);
    ou = S(0.0);
    ov = S(0.0);
  } else if (val(h) < g.desing_eps) {
    if (val(h) <= g.depth_cutoff) {
      oe = -Hm + S(g.depth_cutoff);
      ou = S(0.0);
      ov = S(0.0);
    } else {
      oe = e;
      ou = desing(h, hu, g.desing_eps);
      ov = desing(h, hv, g.desing_eps);
    }
  } else {
    oe = e;
    ou = qdiv(hu, h);
    ov = qdiv(hv, h);
  }
}

// F1: conversion of the stage input.  `in_*` are the caller's arrays (pitch
// sp) of type S.
template <class S>
GPO_HD void f1_cell(Geo g, const S* in_e, const S* in_u, const S* in_v, const S* Hi, Fwd<S> w, int* flag, int r, int c) {
  if (c >= g.nx + 4 || r >= g.ny + 4) return;
  const Src s = map_cell(g, r, c);
  const size_t k = (size_t)s.r * g.sp + s.c;
  S e, u, v;
  conv_v(g, in_e[k], in_u[k] * S(s.shu), in_v[k] * S(s.shv), cell_bm(g, Hi, r, c), e, u, v);
  const size_t o = (size_t)r * g.W + c;
  w.ce[o] = e;
  w.cu[o] = u;
  w.cv[o] = v;
}

// minmodSlope / minmodRaw (common.cu:532-551)
template <class S> GPO_HD inline S minmod_raw(S b, S c, S f) {
  const S sb = sgn1(b), sc = sgn1(c), sf = sgn1(f);
  return S(0.25) * sb * (sb + sc) * (sc + sf) * tmin(tmin(tabs(b), tabs(c)), tabs(f));
}
template <class S> GPO_HD inline S minmod_slope(S l, S c, S r, S theta) {
  return minmod_raw((c - l) * theta, (r - l) * S(0.5), (r - c) * theta);
}

// The argument minmod_raw(b, c, f) passes on, for its reverse: 0, 1, 2 for
// b, c, f (the first of equal magnitudes), 3 when the signs differ and the
// result is the constant 0.  Where it is not 3 the three signs agree, so
// minmod_raw = s * min|.| with s the common sign, and d out / d arg = s * s = 1.
template <class S> GPO_HD inline int minmod_sel(S b, S c, S f) {
  const S sb = sgn1(b), sc = sgn1(c), sf = sgn1(f);
  const S s = S(0.25) * sb * (sb + sc) * (sc + sf);
  if (val(s) == 0) return 3;
  const S ab = tabs(b), ac = tabs(c), af = tabs(f);
  int sel = (val(ac) < val(ab)) ? 1 : 0;
  const S m1 = sel ? ac : ab;
  if (val(af) < val(m1)) sel = 2;
  return sel;
}
// The same on sign bits and magnitudes (tabs(x) = |x|, copysign(1, x) has
// x's sign bit, and s = 0 exactly when two neighbouring signs differ).
template <> GPO_HD inline int minmod_sel<float>(float b, float c, float f) {
  const bool nb = std::signbit(b), nc = std::signbit(c), nf = std::signbit(f);
  if (nb != nc || nc != nf) return 3;
  const float ab = std::fabs(b), ac = std::fabs(c), af = std::fabs(f);
  int sel = (ac < ab) ? 1 : 0;
  const float m1 = sel ? ac : ab;
  if (af < m1) sel = 2;
  return sel;
}
template <class S> GPO_HD inline int minmod_slope_sel(S l, S c, S r, S theta) {
  return minmod_sel((c - l) * theta, (r - l) * S(0.5), (r - c) * theta);
}

// Selections of a slope evaluation, for its reverse (xslope_rev_code,
// yslope_rev_code): bits 0-1, 2-3 and 4-5 hold minmod_sel of the two
// limited slopes and of the potential-energy slope before its adjustment,
// bit 6 is set where the low-side depth test failed (hw, hs <= 0) and bit 7
// where the high-side one did (he, hn <= 0).
constexpr int kSelLowDry = 64, kSelHighDry = 128;

// x-slopes at (r, s): CDKLM16_kernel.cu:673-737 with adjustSlopes_x (:159-203);
// `code`, when given, receives the selections.
template <class S>
GPO_HD void xslope_v(const Geo& g, const Sten<S>& st, int s, S& q0, S& q1, S& K, int* code = nullptr) {
  const S le = st.e0, ce = st.e1, re = st.e2;
  const S lu = st.u0, cu = st.u1, ru = st.u2;
  S lv = st.v0, cv = st.v1, rv = st.v2;
  const S th = S(g.theta);
  q0 = minmod_slope(lu, cu, ru, th);
  q1 = minmod_slope(lv, cv, rv, th);
  if (g.bc_w == 1) {
    if (s < 3) lv = -lv;
    if (s < 2) cv = -cv;
  }
  if (g.bc_e == 1) {
    if (s > g.nx) rv = -rv;
    if (s > g.nx + 1) cv = -cv;
  }
  const S nx_ = S(g.north_x), ny_ = S(g.north_y), G = S(g.g), Vc = S(g.dx) / (S(2.0) * G);
  const S lfv = (nx_ * lu + ny_ * lv) * st.F0;
  const S cfv = (nx_ * cu + ny_ * cv) * st.F1;
  const S rfv = (nx_ * ru + ny_ * rv) * st.F2;
  const S b = th * G * (ce - le - Vc * (cfv + lfv));
  const S c = S(0.5) * G * (re - le - Vc * (rfv + S(2.0) * cfv + lfv));
  const S f = th * G * (re - ce - Vc * (cfv + rfv));
  S k = minmod_raw(b, c, f);
  S v = st.v1;
  if (g.bc_w == 1 && s < 2) v = -v;
  if (g.bc_e == 1 && s > g.nx + 2) v = -v;
  const S g2 = S(2.0) * G;
  const S dxfv = S(g.dx) * st.F1 * v;
  const S Hw = S(0.5) * (st.Ha + st.Hb);
  const S He = S(0.5) * (st.Hc + st.Hd);
  const S hw = ce + Hw - qdiv(k + dxfv, g2);
  const S he = ce + He + qdiv(k + dxfv, g2);
  if (code)
    *code = minmod_slope_sel(lu, cu, ru, th) | (minmod_slope_sel(st.v0, st.v1, st.v2, th) << 2) |
            (minmod_sel(b, c, f) << 4) | ((val(hw) > 0) ? 0 : kSelLowDry) | ((val(he) > 0) ? 0 : kSelHighDry);
  k = (val(hw) > 0) ? k : -dxfv + g2 * (ce + Hw);
  k = (val(he) > 0) ? k : -dxfv - g2 * (ce + He);
  K = k;
}

template <class S>
GPO_HD void xslope(const Geo& g, const Fwd<S>& w, const S* Hi, const S* fc, int r, int s, S& q0, S& q1, S& K) {
  xslope_v(g, xsten(g, w, Hi, fc, r, s), s, q0, q1, K);
}

// y-slopes at (s, c): CDKLM16_kernel.cu:779-842 with adjustSlopes_y (:210-254);
// `code`, when given, receives the selections.
template <class S>
GPO_HD void yslope_v(const Geo& g, const Sten<S>& st, int s, S& q0, S& q1, S& L, int* code = nullptr) {
  const S le = st.e0, ce = st.e1, ue = st.e2;
  S lu = st.u0, cu = st.u1, uu = st.u2;
  const S lv = st.v0, cv = st.v1, uv = st.v2;
  const S th = S(g.theta);
  q0 = minmod_slope(lu, cu, uu, th);
  q1 = minmod_slope(lv, cv, uv, th);
  if (g.bc_s == 1) {
    if (s < 3) lu = -lu;
    if (s < 2) cu = -cu;
  }
  if (g.bc_n == 1) {
    if (s > g.ny) uu = -uu;
    if (s > g.ny + 1) cu = -cu;
  }
  const S ex = S(g.north_y), ey = S(-g.north_x), G = S(g.g), Uc = S(g.dy) / (S(2.0) * G);
  const S lfu = (ex * lu + ey * lv) * st.F0;
  const S cfu = (ex * cu + ey * cv) * st.F1;
  const S ufu = (ex * uu + ey * uv) * st.F2;
  const S b = th * G * (ce - le + Uc * (cfu + lfu));
  const S cc = S(0.5) * G * (ue - le + Uc * (ufu + S(2.0) * cfu + lfu));
  const S f = th * G * (ue - ce + Uc * (cfu + ufu));
  S l = minmod_raw(b, cc, f);
  S u = st.u1;
  if (g.bc_s == 1 && s < 2) u = -u;
  if (g.bc_n == 1 && s > g.ny + 2) u = -u;
  const S g2 = S(2.0) * G;
  const S dyfu = S(g.dy) * st.F1 * u;
  const S Hs = S(0.5) * (st.Ha + st.Hb);
  const S Hn = S(0.5) * (st.Hc + st.Hd);
  const S hs = ce + Hs - qdiv(l - dyfu, g2);
  const S hn = ce + Hn + qdiv(l - dyfu, g2);
  if (code)
    *code = minmod_slope_sel(st.u0, st.u1, st.u2, th) | (minmod_slope_sel(lv, cv, uv, th) << 2) |
            (minmod_sel(b, cc, f) << 4) | ((val(hs) > 0) ? 0 : kSelLowDry) | ((val(hn) > 0) ? 0 : kSelHighDry);
  l = (val(hs) > 0) ? l : dyfu + g2 * (ce + Hs);
  l = (val(hn) > 0) ? l : dyfu - g2 * (ce + Hn);
  L = l;
}

template <class S>
GPO_HD void yslope(const Geo& g, const Fwd<S>& w, const S* Hi, const S* fc, int s, int c, S& q0, S& q1, S& L) {
  yslope_v(g, ysten(g, w, Hi, fc, s, c), s, q0, q1, L);
}

template <class S>
GPO_HD void f2_cell(Geo g, const S* Hi, const S* fc, Fwd<S> w, int r, int c) {
  if (c >= g.nx + 4 || r >= g.ny + 4) return;
  const size_t o = (size_t)r * g.W + c;
  S a, b, k;
  if (r >= 2 && r <= g.ny + 1 && c >= 1 && c <= g.nx + 2) {
    xslope(g, w, Hi, fc, r, c, a, b, k);
    w.qx0[o] = a, w.qx1[o] = b, w.kx[o] = k;
  }
  if (r >= 1 && r <= g.ny + 2 && c >= 2 && c <= g.nx + 1) {
    yslope(g, w, Hi, fc, r, c, a, b, k);
    w.qy0[o] = a, w.qy1[o] = b, w.ly[o] = k;
  }
}

// CDKLM16_F_func (CDKLM16_kernel.cu:82-91) and CDKLM16_flux (:104-151), the
// wet-wet case (the dry face cases of :318-338 do not arise for wet cells).
template <class S>
GPO_HD void flux_fwd(const Geo& g, const S* Qm, const S* Qp, S H, S* F) {
  const S G = S(g.g);
  const S hp = Qp[0] + H, hm = Qm[0] + H;
  S Fp0 = S(0.0), Fp1 = S(0.0), Fp2 = S(0.0), Fm0 = S(0.0), Fm1 = S(0.0), Fm2 = S(0.0);
  S up = S(0.0), cp = S(0.0), um = S(0.0), cm = S(0.0);
  if (val(hp) > g.depth_cutoff) {
    Fp0 = hp * Qp[1];
    Fp1 = hp * Qp[1] * Qp[1] + S(0.5) * G * hp * hp;
    Fp2 = hp * Qp[1] * Qp[2];
    up = Qp[1];
    cp = tsqrt(G * hp);
  }
  if (val(hm) > g.depth_cutoff) {
    Fm0 = hm * Qm[1];
    Fm1 = hm * Qm[1] * Qm[1] + S(0.5) * G * hm * hm;
    Fm2 = hm * Qm[1] * Qm[2];
    um = Qm[1];
    cm = tsqrt(G * hm);
  }
  const S am = tmin(tmin(um - cm, up - cp), S(0.0));
  const S ap = tmax(tmax(um + cm, up + cp), S(0.0));
  if (fabs((double)val(ap - am)) < g.flux_slope_eps) {
    F[0] = F[1] = F[2] = S(0.0);
    return;
  }
  const S fb = S(g.flux_balancer);
  F[0] = ((ap * Fm0 - am * Fp0) + ap * am * (Qp[0] - Qm[0])) / (ap - am);
  F[1] = ((ap * Fm1 - am * Fp1) + ap * am * (Fp0 - Fm0)) / (ap - am);
  S z = (val(Qm[1]) > -val(Qp[1])) ? fb * Fm2 : fb * Fp2;
  F[2] = z + (S(1.0) - fb) * (((ap * Fm2 - am * Fp2) + ap * am * (hp * Qp[2] - hm * Qm[2])) / (ap - am));
}

// x-face between (r, s) and (r, s+1): CDKLM16_kernel.cu:259-340 (computeFFaceFlux)
template <class S>
GPO_HD void xface(const Geo& g, const Fwd<S>& w, const S* Hi, const S* fc, int r, int s, S* F, int* flag) {
  const size_t m = (size_t)r * g.W + s, p = m + 1;
  const S G = S(g.g), g2 = S(2.0) * G;
  const S ebp = w.ce[p], ebm = w.ce[m], up = w.cu[p], um = w.cu[m];
  S vp = w.cv[p], vm = w.cv[m];
  const S Rpx = up - S(0.5) * w.qx0[p], Rpy = vp - S(0.5) * w.qx1[p];
  const S Rmx = um + S(0.5) * w.qx0[m], Rmy = vm + S(0.5) * w.qx1[m];
  const S H = S(0.5) * (corner(g, Hi, r, s + 1) + corner(g, Hi, r + 1, s + 1));
  if (g.bc_w == 1 && s + 1 == 2) vm = -vm;
  if (g.bc_e == 1 && s + 1 == g.nx + 2) vp = -vp;
  const S nx_ = S(g.north_x), ny_ = S(g.north_y);
  const S vpn = up * nx_ + vp * ny_, vmn = um * nx_ + vm * ny_;
  const S ap = ebp - (w.kx[p] + S(g.dx) * fcor(g, fc, r, s + 1) * vpn) / g2;
  const S am = ebm + (w.kx[m] + S(g.dx) * fcor(g, fc, r, s) * vmn) / g2;
  const bool dm = val(ebm) == (double)kDry, dp = val(ebp) == (double)kDry;
  if (dm || dp) {  // land on one or both sides (CDKLM16_kernel.cu:318-338): the wet side's pressure only
    F[0] = F[1] = F[2] = S(0.0);
    if (!(dm && dp)) {
      const S eta = dm ? tmax(-H, ap) : tmax(-H, am);
      F[1] = S(0.5) * G * (eta + H) * (eta + H);
    }
    return;
  }
  if (flag && (!(val(ap) > val(-H)) || !(val(am) > val(-H)))) GPO_FLAG(flag, 8);
  const S Qp[3] = {tmax(-H, ap), Rpx, Rpy}, Qm[3] = {tmax(-H, am), Rmx, Rmy};
  flux_fwd(g, Qm, Qp, H, F);
}

// y-face between (s, c) and (s+1, c): CDKLM16_kernel.cu:345-416 (computeGFaceFlux)
template <class S>
GPO_HD void yface(const Geo& g, const Fwd<S>& w, const S* Hi, const S* fc, int s, int c, S* G3, int* flag) {
  const size_t m = (size_t)s * g.W + c, p = m + g.W;
  const S G = S(g.g), g2 = S(2.0) * G;
  const S ebp = w.ce[p], ebm = w.ce[m], vp = w.cv[p], vm = w.cv[m];
  S up = w.cu[p], um = w.cu[m];
  const S Rpx = up - S(0.5) * w.qy0[p], Rpy = vp - S(0.5) * w.qy1[p];
  const S Rmx = um + S(0.5) * w.qy0[m], Rmy = vm + S(0.5) * w.qy1[m];
  const S H = S(0.5) * (corner(g, Hi, s + 1, c) + corner(g, Hi, s + 1, c + 1));
  if (g.bc_s == 1 && s + 1 == 2) um = -um;
  if (g.bc_n == 1 && s + 1 == g.ny + 2) up = -up;
  const S ex = S(g.north_y), ey = S(-g.north_x);
  const S upe = up * ex + vp * ey, ume = um * ex + vm * ey;
  const S ap = ebp - (w.ly[p] - S(g.dy) * fcor(g, fc, s + 1, c) * upe) / g2;
  const S am = ebm + (w.ly[m] - S(g.dy) * fcor(g, fc, s, c) * ume) / g2;
  const bool dm = val(ebm) == (double)kDry, dp = val(ebp) == (double)kDry;
  if (dm || dp) {  // land (CDKLM16_kernel.cu:394-416): the wet side's pressure in the normal-momentum flux
    G3[0] = G3[1] = G3[2] = S(0.0);
    if (!(dm && dp)) {
      const S eta = dm ? tmax(-H, ap) : tmax(-H, am);
      G3[2] = S(0.5) * G * (eta + H) * (eta + H);
    }
    return;
  }
  if (flag && (!(val(ap) > val(-H)) || !(val(am) > val(-H)))) GPO_FLAG(flag, 8);
  const S Qp[3] = {tmax(-H, ap), Rpy, Rpx}, Qm[3] = {tmax(-H, am), Rmy, Rmx};
  S fl[3];
  flux_fwd(g, Qm, Qp, H, fl);
  G3[0] = fl[0];
  G3[1] = fl[2];
  G3[2] = fl[1];
}

template <class S>
GPO_HD void f3_cell(Geo g, const S* Hi, const S* fc, Fwd<S> w, int* flag, int r, int c) {
  if (c >= g.nx + 4 || r >= g.ny + 4) return;
  const size_t o = (size_t)r * g.W + c;
  if (r >= 2 && r <= g.ny + 1 && c >= 1 && c <= g.nx + 1) {
    S F[3];
    xface(g, w, Hi, fc, r, c, F, flag);
    w.fx[o] = F[0];
    w.fy[o] = F[1];
    w.fz[o] = F[2];
  }
  if (r >= 1 && r <= g.ny + 1 && c >= 2 && c <= g.nx + 1) {
    S G3[3];
    yface(g, w, Hi, fc, r, c, G3, flag);
    w.gx[o] = G3[0];
    w.gy[o] = G3[1];
    w.gz[o] = G3[2];
  }
}

// F4: sources and the RK2 update at interior cells (CDKLM16_kernel.cu:869-1089),
// no friction and no forcing.  qn is read before out is written (may alias).
template <class S>
GPO_HD void f4_cell(Geo g, S dt, int rk_step, const S* Hi, const S* fc, Fwd<S> w, const S* qn_e,
                          const S* qn_u, const S* qn_v, S* out_e, S* out_u, S* out_v, int* flag, int r, int c) {
  if (c > g.nx + 1 || r > g.ny + 1) return;
  const size_t o = (size_t)r * g.W + c;
  const S G = S(g.g), g2 = S(2.0) * G, dx = S(g.dx), dy = S(g.dy);
  const S hm = cell_depth(g, Hi, r, c);
  const S R0 = w.ce[o], R1 = w.cu[o], R2 = w.cv[o];
  S hu = S(0.0), hv = S(0.0);
  if (val(R0 + hm) > g.depth_cutoff) {
    hu = R1 * (R0 + hm);
    hv = R2 * (R0 + hm);
  }
  const S f = fcor(g, fc, r, c);
  S fdx = (w.fx[o] - w.fx[o - 1]) / dx;
  S fdy = (w.fy[o] - w.fy[o - 1]) / dx;
  S fdz = (w.fz[o] - w.fz[o - 1]) / dx;
  const S eta_west = R0 - (w.kx[o] + dx * f * R2) / g2;
  const S eta_east = R0 + (w.kx[o] + dx * f * R2) / g2;
  fdx = fdx + (w.gx[o] - w.gx[o - g.W]) / dy;
  fdy = fdy + (w.gy[o] - w.gy[o - g.W]) / dy;
  fdz = fdz + (w.gz[o] - w.gz[o - g.W]) / dy;
  const S eta_south = R0 - (w.ly[o] - dy * f * R1) / g2;
  const S eta_north = R0 + (w.ly[o] - dy * f * R1) / g2;
  S st1 = S(0.0), st2 = S(0.0);
  const S h = R0 + hm;
  if (val(h) >= g.depth_cutoff && val(R0) != (double)kDry) {
    const S RHxp = S(0.5) * (corner(g, Hi, r, c + 1) + corner(g, Hi, r + 1, c + 1));
    const S RHxm = S(0.5) * (corner(g, Hi, r, c) + corner(g, Hi, r + 1, c));
    const S RHyp = S(0.5) * (corner(g, Hi, r + 1, c) + corner(g, Hi, r + 1, c + 1));
    const S RHym = S(0.5) * (corner(g, Hi, r, c) + corner(g, Hi, r, c + 1));
    const S nx_ = S(g.north_x), ny_ = S(g.north_y), ex = ny_, ey = -nx_;
    const S hu_east = hu * ex + hv * ey;
    const S hv_north = hu * nx_ + hv * ny_;
    const S hu_east_cor = f * hv_north;
    const S hv_north_cor = -f * hu_east;
    const S upx = -nx_, upy = ny_, rtx = upy, rty = -upx;
    const S hu_cor = rtx * hu_east_cor + rty * hv_north_cor;
    const S hv_cor = upx * hu_east_cor + upy * hv_north_cor;
    const S eta_we = S(0.5) * (eta_west + eta_east);
    st1 = hu_cor + (G * (eta_we + hm) * (RHxp - RHxm)) / dx;
    const S eta_sn = S(0.5) * (eta_north + eta_south);
    st2 = hv_cor + (G * (eta_sn + hm) * (RHyp - RHym)) / dy;
  }
  const S L1 = -fdx, L2 = -fdy + st1, L3 = -fdz + st2;
  S ue, uhu, uhv;
  const size_t k = (size_t)r * g.sp + c;
  if (rk_step == 0) {
    ue = R0 + dt * L1;
    uhu = hu + dt * L2;
    uhv = hv + dt * L3;
  } else {
    ue = S(0.5) * (qn_e[k] + (R0 + dt * L1));
    uhu = S(0.5) * (qn_u[k] + (hu + dt * L2));
    uhv = S(0.5) * (qn_v[k] + (hv + dt * L3));
  }
  if (val(ue + hm) <= g.depth_cutoff) {
    if (flag && val(R0) != (double)kDry) GPO_FLAG(flag, 2);  // a wet cell dried (land cells always clamp)
    ue = -hm + S(g.depth_cutoff);
    uhu = S(0.0);
    uhv = S(0.0);
  }
  out_e[k] = tmax(-hm + S(g.depth_cutoff), ue);
  out_u[k] = uhu;
  out_v[k] = uhv;
}

// ----------------------------------------------------------------------------- reverse stages

// Face sensitivities: for the reconstructed state on each side of a face,
// those of (eta, first, second flux component), the eta one already zeroed
// where the -H clamp was active; then the per-cell sums formed from them:
// every sensitivity of a converted cell's (eta, u, v) is px[] of that cell
// (its own terms plus the x-slope terms of its west and east neighbours) plus
// yp[] of its south neighbour and ym[] of its north neighbour.
template <class T> struct Rev {
  GPO_GLOBAL T* xl[3];  // x-face s, side of cell s: (eta, u, v)
  GPO_GLOBAL T* xr[3];  // x-face s, side of cell s+1
  GPO_GLOBAL T* yl[3];  // y-face s, side of cell s: (eta, v, u)
  GPO_GLOBAL T* yr[3];  // y-face s, side of cell s+1
  GPO_GLOBAL T* px[3];  // own terms of the cell plus its x-neighbours' x-slope terms
  GPO_GLOBAL T* ym[3];  // y-slopes of row s -> row s-1
  GPO_GLOBAL T* yp[3];  // y-slopes of row s -> row s+1
};

// Workspace arrays of (ny+4) x (nx+4) elements: the 9 converted/slope arrays
// of Fwd, then the 12 face and 9 per-cell arrays of Rev.
constexpr int kAdjWorkArrays = 9 + 12 + 9;

template <class T> inline Rev<T> carve_rev(void* ws, int nx, int ny) {
  const size_t n = (size_t)(nx + 4) * (ny + 4);
  GPO_GLOBAL T* p = (GPO_GLOBAL T*)ws + 9 * n;  // after conv (3) and slopes (6) of Fwd<T>
  Rev<T> v;
  for (int i = 0; i < 3; ++i) {
    v.xl[i] = p + n * i;
    v.xr[i] = p + n * (3 + i);
    v.yl[i] = p + n * (6 + i);
    v.yr[i] = p + n * (9 + i);
    v.px[i] = p + n * (12 + i);
    v.ym[i] = p + n * (15 + i);
    v.yp[i] = p + n * (18 + i);
  }
  return v;
}

// Adjoint of flux_fwd: lQm, lQp += d(F . lF)/dQ.
template <class T>
GPO_HD void flux_rev(const Geo& g, const T* Qm, const T* Qp, T H, const T* lF, T* lQm, T* lQp) {
  const T G = T(g.g);
  const T hp = Qp[0] + H, hm = Qm[0] + H;
  const bool wp = hp > T(g.depth_cutoff), wm = hm > T(g.depth_cutoff);
  T Fp0 = 0, Fp1 = 0, Fp2 = 0, Fm0 = 0, Fm1 = 0, Fm2 = 0, up = 0, cp = 0, um = 0, cm = 0;
  if (wp) {
    Fp0 = hp * Qp[1];
    Fp1 = hp * Qp[1] * Qp[1] + T(0.5) * G * hp * hp;
    Fp2 = hp * Qp[1] * Qp[2];
    up = Qp[1];
    cp = tsqrt(G * hp);
  }
  if (wm) {
    Fm0 = hm * Qm[1];
    Fm1 = hm * Qm[1] * Qm[1] + T(0.5) * G * hm * hm;
    Fm2 = hm * Qm[1] * Qm[2];
    um = Qm[1];
    cm = tsqrt(G * hm);
  }
  // am = min(min(um - cm, up - cp), 0), ap = max(max(um + cm, up + cp), 0): first on ties
  const T am1 = um - cm, am2 = up - cp, ap1 = um + cm, ap2 = up + cp;
  const int am_sel = (am2 < am1) ? ((T(0) < am2) ? 0 : 2) : ((T(0) < am1) ? 0 : 1);
  const int ap_sel = (ap1 < ap2) ? ((ap2 < T(0)) ? 0 : 2) : ((ap1 < T(0)) ? 0 : 1);
  const T am = am_sel == 1 ? am1 : (am_sel == 2 ? am2 : T(0));
  const T ap = ap_sel == 1 ? ap1 : (ap_sel == 2 ? ap2 : T(0));
  if (fabs((double)(ap - am)) < g.flux_slope_eps) return;
  const T fb = T(g.flux_balancer), d = ap - am;
  const T dq = Qp[0] - Qm[0], dfx = Fp0 - Fm0, dhv = hp * Qp[2] - hm * Qm[2];
  const T Nx = (ap * Fm0 - am * Fp0) + ap * am * dq;
  const T Ny = (ap * Fm1 - am * Fp1) + ap * am * dfx;
  const T Z = (ap * Fm2 - am * Fp2) + ap * am * dhv;
  const T rd = qdiv(T(1), d);
  const T lNx = lF[0] * rd, lNy = lF[1] * rd, lZ = (T(1) - fb) * lF[2] * rd;
  T ld = -(lNx * Nx + lNy * Ny + lZ * Z) * rd;
  T lap = lNx * (Fm0 + am * dq) + lNy * (Fm1 + am * dfx) + lZ * (Fm2 + am * dhv) + ld;
  T lam = lNx * (-Fp0 + ap * dq) + lNy * (-Fp1 + ap * dfx) + lZ * (-Fp2 + ap * dhv) - ld;
  T lFm0 = lNx * ap - lNy * ap * am, lFp0 = -lNx * am + lNy * ap * am;
  T lFm1 = lNy * ap, lFp1 = -lNy * am;
  T lFm2 = lZ * ap, lFp2 = -lZ * am;
  if (Qm[1] > -Qp[1]) lFm2 += fb * lF[2];
  else lFp2 += fb * lF[2];
  T lhp = lZ * ap * am * Qp[2], lhm = -lZ * ap * am * Qm[2];
  T lQp0 = lNx * ap * am, lQm0 = -lNx * ap * am;
  T lvp = lZ * ap * am * hp, lvm = -lZ * ap * am * hm;
  // wave speeds: um, cm (up, cp) are constants 0 on a dry side
  T lup = 0, lum = 0, lcp = 0, lcm = 0;
  if (ap_sel == 1) { lum += lap; lcm += lap; }
  else if (ap_sel == 2) { lup += lap; lcp += lap; }
  if (am_sel == 1) { lum += lam; lcm -= lam; }
  else if (am_sel == 2) { lup += lam; lcp -= lam; }
  if (!wp) lup = lcp = 0;
  if (!wm) lum = lcm = 0;
  if (wp) {
    lhp += lFp0 * Qp[1] + lFp1 * (Qp[1] * Qp[1] + G * hp) + lFp2 * Qp[1] * Qp[2];
    lup += lFp0 * hp + lFp1 * T(2) * hp * Qp[1] + lFp2 * hp * Qp[2];
    lvp += lFp2 * hp * Qp[1];
    if (cp > T(0)) lhp += qdiv(lcp * G, T(2) * cp);
  }
  if (wm) {
    lhm += lFm0 * Qm[1] + lFm1 * (Qm[1] * Qm[1] + G * hm) + lFm2 * Qm[1] * Qm[2];
    lum += lFm0 * hm + lFm1 * T(2) * hm * Qm[1] + lFm2 * hm * Qm[2];
    lvm += lFm2 * hm * Qm[1];
    if (cm > T(0)) lhm += qdiv(lcm * G, T(2) * cm);
  }
  lQp[0] += lQp0 + lhp;
  lQp[1] += lup;
  lQp[2] += lvp;
  lQm[0] += lQm0 + lhm;
  lQm[1] += lum;
  lQm[2] += lvm;
}

// Sensitivity of the interior cell's update outputs, after the depth clamp.
template <class T> struct CellAdj {
  T e, u, v;  // lambda of (ue, uhu, uhv)
};
template <class T>
GPO_HD inline CellAdj<T> out_adj(const Geo& g, const T* Hi, const T* out_e, const T* lo_e,
                                              const T* lo_u, const T* lo_v, int r, int c) {
  CellAdj<T> a{0, 0, 0};
  if (r < 2 || r > g.ny + 1 || c < 2 || c > g.nx + 1) return a;
  const size_t k = (size_t)r * g.sp + c;
  const T hm = cell_depth(g, Hi, r, c);
  if (out_e[k] == -hm + T(g.depth_cutoff)) return a;  // clamped to the dry state: no sensitivity
  a.e = lo_e[k];
  a.u = lo_u[k];
  a.v = lo_v[k];
  return a;
}

// Face sensitivities of one face: those of the reconstructed (eta, first,
// second) state on its minus (m) and plus (p) side, the eta one zeroed where
// the -H clamp was active.
template <class T> struct FaceSens {
  T m[3], p[3];
};

// A1: reverse of the face fluxes.  x-face c of row r (between cells c and
// c+1) from loaded values: the two cells' update sensitivities (out_adj),
// converted states and x-slopes (q0, q1, K), the face depth H and the
// coriolis values.
template <class T>
GPO_HD FaceSens<T> xface_rev_v(const Geo& g, T dt, int rk_step, const CellAdj<T>& am, const CellAdj<T>& ap, T ebm,
                               T um, T vm0, T ebp, T up, T vp0, T qm0, T qm1, T km, T qp0, T qp1, T kp, T H, T fm,
                               T fp, int c) {
  const T coef = (rk_step == 0 ? T(1) : T(0.5)) * dt;
  const T G = T(g.g), g2 = T(2) * G;
  // lambda of the three x-fluxes at face c: L_k = -(F_c - F_{c-1})/dx - ...
  const T cd = coef / T(g.dx);
  const T lF[3] = {cd * (ap.e - am.e), cd * (ap.u - am.u), cd * (ap.v - am.v)};
  const T Rpx = up - T(0.5) * qp0, Rpy = vp0 - T(0.5) * qp1;
  const T Rmx = um + T(0.5) * qm0, Rmy = vm0 + T(0.5) * qm1;
  const T sm = (g.bc_w == 1 && c + 1 == 2) ? T(-1) : T(1);
  const T sp = (g.bc_e == 1 && c + 1 == g.nx + 2) ? T(-1) : T(1);
  const T nx_ = T(g.north_x), ny_ = T(g.north_y);
  const T vpn = up * nx_ + (sp * vp0) * ny_, vmn = um * nx_ + (sm * vm0) * ny_;
  const T argp = ebp - qdiv(kp + T(g.dx) * fp * vpn, g2);
  const T argm = ebm + qdiv(km + T(g.dx) * fm * vmn, g2);
  const T Qp[3] = {(-H < argp) ? argp : -H, Rpx, Rpy}, Qm[3] = {(-H < argm) ? argm : -H, Rmx, Rmy};
  T lQp[3] = {0, 0, 0}, lQm[3] = {0, 0, 0};
  const bool dm = ebm == T(kDry), dp = ebp == T(kDry);
  if (dm || dp) {  // land: F.y = g/2 (eta + H)^2 of the wet side's reconstruction
    if (!(dm && dp)) {
      T* lw = dm ? lQp : lQm;
      const T eta = dm ? Qp[0] : Qm[0];
      lw[0] = G * (eta + H) * lF[1];
    }
  } else {
    flux_rev(g, Qm, Qp, H, lF, lQm, lQp);
  }
  FaceSens<T> o;
  o.m[0] = (-H < argm) ? lQm[0] : T(0);
  o.m[1] = lQm[1];
  o.m[2] = lQm[2];
  o.p[0] = (-H < argp) ? lQp[0] : T(0);
  o.p[1] = lQp[1];
  o.p[2] = lQp[2];
  return o;
}
template <class T>
GPO_HD void a1_xface(const Geo& g, T dt, int rk_step, const T* Hi, const T* fc, const Fwd<T>& w, const Rev<T>& rv,
                     const T* out_e, const T* lo_e, const T* lo_u, const T* lo_v, int r, int c, T qm0, T qm1, T km,
                     T qp0, T qp1, T kp) {
  const size_t o = (size_t)r * g.W + c, p = o + 1;
  const CellAdj<T> am = out_adj(g, Hi, out_e, lo_e, lo_u, lo_v, r, c);
  const CellAdj<T> ap = out_adj(g, Hi, out_e, lo_e, lo_u, lo_v, r, c + 1);
  const T H = T(0.5) * (corner(g, Hi, r, c + 1) + corner(g, Hi, r + 1, c + 1));
  const FaceSens<T> fs = xface_rev_v(g, dt, rk_step, am, ap, w.ce[o], w.cu[o], w.cv[o], w.ce[p], w.cu[p], w.cv[p],
                                     qm0, qm1, km, qp0, qp1, kp, H, fcor(g, fc, r, c), fcor(g, fc, r, c + 1), c);
  for (int i = 0; i < 3; ++i) rv.xl[i][o] = fs.m[i], rv.xr[i][o] = fs.p[i];
}
// y-face r of column c (between rows r and r+1) from loaded values (the
// reconstructed state is stored as (eta, v, u)).
template <class T>
GPO_HD FaceSens<T> yface_rev_v(const Geo& g, T dt, int rk_step, const CellAdj<T>& am, const CellAdj<T>& ap, T ebm,
                               T um0, T vm, T ebp, T up0, T vp, T qm0, T qm1, T lm, T qp0, T qp1, T lp, T H, T fm,
                               T fp, int r) {
  const T coef = (rk_step == 0 ? T(1) : T(0.5)) * dt;
  const T G = T(g.g), g2 = T(2) * G;
  const T cd = coef / T(g.dy);
  const T lG[3] = {cd * (ap.e - am.e), cd * (ap.u - am.u), cd * (ap.v - am.v)};
  const T lfl[3] = {lG[0], lG[2], lG[1]};  // G = (fl.x, fl.z, fl.y)
  const T Rpx = up0 - T(0.5) * qp0, Rpy = vp - T(0.5) * qp1;
  const T Rmx = um0 + T(0.5) * qm0, Rmy = vm + T(0.5) * qm1;
  const T sm = (g.bc_s == 1 && r + 1 == 2) ? T(-1) : T(1);
  const T sp = (g.bc_n == 1 && r + 1 == g.ny + 2) ? T(-1) : T(1);
  const T ex = T(g.north_y), ey = T(-g.north_x);
  const T upe = (sp * up0) * ex + vp * ey, ume = (sm * um0) * ex + vm * ey;
  const T argp = ebp - qdiv(lp - T(g.dy) * fp * upe, g2);
  const T argm = ebm + qdiv(lm - T(g.dy) * fm * ume, g2);
  const T Qp[3] = {(-H < argp) ? argp : -H, Rpy, Rpx}, Qm[3] = {(-H < argm) ? argm : -H, Rmy, Rmx};
  T lQp[3] = {0, 0, 0}, lQm[3] = {0, 0, 0};
  const bool dm = ebm == T(kDry), dp = ebp == T(kDry);
  if (dm || dp) {  // land: G.z = g/2 (eta + H)^2 of the wet side's reconstruction
    if (!(dm && dp)) {
      T* lw = dm ? lQp : lQm;
      const T eta = dm ? Qp[0] : Qm[0];
      lw[0] = G * (eta + H) * lG[2];
    }
  } else {
    flux_rev(g, Qm, Qp, H, lfl, lQm, lQp);
  }
  FaceSens<T> o;
  o.m[0] = (-H < argm) ? lQm[0] : T(0);
  o.m[1] = lQm[1];
  o.m[2] = lQm[2];
  o.p[0] = (-H < argp) ? lQp[0] : T(0);
  o.p[1] = lQp[1];
  o.p[2] = lQp[2];
  return o;
}
template <class T>
GPO_HD void a1_yface(const Geo& g, T dt, int rk_step, const T* Hi, const T* fc, const Fwd<T>& w, const Rev<T>& rv,
                     const T* out_e, const T* lo_e, const T* lo_u, const T* lo_v, int r, int c, T qm0, T qm1, T lm,
                     T qp0, T qp1, T lp) {
  const size_t o = (size_t)r * g.W + c, p = o + g.W;
  const CellAdj<T> am = out_adj(g, Hi, out_e, lo_e, lo_u, lo_v, r, c);
  const CellAdj<T> ap = out_adj(g, Hi, out_e, lo_e, lo_u, lo_v, r + 1, c);
  const T H = T(0.5) * (corner(g, Hi, r + 1, c) + corner(g, Hi, r + 1, c + 1));
  const FaceSens<T> fs = yface_rev_v(g, dt, rk_step, am, ap, w.ce[o], w.cu[o], w.cv[o], w.ce[p], w.cu[p], w.cv[p],
                                     qm0, qm1, lm, qp0, qp1, lp, H, fcor(g, fc, r, c), fcor(g, fc, r + 1, c), r);
  for (int i = 0; i < 3; ++i) rv.yl[i][o] = fs.m[i], rv.yr[i][o] = fs.p[i];
}

// Which faces exist: x-faces c in [1, nx+1] of interior rows, y-faces r in
// [1, ny+1] of interior columns; x-slopes at c in [1, nx+2] of interior rows,
// y-slopes at r in [1, ny+2] of interior columns.
GPO_HD inline bool has_xface(const Geo& g, int r, int c) { return r >= 2 && r <= g.ny + 1 && c >= 1 && c <= g.nx + 1; }
GPO_HD inline bool has_yface(const Geo& g, int r, int c) { return r >= 1 && r <= g.ny + 1 && c >= 2 && c <= g.nx + 1; }
GPO_HD inline bool has_xslope(const Geo& g, int r, int c) { return r >= 2 && r <= g.ny + 1 && c >= 1 && c <= g.nx + 2; }
GPO_HD inline bool has_yslope(const Geo& g, int r, int c) { return r >= 1 && r <= g.ny + 2 && c >= 2 && c <= g.nx + 1; }

// Both faces of cell (r, c) with the slopes stored by F2.
template <class T>
GPO_HD void a1_cell(Geo g, T dt, int rk_step, const T* Hi, const T* fc, Fwd<T> w, Rev<T> rv, const T* out_e,
                    const T* lo_e, const T* lo_u, const T* lo_v, int r, int c) {
  if (c >= g.nx + 4 || r >= g.ny + 4) return;
  const size_t o = (size_t)r * g.W + c;
  if (has_xface(g, r, c))
    a1_xface(g, dt, rk_step, Hi, fc, w, rv, out_e, lo_e, lo_u, lo_v, r, c, w.qx0[o], w.qx1[o], w.kx[o], w.qx0[o + 1],
             w.qx1[o + 1], w.kx[o + 1]);
  if (has_yface(g, r, c))
    a1_yface(g, dt, rk_step, Hi, fc, w, rv, out_e, lo_e, lo_u, lo_v, r, c, w.qy0[o], w.qy1[o], w.ly[o],
             w.qy0[o + g.W], w.qy1[o + g.W], w.ly[o + g.W]);
}

// minmod_raw reverse from its selection (minmod_sel): the sensitivity goes
// to the selected argument.
template <class T> GPO_HD inline void sel_rev(int sel, T lout, T& lb, T& lc, T& lf) {
  lb += (sel == 0) ? lout : T(0);
  lc += (sel == 1) ? lout : T(0);
  lf += (sel == 2) ? lout : T(0);
}

// Reverse of the x-slope computation at (r, s) given the slope sensitivities
// and the selections of its forward evaluation (xslope_v's code); Fl, Fc, Fr
// are the coriolis values of cells s-1, s, s+1.  Adds the contributions to
// those cells (index 0, 1, 2) into le/lu/lv.
template <class T>
GPO_HD void xslope_rev_code(const Geo& g, T Fl, T Fc, T Fr, int s, int code, T lq0, T lq1, T lK, T* le, T* lu,
                            T* lv) {
  const T th = T(g.theta);
  {  // q0 = minmod_slope(lu, cu, ru), q1 = minmod_slope(lv, cv, rv)
    T lb = 0, lc = 0, lf = 0;
    sel_rev(code & 3, lq0, lb, lc, lf);
    lu[1] += th * lb - th * lf;
    lu[0] += -th * lb - T(0.5) * lc;
    lu[2] += T(0.5) * lc + th * lf;
    lb = lc = lf = 0;
    sel_rev((code >> 2) & 3, lq1, lb, lc, lf);
    lv[1] += th * lb - th * lf;
    lv[0] += -th * lb - T(0.5) * lc;
    lv[2] += T(0.5) * lc + th * lf;
  }
  T sl = 1, sc = 1, sr = 1;
  if (g.bc_w == 1) {
    if (s < 3) sl = -1;
    if (s < 2) sc = -1;
  }
  if (g.bc_e == 1) {
    if (s > g.nx) sr = -1;
    if (s > g.nx + 1) sc = -1;
  }
  const T nx_ = T(g.north_x), ny_ = T(g.north_y), G = T(g.g), Vc = T(g.dx) / (T(2) * G), g2 = T(2) * G;
  // K = minmod_raw(b, c, f) of the potential-energy differences, then replaced
  // where a face depth hw (low side) or he (high side) is not positive
  T lK1 = lK, ldxfv = 0, lce = 0;
  if (code & kSelHighDry) {
    ldxfv -= lK;
    lce -= g2 * lK;
    lK1 = 0;
  }
  T lK0 = lK1;
  if (code & kSelLowDry) {
    ldxfv -= lK1;
    lce += g2 * lK1;
    lK0 = 0;
  }
  const T sa = ((g.bc_w == 1 && s < 2) || (g.bc_e == 1 && s > g.nx + 2)) ? T(-1) : T(1);
  lv[1] += sa * T(g.dx) * Fc * ldxfv;
  T lb = 0, lc = 0, lf = 0;
  sel_rev((code >> 4) & 3, lK0, lb, lc, lf);
  lce += th * G * lb - th * G * lf;
  const T lle = -th * G * lb - T(0.5) * G * lc;
  const T lre = T(0.5) * G * lc + th * G * lf;
  const T lcfv = -th * G * Vc * lb - G * Vc * lc - th * G * Vc * lf;
  const T llfv = -th * G * Vc * lb - T(0.5) * G * Vc * lc;
  const T lrfv = -T(0.5) * G * Vc * lc - th * G * Vc * lf;
  le[0] += lle;
  le[1] += lce;
  le[2] += lre;
  lu[0] += nx_ * Fl * llfv;
  lv[0] += sl * ny_ * Fl * llfv;
  lu[1] += nx_ * Fc * lcfv;
  lv[1] += sc * ny_ * Fc * lcfv;
  lu[2] += nx_ * Fr * lrfv;
  lv[2] += sr * ny_ * Fr * lrfv;
}

// The same from the stencil: the forward evaluation gives the selections.
template <class T>
GPO_HD void xslope_rev_v(const Geo& g, const Sten<T>& st, int s, T lq0, T lq1, T lK, T* le, T* lu, T* lv) {
  T q0, q1, K;
  int code = 0;
  xslope_v(g, st, s, q0, q1, K, &code);
  xslope_rev_code(g, st.F0, st.F1, st.F2, s, code, lq0, lq1, lK, le, lu, lv);
}

template <class T>
GPO_HD void xslope_rev(const Geo& g, const Fwd<T>& w, const T* Hi, const T* fc, int r, int s, T lq0, T lq1, T lK,
                       T* le, T* lu, T* lv) {
  xslope_rev_v(g, xsten(g, w, Hi, fc, r, s), s, lq0, lq1, lK, le, lu, lv);
}

// Reverse of the y-slope computation at (s, c) from its selections (yslope_v's
// code); Fl, Fc, Fu are the coriolis values of rows s-1, s, s+1.
template <class T>
GPO_HD void yslope_rev_code(const Geo& g, T Fl, T Fc, T Fu, int s, int code, T lq0, T lq1, T lL, T* le, T* lu,
                            T* lv) {
  const T th = T(g.theta);
  {
    T lb = 0, lc = 0, lf = 0;
    sel_rev(code & 3, lq0, lb, lc, lf);
    lu[1] += th * lb - th * lf;
    lu[0] += -th * lb - T(0.5) * lc;
    lu[2] += T(0.5) * lc + th * lf;
    lb = lc = lf = 0;
    sel_rev((code >> 2) & 3, lq1, lb, lc, lf);
    lv[1] += th * lb - th * lf;
    lv[0] += -th * lb - T(0.5) * lc;
    lv[2] += T(0.5) * lc + th * lf;
  }
  T sl = 1, sc = 1, su = 1;
  if (g.bc_s == 1) {
    if (s < 3) sl = -1;
    if (s < 2) sc = -1;
  }
  if (g.bc_n == 1) {
    if (s > g.ny) su = -1;
    if (s > g.ny + 1) sc = -1;
  }
  const T ex = T(g.north_y), ey = T(-g.north_x), G = T(g.g), Uc = T(g.dy) / (T(2) * G), g2 = T(2) * G;
  T lL1 = lL, ldyfu = 0, lce = 0;
  if (code & kSelHighDry) {
    ldyfu += lL;
    lce -= g2 * lL;
    lL1 = 0;
  }
  T lL0 = lL1;
  if (code & kSelLowDry) {
    ldyfu += lL1;
    lce += g2 * lL1;
    lL0 = 0;
  }
  const T sa = ((g.bc_s == 1 && s < 2) || (g.bc_n == 1 && s > g.ny + 2)) ? T(-1) : T(1);
  lu[1] += sa * T(g.dy) * Fc * ldyfu;
  T lb = 0, lc = 0, lf = 0;
  sel_rev((code >> 4) & 3, lL0, lb, lc, lf);
  lce += th * G * lb - th * G * lf;
  const T lle = -th * G * lb - T(0.5) * G * lc;
  const T lue = T(0.5) * G * lc + th * G * lf;
  const T lcfu = th * G * Uc * lb + G * Uc * lc + th * G * Uc * lf;
  const T llfu = th * G * Uc * lb + T(0.5) * G * Uc * lc;
  const T lufu = T(0.5) * G * Uc * lc + th * G * Uc * lf;
  le[0] += lle;
  le[1] += lce;
  le[2] += lue;
  lu[0] += sl * ex * Fl * llfu;
  lv[0] += ey * Fl * llfu;
  lu[1] += sc * ex * Fc * lcfu;
  lv[1] += ey * Fc * lcfu;
  lu[2] += su * ex * Fu * lufu;
  lv[2] += ey * Fu * lufu;
}

template <class T>
GPO_HD void yslope_rev_v(const Geo& g, const Sten<T>& st, int s, T lq0, T lq1, T lL, T* le, T* lu, T* lv) {
  T q0, q1, L;
  int code = 0;
  yslope_v(g, st, s, q0, q1, L, &code);
  yslope_rev_code(g, st.F0, st.F1, st.F2, s, code, lq0, lq1, lL, le, lu, lv);
}

template <class T>
GPO_HD void yslope_rev(const Geo& g, const Fwd<T>& w, const T* Hi, const T* fc, int s, int c, T lq0, T lq1, T lL,
                       T* le, T* lu, T* lv) {
  yslope_rev_v(g, ysten(g, w, Hi, fc, s, c), s, lq0, lq1, lL, le, lu, lv);
}

// A2: per converted cell (r, c): the reverse of its update and sources, its
// face parts, and the reverse of its own x- and y-slopes, split into the part
// for the cell itself and the parts for its four neighbours.  Zero outside
// the (ny+4) x (nx+4) grid and wherever nothing applies.
template <class T> struct CellParts {
  T own[3], xm[3], xp[3], ym[3], yp[3];
};
// A2 on loaded values: the cell's four corner depths (H_rc with r, c the
// offsets), its coriolis value F, its x- and y-slope stencils, its update
// sensitivity a (out_adj's rule), the sensitivities of its side of x-face c
// (xl), of x-face c-1 (xr), of y-face r (yl) and of y-face r-1 (yr).  With
// SEL the slope reverses take the selections of the cell's forward slope
// evaluations (xcode, ycode) instead of evaluating them again.
template <class T, bool SEL = false>
GPO_HD CellParts<T> a2_parts_v(const Geo& g, T dt, int rk_step, int r, int c, T H00, T H10, T H01, T H11, T F,
                               const Sten<T>& sx, const Sten<T>& sy, const CellAdj<T>& a, const T* xl, const T* xr,
                               const T* yl, const T* yr, int xcode = 0, int ycode = 0) {
  CellParts<T> out;
  for (int i = 0; i < 3; ++i) out.own[i] = out.xm[i] = out.xp[i] = out.ym[i] = out.yp[i] = T(0);
  if (c < 0 || r < 0 || c >= g.nx + 4 || r >= g.ny + 4) return out;
  const bool row_in = r >= 2 && r <= g.ny + 1, col_in = c >= 2 && c <= g.nx + 1;
  const bool left = row_in && c >= 1 && c <= g.nx + 1, right = row_in && c - 1 >= 1 && c - 1 <= g.nx + 1;
  const bool low = col_in && r >= 1 && r <= g.ny + 1, high = col_in && r - 1 >= 1 && r - 1 <= g.ny + 1;
  const bool upd = row_in && col_in;
  T lR0 = 0, lR1 = 0, lR2 = 0;
  T* xm = out.xm;
  T* xp = out.xp;
  T* ym = out.ym;
  T* yp = out.yp;
  const T G = T(g.g);
  if (upd) {
    // update + sources (CDKLM16_kernel.cu:869-1089), reversed
    const T hm = T(0.25) * (H00 + H10 + H01 + H11);
    const T coef = rk_step == 0 ? T(1) : T(0.5);
    const T R0 = sx.e1, R1 = sx.u1, R2 = sx.v1;
    lR0 += coef * a.e;
    T lhu = coef * a.u, lhv = coef * a.v;
    const T lst1 = coef * dt * a.u, lst2 = coef * dt * a.v;
    const T h = R0 + hm;
    if (h >= T(g.depth_cutoff) && R0 != T(kDry)) {
      const T RHxp = T(0.5) * (H01 + H11);
      const T RHxm = T(0.5) * (H00 + H10);
      const T RHyp = T(0.5) * (H10 + H11);
      const T RHym = T(0.5) * (H00 + H01);
      // bathymetry terms: eta_we = (eta_west + eta_east)/2 depends on R0 only
      lR0 += G * (RHxp - RHxm) * (lst1 * T(1.0 / g.dx));
      lR0 += G * (RHyp - RHym) * (lst2 * T(1.0 / g.dy));
      // Coriolis terms (linear in hu, hv)
      const T f = F;
      const T nx_ = T(g.north_x), ny_ = T(g.north_y), ex = ny_, ey = -nx_;
      const T upx = -nx_, upy = ny_, rtx = upy, rty = -upx;
      const T lhu_east_cor = rtx * lst1 + upx * lst2;
      const T lhv_north_cor = rty * lst1 + upy * lst2;
      const T lhv_north = f * lhu_east_cor;
      const T lhu_east = -f * lhv_north_cor;
      lhu += ex * lhu_east + nx_ * lhv_north;
      lhv += ey * lhu_east + ny_ * lhv_north;
    }
    if (h > T(g.depth_cutoff)) {  // hu = R1 h, hv = R2 h
      lR1 += lhu * h;
      lR2 += lhv * h;
      lR0 += lhu * R1 + lhv * R2;
    }
  }
  if (row_in) {
    // face parts: the cell is the left side of x-face c and the right side of
    // x-face c-1; the reconstruction there is eta -/+ (K + dx f v_n)/(2g),
    // u -/+ q0/2, v -/+ q1/2, with v mirrored at a wall's first ghost cell
    T lq0 = 0, lq1 = 0, lK = 0;
    if (left || right) {
      const T rg2 = T(0.5 / g.g), f = F, nx_ = T(g.north_x), ny_ = T(g.north_y);
      const T fdx = T(g.dx) * f * rg2;
      if (left) {
        const T a = xl[0], b = xl[1], d = xl[2];
        const T sm = (g.bc_w == 1 && c == 1) ? T(-1) : T(1);
        const T lvn = a * fdx;
        lR0 += a;
        lR1 += nx_ * lvn + b;
        lR2 += sm * ny_ * lvn + d;
        lq0 += T(0.5) * b;
        lq1 += T(0.5) * d;
        lK += a * rg2;
      }
      if (right) {
        const T a = xr[0], b = xr[1], d = xr[2];
        const T sp = (g.bc_e == 1 && c == g.nx + 2) ? T(-1) : T(1);
        const T lvn = -a * fdx;
        lR0 += a;
        lR1 += nx_ * lvn + b;
        lR2 += sp * ny_ * lvn + d;
        lq0 -= T(0.5) * b;
        lq1 -= T(0.5) * d;
        lK -= a * rg2;
      }
    }
    if ((left || right) && !(lq0 == T(0) && lq1 == T(0) && lK == T(0))) {
      T e3[3] = {0, 0, 0}, u3[3] = {0, 0, 0}, v3[3] = {0, 0, 0};
      if constexpr (SEL)
        xslope_rev_code(g, sx.F0, sx.F1, sx.F2, c, xcode, lq0, lq1, lK, e3, u3, v3);
      else
        xslope_rev_v(g, sx, c, lq0, lq1, lK, e3, u3, v3);
      xm[0] = e3[0], xm[1] = u3[0], xm[2] = v3[0];
      lR0 += e3[1], lR1 += u3[1], lR2 += v3[1];
      xp[0] = e3[2], xp[1] = u3[2], xp[2] = v3[2];
    }
  }
  if (col_in) {
    // y-faces r (cell below) and r-1 (cell above): eta -/+ (L - dy f u_e)/(2g),
    // v -/+ q1/2, u -/+ q0/2; stored as (eta, v, u)
    T lq0 = 0, lq1 = 0, lL = 0;
    if (low || high) {
      const T rg2 = T(0.5 / g.g), f = F, ex = T(g.north_y), ey = T(-g.north_x);
      const T fdy = T(g.dy) * f * rg2;
      if (low) {
        const T a = yl[0], bv = yl[1], bu = yl[2];
        const T sm = (g.bc_s == 1 && r == 1) ? T(-1) : T(1);
        const T lue = -a * fdy;
        lR0 += a;
        lR1 += sm * ex * lue + bu;
        lR2 += ey * lue + bv;
        lq0 += T(0.5) * bu;
        lq1 += T(0.5) * bv;
        lL += a * rg2;
      }
      if (high) {
        const T a = yr[0], bv = yr[1], bu = yr[2];
        const T sp = (g.bc_n == 1 && r == g.ny + 2) ? T(-1) : T(1);
        const T lue = a * fdy;
        lR0 += a;
        lR1 += sp * ex * lue + bu;
        lR2 += ey * lue + bv;
        lq0 -= T(0.5) * bu;
        lq1 -= T(0.5) * bv;
        lL -= a * rg2;
      }
    }
    if ((low || high) && !(lq0 == T(0) && lq1 == T(0) && lL == T(0))) {
      T e3[3] = {0, 0, 0}, u3[3] = {0, 0, 0}, v3[3] = {0, 0, 0};
      if constexpr (SEL)
        yslope_rev_code(g, sy.F0, sy.F1, sy.F2, r, ycode, lq0, lq1, lL, e3, u3, v3);
      else
        yslope_rev_v(g, sy, r, lq0, lq1, lL, e3, u3, v3);
      ym[0] = e3[0], ym[1] = u3[0], ym[2] = v3[0];
      lR0 += e3[1], lR1 += u3[1], lR2 += v3[1];
      yp[0] = e3[2], yp[1] = u3[2], yp[2] = v3[2];
    }
  }
  out.own[0] = lR0;
  out.own[1] = lR1;
  out.own[2] = lR2;
  return out;
}

template <class T>
GPO_HD CellParts<T> a2_parts(const Geo& g, T dt, int rk_step, const T* Hi, const T* fc, const Fwd<T>& w,
                             const Rev<T>& rv, const T* out_e, const T* lo_e, const T* lo_u, const T* lo_v, int r,
                             int c) {
  if (c < 0 || r < 0 || c >= g.nx + 4 || r >= g.ny + 4) {
    CellParts<T> out;
    for (int i = 0; i < 3; ++i) out.own[i] = out.xm[i] = out.xp[i] = out.ym[i] = out.yp[i] = T(0);
    return out;
  }
  const size_t o = (size_t)r * g.W + c;
  const bool row_in = r >= 2 && r <= g.ny + 1, col_in = c >= 2 && c <= g.nx + 1;
  const bool left = row_in && c >= 1 && c <= g.nx + 1, right = row_in && c - 1 >= 1 && c - 1 <= g.nx + 1;
  const bool low = col_in && r >= 1 && r <= g.ny + 1, high = col_in && r - 1 >= 1 && r - 1 <= g.ny + 1;
  // Every value the cell may use is loaded here, before any branch on the
  // data, so the loads are in flight together; where a value is not used its
  // load reads the cell itself.
  const bool xs = left || right, ys = low || high;
  const int cw = xs ? c - 1 : c, ce_ = xs ? c + 1 : c, rs = ys ? r - 1 : r, rn = ys ? r + 1 : r;
  const size_t ow = (size_t)r * g.W + cw, oe = (size_t)r * g.W + ce_;
  const size_t os = (size_t)rs * g.W + c, on = (size_t)rn * g.W + c;
  const T H00 = corner(g, Hi, r, c), H10 = corner(g, Hi, r + 1, c);
  const T H01 = corner(g, Hi, r, c + 1), H11 = corner(g, Hi, r + 1, c + 1);
  const T F = fcor(g, fc, r, c);
  const Sten<T> sx = {w.ce[ow], w.ce[o], w.ce[oe], w.cu[ow], w.cu[o], w.cu[oe], w.cv[ow], w.cv[o], w.cv[oe],
                      fcor(g, fc, r, cw), F, fcor(g, fc, r, ce_), H00, H10, H01, H11};
  const Sten<T> sy = {w.ce[os], sx.e1, w.ce[on], w.cu[os], sx.u1, w.cu[on], w.cv[os], sx.v1, w.cv[on],
                      fcor(g, fc, rs, c), F, fcor(g, fc, rn, c), H00, H01, H10, H11};
  const size_t oxr = right ? o - 1 : o, oyr = high ? o - g.W : o;
  const T xl[3] = {rv.xl[0][o], rv.xl[1][o], rv.xl[2][o]};
  const T xr[3] = {rv.xr[0][oxr], rv.xr[1][oxr], rv.xr[2][oxr]};
  const T yl[3] = {rv.yl[0][o], rv.yl[1][o], rv.yl[2][o]};
  const T yr[3] = {rv.yr[0][oyr], rv.yr[1][oyr], rv.yr[2][oyr]};
  const size_t k = (size_t)r * g.sp + c;
  const T oute = out_e[k], loe = lo_e[k], lou = lo_u[k], lov = lo_v[k];
  // out_adj's rule on the loaded values (it applies only where the update does)
  const T hm = T(0.25) * (H00 + H10 + H01 + H11);
  CellAdj<T> a{0, 0, 0};
  if (!(oute == -hm + T(g.depth_cutoff))) a = {loe, lou, lov};
  return a2_parts_v(g, dt, rk_step, r, c, H00, H10, H01, H11, F, sx, sy, a, xl, xr, yl, yr);
}

// px of a cell from the parts of its west neighbour, itself and its east neighbour.
template <class T>
GPO_HD inline void store_parts(const Geo& g, const Rev<T>& rv, int r, int c, const T* xp_west, const CellParts<T>& self,
                               const T* xm_east) {
  const size_t o = (size_t)r * g.W + c;
  for (int i = 0; i < 3; ++i) {
    rv.px[i][o] = self.own[i] + xp_west[i] + xm_east[i];
    rv.ym[i][o] = self.ym[i];
    rv.yp[i][o] = self.yp[i];
  }
}

// The reverse of the conversion on loaded values: the gathered sensitivity
// (l0, l1, l2) of a converted cell with input eta e, depth bm and velocities
// (u, v) -> that of its loaded (eta, hu, hv); shu/shv the signs of a
// wall-mirrored momentum.  u = hu / den with den = h, or below eps the
// desingularised a = h^2 / (2 eps) + eps / 2 (between the depth cutoff and
// eps, min(a, eps) = a and max(a, |h|) = a), d den / dh = 1 or h / eps.
template <class T>
GPO_HD inline void conv_rev_vals(const Geo& g, T l0, T l1, T l2, T e, T bm, T u, T v, float shu, float shv, T& oe,
                                 T& ou, T& ov) {
  const T h = e + bm;
  oe = ou = ov = T(0);
  if (land_cell(g, bm) || !(h > T(g.depth_cutoff))) return;
  const bool shallow = h < T(g.desing_eps);
  const T den = shallow ? h * h * T(0.5 / g.desing_eps) + T(0.5 * g.desing_eps) : h;
  const T dh = shallow ? h * T(1.0 / g.desing_eps) : T(1);
  ou = qdiv(l1, den) * T(shu);
  ov = qdiv(l2, den) * T(shv);
  oe = l0 - qdiv((l1 * u + l2 * v) * dh, den);
}

// Sensitivity of the loaded stage input of converted cell (r, c) (any cell of
// the grid): the gathered sensitivity of its (eta, u, v) through the
// conversion (wet branch), with the sign of a wall-mirrored momentum.
template <class T>
GPO_HD void conv_rev(const Geo& g, const T* Hi, const Fwd<T>& w, const Rev<T>& rv, const T* in_e, int r, int c, T& oe,
                     T& ou, T& ov) {
  const size_t o = (size_t)r * g.W + c;
  T l0 = rv.px[0][o], l1 = rv.px[1][o], l2 = rv.px[2][o];
  if (r >= 1) {
    l0 += rv.yp[0][o - g.W];
    l1 += rv.yp[1][o - g.W];
    l2 += rv.yp[2][o - g.W];
  }
  if (r <= g.ny + 2) {
    l0 += rv.ym[0][o + g.W];
    l1 += rv.ym[1][o + g.W];
    l2 += rv.ym[2][o + g.W];
  }
  const Src src = map_cell(g, r, c);
  conv_rev_vals(g, l0, l1, l2, in_e[(size_t)src.r * g.sp + src.c], cell_bm(g, Hi, r, c), w.cu[o], w.cv[o], src.shu,
                src.shv, oe, ou, ov);
}

// A ghost cell of row r or column c may load from interior cell (r, c) only
// in the two columns (rows) next to each side.
GPO_HD inline bool may_receive_fold(const Geo& g, int r, int c) {
  return c <= 3 || c >= g.nx || r <= 3 || r >= g.ny;
}

// A3: the sensitivity of interior input cell (r, c): its own converted cell
// plus the ghost cells loaded from it; written to lambda_in, and for the
// second RK stage the direct Q^n term to lambda_qn.  lq_mode 0 adds to
// lambda_qn, 1 overwrites it.  lambda_in may alias lambda_out (every value
// of the cell is loaded before anything is stored); lambda_qn must not alias
// lambda_in.  The cell is interior, so it is its own stencil cell and both
// of its y-neighbours' parts exist: all of its loads are issued together,
// then conv_rev's arithmetic runs on them (the same operations in the same
// order as conv_rev).
template <class T>
GPO_HD void a3_cell(Geo g, int rk_step, int accumulate, int lq_mode, const T* Hi, const Fwd<T>& w, const Rev<T>& rv,
                    const T* in_e, const T* out_e, const T* lo_e, const T* lo_u, const T* lo_v, T* li_e, T* li_u,
                    T* li_v, T* lq_e, T* lq_u, T* lq_v, int r, int c) {
  if (c > g.nx + 1 || r > g.ny + 1) return;
  const size_t o = (size_t)r * g.W + c, k = (size_t)r * g.sp + c;
  const T p0 = rv.px[0][o], p1 = rv.px[1][o], p2 = rv.px[2][o];
  const T s0 = rv.yp[0][o - g.W], s1 = rv.yp[1][o - g.W], s2 = rv.yp[2][o - g.W];
  const T n0 = rv.ym[0][o + g.W], n1 = rv.ym[1][o + g.W], n2 = rv.ym[2][o + g.W];
  const T ein = in_e[k];
  const T ha = corner(g, Hi, r, c), hb = corner(g, Hi, r + 1, c);
  const T hc = corner(g, Hi, r, c + 1), hd = corner(g, Hi, r + 1, c + 1);
  const T u = w.cu[o], v = w.cv[o];
  const bool with_lq = rk_step == 1 && lq_e;
  T oute = T(0), loe = T(0), lou = T(0), lov = T(0), qe = T(0), qu = T(0), qv = T(0);
  if (with_lq) {
    oute = out_e[k], loe = lo_e[k], lou = lo_u[k], lov = lo_v[k];
    if (!lq_mode) qe = lq_e[k], qu = lq_u[k], qv = lq_v[k];
  }
  T ie = T(0), iu = T(0), iv = T(0);
  if (accumulate) ie = li_e[k], iu = li_u[k], iv = li_v[k];

  T e, uu, vv;
  conv_rev_vals(g, (p0 + s0) + n0, (p1 + s1) + n1, (p2 + s2) + n2, ein, bm_of(g, ha, hb, hc, hd), u, v, 1.0f, 1.0f,
                e, uu, vv);
  if (may_receive_fold(g, r, c)) {
    const int gc[4] = {0, 1, g.nx + 2, g.nx + 3};
    for (int i = 0; i < 4; ++i) {
      const Src s = map_cell(g, r, gc[i]);
      if (s.r == r && s.c == c && !(gc[i] == c)) {
        T ge, gu, gv;
        conv_rev(g, Hi, w, rv, in_e, r, gc[i], ge, gu, gv);
        e += ge;
        uu += gu;
        vv += gv;
      }
    }
    const int gr[4] = {0, 1, g.ny + 2, g.ny + 3};
    for (int i = 0; i < 4; ++i) {
      const Src s = map_cell(g, gr[i], c);
      if (s.r == r && s.c == c && !(gr[i] == r)) {
        T ge, gu, gv;
        conv_rev(g, Hi, w, rv, in_e, gr[i], c, ge, gu, gv);
        e += ge;
        uu += gu;
        vv += gv;
      }
    }
  }
  if (with_lq) {  // out_adj's rule on the loaded values
    const T hm = T(0.25) * (ha + hb + hc + hd);
    T ae = T(0), au = T(0), av = T(0);
    if (!(oute == -hm + T(g.depth_cutoff))) ae = loe, au = lou, av = lov;
    if (lq_mode) {
      lq_e[k] = T(0.5) * ae;
      lq_u[k] = T(0.5) * au;
      lq_v[k] = T(0.5) * av;
    } else {
      lq_e[k] = qe + T(0.5) * ae;
      lq_u[k] = qu + T(0.5) * au;
      lq_v[k] = qv + T(0.5) * av;
    }
  }
  if (accumulate) {
    li_e[k] = ie + e;
    li_u[k] = iu + uu;
    li_v[k] = iv + vv;
  } else {
    li_e[k] = e;
    li_u[k] = uu;
    li_v[k] = vv;
  }
}

}  // namespace adj
}  // namespace gpo
