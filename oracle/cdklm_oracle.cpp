// CPU oracle for the CDKLM16 hot path -- TEST INFRASTRUCTURE ONLY.
//
// This file restates, per cell and without tiling, the arithmetic of the
// reference GPU Ocean CDKLM16 path so that the HIP kernels can be checked
// against it.  It is compiled with g++ -O2 -fopenmp -ffp-contract=off
// -fno-fast-math, templated on float (the parity oracle and the timed CPU
// baseline) and double (the accuracy reference).  Only tests/, the smoke()
// entry and bench.py's cpu_baseline leg may load it; the product never does.
//
// Every function cites the reference file:line it restates.  Paths are
// relative to the reference repository root (src/gpuocean/...).
//
// Semantics decisions that differ from a literal reading of the tiled
// reference (documented so the HIP kernels follow the same rule):
//  * Wall ghost cells are mirrored for every interior row/column before the
//    stencil (CDKLM16_kernel.cu:419-470).  The reference mirrors inside each
//    thread block's shared memory, so when a block boundary falls exactly
//    between cell NX and NX+1 the neighbouring block sees the unmirrored
//    global ghost values.  That block-shape artefact is not reproduced.
//  * The RK3 middle stage writes its result into a separate buffer instead
//    of racing with the reads of its own input (CDKLM16_kernel.cu:1076-1083).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle.h"

namespace {

// CDKLM16_kernel.cu:34-35, initHm_kernel.cu:27-28, max_dt.cu:30-35
constexpr float kDryFlag = 1.0e-30f;
constexpr float kDryEps = 1.0e-3f;
constexpr float kMaxDtFltMax = 100000.0f;

template <typename T> inline T sgn1(T x) { return std::copysign(T(1), x); }

// common.cu:532-551 (minmodSlope / minmodRaw), restated literally
template <typename T> inline T minmod_raw(T backward, T central, T forward) {
  return T(0.25) * sgn1(backward) * (sgn1(backward) + sgn1(central)) *
         (sgn1(central) + sgn1(forward)) *
         std::min(std::min(std::fabs(backward), std::fabs(central)), std::fabs(forward));
}
template <typename T> inline T minmod_slope(T left, T center, T right, T theta) {
  const T backward = (center - left) * theta;
  const T central = (right - left) * T(0.5);
  const T forward = (right - center) * theta;
  return minmod_raw(backward, central, forward);
}

// common.cu:664-666.  The reference adds the double literal 0.5*eps, so in the
// float build the sum is formed in double and narrowed by fminf.
inline float desingularize(float h, float hu, float eps) {
  const float a = h * h / (2.0f * eps);
  const float s = static_cast<float>(static_cast<double>(a) + 0.5 * static_cast<double>(eps));
  return hu / std::fmax(std::fmin(s, eps), std::fabs(h));
}
inline double desingularize(double h, double hu, double eps) {
  return hu / std::fmax(std::fmin(h * h / (2.0 * eps) + 0.5 * eps, eps), std::fabs(h));
}

template <typename T> inline T saturate(T x) { return std::min(std::max(x, T(0)), T(1)); }

// interpolation.cu:137-190 (bilinear_interpolation, tex2D-style, clamped)
template <typename T>
T bilinear(const T* data, int data_nx, int data_ny, T norm_x, T norm_y) {
  norm_x = saturate(norm_x - T(0.5) / T(data_nx));
  norm_y = saturate(norm_y - T(0.5) / T(data_ny));
  const T x = norm_x * T(data_nx);
  const T y = norm_y * T(data_ny);
  int x0 = static_cast<int>(std::floor(x));
  int y0 = static_cast<int>(std::floor(y));
  const T dx = x - T(x0);
  const T dy = y - T(y0);
  x0 = std::min(x0, data_nx - 1);
  y0 = std::min(y0, data_ny - 1);
  const int x1 = std::min(x0 + 1, data_nx - 1);
  const int y1 = std::min(y0 + 1, data_ny - 1);
  const T d00 = data[y0 * data_nx + x0];
  const T d01 = data[y0 * data_nx + x1];
  const T d10 = data[y1 * data_nx + x0];
  const T d11 = data[y1 * data_nx + x1];
  return d00 + dx * (d01 - d00) + dy * (d10 - d00) + dx * dy * (d00 + d11 - d01 - d10);
}

template <typename T> struct P {
  int nx, ny, rk_order, bc_n, bc_e, bc_s, bc_w, one_dimensional;
  T dx, dy, g, r, theta, desing_eps, flux_slope_eps, depth_cutoff, flux_balancer;
  T land_value, north_x, north_y, wind_stress_factor, rho_o;
  explicit P(const gpo_oracle_params* q)
      : nx(q->nx), ny(q->ny), rk_order(q->rk_order), bc_n(q->bc_north), bc_e(q->bc_east),
        bc_s(q->bc_south), bc_w(q->bc_west), one_dimensional(q->one_dimensional), dx(T(q->dx)),
        dy(T(q->dy)), g(T(q->g)), r(T(q->r)), theta(T(q->theta)), desing_eps(T(q->desing_eps)),
        flux_slope_eps(T(q->flux_slope_eps)), depth_cutoff(T(q->depth_cutoff)),
        flux_balancer(T(q->flux_balancer)), land_value(T(q->land_value)), north_x(T(q->north_x)),
        north_y(T(q->north_y)), wind_stress_factor(T(q->wind_stress_factor)), rho_o(T(q->rho_o)) {}
};

template <typename T> struct V3 { T x, y, z; };

// CDKLM16_kernel.cu:82-91
template <typename T> inline V3<T> f_func(const V3<T>& Q, T H, T g) {
  const T h = Q.x + H;
  return {h * Q.y, h * Q.y * Q.y + T(0.5) * g * h * h, h * Q.y * Q.z};
}

// CDKLM16_kernel.cu:104-151
template <typename T>
V3<T> cdklm_flux(const P<T>& p, const V3<T>& Qm, const V3<T>& Qp, T H_face) {
  const T hp = Qp.x + H_face;
  const T hm = Qm.x + H_face;
  V3<T> Fp{0, 0, 0}, Fm{0, 0, 0};
  T up = 0, cp = 0, um = 0, cm = 0;
  if (hp > p.depth_cutoff) {
    Fp = f_func(Qp, H_face, p.g);
    up = Qp.y;
    cp = std::sqrt(p.g * hp);
  }
  if (hm > p.depth_cutoff) {
    Fm = f_func(Qm, H_face, p.g);
    um = Qm.y;
    cm = std::sqrt(p.g * hm);
  }
  const T am = std::min(std::min(um - cm, up - cp), T(0));
  const T ap = std::max(std::max(um + cm, up + cp), T(0));
  if (std::fabs(ap - am) < p.flux_slope_eps) return {0, 0, 0};
  V3<T> F;
  F.x = ((ap * Fm.x - am * Fp.x) + ap * am * (Qp.x - Qm.x)) / (ap - am);
  F.y = ((ap * Fm.y - am * Fp.y) + ap * am * (Fp.x - Fm.x)) / (ap - am);
  F.z = (Qm.y > -Qp.y) ? p.flux_balancer * Fm.z : p.flux_balancer * Fp.z;
  F.z += (T(1) - p.flux_balancer) *
         (((ap * Fm.z - am * Fp.z) + ap * am * (hp * Qp.z - hm * Qm.z)) / (ap - am));
  return F;
}

struct Forcing {
  const float* wx_cur;
  const float* wx_next;
  const float* wy_cur;
  const float* wy_next;
  int wx_nx, wx_ny, wy_nx, wy_ny;
  float wind_t;
  const float* p_cur;
  const float* p_next;
  int p_nx, p_ny;
  float p_t;
};

// Working buffers for one substep, sized (ny+4) x (nx+4) (cells) and
// (ny+5) x (nx+5) (corners).
template <typename T> struct Work {
  int W = 0, H = 0, WI = 0;
  std::vector<T> R0, R1, R2, Hl, Hm, Qx0, Qx1, Kx, Qy0, Qy1, Ly, Fx, Fy, Fz, Gx, Gy, Gz, fcell;
  std::vector<T> wind_cur_x, wind_next_x, wind_cur_y, wind_next_y, pres_cur, pres_next;
  void resize(int nx, int ny) {
    W = nx + 4;
    H = ny + 4;
    WI = nx + 5;
    const size_t n = size_t(W) * H, ni = size_t(WI) * (ny + 5);
    for (auto* v : {&R0, &R1, &R2, &Hm, &Qx0, &Qx1, &Kx, &Qy0, &Qy1, &Ly, &Fx, &Fy, &Fz, &Gx, &Gy, &Gz,
                    &fcell})
      v->assign(n, T(0));
    Hl.assign(ni, T(0));
  }
};

template <typename T> std::vector<T> to_T(const float* a, int n) {
  std::vector<T> v(n);
  for (int i = 0; i < n; ++i) v[i] = T(a[i]);
  return v;
}

// external_forcing.cu:17-29 (windStress); the reference passes ti+0.5 as a
// double literal that is narrowed to the float parameter.
template <typename T>
T wind_stress(const T* cur, const T* next, T wt, T ti, T tj, int nx4, int ny4, int dnx, int dny) {
  const T s = ti / T(nx4);
  const T t = tj / T(ny4);
  const T c = bilinear(cur, dnx, dny, s, t);
  const T n = bilinear(next, dnx, dny, s, t);
  return wt * n + (T(1) - wt) * c;
}

// external_forcing.cu:40-68 (atmospheric_pressure_central_diff)
template <typename T>
void pressure_diff(const T* cur, const T* next, T pt, T ti, T tj, int nx4, int ny4, int dnx, int dny, T* ox,
                   T* oy) {
  const T s = ti / T(nx4);
  const T t = tj / T(ny4);
  const T dx = T(1) / T(nx4);
  const T dy = T(1) / T(ny4);
  const T cn = bilinear(cur, dnx, dny, s, t + dy), nn = bilinear(next, dnx, dny, s, t + dy);
  const T cs = bilinear(cur, dnx, dny, s, t - dy), ns = bilinear(next, dnx, dny, s, t - dy);
  const T ce = bilinear(cur, dnx, dny, s + dx, t), ne = bilinear(next, dnx, dny, s + dx, t);
  const T cw = bilinear(cur, dnx, dny, s - dx, t), nw = bilinear(next, dnx, dny, s - dx, t);
  const T pn = pt * nn + (T(1) - pt) * cn;
  const T ps = pt * ns + (T(1) - pt) * cs;
  const T pe = pt * ne + (T(1) - pt) * ce;
  const T pw = pt * nw + (T(1) - pt) * cw;
  *ox = pe - pw;
  *oy = pn - ps;
}

// One RK substep of cdklm_swe_2D (CDKLM16_kernel.cu:475-1090).
// in_*  : Q at the substep input (the kernel's eta0/hu0/hv0 arguments)
// out_* : interior written (the kernel's eta1/hu1/hv1 arguments); for the
//         second/third RK stage Q^n is read from out_* first.
// stage_out_* : only for RK3 step 1, where the reference writes its result
//               into the input buffer (see file header).
template <typename T>
void substep(const P<T>& p, Work<T>& w, const T* in_eta, const T* in_hu, const T* in_hv, T* out_eta,
             T* out_hu, T* out_hv, T* stage_eta, T* stage_hu, T* stage_hv, const T* Hi, const T* Bm,
             const T* fcell, T dt, int rk_step, const Forcing* fo) {
  const int NX = p.nx, NY = p.ny, W = w.W, WI = w.WI;
  auto at = [W](int r, int c) { return size_t(r) * W + c; };
  auto ai = [WI](int r, int c) { return size_t(r) * WI + c; };
  const T g2 = T(2) * p.g;
  const T nx_ = p.north_x, ny_ = p.north_y;
  const T ex = ny_, ey = -nx_;  // east = (north.y, -north.x), CDKLM16_kernel.cu:745

  T* R0 = w.R0.data();
  T* R1 = w.R1.data();
  T* R2 = w.R2.data();
  T* Hl = w.Hl.data();
  T* Hm = w.Hm.data();

  // Shared-memory staging: R = Q, Hi with land corners -> DRY_FLAG
  // (CDKLM16_kernel.cu:552-587).
#pragma omp parallel for schedule(static)
  for (int r = 0; r < NY + 4; ++r)
    for (int c = 0; c < NX + 4; ++c) {
      R0[at(r, c)] = in_eta[at(r, c)];
      R1[at(r, c)] = in_hu[at(r, c)];
      R2[at(r, c)] = in_hv[at(r, c)];
    }
#pragma omp parallel for schedule(static)
  for (int r = 0; r < NY + 5; ++r)
    for (int c = 0; c < NX + 5; ++c) {
      T v = Hi[ai(r, c)];
      if (std::fabs(v - p.land_value) < T(kDryEps)) v = T(kDryFlag);
      Hl[ai(r, c)] = v;
    }

  // handleWallBC (CDKLM16_kernel.cu:419-470)
  if (p.bc_n == 1)
    for (int c = 2; c <= NX + 1; ++c) {
      const int j = NY + 1;
      R0[at(j + 1, c)] = R0[at(j, c)];
      R1[at(j + 1, c)] = R1[at(j, c)];
      R2[at(j + 1, c)] = -R2[at(j, c)];
      R0[at(j + 2, c)] = R0[at(j - 1, c)];
      R1[at(j + 2, c)] = R1[at(j - 1, c)];
      R2[at(j + 2, c)] = -R2[at(j - 1, c)];
    }
  if (p.bc_s == 1)
    for (int c = 2; c <= NX + 1; ++c) {
      const int j = 2;
      R0[at(j - 1, c)] = R0[at(j, c)];
      R1[at(j - 1, c)] = R1[at(j, c)];
      R2[at(j - 1, c)] = -R2[at(j, c)];
      R0[at(j - 2, c)] = R0[at(j + 1, c)];
      R1[at(j - 2, c)] = R1[at(j + 1, c)];
      R2[at(j - 2, c)] = -R2[at(j + 1, c)];
    }
  if (p.bc_e == 1)
    for (int r = 2; r <= NY + 1; ++r) {
      const int i = NX + 1;
      R0[at(r, i + 1)] = R0[at(r, i)];
      R1[at(r, i + 1)] = -R1[at(r, i)];
      R2[at(r, i + 1)] = R2[at(r, i)];
      R0[at(r, i + 2)] = R0[at(r, i - 1)];
      R1[at(r, i + 2)] = -R1[at(r, i - 1)];
      R2[at(r, i + 2)] = R2[at(r, i - 1)];
    }
  if (p.bc_w == 1)
    for (int r = 2; r <= NY + 1; ++r) {
      const int i = 2;
      R0[at(r, i - 1)] = R0[at(r, i)];
      R1[at(r, i - 1)] = -R1[at(r, i)];
      R2[at(r, i - 1)] = R2[at(r, i)];
      R0[at(r, i - 2)] = R0[at(r, i + 1)];
      R1[at(r, i - 2)] = -R1[at(r, i + 1)];
      R2[at(r, i - 2)] = R2[at(r, i + 1)];
    }

  // Per-cell Hm from the shared-memory corners (CDKLM16_kernel.cu:609) and the
  // conversion to (eta, u, v) using the global Bm (CDKLM16_kernel.cu:614-655).
#pragma omp parallel for schedule(static)
  for (int r = 0; r < NY + 4; ++r)
    for (int c = 0; c < NX + 4; ++c) {
      Hm[at(r, c)] = T(0.25) * (Hl[ai(r, c)] + Hl[ai(r + 1, c)] + Hl[ai(r, c + 1)] + Hl[ai(r + 1, c + 1)]);
      const T local_Hm = Bm[at(r, c)];
      const T h = R0[at(r, c)] + local_Hm;
      if (std::fabs(local_Hm - p.land_value) <= T(kDryEps)) {
        R0[at(r, c)] = T(kDryFlag);
        R1[at(r, c)] = 0;
        R2[at(r, c)] = 0;
      } else if (h < p.desing_eps) {
        if (h <= p.depth_cutoff) {
          R0[at(r, c)] = -local_Hm + p.depth_cutoff;
          R1[at(r, c)] = 0;
          R2[at(r, c)] = 0;
        } else {
          R1[at(r, c)] = desingularize(h, R1[at(r, c)], p.desing_eps);
          R2[at(r, c)] = desingularize(h, R2[at(r, c)], p.desing_eps);
        }
      } else {
        R1[at(r, c)] /= h;
        R2[at(r, c)] /= h;
      }
    }

  auto F = [&](int r, int c) -> T { return fcell[at(r, c)]; };

  // x-slopes + adjustSlopes_x for rows 2..NY+1, columns 1..NX+2
  // (CDKLM16_kernel.cu:673-737, 159-203)
  T* Qx0 = w.Qx0.data();
  T* Qx1 = w.Qx1.data();
  T* Kx = w.Kx.data();
  const T Vc = p.dx / (T(2) * p.g);
#pragma omp parallel for schedule(static)
  for (int r = 2; r <= NY + 1; ++r)
    for (int c = 1; c <= NX + 2; ++c) {
      const T le = R0[at(r, c - 1)], ce = R0[at(r, c)], re = R0[at(r, c + 1)];
      const T lu = R1[at(r, c - 1)], cu = R1[at(r, c)], ru = R1[at(r, c + 1)];
      T lv = R2[at(r, c - 1)], cv = R2[at(r, c)], rv = R2[at(r, c + 1)];
      Qx0[at(r, c)] = minmod_slope(lu, cu, ru, p.theta);
      Qx1[at(r, c)] = minmod_slope(lv, cv, rv, p.theta);
      const int gid = c;
      if (p.bc_w == 1) {
        if (gid < 3) lv = -lv;
        if (gid < 2) cv = -cv;
      }
      if (p.bc_e == 1) {
        if (gid > NX) rv = -rv;
        if (gid > NX + 1) cv = -cv;
      }
      const T lfv = (nx_ * lu + ny_ * lv) * F(r, c - 1);
      const T cfv = (nx_ * cu + ny_ * cv) * F(r, c);
      const T rfv = (nx_ * ru + ny_ * rv) * F(r, c + 1);
      const T backward = p.theta * p.g * (ce - le - Vc * (cfv + lfv));
      const T central = T(0.5) * p.g * (re - le - Vc * (rfv + T(2) * cfv + lfv));
      const T forward = p.theta * p.g * (re - ce - Vc * (cfv + rfv));
      T K = minmod_raw(backward, central, forward);
      // adjustSlopes_x
      T v = R2[at(r, c)];
      if (p.bc_w == 1 && c < 2) v = -v;
      if (p.bc_e == 1 && c > NX + 2) v = -v;
      const T dxfv = p.dx * F(r, c) * v;
      const T H_west = T(0.5) * (Hl[ai(r, c)] + Hl[ai(r + 1, c)]);
      const T H_east = T(0.5) * (Hl[ai(r, c + 1)] + Hl[ai(r + 1, c + 1)]);
      const T h_west = ce + H_west - (K + dxfv) / g2;
      const T h_east = ce + H_east + (K + dxfv) / g2;
      K = (h_west > 0) ? K : -dxfv + g2 * (ce + H_west);
      K = (h_east > 0) ? K : -dxfv - g2 * (ce + H_east);
      Kx[at(r, c)] = K;
    }

  // x-face fluxes, face between columns c and c+1 (CDKLM16_kernel.cu:259-340)
#pragma omp parallel for schedule(static)
  for (int r = 2; r <= NY + 1; ++r)
    for (int c = 1; c <= NX + 1; ++c) {
      const T ebp = R0[at(r, c + 1)], ebm = R0[at(r, c)];
      const T up = R1[at(r, c + 1)], um = R1[at(r, c)];
      T vp = R2[at(r, c + 1)], vm = R2[at(r, c)];
      V3<T> out{0, 0, 0};
      if (!(ebm == T(kDryFlag) && ebp == T(kDryFlag))) {
        const T Rpx = up - T(0.5) * Qx0[at(r, c + 1)], Rpy = vp - T(0.5) * Qx1[at(r, c + 1)];
        const T Rmx = um + T(0.5) * Qx0[at(r, c)], Rmy = vm + T(0.5) * Qx1[at(r, c)];
        const T H_face = T(0.5) * (Hl[ai(r, c + 1)] + Hl[ai(r + 1, c + 1)]);
        const T Kp = Kx[at(r, c + 1)], Km = Kx[at(r, c)];
        if (p.bc_w == 1 && c + 1 == 2) vm = -vm;
        if (p.bc_e == 1 && c + 1 == NX + 2) vp = -vp;
        const T vpn = up * nx_ + vp * ny_;
        const T vmn = um * nx_ + vm * ny_;
        const T etap = std::fmax(-H_face, ebp - (Kp + p.dx * F(r, c + 1) * vpn) / g2);
        const T etam = std::fmax(-H_face, ebm + (Km + p.dx * F(r, c) * vmn) / g2);
        const V3<T> Qp{etap, Rpx, Rpy}, Qm{etam, Rmx, Rmy};
        if (ebm == T(kDryFlag) && ebp != T(kDryFlag)) {
          out = {0, T(0.5) * p.g * (Qp.x + H_face) * (Qp.x + H_face), 0};
        } else if (ebm != T(kDryFlag) && ebp == T(kDryFlag)) {
          out = {0, T(0.5) * p.g * (Qm.x + H_face) * (Qm.x + H_face), 0};
        } else {
          out = cdklm_flux(p, Qm, Qp, H_face);
        }
      }
      w.Fx[at(r, c)] = out.x;
      w.Fy[at(r, c)] = out.y;
      w.Fz[at(r, c)] = out.z;
    }

  // y-slopes + adjustSlopes_y for rows 1..NY+2, columns 2..NX+1
  // (CDKLM16_kernel.cu:779-842, 210-254)
  T* Qy0 = w.Qy0.data();
  T* Qy1 = w.Qy1.data();
  T* Ly = w.Ly.data();
  const T Uc = p.dy / (T(2) * p.g);
#pragma omp parallel for schedule(static)
  for (int r = 1; r <= NY + 2; ++r)
    for (int c = 2; c <= NX + 1; ++c) {
      const T le = R0[at(r - 1, c)], ce = R0[at(r, c)], ue = R0[at(r + 1, c)];
      T lu = R1[at(r - 1, c)], cu = R1[at(r, c)], uu = R1[at(r + 1, c)];
      const T lv = R2[at(r - 1, c)], cv = R2[at(r, c)], uv = R2[at(r + 1, c)];
      Qy0[at(r, c)] = minmod_slope(lu, cu, uu, p.theta);
      Qy1[at(r, c)] = minmod_slope(lv, cv, uv, p.theta);
      const int gid = r;
      if (p.bc_s == 1) {
        if (gid < 3) lu = -lu;
        if (gid < 2) cu = -cu;
      }
      if (p.bc_n == 1) {
        if (gid > NY) uu = -uu;
        if (gid > NY + 1) cu = -cu;
      }
      const T lfu = (ex * lu + ey * lv) * F(r - 1, c);
      const T cfu = (ex * cu + ey * cv) * F(r, c);
      const T ufu = (ex * uu + ey * uv) * F(r + 1, c);
      const T backward = p.theta * p.g * (ce - le + Uc * (cfu + lfu));
      const T central = T(0.5) * p.g * (ue - le + Uc * (ufu + T(2) * cfu + lfu));
      const T forward = p.theta * p.g * (ue - ce + Uc * (cfu + ufu));
      T L = minmod_raw(backward, central, forward);
      T u = R1[at(r, c)];
      if (p.bc_s == 1 && r < 2) u = -u;
      if (p.bc_n == 1 && r > NY + 2) u = -u;
      const T dyfu = p.dy * F(r, c) * u;
      const T H_south = T(0.5) * (Hl[ai(r, c)] + Hl[ai(r, c + 1)]);
      const T H_north = T(0.5) * (Hl[ai(r + 1, c)] + Hl[ai(r + 1, c + 1)]);
      const T h_south = ce + H_south - (L - dyfu) / g2;
      const T h_north = ce + H_north + (L - dyfu) / g2;
      L = (h_south > 0) ? L : dyfu + g2 * (ce + H_south);
      L = (h_north > 0) ? L : dyfu - g2 * (ce + H_north);
      Ly[at(r, c)] = L;
    }

  // y-face fluxes, face between rows r and r+1 (CDKLM16_kernel.cu:345-416)
#pragma omp parallel for schedule(static)
  for (int r = 1; r <= NY + 1; ++r)
    for (int c = 2; c <= NX + 1; ++c) {
      const T ebp = R0[at(r + 1, c)], ebm = R0[at(r, c)];
      T up = R1[at(r + 1, c)], um = R1[at(r, c)];
      const T vp = R2[at(r + 1, c)], vm = R2[at(r, c)];
      V3<T> out{0, 0, 0};
      if (!(ebm == T(kDryFlag) && ebp == T(kDryFlag))) {
        const T Rpx = up - T(0.5) * Qy0[at(r + 1, c)], Rpy = vp - T(0.5) * Qy1[at(r + 1, c)];
        const T Rmx = um + T(0.5) * Qy0[at(r, c)], Rmy = vm + T(0.5) * Qy1[at(r, c)];
        const T H_face = T(0.5) * (Hl[ai(r + 1, c)] + Hl[ai(r + 1, c + 1)]);
        const T Lp = Ly[at(r + 1, c)], Lm = Ly[at(r, c)];
        if (p.bc_s == 1 && r + 1 == 2) um = -um;
        if (p.bc_n == 1 && r + 1 == NY + 2) up = -up;
        const T upe = up * ex + vp * ey;
        const T ume = um * ex + vm * ey;
        const T etap = std::fmax(-H_face, ebp - (Lp - p.dy * F(r + 1, c) * upe) / g2);
        const T etam = std::fmax(-H_face, ebm + (Lm - p.dy * F(r, c) * ume) / g2);
        const V3<T> Qp{etap, Rpy, Rpx}, Qm{etam, Rmy, Rmx};
        if (ebm == T(kDryFlag) && ebp != T(kDryFlag)) {
          out = {0, 0, T(0.5) * p.g * (Qp.x + H_face) * (Qp.x + H_face)};
        } else if (ebm != T(kDryFlag) && ebp == T(kDryFlag)) {
          out = {0, 0, T(0.5) * p.g * (Qm.x + H_face) * (Qm.x + H_face)};
        } else {
          const V3<T> fl = cdklm_flux(p, Qm, Qp, H_face);
          out = {fl.x, fl.z, fl.y};
        }
      }
      w.Gx[at(r, c)] = out.x;
      w.Gy[at(r, c)] = out.y;
      w.Gz[at(r, c)] = out.z;
    }

  // Source terms and the RK update (CDKLM16_kernel.cu:739-1089)
  const bool has_wind = fo && fo->wx_cur;
  const bool has_pres = fo && fo->p_cur;
#pragma omp parallel for schedule(static)
  for (int r = 2; r <= NY + 1; ++r)
    for (int c = 2; c <= NX + 1; ++c) {
      const size_t k = at(r, c);
      const T hm_c = Hm[k];
      T hu = 0, hv = 0;
      if ((R0[k] + hm_c) > p.depth_cutoff) {
        hu = R1[k] * (R0[k] + hm_c);
        hv = R2[k] * (R0[k] + hm_c);
      }
      const T fc = F(r, c);
      T fdx = (w.Fx[k] - w.Fx[at(r, c - 1)]) / p.dx;
      T fdy = (w.Fy[k] - w.Fy[at(r, c - 1)]) / p.dx;
      T fdz = (w.Fz[k] - w.Fz[at(r, c - 1)]) / p.dx;
      const T eta_west = R0[k] - (Kx[k] + p.dx * fc * R2[k]) / g2;
      const T eta_east = R0[k] + (Kx[k] + p.dx * fc * R2[k]) / g2;
      if (!p.one_dimensional) {
        fdx = fdx + (w.Gx[k] - w.Gx[at(r - 1, c)]) / p.dy;
        fdy = fdy + (w.Gy[k] - w.Gy[at(r - 1, c)]) / p.dy;
        fdz = fdz + (w.Gz[k] - w.Gz[at(r - 1, c)]) / p.dy;
      }
      const T eta_south = R0[k] - (Ly[k] - p.dy * fc * R1[k]) / g2;
      const T eta_north = R0[k] + (Ly[k] - p.dy * fc * R1[k]) / g2;

      T st1 = 0, st2 = 0;
      const T h = R0[k] + hm_c;
      if (h >= p.depth_cutoff && R0[k] != T(kDryFlag)) {
        const T RHxp = T(0.5) * (Hl[ai(r, c + 1)] + Hl[ai(r + 1, c + 1)]);
        const T RHxm = T(0.5) * (Hl[ai(r, c)] + Hl[ai(r + 1, c)]);
        const T RHyp = T(0.5) * (Hl[ai(r + 1, c)] + Hl[ai(r + 1, c + 1)]);
        const T RHym = T(0.5) * (Hl[ai(r, c)] + Hl[ai(r, c + 1)]);
        const T hu_east = hu * ex + hv * ey;
        const T hv_north = hu * nx_ + hv * ny_;
        const T hu_east_cor = fc * hv_north;
        const T hv_north_cor = -fc * hu_east;
        const T upx = -nx_, upy = ny_;
        const T rtx = upy, rty = -upx;
        const T hu_cor = rtx * hu_east_cor + rty * hv_north_cor;
        const T hv_cor = upx * hu_east_cor + upy * hv_north_cor;
        T apx = 0, apy = 0;
        if (has_pres)
          pressure_diff(w.pres_cur.data(), w.pres_next.data(), T(fo->p_t), T(float(c + 0.5)),
                        T(float(r + 0.5)), NX + 4, NY + 4, fo->p_nx, fo->p_ny, &apx, &apy);
        T X = 0, Y = 0;
        if (has_wind) {
          X = p.wind_stress_factor * wind_stress(w.wind_cur_x.data(), w.wind_next_x.data(), T(fo->wind_t),
                                                 T(float(c + 0.5)), T(float(r + 0.5)), NX + 4, NY + 4,
                                                 fo->wx_nx, fo->wx_ny);
          Y = p.wind_stress_factor * wind_stress(w.wind_cur_y.data(), w.wind_next_y.data(), T(fo->wind_t),
                                                 T(float(c + 0.5)), T(float(r + 0.5)), NX + 4, NY + 4,
                                                 fo->wy_nx, fo->wy_ny);
        }
        {
          const T atm_x = -apx * h / (T(2) * p.dx * p.rho_o);
          const T eta_we = T(0.5) * (eta_west + eta_east);
          const T H_x = RHxp - RHxm;
          const T bathy1 = p.g * (eta_we + hm_c) * H_x;
          st1 = X + hu_cor + atm_x + bathy1 / p.dx;
        }
        {
          const T atm_y = -apy * h / (T(2) * p.dy * p.rho_o);
          const T eta_sn = T(0.5) * (eta_north + eta_south);
          const T H_y = RHyp - RHym;
          const T bathy2 = p.g * (eta_sn + hm_c) * H_y;
          st2 = Y + hv_cor + atm_y + bathy2 / p.dy;
        }
      }
      const T L1 = -fdx;
      const T L2 = -fdy + st1;
      const T L3 = -fdz + st2;
      T ue = 0, uhu = 0, uhv = 0;
      if (p.rk_order < 3) {
        T C = 0;
        if (p.r > 0) {
          if (h < p.desing_eps) {
            const T u = desingularize(h, hu, p.desing_eps);
            const T v = desingularize(h, hv, p.desing_eps);
            C = dt * p.r * desingularize(h, std::sqrt(u * u + v * v), p.desing_eps);
          } else {
            const T u = hu / h;
            const T v = hv / h;
            C = dt * p.r * std::sqrt(u * u + v * v) / h;
          }
        }
        if (rk_step == 0) {
          ue = R0[k] + dt * L1;
          uhu = (hu + dt * L2) / (T(1) + C);
          uhv = (hv + dt * L3) / (T(1) + C);
        } else {
          const T ea = out_eta[k], ha = out_hu[k], va = out_hv[k];
          const T eb = T(0.5) * (ea + (R0[k] + dt * L1));
          const T hb = T(0.5) * (ha + (hu + dt * L2));
          const T vb = T(0.5) * (va + (hv + dt * L3));
          ue = eb;
          uhu = hb / (T(1) + T(0.5) * C);
          uhv = vb / (T(1) + T(0.5) * C);
        }
      } else {
        if (rk_step == 0) {
          ue = R0[k] + dt * L1;
          uhu = hu + dt * L2;
          uhv = hv + dt * L3;
        } else if (rk_step == 1) {
          const T ea = out_eta[k], ha = out_hu[k], va = out_hv[k];
          ue = T(0.75) * ea + T(0.25) * (R0[k] + dt * L1);
          uhu = T(0.75) * ha + T(0.25) * (hu + dt * L2);
          uhv = T(0.75) * va + T(0.25) * (hv + dt * L3);
        } else {
          const T ea = out_eta[k], ha = out_hu[k], va = out_hv[k];
          ue = (ea + T(2) * (R0[k] + dt * L1)) / T(3);
          uhu = (ha + T(2) * (hu + dt * L2)) / T(3);
          uhv = (va + T(2) * (hv + dt * L3)) / T(3);
        }
      }
      const T uh = ue + hm_c;
      if (uh <= p.depth_cutoff) {
        ue = -hm_c + p.depth_cutoff;
        uhu = 0;
        uhv = 0;
      }
      T* de = out_eta;
      T* dh = out_hu;
      T* dv = out_hv;
      if (p.rk_order == 3 && rk_step == 1) {
        de = stage_eta;
        dh = stage_hu;
        dv = stage_hv;
      }
      de[k] = std::fmax(-hm_c + p.depth_cutoff, ue);
      dh[k] = uhu;
      dv[k] = uhv;
    }
}

// periodicBoundary_NS / _EW (boundary_kernels.cu:36-110), NS before EW.
template <typename T> void periodic_ns(int nx, int ny, T* a) {
  const int W = nx + 4;
  for (int j = 0; j < ny + 4; ++j) {
    if (!(j < 2 || j > ny + 1)) continue;
    const int opp = (j > ny + 1) ? j - ny : j + ny;
    std::memcpy(a + size_t(j) * W, a + size_t(opp) * W, sizeof(T) * W);
  }
}
template <typename T> void periodic_ew(int nx, int ny, T* a) {
  const int W = nx + 4;
  for (int j = 0; j < ny + 4; ++j)
    for (int i = 0; i < nx + 4; ++i) {
      if (!(i < 2 || i > nx + 1)) continue;
      const int opp = (i > nx + 1) ? i - nx : i + nx;
      a[size_t(j) * W + i] = a[size_t(j) * W + opp];
    }
}

// bilinear_interpolation_3channels (interpolation.cu:197-260)
template <typename T> V3<T> bilinear3(const float* data, int dnx, int dny, T norm_x, T norm_y) {
  norm_x = saturate(norm_x - T(0.5) / T(dnx));
  norm_y = saturate(norm_y - T(0.5) / T(dny));
  const T x = norm_x * T(dnx), y = norm_y * T(dny);
  int x0 = int(std::floor(x)), y0 = int(std::floor(y));
  const T dx = x - T(x0), dy = y - T(y0);
  x0 = std::min(x0, dnx - 1);
  y0 = std::min(y0, dny - 1);
  const int x1 = std::min(x0 + 1, dnx - 1), y1 = std::min(y0 + 1, dny - 1);
  T res[3];
  for (int c = 0; c < 3; ++c) {
    const T d00 = T(data[(y0 * dnx + x0) * 3 + c]);
    const T d01 = T(data[(y0 * dnx + x1) * 3 + c]);
    const T d10 = T(data[(y1 * dnx + x0) * 3 + c]);
    const T d11 = T(data[(y1 * dnx + x1) * 3 + c]);
    res[c] = d00 + dx * (d01 - d00) + dy * (d10 - d00) + dx * dy * (d00 + d11 - d01 - d10);
  }
  return {res[0], res[1], res[2]};
}

// flowRelaxationScheme_NS (boundary_kernels.cu:406-472)
template <typename T>
void frs_ns(const gpo_oracle_params* q, T* h, T* u, T* v, const float* cur, const float* next, int bc_nx,
            float bc_t) {
  const int nx = q->nx, ny = q->ny, W = nx + 4;
  const int sn = q->sponge_north, ss = q->sponge_south;
  for (int tj = 0; tj < ny + 4; ++tj)
    for (int ti = 0; ti < nx + 4; ++ti) {
      const bool in_s = q->bc_south == 3 && tj < ss && tj >= 0;
      const bool in_n = q->bc_north == 3 && tj > ny + 4 - 1 - sn && tj < ny + 4;
      if (!(in_s || in_n)) continue;
      const T j = (tj <= ss) ? T(tj - 1) : T((ny + 4 - 2) - tj);
      const T alpha = T(1) - std::tanh(std::fmax(T(0), j) / T(2));
      const T s = T(ti) / T(float(nx + 4));
      const T t = (tj < ny / 2.0) ? T(0.25) : T(0.75);
      const V3<T> c = bilinear3<T>(cur, bc_nx, 2, s, t);
      const V3<T> n = bilinear3<T>(next, bc_nx, 2, s, t);
      const T bt = T(bc_t);
      const T eh = bt * n.x + (T(1) - bt) * c.x;
      const T eu = bt * n.y + (T(1) - bt) * c.y;
      const T ev = bt * n.z + (T(1) - bt) * c.z;
      const size_t k = size_t(tj) * W + ti;
      h[k] = (T(1) - alpha) * h[k] + alpha * eh;
      u[k] = (T(1) - alpha) * u[k] + alpha * eu;
      v[k] = (T(1) - alpha) * v[k] + alpha * ev;
    }
}

// flowRelaxationScheme_EW (boundary_kernels.cu:477-540)
template <typename T>
void frs_ew(const gpo_oracle_params* q, T* h, T* u, T* v, const float* cur, const float* next, int bc_ny,
            float bc_t) {
  const int nx = q->nx, ny = q->ny, W = nx + 4;
  const int se = q->sponge_east, sw = q->sponge_west;
  for (int tj = 0; tj < ny + 4; ++tj)
    for (int ti = 0; ti < nx + 4; ++ti) {
      const bool in_w = q->bc_west == 3 && ti < sw && ti >= 0;
      const bool in_e = q->bc_east == 3 && ti > nx + 4 - 1 - se && ti < nx + 4;
      if (!(in_w || in_e)) continue;
      const T j = (ti <= sw) ? T(ti - 1) : T((nx + 4 - 2) - ti);
      const T alpha = T(1) - std::tanh(std::fmax(T(0), j) / T(2));
      const T s = (ti < nx / 2.0) ? T(0.25) : T(0.75);
      const T t = T(tj) / T(float(ny + 4));
      const V3<T> c = bilinear3<T>(cur, 2, bc_ny, s, t);
      const V3<T> n = bilinear3<T>(next, 2, bc_ny, s, t);
      const T bt = T(bc_t);
      const T eh = bt * n.x + (T(1) - bt) * c.x;
      const T eu = bt * n.y + (T(1) - bt) * c.y;
      const T ev = bt * n.z + (T(1) - bt) * c.z;
      const size_t k = size_t(tj) * W + ti;
      h[k] = (T(1) - alpha) * h[k] + alpha * eh;
      u[k] = (T(1) - alpha) * u[k] + alpha * eu;
      v[k] = (T(1) - alpha) * v[k] + alpha * ev;
    }
}

// linearInterpolation_NS (boundary_kernels.cu:265-325): on a code-4 side the
// rows strictly between the outer ghost row and the row `sponge` cells in are
// interpolated linearly between those two rows, which are kept.
template <typename T> void lin_ns(const gpo_oracle_params* q, T* h, T* u, T* v) {
  const int nx = q->nx, ny = q->ny, W = nx + 4, top = ny + 4 - 1;
  const int sn = q->sponge_north, ss = q->sponge_south;
  for (int tj = 0; tj < ny + 4; ++tj)
    for (int ti = 0; ti < nx + 4; ++ti) {
      const bool in = ((q->bc_south == 4 && tj < ss && tj > 0) || (q->bc_north == 4 && tj > top - sn && tj < top)) &&
                      ti > 0 && ti < nx + 4 - 1;
      if (!in) continue;
      int inner = ss, outer = 0;
      if (tj > ss) inner = top - sn, outer = top;
      const float ratio = (float)(tj - outer) / (float)(inner - outer);
      for (T* a : {h, u, v}) {
        const T vi = a[size_t(inner) * W + ti], vo = a[size_t(outer) * W + ti];
        a[size_t(tj) * W + ti] = vo + T(ratio) * (vi - vo);
      }
    }
}

// linearInterpolation_EW (boundary_kernels.cu:330-384), by columns.
template <typename T> void lin_ew(const gpo_oracle_params* q, T* h, T* u, T* v) {
  const int nx = q->nx, ny = q->ny, W = nx + 4, right = nx + 4 - 1;
  const int se = q->sponge_east, sw = q->sponge_west;
  for (int tj = 0; tj < ny + 4; ++tj)
    for (int ti = 0; ti < nx + 4; ++ti) {
      const bool in = ((q->bc_west == 4 && ti < sw && ti > 0) || (q->bc_east == 4 && ti > right - se && ti < right)) &&
                      tj > 0 && tj < ny + 4 - 1;
      if (!in) continue;
      int inner = sw, outer = 0;
      if (ti > sw) inner = right - se, outer = right;
      const float ratio = (float)(ti - outer) / (float)(inner - outer);
      for (T* a : {h, u, v}) {
        const T vi = a[size_t(tj) * W + inner], vo = a[size_t(tj) * W + outer];
        a[size_t(tj) * W + ti] = vo + T(ratio) * (vi - vo);
      }
    }
}

// BoundaryConditionsArakawaA.boundaryCondition (utils/Common.py:1310-1330):
// north/south then east/west; on each axis periodic, or else flow relaxation
// and then linear interpolation for the sides that have them.
template <typename T>
void apply_bc(const gpo_oracle_params* q, T* h, T* u, T* v, const gpo_oracle_frs* frs) {
  if (q->bc_north == 2) {
    periodic_ns(q->nx, q->ny, h);
    periodic_ns(q->nx, q->ny, u);
    periodic_ns(q->nx, q->ny, v);
  } else {
    if ((q->bc_north == 3 || q->bc_south == 3) && frs) frs_ns(q, h, u, v, frs->ns_cur, frs->ns_next, frs->ns_nx, frs->t);
    if (q->bc_north == 4 || q->bc_south == 4) lin_ns(q, h, u, v);
  }
  if (q->bc_east == 2) {
    periodic_ew(q->nx, q->ny, h);
    periodic_ew(q->nx, q->ny, u);
    periodic_ew(q->nx, q->ny, v);
  } else {
    if ((q->bc_east == 3 || q->bc_west == 3) && frs) frs_ew(q, h, u, v, frs->ew_cur, frs->ew_next, frs->ew_ny, frs->t);
    if (q->bc_east == 4 || q->bc_west == 4) lin_ew(q, h, u, v);
  }
}

// coriolisF (CDKLM16_kernel.cu:48-57) evaluated at every cell, ghosts included.
template <typename T> void fill_fcell(const gpo_oracle_params* q, std::vector<T>& fc) {
  const int W = q->nx + 4, H = q->ny + 4;
  if (!q->f_arr) {
    std::fill(fc.begin(), fc.end(), T(q->f));
    return;
  }
  std::vector<T> arr = to_T<T>(q->f_arr, q->f_nx * q->f_ny);
  for (int j = 0; j < H; ++j)
    for (int i = 0; i < W; ++i) {
      const T s = (T(i) + T(0.5)) / (T(q->nx) + T(4));
      const T t = (T(j) + T(0.5)) / (T(q->ny) + T(4));
      fc[size_t(j) * W + i] = bilinear(arr.data(), q->f_nx, q->f_ny, s, t);
    }
}

template <typename T>
void load_forcing(Work<T>& w, const Forcing* fo) {
  if (!fo) return;
  if (fo->wx_cur) {
    w.wind_cur_x = to_T<T>(fo->wx_cur, fo->wx_nx * fo->wx_ny);
    w.wind_next_x = to_T<T>(fo->wx_next, fo->wx_nx * fo->wx_ny);
    w.wind_cur_y = to_T<T>(fo->wy_cur, fo->wy_nx * fo->wy_ny);
    w.wind_next_y = to_T<T>(fo->wy_next, fo->wy_nx * fo->wy_ny);
  }
  if (fo->p_cur) {
    w.pres_cur = to_T<T>(fo->p_cur, fo->p_nx * fo->p_ny);
    w.pres_next = to_T<T>(fo->p_next, fo->p_nx * fo->p_ny);
  }
}

// CDKLM16.step RK loop (SWEsimulators/CDKLM16.py:493-579) for n_steps steps of
// a fixed dt, without noise or drifters.  h0/h1 are the double buffer.
template <typename T>
void run(const gpo_oracle_params* q, T* h0e, T* h0u, T* h0v, T* h1e, T* h1u, T* h1v, const T* Hi, const T* Bm,
         double dt_in, int n_steps, int apply_initial_bc, const gpo_oracle_forcing* forcing,
         const gpo_oracle_frs* frs) {
  P<T> p(q);
  Work<T> w;
  w.resize(q->nx, q->ny);
  fill_fcell(q, w.fcell);
  Forcing fo{};
  const Forcing* fop = nullptr;
  if (forcing) {
    fo = Forcing{forcing->wx_cur, forcing->wx_next, forcing->wy_cur, forcing->wy_next,
                 forcing->wx_nx,  forcing->wx_ny,   forcing->wy_nx,  forcing->wy_ny,
                 forcing->wind_t, forcing->p_cur,   forcing->p_next, forcing->p_nx,
                 forcing->p_ny,   forcing->p_t};
    fop = &fo;
    load_forcing(w, fop);
  }
  const T dt = T(dt_in);
  const size_t n = size_t(q->nx + 4) * (q->ny + 4);
  std::vector<T> se, su, sv;
  if (q->rk_order == 3) {
    se.assign(n, 0);
    su.assign(n, 0);
    sv.assign(n, 0);
  }
  if (apply_initial_bc) apply_bc(q, h0e, h0u, h0v, frs);
  for (int s = 0; s < n_steps; ++s) {
    if (q->rk_order == 2) {
      substep(p, w, h0e, h0u, h0v, h1e, h1u, h1v, (T*)nullptr, (T*)nullptr, (T*)nullptr, Hi, Bm, w.fcell.data(), dt, 0, fop);
      apply_bc(q, h1e, h1u, h1v, frs);
      substep(p, w, h1e, h1u, h1v, h0e, h0u, h0v, (T*)nullptr, (T*)nullptr, (T*)nullptr, Hi, Bm, w.fcell.data(), dt, 1, fop);
    } else if (q->rk_order == 1) {
      substep(p, w, h0e, h0u, h0v, h1e, h1u, h1v, (T*)nullptr, (T*)nullptr, (T*)nullptr, Hi, Bm, w.fcell.data(), dt, 0, fop);
      // The reference swaps the buffer roles (CDKLM16.py:539); copying keeps
      // the caller's h0 as the latest state.
      std::memcpy(h0e, h1e, sizeof(T) * n);
      std::memcpy(h0u, h1u, sizeof(T) * n);
      std::memcpy(h0v, h1v, sizeof(T) * n);
    } else {
      substep(p, w, h0e, h0u, h0v, h1e, h1u, h1v, (T*)nullptr, (T*)nullptr, (T*)nullptr, Hi, Bm, w.fcell.data(), dt, 0, fop);
      apply_bc(q, h1e, h1u, h1v, frs);
      // stage 1 reads Q^(1) from h1 and Q^n from h0, writes Q^(2) to the
      // stage buffer which then takes h1's role.
      std::memcpy(se.data(), h1e, sizeof(T) * n);
      std::memcpy(su.data(), h1u, sizeof(T) * n);
      std::memcpy(sv.data(), h1v, sizeof(T) * n);
      substep(p, w, h1e, h1u, h1v, h0e, h0u, h0v, se.data(), su.data(), sv.data(), Hi, Bm, w.fcell.data(), dt, 1,
              fop);
      std::memcpy(h1e, se.data(), sizeof(T) * n);
      std::memcpy(h1u, su.data(), sizeof(T) * n);
      std::memcpy(h1v, sv.data(), sizeof(T) * n);
      apply_bc(q, h1e, h1u, h1v, frs);
      substep(p, w, h1e, h1u, h1v, h0e, h0u, h0v, (T*)nullptr, (T*)nullptr, (T*)nullptr, Hi, Bm, w.fcell.data(), dt, 2, fop);
    }
    apply_bc(q, h0e, h0u, h0v, frs);
  }
}

// Bathymetry._boundaryConditions (utils/Common.py:1538-1561) and the corner
// kernels it launches (boundary_kernels.cu:122-246), then initBm
// (initHm_kernel.cu:38-106).  Hi is modified in place.
template <typename T>
void init_bathymetry(const gpo_oracle_params* q, T* Hi, T* Bm) {
  const int nx = q->nx, ny = q->ny, WI = nx + 5, W = nx + 4;
  auto ai = [WI](int r, int c) { return size_t(r) * WI + c; };
  const int halo = 2;
  // north-south
  if (q->bc_north == 2 && q->bc_south == 2) {
    for (int tj = 0; tj < ny + 5; ++tj) {
      if (!(tj < halo || tj > ny + halo - 1)) continue;
      const int opp = (tj > ny + halo - 1) ? tj - ny : tj + ny;
      for (int ti = 0; ti < nx + 5; ++ti) Hi[ai(tj, ti)] = Hi[ai(opp, ti)];
    }
  } else if (q->bc_north == 1 && q->bc_south == 1) {
    for (int ti = 0; ti < nx + 5; ++ti) {
      for (int j = 0; j < halo; ++j) Hi[ai(j, ti)] = Hi[ai(2 * halo - j, ti)];
      for (int j = 0; j < halo; ++j) Hi[ai(ny + 2 * halo - j, ti)] = Hi[ai(ny + j, ti)];
    }
  }
  // east-west; the closed branch tests north/south as the reference does
  // (utils/Common.py:1557).
  if (q->bc_east == 2 && q->bc_west == 2) {
    for (int tj = 0; tj < ny + 5; ++tj)
      for (int ti = 0; ti < nx + 5; ++ti) {
        if (!(ti < halo || ti > nx + halo - 1)) continue;
        const int opp = (ti > nx + halo - 1) ? ti - nx : ti + nx;
        Hi[ai(tj, ti)] = Hi[ai(tj, opp)];
      }
  } else if (q->bc_north == 1 && q->bc_south == 1) {
    for (int tj = 0; tj < ny + 5; ++tj) {
      for (int i = 0; i < halo; ++i) Hi[ai(tj, i)] = Hi[ai(tj, 2 * halo - i)];
      for (int i = 0; i < halo; ++i) Hi[ai(tj, nx + 2 * halo - i)] = Hi[ai(tj, nx + i)];
    }
  }
  const T land = T(q->land_value);
  for (int tj = 0; tj < ny + 4; ++tj)
    for (int ti = 0; ti < nx + 4; ++ti) {
      T c4[4] = {Hi[ai(tj, ti)], Hi[ai(tj + 1, ti)], Hi[ai(tj, ti + 1)], Hi[ai(tj + 1, ti + 1)]};
      for (T& x : c4)
        if (std::fabs(x - land) <= T(kDryEps)) x = T(kDryFlag);
      const T wet = T(4) - T(c4[0] == T(kDryFlag)) - T(c4[1] == T(kDryFlag)) - T(c4[2] == T(kDryFlag)) -
                    T(c4[3] == T(kDryFlag));
      Bm[size_t(tj) * W + ti] = (wet == T(0)) ? land : (c4[0] + c4[1] + c4[2] + c4[3]) / wet;
    }
}

// per_block_max_dt + max_dt_reduction (max_dt.cu:42-195): minimum over the
// interior of the CFL time step, dry and land cells giving 1e5.
template <typename T>
T max_dt(const gpo_oracle_params* q, const T* eta, const T* hu, const T* hv, const T* Bm) {
  const int nx = q->nx, ny = q->ny, W = nx + 4;
  const T dx = T(q->dx), dy = T(q->dy), g = T(q->g), land = T(q->land_value);
  T best = T(kMaxDtFltMax);
  for (int tj = 2; tj < ny + 2; ++tj)
    for (int ti = 2; ti < nx + 2; ++ti) {
      const size_t k = size_t(tj) * W + ti;
      const T H = Bm[k];
      const T h = eta[k] + H;
      T d;
      if (h <= 0 || std::fabs(H - land) < T(kDryEps)) {
        d = T(kMaxDtFltMax);
      } else {
        const T u = hu[k] / h, v = hv[k] / h;
        const T gw = std::sqrt(g * h);
        d = T(0.25) * dx / std::fabs(u + gw);
        d = std::min(d, T(0.25) * dx / std::fabs(u - gw));
        d = std::min(d, T(0.25) * dy / std::fabs(v + gw));
        d = std::min(d, T(0.25) * dy / std::fabs(v - gw));
      }
      best = std::min(best, d);
    }
  return best;
}

}  // namespace

extern "C" {

void gpo_oracle_init_bathymetry_f32(const gpo_oracle_params* q, float* Hi, float* Bm) {
  init_bathymetry<float>(q, Hi, Bm);
}
void gpo_oracle_init_bathymetry_f64(const gpo_oracle_params* q, double* Hi, double* Bm) {
  init_bathymetry<double>(q, Hi, Bm);
}

void gpo_oracle_run_f32(const gpo_oracle_params* q, float* h0e, float* h0u, float* h0v, float* h1e, float* h1u,
                        float* h1v, const float* Hi, const float* Bm, double dt, int n_steps, int apply_initial_bc,
                        const gpo_oracle_forcing* forcing, const gpo_oracle_frs* frs) {
  run<float>(q, h0e, h0u, h0v, h1e, h1u, h1v, Hi, Bm, dt, n_steps, apply_initial_bc, forcing, frs);
}
void gpo_oracle_run_f64(const gpo_oracle_params* q, double* h0e, double* h0u, double* h0v, double* h1e,
                        double* h1u, double* h1v, const double* Hi, const double* Bm, double dt, int n_steps,
                        int apply_initial_bc, const gpo_oracle_forcing* forcing, const gpo_oracle_frs* frs) {
  run<double>(q, h0e, h0u, h0v, h1e, h1u, h1v, Hi, Bm, dt, n_steps, apply_initial_bc, forcing, frs);
}

void gpo_oracle_substep_f32(const gpo_oracle_params* q, const float* ie, const float* iu, const float* iv,
                            float* oe, float* ou, float* ov, const float* Hi, const float* Bm, float dt,
                            int rk_step) {
  P<float> p(q);
  Work<float> w;
  w.resize(q->nx, q->ny);
  fill_fcell(q, w.fcell);
  substep(p, w, ie, iu, iv, oe, ou, ov, (float*)nullptr, (float*)nullptr, (float*)nullptr, Hi, Bm, w.fcell.data(), dt, rk_step, nullptr);
}

void gpo_oracle_apply_bc_f32(const gpo_oracle_params* q, float* h, float* u, float* v, const gpo_oracle_frs* frs) {
  apply_bc<float>(q, h, u, v, frs);
}
void gpo_oracle_apply_bc_f64(const gpo_oracle_params* q, double* h, double* u, double* v,
                             const gpo_oracle_frs* frs) {
  apply_bc<double>(q, h, u, v, frs);
}

float gpo_oracle_max_dt_f32(const gpo_oracle_params* q, const float* eta, const float* hu, const float* hv,
                            const float* Bm) {
  return max_dt<float>(q, eta, hu, hv, Bm);
}
double gpo_oracle_max_dt_f64(const gpo_oracle_params* q, const double* eta, const double* hu, const double* hv,
                             const double* Bm) {
  return max_dt<double>(q, eta, hu, hv, Bm);
}

void gpo_oracle_fcell_f32(const gpo_oracle_params* q, float* out) {
  std::vector<float> fc(size_t(q->nx + 4) * (q->ny + 4));
  fill_fcell(q, fc);
  std::memcpy(out, fc.data(), sizeof(float) * fc.size());
}

int gpo_oracle_num_threads(void) {
#ifdef _OPENMP
  int n = 1;
#pragma omp parallel
  {
#pragma omp single
    n = omp_get_num_threads();
  }
  return n;
#else
  return 1;
#endif
}

}  // extern "C"
