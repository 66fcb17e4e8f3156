// Host build of the differentiable substep stages (gpuocean_amd/csrc/
// adjoint_core.h) for the CPU tests: the same per-cell code the gfx950
// kernels run, looped over the grid on one host thread.  Test
// infrastructure only.  Arrays are dense: state rows of nx+4 doubles, Hi
// rows of nx+5, f_cells rows of nx+4.
#include <cstring>
#include <vector>

#include "../../gpuocean_amd/csrc/adjoint_core.h"
#include "../../include/gpuocean_amd.h"

using namespace gpo::adj;

namespace {

int make_geo(const gpo_cdklm16_params* p, uint32_t bc, Geo* g) {
  if (!p || p->nx < 2 || p->ny < 2 || p->rk_order != 2 || p->r != 0.0f) return -1;
  g->nx = p->nx;
  g->ny = p->ny;
  g->W = p->nx + 4;
  g->sp = p->nx + 4;
  g->hp = p->nx + 5;
  g->fp = p->nx + 4;
  g->bc_n = (bc >> 24) & 0xFF;
  g->bc_s = (bc >> 16) & 0xFF;
  g->bc_e = (bc >> 8) & 0xFF;
  g->bc_w = bc & 0xFF;
  g->dx = p->dx;
  g->dy = p->dy;
  g->g = p->g;
  g->f = p->f;
  g->theta = p->theta;
  g->desing_eps = p->desing_eps;
  g->depth_cutoff = p->depth_cutoff;
  g->flux_slope_eps = p->flux_slope_eps;
  g->flux_balancer = p->flux_balancer;
  g->north_x = p->north_x;
  g->north_y = p->north_y;
  g->has_land = p->has_land ? 1 : 0;
  g->land = p->land_value;
  return 0;
}

// Hi with land corners flagged (kDry), as the library's entry points prepare
// it before the stages run.
std::vector<double> flagged_hi(const Geo& g, const double* Hi) {
  std::vector<double> out((size_t)(g.nx + 5) * (g.ny + 5));
  for (int r = 0; r < g.ny + 5; ++r)
    for (int c = 0; c < g.nx + 5; ++c) out[(size_t)r * (g.nx + 5) + c] = corner_land(g, Hi, r, c);
  return out;
}

template <class S> Fwd<S> carve_vec(std::vector<S>& ws, const Geo& g) {
  ws.assign((size_t)kAdjWorkArrays * (g.nx + 4) * (g.ny + 4), S(0.0));
  return carve<S>(ws.data(), g.nx, g.ny);
}

template <class S>
void forward(const Geo& g, S dt, int rk, const S* ie, const S* iu, const S* iv, const S* qe, const S* qu,
             const S* qv, S* oe, S* ou, S* ov, const S* Hi, const S* fc, int* flag) {
  std::vector<S> ws;
  Fwd<S> w = carve_vec(ws, g);
  const int H = g.ny + 4, W = g.nx + 4;
  for (int r = 0; r < H; ++r)
    for (int c = 0; c < W; ++c) f1_cell(g, ie, iu, iv, Hi, w, flag, r, c);
  for (int r = 0; r < H; ++r)
    for (int c = 0; c < W; ++c) f2_cell(g, Hi, fc, w, r, c);
  for (int r = 0; r < H; ++r)
    for (int c = 0; c < W; ++c) f3_cell(g, Hi, fc, w, flag, r, c);
  for (int r = 2; r <= g.ny + 1; ++r)
    for (int c = 2; c <= g.nx + 1; ++c) f4_cell(g, dt, rk, Hi, fc, w, qe, qu, qv, oe, ou, ov, flag, r, c);
}

}  // namespace

extern "C" {

int gpo_host_substep_f64(const gpo_cdklm16_params* p, double dt, int rk, const double* ie, const double* iu,
                         const double* iv, const double* qe, const double* qu, const double* qv, double* oe,
                         double* ou, double* ov, const double* Hi, const double* fc, uint32_t bc, int* flag) {
  Geo g;
  if (make_geo(p, bc, &g)) return -1;
  const std::vector<double> hf = flagged_hi(g, Hi);
  forward<double>(g, dt, rk, ie, iu, iv, qe, qu, qv, oe, ou, ov, hf.data(), fc, flag);
  return 0;
}

int gpo_host_substep_tlm_f64(const gpo_cdklm16_params* p, double dt, int rk, const double* ie, const double* iu,
                             const double* iv, const double* die, const double* diu, const double* div,
                             const double* qe, const double* qu, const double* qv, const double* dqe,
                             const double* dqu, const double* dqv, double* oe, double* ou, double* ov, double* doe,
                             double* dou, double* dov, const double* Hi, const double* fc, uint32_t bc, int* flag) {
  Geo g;
  if (make_geo(p, bc, &g)) return -1;
  const size_t n = (size_t)(g.nx + 4) * (g.ny + 4), nh = (size_t)(g.nx + 5) * (g.ny + 5);
  std::vector<Dual> a(n), b(n), c(n), qa(n), qb(n), qc(n), oa(n), ob(n), oc(n), hi(nh), f(n);
  for (size_t k = 0; k < n; ++k) {
    a[k] = Dual(ie[k], die[k]);
    b[k] = Dual(iu[k], diu[k]);
    c[k] = Dual(iv[k], div[k]);
    if (rk == 1) {
      qa[k] = Dual(qe[k], dqe[k]);
      qb[k] = Dual(qu[k], dqu[k]);
      qc[k] = Dual(qv[k], dqv[k]);
    }
    if (fc) f[k] = Dual(fc[k]);
    oa[k] = ob[k] = oc[k] = Dual(0.0);
  }
  const std::vector<double> hf = flagged_hi(g, Hi);
  for (size_t k = 0; k < nh; ++k) hi[k] = Dual(hf[k]);
  forward<Dual>(g, Dual(dt), rk, a.data(), b.data(), c.data(), qa.data(), qb.data(), qc.data(), oa.data(),
                ob.data(), oc.data(), hi.data(), fc ? f.data() : nullptr, flag);
  for (int r = 2; r <= g.ny + 1; ++r)
    for (int cc = 2; cc <= g.nx + 1; ++cc) {
      const size_t k = (size_t)r * (g.nx + 4) + cc;
      oe[k] = oa[k].v;
      ou[k] = ob[k].v;
      ov[k] = oc[k].v;
      doe[k] = oa[k].d;
      dou[k] = ob[k].d;
      dov[k] = oc[k].d;
    }
  return 0;
}

int gpo_host_substep_adj_f64(const gpo_cdklm16_params* p, double dt, int rk, const double* ie, const double* iu,
                             const double* iv, const double* out_e, const double* le, const double* lu,
                             const double* lv, double* lie, double* liu, double* liv, double* lqe, double* lqu,
                             double* lqv, int accumulate, const double* Hi, const double* fc, uint32_t bc,
                             int* flag) {
  Geo g;
  if (make_geo(p, bc, &g)) return -1;
  const std::vector<double> hf = flagged_hi(g, Hi);
  Hi = hf.data();
  std::vector<double> ws;
  Fwd<double> w = carve_vec(ws, g);
  Rev<double> rv = carve_rev<double>(ws.data(), g.nx, g.ny);
  const int H = g.ny + 4, W = g.nx + 4;
  for (int r = 0; r < H; ++r)
    for (int c = 0; c < W; ++c) f1_cell(g, ie, iu, iv, Hi, w, flag, r, c);
  for (int r = 0; r < H; ++r)
    for (int c = 0; c < W; ++c) f2_cell(g, Hi, fc, w, r, c);
  for (int r = 0; r < H; ++r)
    for (int c = 0; c < W; ++c) a1_cell(g, dt, rk, Hi, fc, w, rv, out_e, le, lu, lv, r, c);
  for (int r = 0; r < H; ++r) {
    std::vector<CellParts<double>> row(W + 2);  // cells -1 .. W
    for (int c = -1; c <= W; ++c) row[c + 1] = a2_parts(g, dt, rk, Hi, fc, w, rv, out_e, le, lu, lv, r, c);
    for (int c = 0; c < W; ++c) store_parts(g, rv, r, c, row[c].xp, row[c + 1], row[c + 2].xm);
  }
  for (int r = 2; r <= g.ny + 1; ++r)
    for (int c = 2; c <= g.nx + 1; ++c)
      a3_cell(g, rk, accumulate, 0, Hi, w, rv, ie, out_e, le, lu, lv, lie, liu, liv, lqe, lqu, lqv, r, c);
  return 0;
}

}  // extern "C"
