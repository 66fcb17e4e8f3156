// Boundary-condition, bathymetry, Coriolis-texture and CFL kernels of the
// CDKLM16 path, plus the C ABI entry points that dispatch them.  Built with
// -ffp-contract=off so their float arithmetic matches the host oracle.

#include <cstdio>
#include <mutex>
#include <string>
#include <unordered_map>

#include "boundary_cells.h"
#include "cdklm_common.h"

namespace gpo {

static thread_local std::string g_last_error;
void set_error(const char* msg) { g_last_error = msg ? msg : ""; }

namespace {

inline int ceil_div(int a, int b) { return (a + b - 1) / b; }

__device__ __forceinline__ float* row_ptr(float* base, int pitch_bytes, int j) {
  return (float*)((char*)base + (size_t)pitch_bytes * j);
}
__device__ __forceinline__ const float* row_ptr(const float* base, int pitch_bytes, int j) {
  return (const float*)((const char*)base + (size_t)pitch_bytes * j);
}

__global__ void periodic_ns(int nx, int ny, float* h, float* u, float* v, int pitch) {
  const int ti = blockIdx.x * blockDim.x + threadIdx.x;
  if (ti >= nx + 4) return;
  cells::periodic_ns_cell(nx, ny, h, u, v, pitch, ti, blockIdx.y);
}

__global__ void periodic_ew(int nx, int ny, float* h, float* u, float* v, int pitch) {
  const int tj = blockIdx.x * blockDim.x + threadIdx.x;
  if (tj >= ny + 4) return;
  cells::periodic_ew_cell(nx, ny, h, u, v, pitch, tj, blockIdx.y);
}

__global__ void frs_ns(int bc_n, int bc_s, int nx, int ny, int sn, int ss, float* h, float* u, float* v, int pitch,
                       const float* cur, const float* next, int bc_nx, float t_) {
  const int ti = blockIdx.x * blockDim.x + threadIdx.x;
  const int tj = blockIdx.y * blockDim.y + threadIdx.y;
  if (ti >= nx + 4 || tj >= ny + 4) return;
  cells::frs_ns_cell(bc_n, bc_s, nx, ny, sn, ss, h, u, v, pitch, cur, next, bc_nx, t_, ti, tj);
}

__global__ void frs_ew(int bc_e, int bc_w, int nx, int ny, int se, int sw, float* h, float* u, float* v, int pitch,
                       const float* cur, const float* next, int bc_ny, float t_) {
  const int ti = blockIdx.x * blockDim.x + threadIdx.x;
  const int tj = blockIdx.y * blockDim.y + threadIdx.y;
  if (ti >= nx + 4 || tj >= ny + 4) return;
  cells::frs_ew_cell(bc_e, bc_w, nx, ny, se, sw, h, u, v, pitch, cur, next, bc_ny, t_, ti, tj);
}

// periodic_boundary_intersections_NS / closed_boundary_intersections_NS
// (boundary_kernels.cu:122-147, 188-212): one thread per corner column.
__global__ void lin_ns(int bc_n, int bc_s, int nx, int ny, int sn, int ss, float* h, float* u, float* v, int pitch) {
  const int ti = blockIdx.x * blockDim.x + threadIdx.x, tj = blockIdx.y * blockDim.y + threadIdx.y;
  if (ti >= nx + 4 || tj >= ny + 4) return;
  cells::lin_ns_cell(bc_n, bc_s, nx, ny, sn, ss, h, u, v, pitch, ti, tj);
}

__global__ void lin_ew(int bc_e, int bc_w, int nx, int ny, int se, int sw, float* h, float* u, float* v, int pitch) {
  const int ti = blockIdx.x * blockDim.x + threadIdx.x, tj = blockIdx.y * blockDim.y + threadIdx.y;
  if (ti >= nx + 4 || tj >= ny + 4) return;
  cells::lin_ew_cell(bc_e, bc_w, nx, ny, se, sw, h, u, v, pitch, ti, tj);
}

__global__ void hi_ns(int nx, int ny, int periodic, float* Hi, int pitch) {
  const int ti = blockIdx.x * blockDim.x + threadIdx.x;
  if (ti >= nx + 5) return;
  if (periodic) {
    // ghost rows 0,1 <- ny, ny+1 ; rows ny+2..ny+4 <- 2..4
    for (int tj = 0; tj < ny + 5; ++tj) {
      if (!(tj < 2 || tj > ny + 1)) continue;
      const int opp = (tj > ny + 1) ? tj - ny : tj + ny;
      row_ptr(Hi, pitch, tj)[ti] = row_ptr(Hi, pitch, opp)[ti];
    }
  } else {
    for (int j = 0; j < 2; ++j) row_ptr(Hi, pitch, j)[ti] = row_ptr(Hi, pitch, 4 - j)[ti];
    for (int j = 0; j < 2; ++j) row_ptr(Hi, pitch, ny + 4 - j)[ti] = row_ptr(Hi, pitch, ny + j)[ti];
  }
}

// periodic_boundary_intersections_EW / closed_boundary_intersections_EW
// (boundary_kernels.cu:152-179, 216-246): one thread per corner row.
__global__ void hi_ew(int nx, int ny, int periodic, float* Hi, int pitch) {
  const int tj = blockIdx.x * blockDim.x + threadIdx.x;
  if (tj >= ny + 5) return;
  float* r = row_ptr(Hi, pitch, tj);
  if (periodic) {
    for (int ti = 0; ti < nx + 5; ++ti) {
      if (!(ti < 2 || ti > nx + 1)) continue;
      const int opp = (ti > nx + 1) ? ti - nx : ti + nx;
      r[ti] = r[opp];
    }
  } else {
    for (int i = 0; i < 2; ++i) r[i] = r[4 - i];
    for (int i = 0; i < 2; ++i) r[nx + 4 - i] = r[nx + i];
  }
}

// initBm (initHm_kernel.cu:38-106)
__global__ void init_bm(int nx, int ny, float land, const float* Hi, int hpitch, float* Bm, int bpitch) {
  const int ti = blockIdx.x * blockDim.x + threadIdx.x;
  const int tj = blockIdx.y * blockDim.y + threadIdx.y;
  if (ti >= nx + 4 || tj >= ny + 4) return;
  auto fl = [land](float x) { return (fabsf(x - land) <= kDryEps) ? kDryFlag : x; };
  const float a = fl(row_ptr(Hi, hpitch, tj)[ti]);
  const float b = fl(row_ptr(Hi, hpitch, tj + 1)[ti]);
  const float c = fl(row_ptr(Hi, hpitch, tj)[ti + 1]);
  const float d = fl(row_ptr(Hi, hpitch, tj + 1)[ti + 1]);
  const float wet = 4.0f - (float)(a == kDryFlag) - (float)(b == kDryFlag) - (float)(c == kDryFlag) -
                    (float)(d == kDryFlag);
  row_ptr(Bm, bpitch, tj)[ti] = (wet == 0.0f) ? land : (a + b + c + d) / wet;
}

__device__ float bilinear1(const float* data, int dnx, int dny, float nx_, float ny_) {
  nx_ = fminf(fmaxf(nx_ - 0.5f / (float)dnx, 0.0f), 1.0f);
  ny_ = fminf(fmaxf(ny_ - 0.5f / (float)dny, 0.0f), 1.0f);
  const float x = nx_ * (float)dnx, y = ny_ * (float)dny;
  int x0 = (int)floorf(x), y0 = (int)floorf(y);
  const float fx = x - (float)x0, fy = y - (float)y0;
  x0 = min(x0, dnx - 1);
  y0 = min(y0, dny - 1);
  const int x1 = min(x0 + 1, dnx - 1), y1 = min(y0 + 1, dny - 1);
  const float d00 = data[y0 * dnx + x0], d01 = data[y0 * dnx + x1];
  const float d10 = data[y1 * dnx + x0], d11 = data[y1 * dnx + x1];
  return d00 + fx * (d01 - d00) + fy * (d10 - d00) + fx * fy * (d00 + d11 - d01 - d10);
}

// coriolisF at every cell (CDKLM16_kernel.cu:48-57)
__global__ void coriolis_cells(int nx, int ny, const float* tex, int fnx, int fny, float* out, int pitch) {
  const int ti = blockIdx.x * blockDim.x + threadIdx.x;
  const int tj = blockIdx.y * blockDim.y + threadIdx.y;
  if (ti >= nx + 4 || tj >= ny + 4) return;
  const float s = ((float)ti + 0.5f) / ((float)nx + 4.0f);
  const float t = ((float)tj + 0.5f) / ((float)ny + 4.0f);
  row_ptr(out, pitch, tj)[ti] = bilinear1(tex, fnx, fny, s, t);
}

__global__ void fill_scalar(float* p, float v) { *p = v; }

// Exhaustive check that one Markstein step q + (a - c q) y with y = RN(1/c)
// equals IEEE a / c for every a with |a| in [2^-96, 2^100] (both signs).  For
// c > 0 the IEEE quotient has the sign of a, so the step kernel's copysign of
// the result onto a changes nothing on that set and turns the +0 that the
// step gives for a = -0 into IEEE's -0; +0 gives +0 either way.
constexpr uint64_t kDivcDomain = 2ull * (0x62000000ull + 1ull);
__global__ void __launch_bounds__(256) verify_divc(float c, float y, unsigned int* bad) {
  const uint64_t base = ((uint64_t)blockIdx.x * 256u + threadIdx.x) * 64u;
  unsigned int local = 0;
  for (int k = 0; k < 64; ++k) {
    const uint64_t i = base + (uint64_t)k;
    if (i >= kDivcDomain) break;
    const uint32_t bits = (0x0F800000u + (uint32_t)(i >> 1)) | ((uint32_t)(i & 1u) << 31);
    const float a = __uint_as_float(bits);
    const float ref = a / c;
    const float q = a * y;
    const float r = __builtin_fmaf(-c, q, a);
    const float q1 = __builtin_fmaf(r, y, q);
    local += (__float_as_uint(q1) != __float_as_uint(ref)) ? 1u : 0u;
  }
  if (local) atomicAdd(bad, local);
}

// Exhaustive check that sqrt_rsq equals the IEEE square root for every float
// in [2^-96, FLT_MAX] (bit patterns 0x0F800000 .. 0x7F7FFFFF).
constexpr uint32_t kSqrtLo = 0x0F800000u, kSqrtHi = 0x7F7FFFFFu;
__global__ void __launch_bounds__(256) verify_sqrt(unsigned int* bad) {
  const uint64_t base = ((uint64_t)blockIdx.x * 256u + threadIdx.x) * 64u;
  unsigned int local = 0;
  for (int k = 0; k < 64; ++k) {
    const uint64_t i = base + (uint64_t)k;
    if (i > (uint64_t)(kSqrtHi - kSqrtLo)) break;
    const float x = __uint_as_float(kSqrtLo + (uint32_t)i);
    local += (__float_as_uint(sqrt_rsq(x)) != __float_as_uint(__builtin_sqrtf(x))) ? 1u : 0u;
  }
  if (local) atomicAdd(bad, local);
}

// per_block_max_dt + max_dt_reduction (max_dt.cu:42-195).  Positive floats
// order like their bit patterns, so the global minimum is an integer atomic.
__global__ void __launch_bounds__(256) max_dt_kernel(int nx, int ny, float dx, float dy, float g, float land,
                                                     const float* eta, const float* hu, const float* hv, int pitch,
                                                     const float* Bm, int bpitch, float* out) {
  const int ti = blockIdx.x * 64 + (threadIdx.x & 63) + 2;
  const int tj = blockIdx.y * 4 + (threadIdx.x >> 6) + 2;
  float d = kMaxDtCap;
  if (ti < nx + 2 && tj < ny + 2) {
    const float e = row_ptr(eta, pitch, tj)[ti];
    const float H = row_ptr(Bm, bpitch, tj)[ti];
    const float h = e + H;
    if (!(h <= 0.0f || fabsf(H - land) < kDryEps)) {
      const float u = row_ptr(hu, pitch, tj)[ti] / h;
      const float v = row_ptr(hv, pitch, tj)[ti] / h;
      const float gw = sqrtf(g * h);
      d = 0.25f * dx / fabsf(u + gw);
      d = fminf(d, 0.25f * dx / fabsf(u - gw));
      d = fminf(d, 0.25f * dy / fabsf(v + gw));
      d = fminf(d, 0.25f * dy / fabsf(v - gw));
    }
  }
  for (int off = 32; off > 0; off >>= 1) d = fminf(d, __shfl_xor(d, off, 64));
  __shared__ float wmin[4];
  if ((threadIdx.x & 63) == 0) wmin[threadIdx.x >> 6] = d;
  __syncthreads();
  if (threadIdx.x == 0) {
    const float m = fminf(fminf(wmin[0], wmin[1]), fminf(wmin[2], wmin[3]));
    atomicMin((unsigned int*)out, __float_as_uint(m));
  }
}

std::mutex g_divc_mutex;
std::unordered_map<uint32_t, bool> g_divc_cache;

int check(hipError_t e, const char* what);

// Markstein division by c verified exhaustively (cached per value of c).
int divc_verified(float c, hipStream_t s, bool* ok) {
  const uint32_t key = __builtin_bit_cast(uint32_t, c);
  {
    std::lock_guard<std::mutex> lk(g_divc_mutex);
    auto it = g_divc_cache.find(key);
    if (it != g_divc_cache.end()) {
      *ok = it->second;
      return 0;
    }
  }
  *ok = false;
  if (!(c > 0.0f) || c > 1e30f || c < 1e-30f) return 0;
  unsigned int* d_bad = nullptr;
  int rc = check(hipMalloc(&d_bad, sizeof(unsigned int)), "gpo_cdklm16_prepare hipMalloc");
  if (rc) return rc;
  unsigned int h_bad = 0;
  rc = check(hipMemcpyAsync(d_bad, &h_bad, sizeof h_bad, hipMemcpyHostToDevice, s), "gpo_cdklm16_prepare copy");
  const float y = 1.0f / c;
  const uint64_t threads = (kDivcDomain + 63) / 64;
  const unsigned int blocks = (unsigned int)((threads + 255) / 256);
  if (!rc) {
    hipLaunchKernelGGL(verify_divc, dim3(blocks), dim3(256), 0, s, c, y, d_bad);
    rc = check(hipGetLastError(), "gpo_cdklm16_prepare verify launch");
  }
  if (!rc) rc = check(hipMemcpyAsync(&h_bad, d_bad, sizeof h_bad, hipMemcpyDeviceToHost, s), "gpo_cdklm16_prepare copy");
  if (!rc) rc = check(hipStreamSynchronize(s), "gpo_cdklm16_prepare sync");
  (void)hipFree(d_bad);
  if (rc) return rc;
  *ok = (h_bad == 0);
  std::lock_guard<std::mutex> lk(g_divc_mutex);
  g_divc_cache[key] = *ok;
  return 0;
}

int check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    char buf[256];
    snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
    set_error(buf);
    return (int)e;
  }
  return 0;
}

int bad(const char* msg) {
  set_error(msg);
  return GPO_ERR_BAD_ARGUMENT;
}

}  // namespace
}  // namespace gpo

namespace gpo {

// Wavefronts of a persistent grid with `wps` waves on each SIMD of the
// current device (4 SIMDs per CU); the CU count is cached per device.
int persistent_waves(int wps) {
  static int cus[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256 * 4 * wps;
  if (cus[dev] == 0) {
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
    cus[dev] = n;
  }
  return cus[dev] * 4 * wps;
}

int make_step_args(const gpo_cdklm16_params* p, float dt, int32_t rk_step, const float* in_eta,
                          const float* in_hu, const float* in_hv, const float* qn_eta, const float* qn_hu,
                          const float* qn_hv, float* out_eta, float* out_hu, float* out_hv, int32_t pitch,
                          const float* Hi, int32_t Hi_pitch, const float* f_cells, const gpo_forcing* forcing,
                          uint32_t boundary_conditions, StepArgs* out) {
  if (!p) return bad("gpo_cdklm16: params is NULL");
  if (p->nx < 1 || p->ny < 1) return bad("gpo_cdklm16: nx and ny must be positive");
  if (!in_eta || !in_hu || !in_hv || !out_eta || !out_hu || !out_hv || !Hi)
    return bad("gpo_cdklm16: NULL state or bathymetry pointer");
  if (pitch % 4 || Hi_pitch % 4) return bad("gpo_cdklm16: pitches must be multiples of 4 bytes");
  if (pitch < 4 * (p->nx + 4) || Hi_pitch < 4 * (p->nx + 5)) return bad("gpo_cdklm16: pitch smaller than the row");
  if (p->rk_order < 1 || p->rk_order > 3) return bad("gpo_cdklm16: rk_order must be 1, 2 or 3");
  if (rk_step < 0 || rk_step >= p->rk_order) return bad("gpo_cdklm16: rk_step out of range");
  if (rk_step > 0 && (!qn_eta || !qn_hu || !qn_hv)) return bad("gpo_cdklm16: Q^n pointers needed");
  if (p->f_is_array && !f_cells) return bad("gpo_cdklm16: f_is_array set but f_cells is NULL");
  if (p->rk_order == 3 && p->r != 0.0f) return bad("gpo_cdklm16: RK3 requires r == 0");
  if (p->chunk_rows < 0 || p->chunk_rows > 4096) return bad("gpo_cdklm16: chunk_rows out of range");
  if (p->waves_per_simd < 0 || p->waves_per_simd > 32) return bad("gpo_cdklm16: waves_per_simd out of range");

  StepArgs& a = *out;
  a = StepArgs{};
  a.nx = p->nx;
  a.ny = p->ny;
  a.pitch = pitch / 4;
  a.hi_pitch = Hi_pitch / 4;
  a.grid_waves = persistent_waves(p->waves_per_simd > 0 ? p->waves_per_simd : kDefaultWavesPerSimd);
  a.n_strips = (p->nx + kStripOut - 1) / kStripOut;
  if (p->chunk_rows > 0) {
    a.chunk_rows = p->chunk_rows;
  } else {
    // Default row chunk: enough (strip, chunk) items to give every wave of the
    // persistent grid one, between 2 and kDefaultChunkRows rows.  A grid with
    // far more rows than that takes taller chunks, about four items per wave
    // (load balance) of up to kLargeChunkRows rows: each re-reads and converts
    // the same 4 halo rows for more output rows (on 8192 x 4096, 24-row chunks
    // run 12 % faster than 6-row ones).
    const long long rows = (long long)a.n_strips * p->ny;
    const long long c = (rows + a.grid_waves - 1) / a.grid_waves;
    if (c <= kDefaultChunkRows) {
      a.chunk_rows = (int)(c < 2 ? 2 : c);
    } else {
      const long long c4 = c / 4;
      a.chunk_rows = (int)(c4 < kDefaultChunkRows ? kDefaultChunkRows : (c4 > kLargeChunkRows ? kLargeChunkRows : c4));
    }
  }
  a.n_chunks = (p->ny + a.chunk_rows - 1) / a.chunk_rows;
  a.bc_n = (boundary_conditions >> 24) & 0xFF;
  a.bc_s = (boundary_conditions >> 16) & 0xFF;
  a.bc_e = (boundary_conditions >> 8) & 0xFF;
  a.bc_w = boundary_conditions & 0xFF;
  a.rk_order = p->rk_order;
  a.rk_step = rk_step;
  a.one_dim = p->one_dimensional;
  a.dt = dt;
  a.dx = p->dx;
  a.dy = p->dy;
  a.g = p->g;
  a.r = p->r;
  a.theta = p->theta;
  a.desing_eps = p->desing_eps;
  a.flux_slope_eps = p->flux_slope_eps;
  a.depth_cutoff = p->depth_cutoff;
  a.flux_balancer = p->flux_balancer;
  a.land_value = p->land_value;
  a.f = p->f;
  a.north_x = p->north_x;
  a.north_y = p->north_y;
  a.wind_stress_factor = p->wind_stress_factor;
  a.rho_o = p->rho_o;
  // Derived constants, formed in float exactly as the reference's constant
  // expressions are (e.g. THETA*GRAV, DX/(2.0f*GRAV)).
  a.g2 = 2.0f * p->g;
  a.inv_g2 = 1.0f / a.g2;
  a.half_g = 0.5f * p->g;
  a.theta_g = p->theta * p->g;
  a.Vc = p->dx / (2.0f * p->g);
  a.Uc = p->dy / (2.0f * p->g);
  a.inv_dx = 1.0f / p->dx;
  a.inv_dy = 1.0f / p->dy;
  a.one_m_fb = 1.0f - p->flux_balancer;
  a.dxf = p->dx * p->f;
  a.dyf = p->dy * p->f;
  a.divc_ok = p->divc_ok & 7;
  a.sqrt_ok = (p->g * p->depth_cutoff) >= 0x1p-96f && (p->divc_ok & 8) != 0;
  a.in_e = in_eta;
  a.in_u = in_hu;
  a.in_v = in_hv;
  a.qn_e = qn_eta ? qn_eta : out_eta;
  a.qn_u = qn_hu ? qn_hu : out_hu;
  a.qn_v = qn_hv ? qn_hv : out_hv;
  a.out_e = out_eta;
  a.out_u = out_hu;
  a.out_v = out_hv;
  a.Hi = Hi;
  a.fcell = f_cells;
  if (forcing && (forcing->wind_x_cur || forcing->pressure_cur)) {
    a.has_forcing = 1;
    a.forcing = *forcing;
  }
  return 0;
}

}  // namespace gpo

using namespace gpo;

namespace {

__global__ void fill_double(double* p, double v) { *p = v; }

// FP64 sum of the interior eta of one row per workgroup (row j+2): wave
// reduction, one atomic per wavefront.
__global__ void __launch_bounds__(256) mass_sum_kernel(int nx, const float* eta, int pitch, double* out) {
  const float* row = (const float*)((const char*)eta + (size_t)pitch * (blockIdx.x + 2));
  double acc = 0.0;
  for (int i = threadIdx.x; i < nx; i += 256) acc += (double)row[i + 2];
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off);
  if ((threadIdx.x & 63) == 0) atomicAdd(out, acc);
}

}  // namespace

extern "C" {

const char* gpo_last_error(void) { return g_last_error.c_str(); }

// sqrt_rsq verified exhaustively (once per process).
int g_sqrt_state = -1;  // -1 unknown, 0 differs somewhere, 1 verified
int sqrt_verified(hipStream_t s, bool* ok) {
  {
    std::lock_guard<std::mutex> lk(g_divc_mutex);
    if (g_sqrt_state >= 0) {
      *ok = g_sqrt_state == 1;
      return 0;
    }
  }
  *ok = false;
  unsigned int* d_bad = nullptr;
  int rc = check(hipMalloc(&d_bad, sizeof(unsigned int)), "gpo_cdklm16_prepare hipMalloc");
  if (rc) return rc;
  unsigned int h_bad = 0;
  rc = check(hipMemcpyAsync(d_bad, &h_bad, sizeof h_bad, hipMemcpyHostToDevice, s), "gpo_cdklm16_prepare copy");
  const uint64_t threads = ((uint64_t)(kSqrtHi - kSqrtLo) + 1 + 63) / 64;
  const unsigned int blocks = (unsigned int)((threads + 255) / 256);
  if (!rc) {
    hipLaunchKernelGGL(verify_sqrt, dim3(blocks), dim3(256), 0, s, d_bad);
    rc = check(hipGetLastError(), "gpo_cdklm16_prepare verify launch");
  }
  if (!rc) rc = check(hipMemcpyAsync(&h_bad, d_bad, sizeof h_bad, hipMemcpyDeviceToHost, s), "gpo_cdklm16_prepare copy");
  if (!rc) rc = check(hipStreamSynchronize(s), "gpo_cdklm16_prepare sync");
  (void)hipFree(d_bad);
  if (rc) return rc;
  *ok = (h_bad == 0);
  std::lock_guard<std::mutex> lk(g_divc_mutex);
  g_sqrt_state = *ok ? 1 : 0;
  return 0;
}

const char* gpo_version(void) { return "gpuocean_amd 0.1 (gfx950)"; }

static hipError_t launch_step(const gpo_cdklm16_params* p, const StepArgs& a, int exact, hipStream_t s) {
  return exact ? launch_cdklm_exact(a, p->has_land, p->f_is_array, p->angle_is_zero, s)
               : launch_cdklm_fast(a, p->has_land, p->f_is_array, p->angle_is_zero, s);
}

int gpo_cdklm16_prepare(gpo_cdklm16_params* p, void* stream) {
  if (!p) return bad("gpo_cdklm16_prepare: params is NULL");
  hipStream_t s = (hipStream_t)stream;
  bool okx = false, oky = false, okg = false, oks = false;
  int rc = divc_verified(p->dx, s, &okx);
  if (!rc) rc = divc_verified(p->dy, s, &oky);
  if (!rc) rc = divc_verified(2.0f * p->g, s, &okg);
  if (!rc) rc = sqrt_verified(s, &oks);
  if (rc) return rc;
  p->divc_ok = (okx ? 1 : 0) | (oky ? 2 : 0) | (okg ? 4 : 0) | (oks ? 8 : 0);
  return 0;
}

int gpo_cdklm16_substep(const gpo_cdklm16_params* p, float dt, int32_t rk_step, const float* in_eta,
                        const float* in_hu, const float* in_hv, const float* qn_eta, const float* qn_hu,
                        const float* qn_hv, float* out_eta, float* out_hu, float* out_hv, int32_t pitch,
                        const float* Hi, int32_t Hi_pitch, const float* f_cells, const gpo_forcing* forcing,
                        uint32_t boundary_conditions, int32_t exact, void* stream) {
  StepArgs a;
  const int rc = make_step_args(p, dt, rk_step, in_eta, in_hu, in_hv, qn_eta, qn_hu, qn_hv, out_eta, out_hu,
                                out_hv, pitch, Hi, Hi_pitch, f_cells, forcing, boundary_conditions, &a);
  if (rc) return rc;
  return check(launch_step(p, a, exact, (hipStream_t)stream), "gpo_cdklm16_substep launch");
}

int gpo_cdklm16_rk2_steps(const gpo_cdklm16_params* p, int32_t n_steps, float dt, float* h0_eta, float* h0_hu,
                          float* h0_hv, float* h1_eta, float* h1_hu, float* h1_hv, int32_t pitch, const float* Hi,
                          int32_t Hi_pitch, const float* f_cells, uint32_t boundary_conditions, int32_t exact,
                          void* stream) {
  if (!p || p->rk_order != 2) return bad("gpo_cdklm16_rk2_steps: rk_order must be 2");
  if (n_steps < 0) return bad("gpo_cdklm16_rk2_steps: negative step count");
  for (int side = 0; side < 4; ++side) {
    const int code = (boundary_conditions >> (8 * side)) & 0xFF;
    if (code != 1 && code != 2) return bad("gpo_cdklm16_rk2_steps: only wall (1) and periodic (2) boundaries");
  }
  StepArgs a0, a1;
  int rc = make_step_args(p, dt, 0, h0_eta, h0_hu, h0_hv, nullptr, nullptr, nullptr, h1_eta, h1_hu, h1_hv, pitch,
                          Hi, Hi_pitch, f_cells, nullptr, boundary_conditions, &a0);
  if (rc) return rc;
  rc = make_step_args(p, dt, 1, h1_eta, h1_hu, h1_hv, h0_eta, h0_hu, h0_hv, h0_eta, h0_hu, h0_hv, pitch, Hi,
                      Hi_pitch, f_cells, nullptr, boundary_conditions, &a1);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  for (int i = 0; i < n_steps; ++i) {
    hipError_t e = launch_step(p, a0, exact, s);
    if (e == hipSuccess) e = launch_step(p, a1, exact, s);
    if (e != hipSuccess) return check(e, "gpo_cdklm16_rk2_steps launch");
  }
  return 0;
}

int gpo_boundary_conditions(int32_t nx, int32_t ny, uint32_t bc, const int32_t sponge[4], float* eta, float* hu,
                            float* hv, int32_t pitch, const float* ns_cur, const float* ns_next, int32_t ns_nx,
                            const float* ew_cur, const float* ew_next, int32_t ew_ny, float frs_t, void* stream) {
  if (nx < 1 || ny < 1 || !eta || !hu || !hv) return bad("gpo_boundary_conditions: bad arguments");
  const int bn = (bc >> 24) & 0xFF, bs = (bc >> 16) & 0xFF, be = (bc >> 8) & 0xFF, bw = bc & 0xFF;
  hipStream_t s = (hipStream_t)stream;
  const int sn = sponge ? sponge[0] : 0, se = sponge ? sponge[1] : 0, ss = sponge ? sponge[2] : 0,
            sw = sponge ? sponge[3] : 0;
  const dim3 grid(ceil_div(nx + 4, 16), ceil_div(ny + 4, 16)), block(16, 16);
  if (bn == 2) {
    hipLaunchKernelGGL(periodic_ns, dim3(ceil_div(nx + 4, 256), 4), dim3(256), 0, s, nx, ny, eta, hu, hv, pitch);
  } else {
    if (bn == 3 || bs == 3) {
      if (!ns_cur || !ns_next || ns_nx < 1) return bad("gpo_boundary_conditions: FRS north/south data missing");
      hipLaunchKernelGGL(frs_ns, grid, block, 0, s, bn, bs, nx, ny, sn, ss, eta, hu, hv, pitch, ns_cur, ns_next, ns_nx,
                         frs_t);
    }
    if (bn == 4 || bs == 4)
      hipLaunchKernelGGL(lin_ns, grid, block, 0, s, bn, bs, nx, ny, sn, ss, eta, hu, hv, pitch);
  }
  if (be == 2) {
    hipLaunchKernelGGL(periodic_ew, dim3(ceil_div(ny + 4, 256), 4), dim3(256), 0, s, nx, ny, eta, hu, hv, pitch);
  } else {
    if (be == 3 || bw == 3) {
      if (!ew_cur || !ew_next || ew_ny < 1) return bad("gpo_boundary_conditions: FRS east/west data missing");
      hipLaunchKernelGGL(frs_ew, grid, block, 0, s, be, bw, nx, ny, se, sw, eta, hu, hv, pitch, ew_cur, ew_next, ew_ny,
                         frs_t);
    }
    if (be == 4 || bw == 4)
      hipLaunchKernelGGL(lin_ew, grid, block, 0, s, be, bw, nx, ny, se, sw, eta, hu, hv, pitch);
  }
  return check(hipGetLastError(), "gpo_boundary_conditions launch");
}

int gpo_bathymetry_init(int32_t nx, int32_t ny, uint32_t bc, float land_value, float* Hi, int32_t Hi_pitch, float* Bm,
                        int32_t Bm_pitch, void* stream) {
  if (nx < 1 || ny < 1 || !Hi || !Bm) return bad("gpo_bathymetry_init: bad arguments");
  const int bn = (bc >> 24) & 0xFF, bs = (bc >> 16) & 0xFF, be = (bc >> 8) & 0xFF, bw = bc & 0xFF;
  hipStream_t s = (hipStream_t)stream;
  if (bn == 2 && bs == 2) {
    hipLaunchKernelGGL(hi_ns, dim3(ceil_div(nx + 5, 256)), dim3(256), 0, s, nx, ny, 1, Hi, Hi_pitch);
  } else if (bn == 1 && bs == 1) {
    hipLaunchKernelGGL(hi_ns, dim3(ceil_div(nx + 5, 256)), dim3(256), 0, s, nx, ny, 0, Hi, Hi_pitch);
  }
  // The closed east-west branch is taken on the north/south codes, as in the
  // reference (utils/Common.py:1557).
  if (be == 2 && bw == 2) {
    hipLaunchKernelGGL(hi_ew, dim3(ceil_div(ny + 5, 256)), dim3(256), 0, s, nx, ny, 1, Hi, Hi_pitch);
  } else if (bn == 1 && bs == 1) {
    hipLaunchKernelGGL(hi_ew, dim3(ceil_div(ny + 5, 256)), dim3(256), 0, s, nx, ny, 0, Hi, Hi_pitch);
  }
  hipLaunchKernelGGL(init_bm, dim3(ceil_div(nx + 4, 16), ceil_div(ny + 4, 16)), dim3(16, 16), 0, s, nx, ny,
                     land_value, (const float*)Hi, Hi_pitch, Bm, Bm_pitch);
  return check(hipGetLastError(), "gpo_bathymetry_init launch");
}

int gpo_coriolis_cells(int32_t nx, int32_t ny, const float* f_tex, int32_t f_nx, int32_t f_ny, float* f_cells,
                       int32_t pitch, void* stream) {
  if (nx < 1 || ny < 1 || !f_tex || !f_cells || f_nx < 1 || f_ny < 1) return bad("gpo_coriolis_cells: bad arguments");
  hipLaunchKernelGGL(coriolis_cells, dim3(ceil_div(nx + 4, 16), ceil_div(ny + 4, 16)), dim3(16, 16), 0,
                     (hipStream_t)stream, nx, ny, f_tex, f_nx, f_ny, f_cells, pitch);
  return check(hipGetLastError(), "gpo_coriolis_cells launch");
}

int gpo_max_dt(int32_t nx, int32_t ny, float dx, float dy, float g, float land_value, const float* eta,
               const float* hu, const float* hv, int32_t pitch, const float* Bm, int32_t Bm_pitch, float* dt_out,
               void* stream) {
  if (nx < 1 || ny < 1 || !eta || !hu || !hv || !Bm || !dt_out) return bad("gpo_max_dt: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(fill_scalar, dim3(1), dim3(1), 0, s, dt_out, kMaxDtCap);
  hipLaunchKernelGGL(max_dt_kernel, dim3(ceil_div(nx, 64), ceil_div(ny, 4)), dim3(256), 0, s, nx, ny, dx, dy, g,
                     land_value, eta, hu, hv, pitch, Bm, Bm_pitch, dt_out);
  return check(hipGetLastError(), "gpo_max_dt launch");
}

int gpo_mass_sum(int32_t nx, int32_t ny, const float* eta, int32_t pitch, double* sum_out, void* stream) {
  if (nx < 1 || ny < 1 || !eta || !sum_out || pitch < 4 * (nx + 4)) return bad("gpo_mass_sum: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(fill_double, dim3(1), dim3(1), 0, s, sum_out, 0.0);
  hipLaunchKernelGGL(mass_sum_kernel, dim3(ny), dim3(256), 0, s, nx, eta, pitch, sum_out);
  return check(hipGetLastError(), "gpo_mass_sum launch");
}

}  // extern "C"
