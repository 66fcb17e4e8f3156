// Observation operators of the CDKLM16 path: member states read at drifter
// positions or at given cells, for every member of an ensemble in one launch.
// Built with -ffp-contract=off: the FP32 and FP64 arithmetic of each mode
// rounds as the reference's numpy or CUDA expression does.

#include <cmath>
#include <cstdio>

#include "cdklm_common.h"

namespace gpo {
namespace {

inline int ceil_div(int a, int b) { return (a + b - 1) / b; }

__device__ __forceinline__ float at(const float* base, int pitch_bytes, int j, int i) {
  return ((const float*)((const char*)base + (size_t)pitch_bytes * j))[i];
}

// numpy's rule for an index into an axis of n interior cells: -n..n-1, the
// negative ones counting from the end.  q is the (possibly fractional) index
// before flooring; false when it lies outside that range or is NaN.
__device__ __forceinline__ bool numpy_index(double q, int n, int* out) {
  if (!(q >= -(double)n && q < (double)n)) return false;
  int i = (int)floor(q);
  if (i < 0) i += n;
  *out = i;
  return true;
}

// One (member, point) per thread: blockIdx.y is the member.
__global__ void observe_members(const gpo_obs_member* members, int n_points, int mode, int ghosted,
                                const double* observed, float mean_depth, double* out, int* n_bad) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= n_points) return;
  const gpo_obs_member m = members[blockIdx.y];
  const int k = mode == GPO_OBS_STATE ? 3 : 2;
  double* o = out + ((size_t)blockIdx.y * n_points + d) * k;
  int i = 0, j = 0;
  bool ok;
  if (m.cells) {
    ok = numpy_index((double)m.cells[2 * d], m.nx, &i) && numpy_index((double)m.cells[2 * d + 1], m.ny, &j);
    i += 2, j += 2;
  } else if (ghosted) {
    // observeUnderlyingFlow (gpu_kernels/observationKernels.cu:63-65): FP32
    // quotient, reference cell 2, every cell of the ghosted array reachable
    const float qx = floorf((float)m.points[2 * d] / (float)m.dx);
    const float qy = floorf((float)m.points[2 * d + 1] / (float)m.dy);
    ok = qx >= -2.0f && qx < (float)(m.nx + 2) && qy >= -2.0f && qy < (float)(m.ny + 2);
    if (ok) i = (int)qx + 2, j = (int)qy + 2;
  } else {
    // np.floor(x/dx) in FP64 on the interior arrays
    // (ensembles/OceanModelEnsemble.py:190-191, BaseOceanStateEnsemble.py:137-138)
    ok = numpy_index(m.points[2 * d] / m.dx, m.nx, &i) && numpy_index(m.points[2 * d + 1] / m.dy, m.ny, &j);
    i += 2, j += 2;
  }
  if (!ok) {
    for (int c = 0; c < k; ++c) o[c] = __builtin_nan("");
    if (n_bad) atomicAdd(n_bad, 1);
    return;
  }
  const float hu = at(m.hu, m.pitch, j, i), hv = at(m.hv, m.pitch, j, i);
  if (mode == GPO_OBS_FLOW) {
    o[0] = hu, o[1] = hv;
  } else if (mode == GPO_OBS_STATE) {
    // ParticleInfo.add_state_sample_from_sim (utils/ParticleInfo.py:96-101)
    o[0] = at(m.eta, m.pitch, j, i), o[1] = hu, o[2] = hv;
  } else if (mode == GPO_OBS_BUOY) {
    // Observation.add_observation_from_sim (utils/Observation.py:157-166):
    // FP32 numpy scalars, H the cell-centre depth
    const float H = at(m.Bm, m.Bm_pitch, j, i), eta = at(m.eta, m.pitch, j, i);
    const float f = H / (H + eta);
    o[0] = hu * f, o[1] = hv * f;
  } else {
    // innovation (ensembles/BaseOceanStateEnsemble.py:145-153): the eta
    // compensation in FP32 (the mean depth is np.mean of the FP32 depths),
    // the difference in FP64
    double c = 1.0;
    if (mean_depth > 0.0f) c = (double)((mean_depth + at(m.eta, m.pitch, j, i)) / mean_depth);
    o[0] = observed[4 * d + 2] * c - (double)hu;
    o[1] = observed[4 * d + 3] * c - (double)hv;
  }
}

// Mirror of observeUnderlyingFlow with its own arguments; positions outside
// the ghosted array give NaN instead of a read outside it.
__global__ void observe_underlying_flow(int nx, int ny, float dx, float dy, int x0, int y0, const float* hu,
                                        int hu_pitch, const float* hv, int hv_pitch, int n, const float* positions,
                                        int pos_pitch, float* obs, int obs_pitch) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const float* p = (const float*)((const char*)positions + (size_t)pos_pitch * t);
  const float qx = floorf(p[0] / dx), qy = floorf(p[1] / dy);
  float* o = (float*)((char*)obs + (size_t)obs_pitch * t);
  const bool ok = qx >= (float)-x0 && qx < (float)(nx + 4 - x0) && qy >= (float)-y0 && qy < (float)(ny + 4 - y0);
  if (!ok) {
    o[0] = o[1] = __builtin_nanf("");
    return;
  }
  const int i = (int)qx + x0, j = (int)qy + y0;
  o[0] = at(hu, hu_pitch, j, i);
  o[1] = at(hv, hv_pitch, j, i);
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

using namespace gpo;

extern "C" {

int gpo_observe(const gpo_obs_member* members, int32_t n_members, int32_t n_points, int32_t mode,
                const double* observed, float mean_depth, double* out, int32_t* n_bad, void* stream) {
  const int base = mode & ~GPO_OBS_GHOSTED, ghosted = (mode & GPO_OBS_GHOSTED) ? 1 : 0;
  if (!members || n_members < 1 || n_members > 65535 || n_points < 0 || !out)
    return bad("gpo_observe: bad arguments");
  if (base != GPO_OBS_FLOW && base != GPO_OBS_STATE && base != GPO_OBS_BUOY && base != GPO_OBS_INNOVATION)
    return bad("gpo_observe: unknown mode");
  if (base == GPO_OBS_INNOVATION && !observed) return bad("gpo_observe: innovation needs the observations");
  if (n_points == 0) return 0;
  hipLaunchKernelGGL(observe_members, dim3(ceil_div(n_points, 256), n_members), dim3(256), 0, (hipStream_t)stream,
                     members, n_points, base, ghosted, observed, mean_depth, out, n_bad);
  return check(hipGetLastError(), "gpo_observe launch");
}

int gpo_observe_underlying_flow(int32_t nx, int32_t ny, float dx, float dy, int32_t x0ref, int32_t y0ref,
                                const float* eta, int32_t eta_pitch, const float* hu, int32_t hu_pitch,
                                const float* hv, int32_t hv_pitch, float H, int32_t n_drifters,
                                const float* positions, int32_t pos_pitch, float* observation, int32_t obs_pitch,
                                void* stream) {
  (void)eta, (void)eta_pitch, (void)H;  // part of the reference signature, not read there either
  if (nx < 1 || ny < 1 || n_drifters < 0 || !hu || !hv || (n_drifters > 0 && (!positions || !observation)) ||
      x0ref < 0 || x0ref > 4 || y0ref < 0 || y0ref > 4)
    return bad("gpo_observe_underlying_flow: bad arguments");
  if (n_drifters == 0) return 0;
  hipLaunchKernelGGL(observe_underlying_flow, dim3(ceil_div(n_drifters, 64)), dim3(64), 0, (hipStream_t)stream, nx,
                     ny, dx, dy, x0ref, y0ref, hu, hu_pitch, hv, hv_pitch, n_drifters, positions, pos_pitch,
                     observation, obs_pitch);
  return check(hipGetLastError(), "gpo_observe_underlying_flow launch");
}

}  // extern "C"
