// The product's FP32 reverse stages (adjoint_stages.inc, adjoint_march.inc)
// built with fast arithmetic: GPO_ADJ_FAST_F32 makes the stages' data-dependent
// divisions and square roots the hardware reciprocal and square root (written
// as builtins, so the same instructions wherever a function is inlined), and
// -ffp-contract=on fuses the multiply and add of one source expression (a
// choice made per expression, so the staged and the fused march kernels,
// which evaluate the same expressions, agree bit for bit).  The forward
// arithmetic the reverse recomputes need not match the step kernel's
// rounding: the gradient's accuracy is set by the FP32 trajectory.

#define GPO_ADJ_FAST_F32 1

#include <hip/hip_runtime.h>

#include "adjoint_core.h"
#include "adjoint_stages.h"

namespace gpo {
namespace adj {
namespace fast32 {

#include "adjoint_stages.inc"
#include "adjoint_march.inc"

void launch_stages(const Geo& g, float dt, int rk, const float* ie, const float* iu, const float* iv,
                   const float* out_e, const float* le, const float* lu, const float* lv, float* lie, float* liu,
                   float* liv, float* lqe, float* lqu, float* lqv, int accumulate, int lq_mode, const float* Hi,
                   const float* fc, void* ws, int* flag, hipStream_t s) {
  stages_t<float>(g, dt, rk, ie, iu, iv, out_e, le, lu, lv, lie, liu, liv, lqe, lqu, lqv, accumulate, lq_mode, Hi, fc,
                  ws, flag, s);
}

void launch_batch_stage(const AdjItem<float>* tab, int max_w, int max_h, int n, bool any_fc, int* flag,
                        hipStream_t s) {
  batch_stage_t<float>(tab, max_w, max_h, n, any_fc, flag, s);
}

void launch_batch_march(const AdjItem<float>* tab, const int4* work, int n_work, int n, int max_fold, bool any_fc,
                        hipStream_t s) {
  if (n_work > 0) {
    const dim3 grid((n_work + 3) / 4);
    if (any_fc)
      hipLaunchKernelGGL(march_rev_b<true>, grid, dim3(256), 0, s, tab, work, n_work);
    else
      hipLaunchKernelGGL(march_rev_b<false>, grid, dim3(256), 0, s, tab, work, n_work);
  }
  if (max_fold > 0) hipLaunchKernelGGL(fold_fix_b, dim3((max_fold + 255) / 256, n), dim3(256), 0, s, tab);
}

void launch_march(const AdjItem<float>& it, int rows, hipStream_t s) {
  const int n_strips = (it.g.nx + 4 + kMarchCols - 1) / kMarchCols;
  const int n_items = n_strips * ((it.g.ny + 4 + rows - 1) / rows);
  const dim3 grid((n_items + 3) / 4);
  if (it.fc)
    hipLaunchKernelGGL(march_rev_1<true>, grid, dim3(256), 0, s, it, n_strips, n_items, rows);
  else
    hipLaunchKernelGGL(march_rev_1<false>, grid, dim3(256), 0, s, it, n_strips, n_items, rows);
  const int nf = fold_cells(it.g.nx, it.g.ny);
  if (nf > 0) hipLaunchKernelGGL(fold_fix_1, dim3((nf + 255) / 256), dim3(256), 0, s, it);
}

void march_work_items(const int* nx, const int* ny, int n, int rows, std::vector<int4>& out) {
  march_work(nx, ny, n, rows, out);
}
int march_fold_cells(int nx, int ny) { return fold_cells(nx, ny); }

}  // namespace fast32
}  // namespace adj
}  // namespace gpo
