// The reverse stages of the differentiable substep as kernels, built twice:
// adjoint.hip with IEEE arithmetic (the FP64 verification build) and
// adjoint_rev32.hip with the hardware reciprocal and square root and
// contraction within expressions (the product's FP32 reverse, whose result is
// a gradient: its accuracy is set by the FP32 trajectory, not by the rounding
// of the reverse arithmetic).  The kernels of each build live in their own
// namespace (exact, fast32) so the two code objects never share a name.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "adjoint_core.h"

namespace gpo {
namespace adj {

// A1 works on blocks of 64 x kSlopeRows cells.
constexpr int kSlopeRows = 8;

// Batched reverse substep: block row z of a launch is member z of a table.
// One member's reverse substep: its geometry, the stage input and output, the
// sensitivities in and out, and its workspace.
template <class T>
struct AdjItem {
  Geo g;
  T dt;
  int rk, accumulate, lq_mode;
  const GPO_GLOBAL T *ie, *iu, *iv, *out_e, *le, *lu, *lv;
  GPO_GLOBAL T *lie, *liu, *liv, *lqe, *lqu, *lqv;
  const GPO_GLOBAL T *Hi, *fc;
  Fwd<T> w;
  Rev<T> rv;
  int frs;  // flow relaxation: lambda in is scaled by the sponge transpose first
  int sn, se, ss, sw;
};


namespace exact {
void launch_stages(const Geo& g, double dt, int rk, const double* ie, const double* iu, const double* iv,
                   const double* out_e, const double* le, const double* lu, const double* lv, double* lie, double* liu,
                   double* liv, double* lqe, double* lqu, double* lqv, int accumulate, int lq_mode, const double* Hi,
                   const double* fc, void* ws, int* flag, hipStream_t s);
}  // namespace exact

namespace fast32 {
// One member's reverse substep (F1, A1, A2, A3 over its grid).
void launch_stages(const Geo& g, float dt, int rk, const float* ie, const float* iu, const float* iv,
                   const float* out_e, const float* le, const float* lu, const float* lv, float* lie, float* liu,
                   float* liv, float* lqe, float* lqu, float* lqv, int accumulate, int lq_mode, const float* Hi,
                   const float* fc, void* ws, int* flag, hipStream_t s);
// The reverse substep of every member of a table (max_w x max_h the largest
// ghost-inclusive grid; any_fc: some member has a coriolis array).
void launch_batch_stage(const AdjItem<float>* tab, int max_w, int max_h, int n, bool any_fc, int* flag,
                        hipStream_t s);
// The same reverse substep as one strip march per work item (adjoint_march.inc)
// followed by the fold of ghost-cell sensitivities (max_fold: the most cells
// next to a side of one member); bit-identical to launch_batch_stage.
void launch_batch_march(const AdjItem<float>* tab, const int4* work, int n_work, int n, int max_fold, bool any_fc,
                        hipStream_t s);
// Output columns of one march wavefront (kMarchCols of adjoint_march.inc).
constexpr int kMarchColsHost = 58;
// One member's reverse substep as the march (the item's lambda_in must not
// alias its lambda_out or lambda_qn), chunks of `rows` rows.
void launch_march(const AdjItem<float>& it, int rows, hipStream_t s);
// Work items of the march: members of sizes nx[m] x ny[m], `rows` rows each.
void march_work_items(const int* nx, const int* ny, int n, int rows, std::vector<int4>& out);
int march_fold_cells(int nx, int ny);
}  // namespace fast32

}  // namespace adj
}  // namespace gpo
