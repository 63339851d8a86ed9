// The default single-doublet FP32 step kernels (SM, 4 waves per SIMD, shared
// in-plane staging; compensated and plain storage), compiled apart from the
// rest of the library so that this translation unit alone can take its own
// code-generation options: gftevolver_amd/build.py compiles it with the
// AMDGPU instruction-level-parallelism machine scheduler.  Scheduling only
// reorders instructions (contraction and rounding are fixed in the source),
// so the fields are bitwise those of the library's other kernels.
// Reference: the per-site update of WilsonLoops/Source/StandardModel.cpp:456-514
// and Gradients/Source/StandardModel.cpp, as for every mom_step instantiation.
#define GFT_KERNELS_ONLY
#include "gft_momentum.hip"

namespace gft {
template __global__ void mom_step<float, 1, 4, true, 2, 4>(StepArgs<float>, Coef);
template __global__ void mom_step<float, 1, 4, false, 2, 4>(StepArgs<float>, Coef);
}  // namespace gft
