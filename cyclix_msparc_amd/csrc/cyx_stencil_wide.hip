// FP64 H-applies over FP32 blocks: the stencil reads and writes complex64
// (or real FP32) and computes in FP64 with the FP64 coefficients, for the
// Rayleigh-Ritz products and residual checks of the mixed-precision path.
#define CYX_STENCIL_KERNELS
#include "cyx_stencil_kernel.hpp"

namespace cyx {
template void launch_step<double, float>(const StepParams<double, float>&, int, int, size_t, bool, bool, bool,
                                         hipStream_t);
template void launch_step<cplx<double>, cplx<float>>(const StepParams<cplx<double>, cplx<float>>&, int, int, size_t,
                                                     bool, bool, bool, hipStream_t);
}  // namespace cyx
