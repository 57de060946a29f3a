// Stencil kernels of one storage type (cplx<float>): instantiates the launchers
// declared in cyx_stencil_kernel.hpp.
#define CYX_STENCIL_KERNELS
#include "cyx_stencil_kernel.hpp"

namespace cyx {
template void launch_step<cplx<float>, cplx<float>>(const StepParams<cplx<float>, cplx<float>>&, int, int, size_t, bool, bool, bool, hipStream_t);
template void launch_batch<cplx<float>>(const BatchEntry<cplx<float>>*, int, int, int, size_t, bool, bool, bool, int, hipStream_t);
}  // namespace cyx
