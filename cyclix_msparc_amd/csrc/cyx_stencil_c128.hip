// Stencil kernels of one storage type (cplx<double>): instantiates the launchers
// declared in cyx_stencil_kernel.hpp.
#define CYX_STENCIL_KERNELS
#include "cyx_stencil_kernel.hpp"

namespace cyx {
template void launch_step<cplx<double>, cplx<double>>(const StepParams<cplx<double>, cplx<double>>&, int, int, size_t, bool, bool, bool, hipStream_t);
template void launch_batch<cplx<double>>(const BatchEntry<cplx<double>>*, int, int, int, size_t, bool, bool, bool, int, hipStream_t);
}  // namespace cyx
