// Stencil kernels of one storage type (double): instantiates the launchers
// declared in cyx_stencil_kernel.hpp.
#define CYX_STENCIL_KERNELS
#include "cyx_stencil_kernel.hpp"

namespace cyx {
template void launch_step<double, double>(const StepParams<double, double>&, int, int, size_t, bool, bool, bool, hipStream_t);
template void launch_batch<double>(const BatchEntry<double>*, int, int, int, size_t, bool, bool, bool, int, hipStream_t);
}  // namespace cyx
