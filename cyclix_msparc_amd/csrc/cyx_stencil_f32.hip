// Stencil kernels of one storage type (float): instantiates the launchers
// declared in cyx_stencil_kernel.hpp.
#define CYX_STENCIL_KERNELS
#include "cyx_stencil_kernel.hpp"

namespace cyx {
template void launch_step<float, float>(const StepParams<float, float>&, int, int, size_t, bool, bool, bool, hipStream_t);
template void launch_batch<float>(const BatchEntry<float>*, int, int, int, size_t, bool, bool, bool, int, hipStream_t);
}  // namespace cyx
