// device/kernels_v4.hip - the kernels of the BVH4 layout (instantiated here only).
#include "device/kernels.h"

namespace cwp {
namespace dev {
template void pickKernels<4>(int, int, int, Kernels&);
}  // namespace dev
}  // namespace cwp
