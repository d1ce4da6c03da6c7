// device/kernels_v2.hip - the kernels of the BVH2 layout (instantiated here only).
#include "device/kernels.h"

namespace cwp {
namespace dev {
template void pickKernels<2>(int, int, int, Kernels&);
}  // namespace dev
}  // namespace cwp
