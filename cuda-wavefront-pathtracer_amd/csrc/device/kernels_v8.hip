// device/kernels_v8.hip - the kernels of the compressed BVH8 layout (instantiated here only).
#include "device/kernels.h"

namespace cwp {
namespace dev {
template void pickKernels<8>(int, int, int, Kernels&);
}  // namespace dev
}  // namespace cwp
