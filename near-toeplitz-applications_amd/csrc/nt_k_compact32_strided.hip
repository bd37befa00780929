// Line-kernel instantiations: fused compact d/dy, d/dz of FP32-stored fields, FP64 result.
#include "nt_launch.h"

namespace nt {
NT_INSTANTIATE(float, double, STRIDED, P_COMPACT, PREC_F64)
NT_INSTANTIATE(float, double, STRIDED, P_COMPACT, PREC_F32R)
}  // namespace nt
