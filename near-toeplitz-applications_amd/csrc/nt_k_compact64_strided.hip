// Line-kernel instantiations: fused compact d/dy, d/dz of FP64 fields (compact/permute.py + rhs + solve + permute).
#include "nt_launch.h"

namespace nt {
NT_INSTANTIATE(double, double, STRIDED, P_COMPACT, PREC_F64)
NT_INSTANTIATE(double, double, STRIDED, P_COMPACT, PREC_F32R)
}  // namespace nt
