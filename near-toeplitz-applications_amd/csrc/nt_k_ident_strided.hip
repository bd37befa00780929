// Line-kernel instantiations: plain solves of strided (y/z) lines, in place.
#include "nt_launch.h"

namespace nt {
NT_INSTANTIATE(double, double, STRIDED, P_IDENT, PREC_F64)
NT_INSTANTIATE(double, double, STRIDED, P_IDENT, PREC_F32R)
}  // namespace nt
