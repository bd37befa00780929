// Line-kernel instantiations: fused compact derivative of FP64 fields
// (compact/rhs.py:8-53 + neato.py solve + compact/permute.py, one pass).
#include "nt_launch.h"

namespace nt {
NT_INSTANTIATE(double, double, CONTIG, P_COMPACT, PREC_F64)
NT_INSTANTIATE(double, double, STRIDED, P_COMPACT, PREC_F64)
NT_INSTANTIATE(double, double, CONTIG, P_COMPACT, PREC_F32R)
NT_INSTANTIATE(double, double, STRIDED, P_COMPACT, PREC_F32R)
}  // namespace nt
