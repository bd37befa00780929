// Line-kernel instantiations: fused compact derivative of FP32-stored fields
// with an FP64 result.
#include "nt_launch.h"

namespace nt {
NT_INSTANTIATE(float, double, CONTIG, P_COMPACT, PREC_F64)
NT_INSTANTIATE(float, double, STRIDED, P_COMPACT, PREC_F64)
NT_INSTANTIATE(float, double, CONTIG, P_COMPACT, PREC_F32R)
NT_INSTANTIATE(float, double, STRIDED, P_COMPACT, PREC_F32R)
}  // namespace nt
