// Line-kernel instantiations: fused compact d/dx of FP64 fields (compact/rhs.py + neato.py solve).
#include "nt_launch.h"

namespace nt {
NT_INSTANTIATE(double, double, CONTIG, P_COMPACT, PREC_F64)
NT_INSTANTIATE(double, double, CONTIG, P_COMPACT, PREC_F32R)
}  // namespace nt
