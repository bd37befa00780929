// Line-kernel instantiations: x-implicit Peaceman-Rachford half steps (adi/rhs.py stencil + neato.py solve).
#include "nt_launch.h"

namespace nt {
NT_INSTANTIATE(double, double, CONTIG, P_ADI, PREC_F64)
NT_INSTANTIATE(double, double, CONTIG, P_ADI, PREC_F32R)
}  // namespace nt
