// Line-kernel instantiations: y-implicit Peaceman-Rachford half steps (no adi/transpose.py pass).
#include "nt_launch.h"

namespace nt {
NT_INSTANTIATE(double, double, STRIDED, P_ADI, PREC_F64)
NT_INSTANTIATE(double, double, STRIDED, P_ADI, PREC_F32R)
}  // namespace nt
