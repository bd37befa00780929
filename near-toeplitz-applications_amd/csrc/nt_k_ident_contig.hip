// Line-kernel instantiations: plain solves of contiguous lines (NearToeplitzSolver.solve, neato.py:168-197).
#include "nt_launch.h"

namespace nt {
NT_INSTANTIATE(double, double, CONTIG, P_IDENT, PREC_F64)
NT_INSTANTIATE(double, double, CONTIG, P_IDENT, PREC_F32R)
}  // namespace nt
