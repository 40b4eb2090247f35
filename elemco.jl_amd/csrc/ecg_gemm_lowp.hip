#include "ecg_gemm.h"

namespace ecg {
void gemm_lowprec(GemmEngine&, const GemmOp&, Prec) { throw Error(-4, "low-precision GEMM: not implemented"); }
}  // namespace ecg
