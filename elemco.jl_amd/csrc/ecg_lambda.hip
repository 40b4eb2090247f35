#include "ecg_lambda.h"

namespace ecg {
void lambda_residual_host(Context&, int, Prec, const double*, const double*, const double*, const double*, double*,
                          double*) {
  throw Error(-4, "lambda_residual: not implemented");
}
void lambda_solve(Context&, int, const SolverOpts&, const double*, const double*, double*, double*, int*, int*) {
  throw Error(-4, "lambda_solve: not implemented");
}
void one_rdm_host(Context&, const double*, const double*, const double*, const double*, double*, double*) {
  throw Error(-4, "one_rdm: not implemented");
}
}  // namespace ecg
