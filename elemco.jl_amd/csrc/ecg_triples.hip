#include "ecg_triples.h"

namespace ecg {
void pert_triples(Context&, bool, Prec, const double*, const double*, const double*, const double*, double*) {
  throw Error(-4, "pert_triples: not implemented");
}
}  // namespace ecg
