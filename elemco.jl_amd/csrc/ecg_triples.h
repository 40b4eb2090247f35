// Perturbative triples (T) and Lambda-(T) for the closed-shell path.
#pragma once
#include "ecg_cc.h"

namespace ecg {
// e = [E(T), E[T]]; host amplitudes (column-major).
void pert_triples(Context& c, bool lambda, Prec pr, const double* T1h, const double* T2h, const double* U1h,
                  const double* U2h, double* e);
}  // namespace ecg
