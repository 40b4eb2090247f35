// Perturbative triples (T) and Lambda-(T) for the closed-shell path.
#pragma once
#include "ecg_cc.h"

namespace ecg {
// e = [E(T), E[T]]; host amplitudes (column-major).  Only triplets first ..
// first+count-1 of the reference's loop order (k, j<=k, i<=j, i=j=k skipped) are
// evaluated (count < 0: all); the energy is a sum over triplets, so partial
// results of disjoint ranges add up.
void pert_triples(Context& c, bool lambda, Prec pr, const double* T1h, const double* T2h, const double* U1h,
                  const double* U2h, long first, long count, double* e);
}  // namespace ecg
