"""(T) at the full o=40, v=360, naux=1200 size, seed 4: FP64 against the oracle on
sampled triplets, and complete FP32 and BF16 hi/lo (T) against the complete FP64 (T).

The amplitudes are the converged CCSD amplitudes of the conditioned system of
these sizes (the literal 0.05 factor width diverges; conditioned_scale keeps
the shapes), from a defect-corrected BF16x3 solve that ends at the FP64 fixed
point (thr 1e-10).

Bars:
* FP64 (T) over triplet ranges at the start and the end of the reference's loop
  order (cc_triples.jl:61-71) equals the oracle's calc_pertT_closed_shell
  restricted to the same triplets (E(T) and E[T]) to 1e-10 relative or 1e-13 Eh
  absolute, whichever is larger: five triplets give a partial sum of only ~7e-6 Eh
  over 5 v^3 = 2.3e8 terms of both signs, and FP64 rounding in another summation
  order (MFMA tiles against BLAS) differs by ~1e-15 Eh (measured 1.3e-15); a
  dropped or mis-indexed term changes the sum by orders of magnitude more;
* complete FP32 and BF16 hi/lo (T): |dE(T)| and |dE[T]| <= 1e-6 Eh against the
  complete FP64 (T) (the target of the mixed path), and the energies differ from
  the FP64 ones (the low-precision products ran).
"""
import numpy as np
import pytest

from gpu_helpers import requires_gpu

pytestmark = [pytest.mark.gpu, requires_gpu]

import elemco_amd as E  # noqa: E402
from elemco_amd.flops import triplet_count  # noqa: E402
from elemco_amd.synthetic import conditioned_scale  # noqa: E402
from oracle import ccsd as O  # noqa: E402
from oracle import triples as OT  # noqa: E402
from oracle.synthetic import make_df_system  # noqa: E402

CFG = (40, 360, 1200, 4)


@pytest.fixture(scope="module")
def c4a():
    o, v, q, seed = CFG
    print(f"\nc4a system o={o} v={v} naux={q} seed={seed} (conditioned)", flush=True)
    EC = E.ECInfo.synthetic(o, v, q, seed, scale=conditioned_scale(o, v, q))
    e, T1, T2, it, conv = EC.ctx.cc_solve(E.ECG_CCSD, E.ECG_BF16X3, E.default_opts(maxit=60))
    print(f"  CCSD (BF16x3, defect-corrected): E={float(e[0]):.12f} it={it} conv={conv}", flush=True)
    assert conv
    out = {"EC": EC, "T1": T1, "T2": T2}
    yield out
    EC.ctx.close()


def test_c4a_fp64_triples_match_oracle_on_sampled_triplets(c4a):
    o, v, q, seed = CFG
    EC, T1, T2 = c4a["EC"], c4a["T1"], c4a["T2"]
    ntrip = triplet_count(o)
    ranges = [(0, 3), (ntrip - 2, 2)]
    ECo = O.ECInfo.from_df(make_df_system(o, v, q, seed, conditioned_scale(o, v, q)))
    O.calc_fock_matrix(ECo)
    trip = list(OT.triplets(o))
    assert len(trip) == ntrip
    et = etb = 0.0
    subset = []
    for first, count in ranges:     # E(T) and E[T] are sums over triplets: ranges add up
        a, b = EC.ctx.pert_triples_range(T1, T2, first, count)
        et, etb = et + a, etb + b
        subset += trip[first:first + count]
    eto, etbo = OT.calc_pertT_closed_shell(ECo, T1, T2, triplet_subset=subset)
    print(f"  {len(subset)} triplets: E(T) {et:.15e} oracle {eto:.15e}; E[T] {etb:.15e} oracle {etbo:.15e}",
          flush=True)
    assert abs(etb - etbo) < max(1e-10 * abs(etbo), 1e-13)
    assert abs(et - eto) < max(1e-10 * abs(eto), 1e-13)


def test_c4a_complete_mixed_triples_match_fp64(c4a):
    EC, T1, T2 = c4a["EC"], c4a["T1"], c4a["T2"]
    res = {}
    for prec in (E.ECG_FP64, E.ECG_FP32, E.ECG_BF16X3):
        res[prec] = EC.ctx.pert_triples(T1, T2, precision=prec)
        print(f"  (T) precision {prec}: E(T) {res[prec][0]:.12f} E[T] {res[prec][1]:.12f}", flush=True)
    et64, etb64 = res[E.ECG_FP64]
    assert np.isfinite(et64) and et64 < 0.0
    for prec in (E.ECG_FP32, E.ECG_BF16X3):
        et, etb = res[prec]
        assert abs(et - et64) <= 1e-6 and abs(etb - etb64) <= 1e-6, (prec, et, et64)
        assert et != et64
