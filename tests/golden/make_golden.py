"""Write oracle golden vectors for the GPU parity tests (run on a CPU host).

    python tests/golden/make_golden.py        # -> tests/golden/c1_oracle.json

The GPU box has no time budget for a CPU oracle at o=20, v=180, naux=600, so
the oracle results for that size are computed once here and committed:

* "c1" (the specified system, factor width 0.05, seed 1): CCSD and DCSD
  residuals at the MP2 amplitudes -- full R1, statistics and seeded samples of
  R2.  (With this width CCSD diverges, so no converged energies exist.)
* "c1s" (same sizes and seed, width 0.05*conditioned_scale): converged CCSD and
  DCSD energies, iteration counts, and (T)/[T] at the converged CCSD amplitudes.
"""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import triples  # noqa: E402
from oracle.ccsd import ECInfo, calc_MP2, calc_cc, calc_cc_resid, calc_fock_matrix, calc_HF_energy  # noqa: E402
from oracle.synthetic import conditioned_scale, make_df_system  # noqa: E402

SIZE = (20, 180, 600, 1)


def r2_checks(R2, nsamp=400, seed=12345):
    """Size-independent summaries of R2[a,b,i,j] plus seeded samples."""
    rng = np.random.default_rng(seed)
    idx = [tuple(int(rng.integers(0, s)) for s in R2.shape) for _ in range(nsamp)]
    w = np.cos(np.arange(R2.size, dtype=np.float64) * 0.001).reshape(R2.shape, order="F")
    return dict(sumsq=float(np.sum(R2 * R2)), total=float(np.sum(R2)), weighted=float(np.sum(R2 * w)),
                samples=[[list(i), float(R2[i])] for i in idx])


def main():
    no, nv, nq, seed = SIZE
    t0 = time.time()
    out = dict(config=dict(nocc=no, nvirt=nv, naux=nq, seed=seed))
    s = make_df_system(no, nv, nq, seed)
    EC = ECInfo.from_df(s)
    calc_fock_matrix(EC)
    out["HF"] = calc_HF_energy(EC)
    out["MP2c"] = calc_MP2(EC)["E"]
    T1, T2 = EC.files["T_vo"].copy(), EC.files["T_vvoo"].copy()
    for meth, dc in (("ccsd", False), ("dcsd", True)):
        R1, R2 = calc_cc_resid(EC, T1, T2, dc=dc)
        out[f"resid_mp2_{meth}"] = dict(R1=R1.ravel(order="F").tolist(), R2=r2_checks(R2))
        print("c1", meth, "residual", time.time() - t0, flush=True)
    del EC, s
    sc = conditioned_scale(no, nv, nq)
    cond = dict(scale=sc)
    s = make_df_system(no, nv, nq, seed, sc)
    for meth, dc in (("dcsd", True), ("ccsd", False)):
        EC2 = ECInfo.from_df(s)
        calc_fock_matrix(EC2)
        cond["HF"] = calc_HF_energy(EC2)
        cond["MP2c"] = calc_MP2(EC2)["E"]
        log = []
        Eh, conv, it = calc_cc(EC2, dc=dc, log=log)
        cond[meth.upper() + "c"] = Eh["E"]
        cond[meth + "_iterations"] = it
        cond[meth + "_converged"] = conv
        cond[meth + "_trajectory"] = [[l[0], l[3], l[4], l[1]] for l in log]
        print("c1s", meth, Eh["E"], it, conv, time.time() - t0, flush=True)
        if not dc:
            et, etb = triples.calc_pertT_closed_shell(EC2)
            cond["(T)"] = et
            cond["[T]"] = etb
            print("c1s (T)", et, etb, time.time() - t0, flush=True)
    out["conditioned"] = cond
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c1_oracle.json")
    with open(path, "w") as f:
        json.dump(out, f)
    print("wrote", path, time.time() - t0)


if __name__ == "__main__":
    main()
