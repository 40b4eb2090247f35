// FCIDUMP input of the closed-shell path (host side): the reader of ElemCo.jl's
// src/integrals/dump.jl for IUHF=0 files, and the exact density-fitting
// factorisation that brings (pq|rs) to the factor form the library takes.
#pragma once
#include <map>
#include <string>
#include <vector>

namespace ecg {

struct Fcidump {
  // header: upper-case key -> values as written (quotes stripped), and the
  // element type the reference gives the key (first value: Int, else Float64,
  // else String; read_header, dump.jl:360-405)
  std::map<std::string, std::vector<std::string>> head;
  std::map<std::string, char> head_type;   // 'i', 'f' or 's'
  int norb = 0;
  double int0 = 0.0;
  std::vector<double> h;     // h[p + n q]
  std::vector<double> eri;   // (pq|rs) at p + n q + n^2 r + n^3 s (chemists' notation)
  bool from_npy = false;

  bool has(const std::string& key) const { return head.count(key) > 0; }
  // headvar(fd, key, Int): first value, or `def` when the key is absent
  long head_int(const std::string& key, long def) const;
  double head_float(const std::string& key, double def) const;
  bool simtra() const { return head_int("ST", 0) > 0; }
};

// read_fcidump (dump.jl:326-346): header, then the integrals from the NPY files named
// by NPY2/NPY1 (dump.jl:442-461, relative to the FCIDUMP's directory; missing files
// fall back to the text integrals as in mmap_integrals, dump.jl:731-745) or from the
// integral lines (dump.jl:573-604).  ".gz" files are read through zlib.
Fcidump read_fcidump(const std::string& path);

// freeze_orbs_in_dump (src/integrals/dumptools.jl:20-62), closed-shell branch: the
// orbitals of `orbs` (1-based) that lie in the occupied space (the lowest NELEC/2) are
// folded into the integrals as a frozen core, those in the virtual space are dropped.
// With c over the frozen occupied orbitals and core_fock[p,q] = sum_c 2 (pq|cc) - (pc|cq)
// (gen_fock over all occupied minus gen_fock over the remaining ones): int0 += 2 sum_c
// h[c,c] + sum_c core_fock[c,c], h <- (h + core_fock)[kept, kept], (pq|rs) <- its kept
// block; NORB, NELEC lose the frozen orbitals and ORBSYM becomes all 1 (modify_header!,
// dump.jl:182-196).  As in the reference, numbers outside 1..NORB and repeats match no
// orbital and are ignored.  Returns the numbers of frozen occupied and virtual orbitals.
struct FreezeCount {
  int ncore = 0, nvirt = 0;
};
FreezeCount freeze_orbs_in_dump(Fcidump& fd, const std::vector<int>& orbs);

// write_fcidump (dump.jl:752-945) for IUHF = 0: "&FCI", the keys NORB, NELEC, MS2, ISYM,
// ORBSYM first, then the other integer, float and string keys (strings in quotes; IUHF,
// ST and III left out when 0), "/"; then "%23.15e %3i %3i %3i %3i" lines: the two-electron
// integrals (pq|rs) with |value| > tol in the reference's loop order (write_integrals2_normal:
// p >= q, p >= r >= s, pair (rs) <= pair (pq); write_integrals2_simtra for ST > 0: all (p,q)
// and (r,s) with r < p, or r = p and s <= q), the one-electron integrals (q <= p, or all
// (p,q) for ST > 0), and the core energy with indices 0 0 0 0.  The reference iterates each
// typed group of the header in its Dict order; here each group is written in key order.
void write_fcidump(const Fcidump& fd, const std::string& path, double tol);

// Exact factorisation (pq|rs) = sum_Q BL[p,q,Q] BR[r,s,Q] of the (pq),(rs)
// supermatrix (symmetric by particle symmetry) by diagonally pivoted LDL^T:
// BL = L sqrt|D|, BR = L sign(D) sqrt|D|.  Pivots stop once the largest remaining
// diagonal is at most rtol times the largest initial one.  Hermitian integrals
// (hermitian = true) pivot on the largest positive diagonal (a Cholesky
// factorisation, BR = BL); a remaining diagonal below -1e-9 of the largest is an
// error.  Non-Hermitian ones pivot on the largest |diagonal|; a remainder above
// max(100 rtol * the largest diagonal, 1e-10 * the largest |(pq|rs)|) is an error (it
// would need 2x2 pivots).
struct DfFactors {
  int naux = 0;
  std::vector<double> BL, BR;   // n x n x naux; BR empty when BR = BL
};
DfFactors df_factorise(const std::vector<double>& eri, int n, bool hermitian, double rtol);

}  // namespace ecg
