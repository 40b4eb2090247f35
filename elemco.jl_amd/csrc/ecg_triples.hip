// Closed-shell (T) and Lambda-(T) on the device.
//
// Restates calc_pertT_closed_shell (src/cc/cc_triples.jl:39-148) and
// calc_LambdapertT_closed_shell (:159-288).  For each triplet i<=j<=k (i=j=k
// skipped; weight 2 if i<j<k, else 1):
//   K[a,b,c]  = 6 particle terms  sum_d T2 . <..|dk>   (v x v^2 x v products)
//             - 6 hole terms      sum_l T2 . <lc|jk>   (v^2 x v x o products)
//   X = (4K - 2K_acb - 2K_cba - 2K_bac + K_cab + K_bca) / (e_i+e_j+e_k-e_a-e_b-e_c)
//   E[T] += w sum K.X   (Lambda: sum KU.X with KU built from U2 and the transposed integrals)
//   IntX[:,i] += w X.<jk|bc>,  IntY[:,i] += w X.T2[b,c,j,k]  (and the j, k analogues)
// E(T) = E[T] + sum T1.IntX + sum f_ov.IntY   (Lambda: 1/2 U1.IntX + f_ov.IntY).
// The twelve products fall into three classes by the index of K that the
// amplitude factor carries (a, b or c).  Each class is one segmented MFMA GEMM
// launch (four K segments: two particle terms of length v, two hole terms of
// length o) whose output orientation lets every operand stream contiguously;
// a transposed copy of the particle integrals and a negated transposed copy of
// the hole integrals make that possible.  K is therefore written three times
// per triplet instead of twelve.  The combine step stages, for each sorted
// triple of 8-wide virtual tiles, the six 8x8x8 boxes of K at the permutations
// of that triple in LDS (K is read once), evaluates every output block that is
// a permutation of the triple, and writes per-block partial sums that a second
// kernel reduces in a fixed order (deterministic, no atomics).
#include "ecg_triples.h"

#include <algorithm>
#include <cmath>

namespace ecg {

namespace {

constexpr int TT = 8;                       // combine tile edge
constexpr int NPART = 6 * TT + 1;           // per-block partial record: 6 vectors x 8 + energy

template <typename KT>
struct CombArgs {
  int nv, no, i, j, k, ntile;
  const KT* K[3];       // v^3, [a,b,c]: K is the sum of the three (one per product class)
  const KT* KU[3];      // Lambda: the second intermediate, likewise; else null
  const double* eo;     // occupied orbital energies
  const double* ev;     // virtual orbital energies
  const double* W;      // <ij|ab> as [a,b,i,j]
  const double* Y;      // T2 (or covariant U2) [a,b,i,j] for IntY
  const int* tri;       // sorted tile triples (A <= B <= C), 3 ints each
  long kz, pz;          // batch member blockIdx.y: triplet (i + y, j, k), K/KU at + y kz, records at + y pz
  double* part;         // [ntile^3][NPART], record of output block at + nt*(bt + nt*ct)
  // per tie pattern (A==B) + 2 (B==C): the distinct output blocks, each as its
  // permutation index and the box slot of each of the six permutations of K
  int nblk[4];
  signed char blk[4][6][7];
};

// padded LDS strides of a box: (9, 76) keeps every access pattern of the combine
// (six permuted box reads, X stores, the dot products) at most 3-way bank-conflicted
constexpr int BX1 = TT + 1, BX2 = 76;

// One workgroup per sorted tile triple (A,B,C).  The six boxes K[perm(A,B,C)]
// (8x8x8; K summed over its three parts) are staged in LDS once and serve every
// distinct output block (at,bt,ct) that is a permutation of (A,B,C): all six
// index permutations of K that the combine of such a block reads lie in those
// boxes, so K is read once per triplet.  The 8x8 slices of W and Y that the
// output blocks multiply X with (36: two tensors, three occupied pairs, six
// ordered pairs of the three tiles) and the tiles' orbital energies are staged
// in the same pass.  Per output block, X is formed (two elements per thread)
// into LDS, and the 48 vector entries of the block (IntX and IntY over a, b
// and c) are 64-term dot products of X with a slice, four lanes per entry and
// a two-step shuffle.  The energy sum K.X is accumulated per thread over the
// whole triple and reduced once; it is recorded with the first output block.
// All sums run in a fixed order (deterministic).  X and the staged W / Y slices are held
// in the type of K (FP32 for the FP32 and BF16 hi/lo paths: 26 instead of 36 KB of LDS, and
// with at most 80 VGPRs six workgroups per CU instead of four: 0.276 against 0.354 ms per
// triplet at o=40 v=360); the sums with W and Y are FP64.
template <typename KT>
__global__ __launch_bounds__(256, 6) void triples_combine_kernel(CombArgs<KT> p) {
  __shared__ KT box[6][TT * BX2];
  __shared__ KT xs[TT * BX2];                // X of the current output block
  __shared__ KT wsl[2][3][6][TT * TT];       // [W|Y][occupied pair][ordered tile pair][x + 8 y]
  __shared__ double esl[3][TT];              // virtual orbital energies of tiles T[0], T[1], T[2]
  __shared__ double ered[4];
  const int nt = p.ntile;
  const long v = p.nv;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int la = lane & 7, lb = lane >> 3;   // element e = la + 8 lb + 64 lc, lc = w and w + 4
  const int T[3] = {p.tri[3 * blockIdx.x], p.tri[3 * blockIdx.x + 1], p.tri[3 * blockIdx.x + 2]};
  const long zb = blockIdx.y;
  const int pi = p.i + (int)zb;
  const KT* K0 = p.K[0] + zb * p.kz;
  const KT* K1 = p.K[1] + zb * p.kz;
  const KT* K2 = p.K[2] + zb * p.kz;
  double* const part = p.part + zb * p.pz;
  // permutation table: which of the three coordinates sits at K index 1,2,3
  const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {2, 1, 0}, {1, 0, 2}, {2, 0, 1}, {1, 2, 0}};
  // ordered pair (p, q), p != q, of tile positions -> 0..5
  auto pair = [](int a, int b) { return 2 * a + (b > a ? b - 1 : b); };
  // occupied pair of each role: 0 -> (j,k), 1 -> (i,k), 2 -> (i,j)
  const long ooff[3] = {v * v * (p.j + (long)p.no * p.k), v * v * (pi + (long)p.no * p.k),
                        v * v * (pi + (long)p.no * p.j)};
  // ---- staging (all global loads of the workgroup)
#pragma unroll
  for (int s = 0; s < 6; ++s)
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int e = t + 256 * u;
      const int x1 = T[perm[s][0]] * TT + e % TT, x2 = T[perm[s][1]] * TT + (e / TT) % TT,
                x3 = T[perm[s][2]] * TT + e / (TT * TT);
      const bool ok = x1 < v && x2 < v && x3 < v;
      const long off = ((long)T[perm[s][0]] + nt * ((long)T[perm[s][1]] + nt * (long)T[perm[s][2]])) * (TT * TT * TT) + e;
      box[s][e % TT + BX1 * ((e / TT) % TT) + BX2 * (e / (TT * TT))] =
          ok ? K0[off] + K1[off] + K2[off] : KT(0);
    }
#pragma unroll
  for (int q = 0; q < 9; ++q) {
    const int u = t + 256 * q;
    const int x = u % TT, y = (u / TT) % TT, pr = (u / (TT * TT)) % 6, role = (u / (6 * TT * TT)) % 3;
    const int wy = u / (18 * TT * TT);
    const int pa = pr / 2, pb = (pr % 2 < pa) ? pr % 2 : pr % 2 + 1;   // inverse of pair()
    const int c1 = T[pa] * TT + x, c2 = T[pb] * TT + y;
    const double* src = wy == 0 ? p.W : p.Y;
    const bool ok = c1 < v && c2 < v;
    (&wsl[0][0][0][0])[u] = ok ? (KT)src[ok ? c1 + v * c2 + ooff[role] : 0] : KT(0);
  }
  if (t < 3 * TT) {
    const int x = T[t / TT] * TT + t % TT;
    esl[t / TT][t % TT] = x < v ? p.ev[x < v ? x : 0] : 0.0;
  }
  const double eijk = p.eo[pi] + p.eo[p.j] + p.eo[p.k];
  __syncthreads();

  // dot-product assignment of the 192 first threads: entry q = t/4 (vector q/8, index q%8), part t%4
  const int dq = t >> 2, dpart = t & 3, dvec = dq >> 3, didx = dq & 7;
  const int drole = dvec % 3, dwy = dvec / 3;
  double esum = 0.0;
  const int pat = (T[0] == T[1] ? 1 : 0) + (T[1] == T[2] ? 2 : 0);
  const int nblk = p.nblk[pat];
  for (int nb = 0; nb < nblk; ++nb) {
    // the nb-th distinct output block: permutation o of the triple, and the box holding
    // permutation q of its coordinates
    const int o = p.blk[pat][nb][0];
    int slot[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) slot[q] = p.blk[pat][nb][1 + q];
    const int tg[3] = {T[perm[o][0]], T[perm[o][1]], T[perm[o][2]]};
    const double* ea = esl[perm[o][0]];
    const double* eb = esl[perm[o][1]];
    const double* ec = esl[perm[o][2]];
    // X of the block (zero outside the virtual range)
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int lc = w + 4 * h;
      const int a = tg[0] * TT + la, b = tg[1] * TT + lb, c = tg[2] * TT + lc;
      KT xv = KT(0);
      if (a < v && b < v && c < v) {
        const int lo[3] = {la, lb, lc};
        KT kv[6];
#pragma unroll
        for (int q = 0; q < 6; ++q) kv[q] = box[slot[q]][lo[perm[q][0]] + BX1 * lo[perm[q][1]] + BX2 * lo[perm[q][2]]];
        const KT num = KT(4) * kv[0] - KT(2) * (kv[1] + kv[2] + kv[3]) + kv[4] + kv[5];
        xv = num / KT(eijk - ea[la] - eb[lb] - ec[lc]);
        double kk = (double)kv[0];
        if (p.KU[0]) {
          const long off = ((long)tg[0] + nt * ((long)tg[1] + nt * (long)tg[2])) * (TT * TT * TT) + la + TT * lb +
                           TT * TT * lc;
          kk = (double)p.KU[0][off + zb * p.kz] + (double)p.KU[1][off + zb * p.kz] + (double)p.KU[2][off + zb * p.kz];
        }
        esum += kk * (double)xv;
      }
      xs[la + BX1 * lb + BX2 * lc] = xv;
    }
    __syncthreads();   // X staged
    double* out = part + ((long)tg[0] + (long)nt * (tg[1] + (long)nt * tg[2])) * NPART;
    if (t < 192) {
      // entry didx of vector dvec: role 0 sums over (b,c), role 1 over (a,c), role 2 over (a,b)
      const int sl = drole == 0 ? pair(perm[o][1], perm[o][2])
                                : (drole == 1 ? pair(perm[o][0], perm[o][2]) : pair(perm[o][0], perm[o][1]));
      const KT* ws = wsl[dwy][drole][sl];
      double sm = 0.0;
#pragma unroll
      for (int u2 = 0; u2 < 2; ++u2)
#pragma unroll
        for (int u1 = 0; u1 < TT; ++u1) {
          const int u = 2 * dpart + u2;   // the slower summed index: this part's two planes
          int xi, wi;
          if (drole == 0) { xi = didx + BX1 * u1 + BX2 * u; wi = u1 + TT * u; }
          else if (drole == 1) { xi = u1 + BX1 * didx + BX2 * u; wi = u1 + TT * u; }
          else { xi = u1 + BX1 * u + BX2 * didx; wi = u1 + TT * u; }
          sm += (double)xs[xi] * (double)ws[wi];
        }
      sm += __shfl_xor(sm, 1, 64);
      sm += __shfl_xor(sm, 2, 64);
      if (dpart == 0) out[dvec * TT + didx] = sm;
    } else if (t == 192 && nb > 0) {
      out[48] = 0.0;   // the triple's energy is recorded with its first output block
    }
    __syncthreads();   // xs is reused by the next output block
  }
  // energy of the triple: lanes, then waves, in a fixed order
  for (int m = 1; m < 64; m <<= 1) esum += __shfl_xor(esum, m, 64);
  if (lane == 0) ered[w] = esum;
  __syncthreads();
  if (t == 0) {
    const int tg0 = T[0], tg1 = T[1], tg2 = T[2];
    part[((long)tg0 + (long)nt * (tg1 + (long)nt * tg2)) * NPART + 48] = (ered[0] + ered[1]) + (ered[2] + ered[3]);
  }
}

// Reduce the per-block partials of the triplets (i0 + z, j, k), z < nb, in a fixed
// order: member by member, and per member as follows.  Workgroups g < 2*ntile:
// vector type = g / ntile (0: IntX, 1: IntY) and virtual tile tx = g % ntile; the
// workgroup adds the i, j and k contributions (roles 0, 1, 2) in turn, so with
// i = j or j = k the two updates of one element are ordered.  Workgroups
// g >= 2*ntile: the energy partials of the c-tile slab ct = g - 2*ntile, added to
// eacc[ct].  Every element has one writer per launch and launches are
// stream-ordered, so the result is deterministic.
constexpr int RED_T = 1024;   // 128 strands per vector element: enough loads in flight for the nt^3 records
__global__ __launch_bounds__(RED_T) void triples_reduce_kernel(int nv, int ntile, int i0, int nb, int j, int k,
                                                               const double* __restrict__ part, long pz,
                                                               double* IntX, double* IntY, double* eacc) {
  __shared__ double red[RED_T];
  const int g = blockIdx.x, t = threadIdx.x;
  const long nt2 = (long)ntile * ntile;
  // the members (i0 + z, j, k) of a batch, in order z = 0, 1, ...
  for (int z = 0; z < nb; ++z) {
    const int i = i0 + z;
    const double w = (i < j && j < k) ? 2.0 : 1.0;
    const double* pm = part + (long)z * pz;
    if (g < 2 * ntile) {
      const int type = g / ntile, tx = g % ntile;
      const int l = t & (TT - 1), s8 = t / TT;      // RED_T / TT strands per vector element
      constexpr int NS = RED_T / TT;
      for (int role = 0; role < 3; ++role) {
        const int q = 3 * type + role;
        double s = 0.0;
        for (long u = s8; u < nt2; u += NS) {
          const int u1 = (int)(u % ntile), u2 = (int)(u / ntile);
          int at, bt, ct;
          if (role == 0) { at = tx; bt = u1; ct = u2; }
          else if (role == 1) { bt = tx; at = u1; ct = u2; }
          else { ct = tx; at = u1; bt = u2; }
          const long blk = at + (long)ntile * (bt + (long)ntile * ct);
          s += pm[blk * NPART + q * TT + l];
        }
        red[t] = s;
        __syncthreads();
        for (int h = NS / 2; h > 0; h >>= 1) {
          if (s8 < h) red[t] += red[t + h * TT];
          __syncthreads();
        }
        const int x = tx * TT + l;
        if (s8 == 0 && x < nv) {
          const int occ = role == 0 ? i : (role == 1 ? j : k);
          double* dst = (type == 0 ? IntX : IntY) + x + (long)nv * occ;
          *dst += w * red[t];
        }
        __syncthreads();
      }
    } else {
      const int ct = g - 2 * ntile;
      const double* src = pm + (long)ct * nt2 * NPART + 48;
      double s = 0.0;
      for (long u = t; u < nt2; u += RED_T) s += src[u * NPART];
      red[t] = s;
      __syncthreads();
      for (int h = RED_T / 2; h > 0; h >>= 1) {
        if (t < h) red[t] += red[t + h];
        __syncthreads();
      }
      if (t == 0) eacc[ct] += w * red[0];
      __syncthreads();
    }
  }
}

Tensor T3d(double* p, long a, long b, long c) { return Tensor(p, {a, b, c}); }
Tensor T4d(double* p, long a, long b, long c, long d) { return Tensor(p, {a, b, c, d}); }

// Integrals of one (T) evaluation: W[x,y,d,m] (v,v,v,o), its transpose
// Wt[x,y,d,m] = W[y,x,d,m], and Hn[c,l,m,n] = -H[l,c,m,n] (v,o,o,o) for the
// hole integrals H[l,c,m,n].
struct TripleInts {
  const double* W;
  const double* Wt;
  const double* Hn;
};

// Packed planes of the (T) operands (BF16 hi/lo, or one FP32 plane for FP32 products),
// packed once per evaluation so that each triplet's three products read them directly
// (no per-triplet packing).  Every family is a stack of blocks with K-contiguous rows:
//   W [m][(x,y)][d] = W[x,y,d,m]     Wt[m][(x,y)][d] = W[y,x,d,m]   (K = v)
//   T3[x][(p,q)][l] = T[p,q,x,l]     T4[x][(p,q)][l] = T[p,q,l,x]   (K = o)
//   Ts[(i,j)][a][d] = T[a,d,i,j]                                     (K = v)
//   Hn[(m,n)][c][l] = -H[l,c,m,n]                                    (K = o)
struct Planes {
  DevBuf* buf = nullptr;     // storage owned by the context (grown on demand)
  bool f32 = false;          // one FP32 plane (f) instead of BF16 hi / lo
  bool f64 = false;          // one FP64 plane (d, pack_f64p_kernel layout)
  uint16_t* hi = nullptr;
  uint16_t* lo = nullptr;
  float* f = nullptr;
  double* d = nullptr;
  long ld = 0;
  int K = 0;                 // data elements of a row (zeros from K to ld)
  // rows of src as in GemmEngine::pack_rows_bf16x3; 256 readable rows are kept after the last
  void make(GemmEngine& g, const double* src, long rows, int d0, long s0, long s1, int K_, long sk) {
    K = K_;
    ld = ((K + 31) / 32) * 32;
    const long rows_alloc = ((rows + 63) / 64) * 64 + 256;
    const long packed = ((rows + 63) / 64) * 64;
    if (f64) {
      buf->ensure(sizeof(double) * rows_alloc * ld);
      d = static_cast<double*>(buf->p);
      ECG_HIP(hipMemsetAsync(d + packed * ld, 0, sizeof(double) * (rows_alloc - packed) * ld, g.stream));
      g.pack_rows_f64(src, rows, d0, s0, s1, K_, sk, (int)ld, d);
      return;
    }
    buf->ensure(sizeof(uint16_t) * 2 * rows_alloc * ld);   // 4 bytes per element either way
    if (f32) {
      f = static_cast<float*>(buf->p);
      ECG_HIP(hipMemsetAsync(f + packed * ld, 0, sizeof(float) * (rows_alloc - packed) * ld, g.stream));
      g.pack_rows_f32(src, rows, d0, s0, s1, K_, sk, (int)ld, f);
      return;
    }
    hi = static_cast<uint16_t*>(buf->p);
    lo = hi + rows_alloc * ld;
    ECG_HIP(hipMemsetAsync(hi + packed * ld, 0, sizeof(uint16_t) * (rows_alloc - packed) * ld, g.stream));
    ECG_HIP(hipMemsetAsync(lo + packed * ld, 0, sizeof(uint16_t) * (rows_alloc - packed) * ld, g.stream));
    g.pack_rows_bf16x3(src, rows, d0, s0, s1, K_, sk, (int)ld, hi, lo);
  }
};
struct PackedInts { Planes W, Wt, Hn; };
struct PackedAmps { Planes T3, T4, Ts; };

// plane kind: 0 BF16 hi/lo, 1 FP32, 2 FP64
void pack_ints(GemmEngine& g, int v, int o, const TripleInts& I, int kind, PackedInts& P) {
  const long vv = (long)v * v, vvv = vv * v;
  P.W.f32 = P.Wt.f32 = P.Hn.f32 = (kind == 1);
  P.W.f64 = P.Wt.f64 = P.Hn.f64 = (kind == 2);
  P.W.make(g, I.W, (long)o * vv, (int)vv, 1, vvv, v, vv);
  P.Wt.make(g, I.Wt, (long)o * vv, (int)vv, 1, vvv, v, vv);
  P.Hn.make(g, I.Hn, (long)o * o * v, v, 1, (long)v * o, o, v);
}

void pack_amps(GemmEngine& g, int v, int o, const double* T, int kind, PackedAmps& P) {
  const long vv = (long)v * v;
  P.T3.f32 = P.T4.f32 = P.Ts.f32 = (kind == 1);
  P.T3.f64 = P.T4.f64 = P.Ts.f64 = (kind == 2);
  P.T3.make(g, T, (long)o * vv, 1 << 30, 1, 0, o, vv * o);
  P.T4.make(g, T, (long)o * vv, (int)vv, 1, vv * o, o, vv);
  P.Ts.make(g, T, (long)o * o * v, v, 1, vv, v, v);
}

// K[a,b,c] (FP32; FP64 from FP64 planes) for the triplets (i0 + z, j, k), z < nb, from
// packed planes (BF16 hi/lo, FP32 or FP64 products): the
// same twelve products as build_K, each class with the pair of K indices carried
// by the integral (or amplitude pair) operand as M (v^2 rows) and the single index
// as N (v), so all three launches have the shape v^2 x v x (2v + 2o), batched over
// z (every operand offset is affine in i).  K is written in 8x8x8 boxes (the
// combine's unit), one contiguous 512-element run per box; member z's parts start
// kz elements after member z-1's.
void build_K_packed(GemmEngine& g, int v, int o, int i0, int nb, int j, int k, const PackedAmps& T,
                    const PackedInts& I, float* const K[3], double* const Kd[3], long kz) {
  const long vv = (long)v * v;
  auto rowsW = [&](const Planes& P, int m) { return m * vv * P.ld; };           // block m of a v^2-row family
  auto rowsS = [&](const Planes& P, int p, int q) { return (p + (long)o * q) * v * P.ld; };   // v-row block
  // operand offsets as functions of i; the batch advances them by their change per unit i
  auto seg = [&](PackedSeg& s, const Planes& A, auto aoff, const Planes& B, auto boff) {
    const long a0 = aoff(i0), b0 = boff(i0);
    if (A.f64) {   // FP64 planes: the plane pointers travel in Ah / Bh (PackedOp::f64)
      s.Ah = reinterpret_cast<const uint16_t*>(A.d + a0); s.Al = nullptr;
      s.Bh = reinterpret_cast<const uint16_t*>(B.d + b0); s.Bl = nullptr;
    } else if (A.f32) {   // FP32 planes: the plane pointers travel in Ah / Bh (PackedOp::f32)
      s.Ah = reinterpret_cast<const uint16_t*>(A.f + a0); s.Al = nullptr;
      s.Bh = reinterpret_cast<const uint16_t*>(B.f + b0); s.Bl = nullptr;
    } else {
      s.Ah = A.hi + a0; s.Al = A.lo + a0;
      s.Bh = B.hi + b0; s.Bl = B.lo + b0;
    }
    s.lda = A.ld; s.a_z = aoff(i0 + 1) - a0;
    s.ldb = B.ld; s.b_z = boff(i0 + 1) - b0;
    s.K = (int)A.ld;
    s.Kt = A.K;
  };
  PackedOp op;
  op.f32 = T.Ts.f32;
  op.f64 = T.Ts.f64;
  auto out = [&](int q) {
    op.C = op.f64 ? Kd[q] : nullptr;
    op.Cf = op.f64 ? nullptr : K[q];
  };
  op.M = (int)vv; op.N = v; op.nseg = 4;
  out(0);
  op.algorithmic_k = 2LL * v + 2LL * o;
  op.batch = nb; op.c_z = kz;
  // class a: m = (b,c), n = a
  seg(op.seg[0], I.W, [&](int) { return rowsW(I.W, k); }, T.Ts, [&](int i) { return rowsS(T.Ts, i, j); });
  seg(op.seg[1], I.Wt, [&](int) { return rowsW(I.Wt, j); }, T.Ts, [&](int i) { return rowsS(T.Ts, i, k); });
  seg(op.seg[2], T.T3, [&](int) { return rowsW(T.T3, j); }, I.Hn, [&](int i) { return rowsS(I.Hn, k, i); });
  seg(op.seg[3], T.T4, [&](int) { return rowsW(T.T4, k); }, I.Hn, [&](int i) { return rowsS(I.Hn, j, i); });
  // terms 1: T^{ij}[a,d] W_k[b,c,d], 3: T^{ik}[a,d] W_j[c,b,d], 11: -H[l,a,k,i] T[b,c,j,l],
  //       12: -H[l,a,j,i] T[b,c,l,k]
  op.box_class = 1; op.box_v = v; op.beta = 0.0;
  g.run_bf16x3_packed(op);
  // class b: m = (a,c), n = b; terms 2: W_k[a,c,d] T^{ji}[b,d], 5: W_i[c,a,d] T^{jk}[b,d],
  //       9: T[a,c,i,l] (-H[l,b,k,j]), 10: T[a,c,l,k] (-H[l,b,i,j])
  seg(op.seg[0], I.W, [&](int) { return rowsW(I.W, k); }, T.Ts, [&](int i) { return rowsS(T.Ts, j, i); });
  seg(op.seg[1], I.Wt, [&](int i) { return rowsW(I.Wt, i); }, T.Ts, [&](int) { return rowsS(T.Ts, j, k); });
  seg(op.seg[2], T.T3, [&](int i) { return rowsW(T.T3, i); }, I.Hn, [&](int) { return rowsS(I.Hn, k, j); });
  seg(op.seg[3], T.T4, [&](int) { return rowsW(T.T4, k); }, I.Hn, [&](int i) { return rowsS(I.Hn, i, j); });
  op.box_class = 2; out(1);
  g.run_bf16x3_packed(op);
  // class c: m = (a,b), n = c; terms 4: W_j[a,b,d] T^{ki}[c,d], 6: W_i[b,a,d] T^{kj}[c,d],
  //       7: T[a,b,i,l] (-H[l,c,j,k]), 8: T[a,b,l,j] (-H[l,c,i,k])
  seg(op.seg[0], I.W, [&](int) { return rowsW(I.W, j); }, T.Ts, [&](int i) { return rowsS(T.Ts, k, i); });
  seg(op.seg[1], I.Wt, [&](int i) { return rowsW(I.Wt, i); }, T.Ts, [&](int) { return rowsS(T.Ts, k, j); });
  seg(op.seg[2], T.T3, [&](int i) { return rowsW(T.T3, i); }, I.Hn, [&](int) { return rowsS(I.Hn, j, k); });
  seg(op.seg[3], T.T4, [&](int) { return rowsW(T.T4, j); }, I.Hn, [&](int i) { return rowsS(I.Hn, i, k); });
  op.box_class = 3; out(2);
  g.run_bf16x3_packed(op);
}

// Transposed and negated-transposed integral copies for build_K.
void make_triple_ints(TensorOps& ops, int v, int o, const double* W, const double* H, DevBuf& Wt, DevBuf& Hn) {
  const int p_yx[4] = {1, 0, 2, 3};
  Wt.ensure(sizeof(double) * (long)v * v * v * o);
  Hn.ensure(sizeof(double) * (long)v * o * o * o);
  Tensor tw = T4d(const_cast<double*>(W), v, v, v, o), twt = T4d(Wt.d(), v, v, v, o);
  ops.permute(tw, p_yx, twt, 1.0, 0.0);
  Tensor th = T4d(const_cast<double*>(H), o, v, o, o), thn = T4d(Hn.d(), v, o, o, o);
  ops.permute(th, p_yx, thn, -1.0, 0.0);
}

}  // namespace

void pert_triples(Context& c, bool lambda, Prec pr, const double* T1h, const double* T2h, const double* U1h,
                  const double* U2h, long first, long count, double* e) {
  const int v = c.nv, o = c.no;
  const long n1 = c.nvo(), n2 = c.nvvoo();
  const long vvv = (long)v * v * v;
  hipStream_t s = c.stream;
  TensorOps& ops = c.ops;
  ops.gemm.stream = s;
  const Blocks& Rb = c.hermitian ? c.L : c.R;
  // the large buffers live in the context between calls (hipMalloc of tens of GB
  // costs far more than a triplet); every size below is explicit
  DevBuf* W = c.tws;
  DevBuf dT1(sizeof(double) * n1);
  DevBuf& dT2 = W[0];
  dT2.ensure(sizeof(double) * n2);
  ECG_HIP(hipMemcpyAsync(dT1.p, T1h, sizeof(double) * n1, hipMemcpyHostToDevice, s));
  ECG_HIP(hipMemcpyAsync(dT2.p, T2h, sizeof(double) * n2, hipMemcpyHostToDevice, s));
  // undressed integrals: vvvo[a,b,c,k] = (ac|bk),  ovoo[l,c,j,k] = (lj|ck)
  DevBuf& vvvo = W[1];
  DevBuf& ovoo = W[2];
  vvvo.ensure(sizeof(double) * vvv * o);
  ovoo.ensure(sizeof(double) * (long)o * v * o * o);
  {
    Tensor tw = T4d(vvvo.d(), v, v, v, o), th = T4d(ovoo.d(), o, v, o, o);
    ops.arena.reset();
    ops.contract_p(Prec::F64, 1.0, T3d(c.L.vv.d(), v, v, c.nq), "acQ", T3d(Rb.vo.d(), v, o, c.nq), "bkQ", 0.0, tw,
                   "abck");
    ops.contract_p(Prec::F64, 1.0, T3d(c.L.oo.d(), o, o, c.nq), "ljQ", T3d(Rb.vo.d(), v, o, c.nq), "ckQ", 0.0, th,
                   "lcjk");
  }
  DevBuf& vvvoT = W[3];
  DevBuf& ovooN = W[4];
  make_triple_ints(ops, v, o, vvvo.d(), ovoo.d(), vvvoT, ovooN);
  const TripleInts IT{vvvo.d(), vvvoT.d(), ovooN.d()};
  // Lambda: covariant U2 = (2U + U[b,a,i,j])/3 (cctools.jl:935-938) and the
  // integrals <ck|ab> as [a,b,c,k] = (ca|kb), <jk|lc> as [l,c,j,k] = (jl|kc)
  DevBuf dU1;
  DevBuf& dU2 = W[5];
  DevBuf& vvvoL = W[6];
  DevBuf& ovooL = W[7];
  DevBuf& vvvoLT = W[8];
  DevBuf& ovooLN = W[9];
  TripleInts IL = IT;
  if (lambda) {
    dU1.alloc(sizeof(double) * n1);
    dU2.ensure(sizeof(double) * n2);
    ECG_HIP(hipMemcpyAsync(dU1.p, U1h, sizeof(double) * n1, hipMemcpyHostToDevice, s));
    DevBuf& raw = W[10];
    raw.ensure(sizeof(double) * n2);
    ECG_HIP(hipMemcpyAsync(raw.p, U2h, sizeof(double) * n2, hipMemcpyHostToDevice, s));
    Tensor tr = T4d(raw.d(), v, v, o, o), tu = T4d(dU2.d(), v, v, o, o);
    const int p_ba[4] = {1, 0, 2, 3};
    ops.axpby(n2, 2.0 / 3.0, raw.d(), 0.0, dU2.d());
    ops.permute(tr, p_ba, tu, 1.0 / 3.0, 1.0);
    if (!c.hermitian) {
      // (ca|kb) = (ac|bk) and (jl|kc) = (lj|ck) hold only for Hermitian real integrals
      vvvoL.ensure(sizeof(double) * vvv * o);
      ovooL.ensure(sizeof(double) * (long)o * v * o * o);
      Tensor tw = T4d(vvvoL.d(), v, v, v, o), th = T4d(ovooL.d(), o, v, o, o);
      ops.arena.reset();
      ops.contract_p(Prec::F64, 1.0, T3d(c.L.vv.d(), v, v, c.nq), "caQ", T3d(Rb.ov.d(), o, v, c.nq), "kbQ", 0.0, tw,
                     "abck");
      ops.contract_p(Prec::F64, 1.0, T3d(c.L.oo.d(), o, o, c.nq), "jlQ", T3d(Rb.ov.d(), o, v, c.nq), "kcQ", 0.0, th,
                     "lcjk");
      make_triple_ints(ops, v, o, vvvoL.d(), ovooL.d(), vvvoLT, ovooLN);
      IL = TripleInts{vvvoL.d(), vvvoLT.d(), ovooLN.d()};
    }
  }
  // every precision runs on operands packed once for all triplets: BF16 hi/lo and FP32
  // planes with K in FP32, FP64 planes with K in FP64
  const bool f64p = (pr == Prec::F64);
  const bool f32 = (pr == Prec::F32);
  const int kind = f32 ? 1 : (f64p ? 2 : 0);
  PackedInts PI, PIL;
  PackedAmps PT, PU;
  PI.W.buf = &W[11]; PI.Wt.buf = &W[12]; PI.Hn.buf = &W[13];
  PT.T3.buf = &W[14]; PT.T4.buf = &W[15]; PT.Ts.buf = &W[16];
  PU.T3.buf = &W[17]; PU.T4.buf = &W[18]; PU.Ts.buf = &W[19];
  PIL.W.buf = &W[20]; PIL.Wt.buf = &W[21]; PIL.Hn.buf = &W[22];
  pack_ints(ops.gemm, v, o, IT, kind, PI);
  pack_amps(ops.gemm, v, o, dT2.d(), kind, PT);
  if (lambda) {
    pack_amps(ops.gemm, v, o, dU2.d(), kind, PU);
    if (!c.hermitian) pack_ints(ops.gemm, v, o, IL, kind, PIL);
  }
  const PackedInts& PIU = (lambda && !c.hermitian) ? PIL : PI;
  const int ntile = (v + TT - 1) / TT;
  const long nblocks = (long)ntile * ntile * ntile;
  // blocks of up to BMAX triplets (i0 + z, j, k) per launch
  const int BMAX = 8;
  // K is stored in whole 8x8x8 boxes (ntile^3 of them, padding zeroed below)
  const long kmem = nblocks * TT * TT * TT;   // elements of one part of one triplet
  const size_t kbytes = (f64p ? sizeof(double) : sizeof(float)) * kmem * BMAX;   // one part, all members
  const long pz = nblocks * NPART;
  DevBuf& K = W[23];
  DevBuf& KU = W[24];
  DevBuf part(sizeof(double) * pz * BMAX);
  K.ensure(3 * kbytes);
  if (lambda) KU.ensure(3 * kbytes);
  ECG_HIP(hipMemsetAsync(K.p, 0, 3 * kbytes, s));
  if (lambda) ECG_HIP(hipMemsetAsync(KU.p, 0, 3 * kbytes, s));
  DevBuf IntX(sizeof(double) * n1), IntY(sizeof(double) * n1), eacc(sizeof(double) * ntile);
  ops.fill(n1, 0.0, IntX.d());
  ops.fill(n1, 0.0, IntY.d());
  ops.fill(ntile, 0.0, eacc.d());
  // sorted tile triples A <= B <= C: one combine workgroup each
  std::vector<int> trih;
  for (int C = 0; C < ntile; ++C)
    for (int B = 0; B <= C; ++B)
      for (int A = 0; A <= B; ++A) { trih.push_back(A); trih.push_back(B); trih.push_back(C); }
  const unsigned ntri = (unsigned)(trih.size() / 3);
  DevBuf tri(sizeof(int) * trih.size());
  ECG_HIP(hipMemcpyAsync(tri.p, trih.data(), tri.bytes, hipMemcpyHostToDevice, s));
  // K (and KU) as three parts, one per product class
  char* kb = static_cast<char*>(K.p);
  char* kub = lambda ? static_cast<char*>(KU.p) : nullptr;
  double* const Kd[3] = {reinterpret_cast<double*>(kb), reinterpret_cast<double*>(kb + kbytes),
                         reinterpret_cast<double*>(kb + 2 * kbytes)};
  float* const Kf[3] = {reinterpret_cast<float*>(kb), reinterpret_cast<float*>(kb + kbytes),
                        reinterpret_cast<float*>(kb + 2 * kbytes)};
  double* const KUd[3] = {reinterpret_cast<double*>(kub), reinterpret_cast<double*>(kub + kbytes),
                          reinterpret_cast<double*>(kub + 2 * kbytes)};
  float* const KUf[3] = {reinterpret_cast<float*>(kub), reinterpret_cast<float*>(kub + kbytes),
                         reinterpret_cast<float*>(kub + 2 * kbytes)};
  CombArgs<double> ca{};
  {
    // distinct output blocks per tie pattern, from a representative triple of each
    const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {2, 1, 0}, {1, 0, 2}, {2, 0, 1}, {1, 2, 0}};
    const int rep[4][3] = {{0, 1, 2}, {0, 0, 1}, {0, 1, 1}, {0, 0, 0}};
    for (int pt = 0; pt < 4; ++pt) {
      const int* T = rep[pt];
      int nb = 0;
      for (int o2 = 0; o2 < 6; ++o2) {
        const int tg[3] = {T[perm[o2][0]], T[perm[o2][1]], T[perm[o2][2]]};
        bool dup = false;
        for (int o3 = 0; o3 < o2; ++o3)
          dup = dup || (T[perm[o3][0]] == tg[0] && T[perm[o3][1]] == tg[1] && T[perm[o3][2]] == tg[2]);
        if (dup) continue;
        ca.blk[pt][nb][0] = (signed char)o2;
        for (int q = 0; q < 6; ++q) {
          const int w0 = tg[perm[q][0]], w1 = tg[perm[q][1]], w2 = tg[perm[q][2]];
          int sl = 0;
          for (int s5 = 5; s5 >= 0; --s5)
            if (T[perm[s5][0]] == w0 && T[perm[s5][1]] == w1 && T[perm[s5][2]] == w2) sl = s5;
          ca.blk[pt][nb][1 + q] = (signed char)sl;
        }
        ++nb;
      }
      ca.nblk[pt] = nb;
    }
  }
  ca.nv = v; ca.no = o; ca.ntile = ntile;
  ca.tri = static_cast<const int*>(tri.p);
  ca.eo = c.eps.d(); ca.ev = c.eps.d() + o;
  ca.W = c.vvoo.d();
  ca.Y = lambda ? dU2.d() : dT2.d();
  ca.part = part.d();
  CombArgs<float> cf{};
  cf.nv = v; cf.no = o; cf.ntile = ntile; cf.tri = ca.tri;
  cf.eo = ca.eo; cf.ev = ca.ev; cf.W = ca.W; cf.Y = ca.Y; cf.part = ca.part;
  ca.kz = kmem; cf.kz = kmem;
  ca.pz = pz; cf.pz = pz;
  std::copy(&ca.nblk[0], &ca.nblk[0] + 4, &cf.nblk[0]);
  std::copy(&ca.blk[0][0][0], &ca.blk[0][0][0] + 4 * 6 * 7, &cf.blk[0][0][0]);
  for (int q = 0; q < 3; ++q) {
    ca.K[q] = Kd[q]; ca.KU[q] = lambda ? KUd[q] : nullptr;
    cf.K[q] = Kf[q]; cf.KU[q] = lambda ? KUf[q] : nullptr;
  }
  const unsigned cgrid = ntri;   // one workgroup per sorted tile triple
  // one block of triplets (i0 + z, j, k), z < nb: products, combine, and (when reduce) the
  // ordered accumulation into IntX, IntY and the energy
  auto run_block = [&](int i0, int nb, int j, int k, bool reduce) {
    build_K_packed(ops.gemm, v, o, i0, nb, j, k, PT, PI, Kf, Kd, kmem);
    if (lambda) build_K_packed(ops.gemm, v, o, i0, nb, j, k, PU, PIU, KUf, KUd, kmem);
    if (f64p) {
      ca.i = i0; ca.j = j; ca.k = k;
      hipLaunchKernelGGL(triples_combine_kernel<double>, dim3(cgrid, (unsigned)nb), dim3(256), 0, s, ca);
    } else {
      cf.i = i0; cf.j = j; cf.k = k;
      hipLaunchKernelGGL(triples_combine_kernel<float>, dim3(cgrid, (unsigned)nb), dim3(256), 0, s, cf);
    }
    if (reduce)
      hipLaunchKernelGGL(triples_reduce_kernel, dim3(3 * ntile), dim3(RED_T), 0, s, v, ntile, i0, nb, j, k,
                         part.d(), pz, IntX.d(), IntY.d(), eacc.d());
    ECG_HIP(hipGetLastError());
  };
  // block size: chosen by measurement on this context (and kept), or from the tile count
  // of one triplet's products for short ranges
  const long ntrip_all = (long)o * (o + 1) * (o + 2) / 6 - o;
  const long nrange = count < 0 ? ntrip_all : std::min<long>(count, std::max<long>(0, ntrip_all - first));
  int nbatch = 1;
  {
    int& tuned = c.triples_batch[kind][lambda ? 1 : 0];
    const long tiles = (((long)v * v + 255) / 256) * ((v + 127) / 128);
    nbatch = (int)std::max<long>(1, std::min<long>(BMAX, 1024 / std::max<long>(1, tiles)));
    if (tuned > 0) {
      nbatch = tuned;
    } else if (nrange >= 64 && o >= 3) {
      // time blocks of 1, 2, 4, 8 triplets (i, o-1, o-1) on the real operands; nothing is
      // accumulated (the reduce is skipped), the main loop overwrites K and the records
      hipEvent_t e0, e1;
      ECG_HIP(hipEventCreate(&e0));
      ECG_HIP(hipEventCreate(&e1));
      const int avail = o - 1;   // i < o - 1 for (j, k) = (o-1, o-1)
      double best = HUGE_VAL;
      for (int nb = 1; nb <= std::min(BMAX, avail); nb *= 2) {
        run_block(0, nb, o - 1, o - 1, false);   // warm-up of this shape
        ECG_HIP(hipEventRecord(e0, s));
        for (int rep = 0; rep < 2; ++rep) run_block(0, nb, o - 1, o - 1, false);
        ECG_HIP(hipEventRecord(e1, s));
        ECG_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        ECG_HIP(hipEventElapsedTime(&ms, e0, e1));
        const double per = ms / (2.0 * nb);
        if (per < 0.97 * best) { best = per; nbatch = nb; }
      }
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      tuned = nbatch;
    }
  }
  c.triples_last_batch = nbatch;
  // loop order of cc_triples.jl:61-71 (k, j <= k, i <= j, i = j = k skipped): the triplets of
  // one (j, k) are consecutive in i, so a block is a run of consecutive triplets
  long g0 = 0;
  const long last = count < 0 ? (1L << 62) : first + count;
  for (int k = 0; k < o; ++k)
    for (int j = 0; j <= k; ++j) {
      const long ni = (j == k) ? j : j + 1;
      const long lo = std::max<long>(g0, first), hi = std::min<long>(g0 + ni, last);
      for (long t0 = lo; t0 < hi; t0 += nbatch) {
        const int nb = (int)std::min<long>(nbatch, hi - t0);
        run_block((int)(t0 - g0), nb, j, k, true);
      }
      g0 += ni;
    }
  // E(T) = E[T] + T1.IntX + f_ov.IntY  (Lambda: 1/2 U1.IntX)
  std::vector<double> hx(n1), hy(n1), hf((size_t)c.n * c.n), he(ntile);
  ECG_HIP(hipMemcpyAsync(hx.data(), IntX.p, IntX.bytes, hipMemcpyDeviceToHost, s));
  ECG_HIP(hipMemcpyAsync(hy.data(), IntY.p, IntY.bytes, hipMemcpyDeviceToHost, s));
  ECG_HIP(hipMemcpyAsync(he.data(), eacc.p, eacc.bytes, hipMemcpyDeviceToHost, s));
  ECG_HIP(hipMemcpyAsync(hf.data(), c.fock.p, hf.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  ECG_HIP(hipStreamSynchronize(s));
  double eb = 0.0;
  for (int t = 0; t < ntile; ++t) eb += he[t];
  const double* t1 = lambda ? U1h : T1h;
  double fac1 = lambda ? 0.5 : 1.0;
  double en = eb;
  for (int i = 0; i < o; ++i)
    for (int a = 0; a < v; ++a) {
      long ai = a + (long)v * i;
      en += fac1 * t1[ai] * hx[ai];
      en += hf[(size_t)(o + a) * c.n + i] * hy[ai];   // f[i, o+a] (occupied row, virtual column)
    }
  e[0] = en;
  e[1] = eb;
  ops.arena.reset();
}

}  // namespace ecg
