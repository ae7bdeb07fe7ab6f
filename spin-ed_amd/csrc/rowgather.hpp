// rowgather.hpp -- the per-row arithmetic of the real matrix-free H*x,
// shared by the single-operator and the batched matvec kernels.
// For representative r (row i) with |Stab r| = s_r:
//   y_i = D(r) x_i + sum_sigma conj(c_sigma) chi(h) sqrt(|Stab r'| / s_r) x_{idx(r')}
// (H|r> = sum c_sigma |sigma>, r' = h.sigma = min orbit(sigma)); terms whose
// orbit has inconsistent phases or whose representative is not in the basis
// are dropped.
#pragma once

#include <hip/hip_runtime.h>

#include "orbit.hpp"

namespace spined {

struct OpArgs {
  const uint32_t *__restrict__ tuple_sites;
  const uint32_t *__restrict__ tuple_n;
  const uint32_t *__restrict__ tuple_off;
  const uint32_t *__restrict__ range;
  const double *__restrict__ diag_re;
  const double *__restrict__ diag_im;
  const uint64_t *__restrict__ flip;
  const double *__restrict__ coef_re;
  const double *__restrict__ coef_im;
  const double *__restrict__ phase_cos; // complex sectors: chi(phase)
  const double *__restrict__ phase_sin;
  int32_t ntuples;
  int32_t ntemplates;
};

template <class T> struct is_cplx : std::false_type {};
template <> struct is_cplx<float2> : std::true_type {};
template <> struct is_cplx<double2> : std::true_type {};

__device__ inline double ld_re(const float *x, uint64_t i) { return static_cast<double>(x[i]); }
__device__ inline double ld_re(const double *x, uint64_t i) { return x[i]; }
__device__ inline double2 ld_c(const float2 *x, uint64_t i) {
  const float2 v = x[i];
  return make_double2(v.x, v.y);
}
__device__ inline double2 ld_c(const double2 *x, uint64_t i) { return x[i]; }
__device__ inline void st_re(float *y, uint64_t i, double v) { y[i] = static_cast<float>(v); }
__device__ inline void st_re(double *y, uint64_t i, double v) { y[i] = v; }
__device__ inline void st_c(float2 *y, uint64_t i, double2 v) {
  y[i] = make_float2(static_cast<float>(v.x), static_cast<float>(v.y));
}
__device__ inline void st_c(double2 *y, uint64_t i, double2 v) { y[i] = v; }

// Bijective block remap: each of the 8 XCDs gets a contiguous range of rows,
// so neighbouring rows (which share much of their gather targets) meet in
// the same L2.
__device__ inline uint64_t xcd_block(uint64_t b, uint64_t nblocks) {
  const uint64_t xcd = b & 7, idx = b >> 3;
  const uint64_t q = nblocks >> 3, r = nblocks & 7;
  return xcd * q + (xcd < r ? xcd : r) + idx;
}

template <int METHOD>
__device__ inline OrbitResult do_orbit(const GenericOrbitArgs &ga, const RingOrbitArgs &ra,
                                       uint64_t s) {
  if constexpr (METHOD == 1)
    return orbit_ring(ra, s);
  else
    return orbit_generic(ga, s);
}

__device__ inline int local_state(uint32_t sites, int n, uint64_t r) {
  int a = 0;
  for (int k = 0; k < n; ++k) a = (a << 1) | static_cast<int>((r >> ((sites >> (8 * k)) & 0xff)) & 1);
  return a;
}

template <int NW>
__device__ inline void mark_range(uint64_t (&act)[NW], uint32_t b, uint32_t e) {
#pragma unroll
  for (int w = 0; w < NW; ++w) {
    const uint32_t lo = 64u * w, hi = lo + 64u;
    const uint32_t s = b > lo ? b : lo, t = e < hi ? e : hi;
    if (s < t) {
      const uint32_t len = t - s;
      const uint64_t m = len == 64 ? ~0ull : (((1ull << len) - 1) << (s - lo));
      act[w] |= m;
    }
  }
}

template <int NW> __device__ inline int pop_first(uint64_t (&act)[NW]) {
  int j = -1;
#pragma unroll
  for (int w = 0; w < NW; ++w) {
    if (j < 0 && act[w]) {
      j = 64 * w + __builtin_ctzll(act[w]);
      act[w] &= act[w] - 1;
    }
  }
  return j;
}

template <int NW> __device__ inline bool any_set(const uint64_t (&act)[NW]) {
  uint64_t o = 0;
#pragma unroll
  for (int w = 0; w < NW; ++w) o |= act[w];
  return o != 0;
}

// Exchange bonds (Heisenberg / XXZ): diagonal dp (parallel) or da
// (antiparallel) per bond, one off-diagonal template per antiparallel bond.
struct BondArgs {
  const uint32_t *__restrict__ sites; // s0 | s1 << 8
  const double *__restrict__ dp;
  const double *__restrict__ da;
  const double *__restrict__ o;
  const uint64_t *__restrict__ flip;
  int32_t nbonds;
  // non-null when every bond joins ring neighbours: lower site of each bond
  // (the ring-orbit gather then takes the hop path, orbit_ring_hop)
  const uint8_t *__restrict__ nn_lo;
  // 1 when bond j joins sites j and j + 1 (nn_lo[j] == j): the lower site
  // is the bond index and the table is not read
  int32_t nn_identity;
};

// sqrt(|Stab r'| / |Stab r|) from a table over the divisors of |G|
struct RatioArgs {
  const uint16_t *__restrict__ div_idx;
  const double *__restrict__ ratio;
  int32_t nd;
};

// Orbit of the row state r with template j applied.  Ring orbits of
// nearest-neighbour bond moves take the hop path from the row's run data
// (computed once per row); everything else the full scan.  Same results.
template <int METHOD, bool BONDS> struct RowOrbit {
  RunInfo ri0{0, 0, 0, false}, ri1{0, 0, 0, false};
  __device__ RowOrbit(const RingOrbitArgs &ra, const BondArgs &bd, uint64_t r) {
    if constexpr (METHOD == 1 && BONDS) {
      if (bd.nn_lo) {
        ri0 = run_info(~r & ra.mask, ra.n, ra.mask);
        if (ra.has_flip) ri1 = run_info(r, ra.n, ra.mask);
      }
    }
  }
  __device__ OrbitResult orbit(const OpArgs &o, const BondArgs &bd, const GenericOrbitArgs &ga,
                               const RingOrbitArgs &ra, uint64_t r, int j) const {
    if constexpr (METHOD == 1 && BONDS) {
      if (bd.nn_lo) return orbit_ring_hop(ra, ri0, ri1, r, bd.nn_identity ? j : bd.nn_lo[j]);
    }
    return do_orbit<METHOD>(ga, ra, r ^ (BONDS ? bd.flip[j] : o.flip[j]));
  }
};

// Index lookups of two orbit representatives and the gathers of their x
// values (ld: index -> value; V{} when absent).  With the hash index the
// confirming reps[k] loads are issued together with the x gathers; a
// fingerprint false positive (rare) falls back to the exact lookup.
template <class V, class LoadX>
__device__ inline void lookup_gather2(const IndexArgs &ix, const OrbitResult &oa,
                                      const OrbitResult &ob, LoadX ld, int64_t &ka, int64_t &kb,
                                      V &xa, V &xb) {
  xa = V{};
  xb = V{};
  if (ix.slots) {
    uint64_t pa = 0, pb = 0;
    ka = oa.valid ? hash_probe(ix, oa.rep, &pa) : -1;
    kb = ob.valid ? hash_probe(ix, ob.rep, &pb) : -1;
    uint64_t ra = oa.rep, rb = ob.rep;
    if (ka >= 0) {
      ra = ix.reps[ka];
      xa = ld(ka);
    }
    if (kb >= 0) {
      rb = ix.reps[kb];
      xb = ld(kb);
    }
    if (ra != oa.rep) {
      ka = hash_find(ix, oa.rep);
      xa = ka >= 0 ? ld(ka) : V{};
    }
    if (rb != ob.rep) {
      kb = hash_find(ix, ob.rep);
      xb = kb >= 0 ? ld(kb) : V{};
    }
    return;
  }
  find_index2(ix, oa.rep, oa.valid, ob.rep, ob.valid, ka, kb);
  if (ka >= 0) xa = ld(ka);
  if (kb >= 0) xb = ld(kb);
}

// Pass A of a row: the diagonal and the active off-diagonal templates (in
// exchange-bond form: the antiparallel bonds), one bit per template.
template <int NW, bool BONDS>
__device__ inline double row_pass_a(const OpArgs &o, const BondArgs &bd, uint64_t r,
                                    uint64_t (&act)[NW]) {
#pragma unroll
  for (int w = 0; w < NW; ++w) act[w] = 0;
  double diag = 0.0;
  if constexpr (BONDS) {
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      const int tb = 64 * w;
      const int te = bd.nbonds < tb + 64 ? bd.nbonds : tb + 64;
      for (int t = tb; t < te; ++t) {
        const uint32_t s = bd.sites[t];
        const uint64_t anti = ((r >> (s & 0xffu)) ^ (r >> (s >> 8))) & 1u;
        diag += anti ? bd.da[t] : bd.dp[t];
        if (bd.o[t] != 0.0) act[w] |= anti << (t - tb);
      }
    }
  } else {
    for (int t = 0; t < o.ntuples; ++t) {
      const int n = static_cast<int>(o.tuple_n[t]);
      const uint32_t off = o.tuple_off[t];
      const int a = local_state(o.tuple_sites[t], n, r);
      diag += o.diag_re[off + a];
      const uint32_t *rg = o.range + off + t + a;
      mark_range<NW>(act, rg[0], rg[1]);
    }
  }
  return diag;
}

// Target index of bond j's move in a symmetry-free weight sector (rank
// index), from the row's own index i.  When the bond joins sites a and a + 1
// one up spin moves across it, and the colex rank sum_k C(p_k, k + 1)
// changes only in that spin's term: + C(a, c) moving up, - C(a, c) moving
// down (Pascal's rule), c = up spins below a.  False, k untouched, for
// sites further apart and, with the spin flip, for a move into the top site
// (the target's representative is then its complement); those take the
// orbit and the full rank.  The target is a sector state (the row is, the
// move keeps the weight and the top site down), so k is its index.
__device__ inline bool rank_hop(const IndexArgs &ix, const GenericOrbitArgs &ga,
                                const BondArgs &bd, uint64_t r, uint64_t i, int j, int64_t &k) {
  const uint32_t st = bd.sites[j];
  const int s0 = static_cast<int>(st & 0xffu), s1 = static_cast<int>(st >> 8);
  const int a = s0 < s1 ? s0 : s1, b = s0 < s1 ? s1 : s0;
  if (b != a + 1 || (ga.has_flip && b == ga.n_sites - 1)) return false;
  const int c = __builtin_popcountll(r & ((1ull << a) - 1));
  const int64_t cac = c == 0 ? 1 : static_cast<int64_t>(ix.rank_tab[(c - 1) * 64 + a]);
  const int64_t t = static_cast<int64_t>(i) + (((r >> a) & 1) ? cac : -cac);
  k = t >= 0 && static_cast<uint64_t>(t) < ix.rank_dim ? t : -1;
  return true;
}

// One row of the real row gather (shared by the single and the batched
// matvec): diagonal + active templates, two orbit scans / searches / gathers
// in flight per iteration, accumulation in template order.
template <int METHOD, int NW, bool BONDS, class XT>
__device__ inline double gather_row_real(const OpArgs &o, const BondArgs &bd, const RatioArgs &rt,
                                         const GenericOrbitArgs &ga, const RingOrbitArgs &ra,
                                         const IndexArgs &ix_in, const uint32_t *__restrict__ stab,
                                         uint64_t i, const XT *__restrict__ x,
                                         const double *__restrict__ bond_o = nullptr) {
  const IndexArgs ix = index_for_method<METHOD>(ix_in);
  if (!bond_o) bond_o = bd.o;
  const uint64_t r = ix.reps[i];
  uint64_t act[NW];
  const double diag = row_pass_a<NW, BONDS>(o, bd, r, act);
  const RowOrbit<METHOD, BONDS> rob(ra, bd, r);
  double acc = diag * ld_re(x, i);
  const uint32_t si = stab[i];
  const double sr = static_cast<double>(si);
  const int row_di = rt.nd ? static_cast<int>(rt.div_idx[si]) : 0;
  // two templates per iteration: two orbit scans, then both searches and
  // gathers in flight together; accumulation stays in template order
  // sqrt(|Stab r'| / |Stab r|) is 1 when the stabilisers have equal size
  // (almost always; no table read then)
  auto coef_of = [&](int j, const OrbitResult &orb) {
    const double f = orb.count == static_cast<int32_t>(si)
                         ? 1.0
                         : (rt.nd ? rt.ratio[rt.div_idx[orb.count] * rt.nd + row_di]
                                  : sqrt(static_cast<double>(orb.count) / sr));
    double c = (BONDS ? bond_o[j] : o.coef_re[j]) * f;
    return orb.phase != 0 ? -c : c;
  };
  while (any_set<NW>(act)) {
    const int ja = pop_first<NW>(act);
    const int jb = pop_first<NW>(act);
    if constexpr (METHOD == 0 && BONDS) {
      if (ix.rank_mode == 1) {
        // symmetry-free weight sector: nearest-neighbour moves from the
        // row's index (rank_hop), the rest through the orbit and the rank
        // (the identity, or the flip, reaches the target once: count 1)
        int64_t ka = -1, kb = -1;
        OrbitResult oa{0, 0, 1, true}, ob{0, 0, 1, true};
        const bool fa = rank_hop(ix, ga, bd, r, i, ja, ka);
        const bool fb = jb < 0 || rank_hop(ix, ga, bd, r, i, jb, kb);
        if (!fa || !fb) {
          if (!fa) oa = rob.orbit(o, bd, ga, ra, r, ja);
          if (!fb) ob = rob.orbit(o, bd, ga, ra, r, jb);
          int64_t k1 = -1, k2 = -1;
          rank_of2(ix, oa.rep, !fa && oa.valid, ob.rep, !fb && ob.valid, k1, k2);
          if (!fa) ka = k1;
          if (!fb) kb = k2;
        }
        const double xa = ka >= 0 ? ld_re(x, static_cast<uint64_t>(ka)) : 0.0;
        const double xb = kb >= 0 ? ld_re(x, static_cast<uint64_t>(kb)) : 0.0;
        if (ka >= 0) acc = fma(coef_of(ja, oa), xa, acc);
        if (kb >= 0) acc = fma(coef_of(jb, ob), xb, acc);
        continue;
      }
    }
    const OrbitResult oa = rob.orbit(o, bd, ga, ra, r, ja);
    OrbitResult ob{0, 0, 0, false};
    if (jb >= 0) ob = rob.orbit(o, bd, ga, ra, r, jb);
    int64_t ka, kb;
    double xa, xb;
    lookup_gather2(ix, oa, ob, [&](int64_t q) { return ld_re(x, static_cast<uint64_t>(q)); }, ka,
                   kb, xa, xb);
    if (ka >= 0) acc = fma(coef_of(ja, oa), xa, acc);
    if (kb >= 0) acc = fma(coef_of(jb, ob), xb, acc);
  }
  return acc;
}


// The row of gather_row_real on L lanes (L = 2 or 4, lane `sub` of the row's
// group of L consecutive lanes): every lane evaluates the diagonal and the
// active template list; in each round lane l takes the l-th of the next L
// templates (orbit, index search, x gather, coefficient), and every lane then
// accumulates the round's terms in template order from the group's
// shuffles.  Same terms, same fma chain: bit-identical to gather_row_real.
template <int METHOD, int NW, bool BONDS, int L, class XT>
__device__ inline double gather_row_real_lanes(const OpArgs &o, const BondArgs &bd,
                                               const RatioArgs &rt, const GenericOrbitArgs &ga,
                                               const RingOrbitArgs &ra, const IndexArgs &ix_in,
                                               const uint32_t *__restrict__ stab, uint64_t i,
                                               const XT *__restrict__ x, const double *bond_o,
                                               int sub) {
  const IndexArgs ix = index_for_method<METHOD>(ix_in);
  if (!bond_o) bond_o = bd.o;
  const uint64_t r = ix.reps[i];
  uint64_t act[NW];
  const double diag = row_pass_a<NW, BONDS>(o, bd, r, act);
  const RowOrbit<METHOD, BONDS> rob(ra, bd, r);
  double acc = diag * ld_re(x, i);
  const uint32_t si = stab[i];
  const double sr = static_cast<double>(si);
  const int row_di = rt.nd ? static_cast<int>(rt.div_idx[si]) : 0;
  auto coef_of = [&](int j, const OrbitResult &orb) {
    const double f = orb.count == static_cast<int32_t>(si)
                         ? 1.0
                         : (rt.nd ? rt.ratio[rt.div_idx[orb.count] * rt.nd + row_di]
                                  : sqrt(static_cast<double>(orb.count) / sr));
    double c = (BONDS ? bond_o[j] : o.coef_re[j]) * f;
    return orb.phase != 0 ? -c : c;
  };
  while (any_set<NW>(act)) {
    int j = -1;
#pragma unroll
    for (int l = 0; l < L; ++l) {
      const int t = pop_first<NW>(act);
      if (l == sub) j = t;
    }
    double c = 0.0, xv = 0.0;
    int have = 0;
    if (j >= 0) {
      int64_t k = -1;
      OrbitResult orb{0, 0, 1, true};
      bool by_rank = false;
      if constexpr (METHOD == 0 && BONDS) {
        // symmetry-free weight sector: the move's target from the row index
        if (ix.rank_mode == 1) {
          by_rank = rank_hop(ix, ga, bd, r, i, j, k);
          if (!by_rank) {
            orb = rob.orbit(o, bd, ga, ra, r, j);
            int64_t k2 = -1;
            rank_of2(ix, orb.rep, orb.valid, 0, false, k, k2);
          }
          by_rank = true;
        }
      }
      if (!by_rank) {
        orb = rob.orbit(o, bd, ga, ra, r, j);
        const OrbitResult none{0, 0, 0, false};
        int64_t kb;
        double xb;
        lookup_gather2(ix, orb, none,
                       [&](int64_t q) { return ld_re(x, static_cast<uint64_t>(q)); }, k, kb,
                       xv, xb);
      } else if (k >= 0) {
        xv = ld_re(x, static_cast<uint64_t>(k));
      }
      if (k >= 0) {
        c = coef_of(j, orb);
        have = 1;
      }
    }
#pragma unroll
    for (int l = 0; l < L; ++l) {
      const double cl = __shfl(c, l, L);
      const double xl = __shfl(xv, l, L);
      if (__shfl(have, l, L)) acc = fma(cl, xl, acc);
    }
  }
  return acc;
}

// K values of one row of a row-interleaved block, loaded as one aligned
// record (16-byte loads).
template <class XT, int K> struct alignas(K * sizeof(XT) < 16 ? K * sizeof(XT) : 16) RowK {
  XT v[K];
};

// Block variant: K vectors stored row-interleaved (x[row * K + c]); one
// orbit scan and one index search per term serve all K columns, and each
// gather reads the K values of one row (one 64-byte line for K = 8 FP64).
// Per column the arithmetic (order and fma pattern) equals gather_row_real.
template <int METHOD, int NW, bool BONDS, int K, class XT>
__device__ inline void gather_row_real_block(const OpArgs &o, const BondArgs &bd,
                                             const RatioArgs &rt, const GenericOrbitArgs &ga,
                                             const RingOrbitArgs &ra, const IndexArgs &ix_in,
                                             const uint32_t *__restrict__ stab, uint64_t i,
                                             const XT *__restrict__ x, double (&acc)[K]) {
  const IndexArgs ix = index_for_method<METHOD>(ix_in);
  using Row = RowK<XT, K>;
  const Row *__restrict__ xr = reinterpret_cast<const Row *>(x);
  const uint64_t r = ix.reps[i];
  uint64_t act[NW];
  const double diag = row_pass_a<NW, BONDS>(o, bd, r, act);
  const RowOrbit<METHOD, BONDS> rob(ra, bd, r);
  {
    const Row xi = xr[i];
#pragma unroll
    for (int c = 0; c < K; ++c) acc[c] = diag * static_cast<double>(xi.v[c]);
  }
  const uint32_t si = stab[i];
  const double sr = static_cast<double>(si);
  const int row_di = rt.nd ? static_cast<int>(rt.div_idx[si]) : 0;
  // sqrt(|Stab r'| / |Stab r|) is 1 when the stabilisers have equal size
  // (almost always; no table read then)
  auto coef_of = [&](int j, const OrbitResult &orb) {
    const double f = orb.count == static_cast<int32_t>(si)
                         ? 1.0
                         : (rt.nd ? rt.ratio[rt.div_idx[orb.count] * rt.nd + row_di]
                                  : sqrt(static_cast<double>(orb.count) / sr));
    double c = (BONDS ? bd.o[j] : o.coef_re[j]) * f;
    return orb.phase != 0 ? -c : c;
  };
  while (any_set<NW>(act)) {
    const int ja = pop_first<NW>(act);
    const int jb = pop_first<NW>(act);
    const OrbitResult oa = rob.orbit(o, bd, ga, ra, r, ja);
    OrbitResult ob{0, 0, 0, false};
    if (jb >= 0) ob = rob.orbit(o, bd, ga, ra, r, jb);
    int64_t ka, kb;
    Row xa, xb;
    lookup_gather2(ix, oa, ob, [&](int64_t q) { return xr[q]; }, ka, kb, xa, xb);
    if (ka >= 0) {
      const double ca = coef_of(ja, oa);
#pragma unroll
      for (int c = 0; c < K; ++c) acc[c] = fma(ca, static_cast<double>(xa.v[c]), acc[c]);
    }
    if (kb >= 0) {
      const double cb = coef_of(jb, ob);
#pragma unroll
      for (int c = 0; c < K; ++c) acc[c] = fma(cb, static_cast<double>(xb.v[c]), acc[c]);
    }
  }
}

} // namespace spined
