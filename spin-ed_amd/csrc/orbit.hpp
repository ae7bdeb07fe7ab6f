// orbit.hpp -- orbit-minimum and index-search primitives shared by every
// kernel.  All functions are __host__ __device__ so the same integer logic
// can be exercised on the CPU (spined_debug_orbit) and compared with the
// oracle without a GPU.
//
// State convention: bit i of a uint64 is site i; a permutation p acts as
// (g.x)_i = x_{p[i]}; the representative of x is min_g g.x.  For each x the
// scan returns
//   rep   = min_g g.x,
//   phase = phase (integer modulo L) of the first element reaching rep,
//   count = #{g : g.x = rep}  (= |Stab(rep)|),
//   valid = all elements reaching rep carry the same phase (the character is
//           trivial on the stabiliser, i.e. rep spans a symmetric state).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace spined {

struct OrbitResult {
  uint64_t rep;
  int32_t phase;
  int32_t count;
  bool valid;
};

__host__ __device__ inline uint64_t site_mask_of(int n) {
  return n >= 64 ? ~0ull : ((1ull << n) - 1);
}

// Reverse the lowest n bits: site i -> site n-1-i.
__host__ __device__ inline uint64_t reverse_n(uint64_t x, int n) {
  const uint64_t r = __builtin_bitreverse64(x);
  return n >= 64 ? r : (r >> (64 - n));
}

__host__ __device__ inline int ctz64(uint64_t v) { return __builtin_ctzll(v); }

__host__ __device__ inline void orbit_update(uint64_t &best, int32_t &ph, int32_t &cnt, bool &ok,
                                             uint64_t y, int32_t p) {
  if (y < best) {
    best = y;
    ph = p;
    cnt = 1;
    ok = true;
  } else if (y == best) {
    cnt += 1;
    ok = ok && (p == ph);
  }
}

// ---------------------------------------------------------------------------
// Generic groups: every permutation is a short program of (mask, shift)
// terms, image = OR_t shift(src & mask_t, s_t), where src is x or its
// n-bit reversal (whichever needs fewer distinct shifts; the reversal is
// computed once per scan).  Spin inversion doubles every image: y and
// y ^ site_mask.
// ---------------------------------------------------------------------------
struct GenericOrbitArgs {
  const uint64_t *__restrict__ masks;   // all terms, element after element
  const int32_t *__restrict__ shifts;   // > 0: left shift, < 0: right shift
  const uint32_t *__restrict__ offsets; // n_perm + 1 term offsets
  const uint32_t *__restrict__ flags;   // bit 0: program reads reverse_n(x)
  const int32_t *__restrict__ phases;   // phase of each permutation (mod L)
  int32_t n_perm;
  int32_t n_sites;
  uint64_t site_mask;
  int32_t has_flip;
  int32_t flip_phase;
  int32_t L;
  // Torus factorisation (torus_l1 > 0): the group holds every translation
  // t_{dx,dy} of an l1 x l2 torus in row-major site order (site = l1 y + x;
  // (t.x) at (x, y) = x at ((x + dx) mod l1, (y + dy) mod l2)), and every
  // element is t_{dx,dy} after one of the n_point elements point_elem[p]
  // (a transversal of the translation cosets).  tp_phase[p * l1 l2 +
  // dy l1 + dx] is that element's phase (not read when trivial);
  // row_lo[dx] / row_hi[dx] select, in every row, the bits that the in-row
  // rotation by dx takes from the right / wraps around from the left.
  const uint32_t *__restrict__ point_elem;
  const int32_t *__restrict__ tp_phase;
  const uint64_t *__restrict__ row_lo;
  const uint64_t *__restrict__ row_hi;
  int32_t torus_l1;
  int32_t torus_l2;
  int32_t n_point;
  int32_t trivial; // every phase is 0 (the flip's included)
};

__host__ __device__ inline uint64_t apply_program(const GenericOrbitArgs &a, int g, uint64_t x,
                                                  uint64_t rx) {
  const uint64_t src = (a.flags[g] & 1u) ? rx : x;
  uint64_t y = 0;
  const uint32_t e = a.offsets[g + 1];
  for (uint32_t t = a.offsets[g]; t < e; ++t) {
    const uint64_t m = src & a.masks[t];
    const int32_t s = a.shifts[t];
    y |= s >= 0 ? (m << s) : (m >> (-s));
  }
  return y;
}

// Orbit over the torus factorisation: the point element's program once,
// then the l1 in-row rotations (two masked shifts each) and, for each, the
// l2 row rotations (one rotation of the word by l1 sites each).  Same
// (rep, count, valid) as the scan over every element; the phase is that of
// an element reaching rep (all of them carry it when valid).
__host__ __device__ inline OrbitResult orbit_torus(const GenericOrbitArgs &a, uint64_t x) {
  const uint64_t rx = reverse_n(x, a.n_sites);
  const int l1 = a.torus_l1, l2 = a.torus_l2, n = a.n_sites;
  const uint64_t mask = a.site_mask;
  uint64_t best = ~0ull;
  int32_t ph = 0, cnt = 0;
  bool ok = true;
  for (int p = 0; p < a.n_point; ++p) {
    const uint64_t y = apply_program(a, static_cast<int>(a.point_elem[p]), x, rx);
    const int32_t *tph = a.tp_phase + p * l1 * l2;
    for (int dx = 0; dx < l1; ++dx) {
      uint64_t z = ((y >> dx) & a.row_lo[dx]) | ((y << ((l1 - dx) & 63)) & a.row_hi[dx]);
      for (int dy = 0; dy < l2; ++dy) {
        const int32_t q = a.trivial ? 0 : tph[dy * l1 + dx];
        if (a.has_flip) {
          // z and z ^ mask differ, so only the smaller can reach or beat
          // best: one update with it counts exactly what two would
          const uint64_t zf = z ^ mask;
          int32_t qf = q + a.flip_phase;
          if (qf >= a.L) qf -= a.L;
          const bool lo = zf < z;
          orbit_update(best, ph, cnt, ok, lo ? zf : z, lo ? qf : q);
        } else {
          orbit_update(best, ph, cnt, ok, z, q);
        }
        z = ((z >> l1) | (z << (n - l1))) & mask;
      }
    }
  }
  return OrbitResult{best, ph, cnt, ok};
}

__host__ __device__ inline OrbitResult orbit_generic(const GenericOrbitArgs &a, uint64_t x) {
  if (a.torus_l1) return orbit_torus(a, x);
  const uint64_t rx = reverse_n(x, a.n_sites);
  uint64_t best = ~0ull;
  int32_t ph = 0, cnt = 0;
  bool ok = true;
  for (int g = 0; g < a.n_perm; ++g) {
    const uint64_t y = apply_program(a, g, x, rx);
    const int32_t p = a.phases[g];
    if (a.has_flip) {
      // y and y ^ mask differ: one update with the smaller (orbit_torus)
      const uint64_t yf = y ^ a.site_mask;
      int32_t pf = p + a.flip_phase;
      if (pf >= a.L) pf -= a.L;
      const bool lo = yf < y;
      orbit_update(best, ph, cnt, ok, lo ? yf : y, lo ? pf : p);
    } else {
      orbit_update(best, ph, cnt, ok, y, p);
    }
  }
  return OrbitResult{best, ph, cnt, ok};
}

// Orbit minimum alone, for sectors whose characters are all trivial (every
// orbit is then valid with phase 0, and |Stab| is the basis' stabiliser
// entry of the representative): the scans of orbit_torus / orbit_generic
// without the count and phase bookkeeping.  Same rep as orbit_generic.
__host__ __device__ inline uint64_t orbit_min_generic(const GenericOrbitArgs &a, uint64_t x) {
  const uint64_t rx = reverse_n(x, a.n_sites);
  const uint64_t mask = a.site_mask;
  uint64_t best = ~0ull;
  if (a.torus_l1) {
    const int l1 = a.torus_l1, l2 = a.torus_l2, n = a.n_sites;
    for (int p = 0; p < a.n_point; ++p) {
      const uint64_t y = apply_program(a, static_cast<int>(a.point_elem[p]), x, rx);
      for (int dx = 0; dx < l1; ++dx) {
        uint64_t z = ((y >> dx) & a.row_lo[dx]) | ((y << ((l1 - dx) & 63)) & a.row_hi[dx]);
        for (int dy = 0; dy < l2; ++dy) {
          uint64_t m = z;
          if (a.has_flip) {
            const uint64_t zf = z ^ mask;
            m = zf < z ? zf : z;
          }
          best = m < best ? m : best;
          z = ((z >> l1) | (z << (n - l1))) & mask;
        }
      }
    }
    return best;
  }
  for (int g = 0; g < a.n_perm; ++g) {
    const uint64_t y = apply_program(a, g, x, rx);
    uint64_t m = y;
    if (a.has_flip) {
      const uint64_t yf = y ^ mask;
      m = yf < y ? yf : y;
    }
    best = m < best ? m : best;
  }
  return best;
}

// Early-exit representative test used by the basis builder: returns the
// stabiliser size if x is a representative whose symmetric state exists,
// else 0.
// rep_test_generic over the torus factorisation (same result).
__host__ __device__ inline uint32_t rep_test_torus(const GenericOrbitArgs &a, uint64_t x) {
  const uint64_t rx = reverse_n(x, a.n_sites);
  const int l1 = a.torus_l1, l2 = a.torus_l2, n = a.n_sites;
  const uint64_t mask = a.site_mask;
  uint32_t cnt = 0;
  for (int p = 0; p < a.n_point; ++p) {
    const uint64_t y = apply_program(a, static_cast<int>(a.point_elem[p]), x, rx);
    const int32_t *tph = a.tp_phase + p * l1 * l2;
    for (int dx = 0; dx < l1; ++dx) {
      uint64_t z = ((y >> dx) & a.row_lo[dx]) | ((y << ((l1 - dx) & 63)) & a.row_hi[dx]);
      for (int dy = 0; dy < l2; ++dy) {
        const int32_t q = a.trivial ? 0 : tph[dy * l1 + dx];
        if (z < x) return 0;
        if (z == x) {
          if (q != 0) return 0;
          ++cnt;
        }
        if (a.has_flip) {
          const uint64_t zf = z ^ mask;
          if (zf < x) return 0;
          if (zf == x) {
            int32_t qf = q + a.flip_phase;
            if (qf >= a.L) qf -= a.L;
            if (qf != 0) return 0;
            ++cnt;
          }
        }
        z = ((z >> l1) | (z << (n - l1))) & mask;
      }
    }
  }
  return cnt;
}

__host__ __device__ inline uint32_t rep_test_generic(const GenericOrbitArgs &a, uint64_t x) {
  if (a.torus_l1) return rep_test_torus(a, x);
  const uint64_t rx = reverse_n(x, a.n_sites);
  uint32_t cnt = 0;
  for (int g = 0; g < a.n_perm; ++g) {
    const uint64_t y = apply_program(a, g, x, rx);
    const int32_t p = a.phases[g];
    if (y < x) return 0;
    if (y == x) {
      if (p != 0) return 0;
      ++cnt;
    }
    if (a.has_flip) {
      const uint64_t yf = y ^ a.site_mask;
      if (yf < x) return 0;
      if (yf == x) {
        int32_t pf = p + a.flip_phase;
        if (pf >= a.L) pf -= a.L;
        if (pf != 0) return 0;
        ++cnt;
      }
    }
  }
  return cnt;
}

// ---------------------------------------------------------------------------
// Rings: the group is the rotations of n sites in natural order, optionally
// with the reflections and spin inversion.  Rotation by k maps u to
// rotr(u, k) ((R_k u)_i = u_{(i+k) mod n}); reflection elements are
// rotations of reverse_n(u).  The minimum image must start (at bit n-1) with
// a longest cyclic run of zeros, so only rotations that align the top of a
// longest zero run (longest one run for the flipped images) with bit n-1
// are candidates: usually 1-4 instead of up to 4n.
// ---------------------------------------------------------------------------
struct RingOrbitArgs {
  const int32_t *__restrict__ phase_tab; // [2][n]: phase of (reflected?, k)
  int32_t n;
  uint64_t mask;
  int32_t has_rev;
  int32_t has_flip;
  int32_t flip_phase;
  int32_t L;
  int32_t trivial; // every phase is 0: the table is not read
};

__host__ __device__ inline uint64_t rotl_n(uint64_t z, int j, int n, uint64_t mask) {
  return j == 0 ? z : (((z << j) | (z >> (n - j))) & mask);
}

__host__ __device__ inline uint64_t rotr_n(uint64_t u, int k, int n, uint64_t mask) {
  return k == 0 ? u : (((u >> k) | (u << (n - k))) & mask);
}

// Rotation by a power of two known at compile time (meaningful for S < n;
// the right shift is kept in range so the unselected value is harmless).
template <int S> __host__ __device__ inline uint64_t rotl_c(uint64_t z, int n, uint64_t mask) {
  return ((z << S) | (z >> ((n - S) & 63))) & mask;
}

// Branch-free variant of longest_run_tops (same result): all doubling
// levels and refinement steps are evaluated with selects, so a wave runs one
// straight-line sequence instead of divergent branches.
__host__ __device__ inline uint64_t longest_run_tops_bf(uint64_t z, int n, uint64_t mask,
                                                        int &len) {
  // doubling: b_k = tops of windows of 2^k ones, valid while 2^k <= n - 1
  const uint64_t b0 = z;
  const uint64_t b1 = n > 2 ? b0 & rotl_c<1>(b0, n, mask) : 0;
  const uint64_t b2 = n > 4 ? b1 & rotl_c<2>(b1, n, mask) : 0;
  const uint64_t b3 = n > 8 ? b2 & rotl_c<4>(b2, n, mask) : 0;
  const uint64_t b4 = n > 16 ? b3 & rotl_c<8>(b3, n, mask) : 0;
  const uint64_t b5 = n > 32 ? b4 & rotl_c<16>(b4, n, mask) : 0;
  // highest non-empty level p (z != 0, z != mask handled by the caller)
  uint64_t cur = b0;
  int l = 1;
  cur = b1 ? b1 : cur;
  l = b1 ? 2 : l;
  cur = b2 ? b2 : cur;
  l = b2 ? 4 : l;
  cur = b3 ? b3 : cur;
  l = b3 ? 8 : l;
  cur = b4 ? b4 : cur;
  l = b4 ? 16 : l;
  cur = b5 ? b5 : cur;
  l = b5 ? 32 : l;
  // refinement a_{l+s} = a_l & rotl(a_s, l) for s = l/2, l/4, ..., 1
  // (s < l, and windows stay below n: a run of n - 1 ones is the maximum)
  const uint64_t bs[5] = {b0, b1, b2, b3, b4};
#pragma unroll
  for (int q = 4; q >= 0; --q) {
    const int s = 1 << q;
    const bool use = s < l && l + s <= n - 1;
    const int sh = use ? l : 1; // keep the shift amounts in range
    const uint64_t t = cur & (((bs[q] << sh) | (bs[q] >> (n - sh))) & mask);
    const bool take = use && t != 0;
    cur = take ? t : cur;
    l = take ? l + s : l;
  }
  len = l;
  return cur;
}

// Positions q that are the top (highest cyclic position) of a longest
// cyclic run of ones in z; len receives that length.  z == 0 yields every
// position with len 0 (all rotations of an all-ones / all-zeros word tie).
__host__ __device__ inline uint64_t longest_run_tops(uint64_t z, int n, uint64_t mask, int &len) {
  if (z == 0) {
    len = 0;
    return mask;
  }
  if (z == mask) {
    len = n;
    return mask;
  }
  return longest_run_tops_bf(z, n, mask, len);
}

__host__ __device__ inline void ring_candidates(const RingOrbitArgs &a, uint64_t u, uint64_t tops,
                                                int len, int32_t fph, uint64_t &best, int32_t &ph,
                                                int32_t &cnt, bool &ok) {
  const int n = a.n;
  const uint64_t ru = a.has_rev ? reverse_n(u, n) : 0;
  while (tops) {
    const int q = ctz64(tops);
    tops &= tops - 1;
    int k = q + 1;
    if (k >= n) k -= n;
    int32_t p = 0;
    if (!a.trivial) {
      p = a.phase_tab[k] + fph;
      if (p >= a.L) p -= a.L;
    }
    orbit_update(best, ph, cnt, ok, rotr_n(u, k, n, a.mask), p);
    if (a.has_rev) {
      int k2 = n - q + len - 1;
      k2 = k2 >= n ? k2 - n : k2; // n - q + len - 1 lies in [len, n + len - 1]
      k2 = k2 >= n ? k2 - n : k2;
      int32_t p2 = 0;
      if (!a.trivial) {
        p2 = a.phase_tab[n + k2] + fph;
        if (p2 >= a.L) p2 -= a.L;
      }
      orbit_update(best, ph, cnt, ok, rotr_n(ru, k2, n, a.mask), p2);
    }
  }
}

// Candidate stage of the ring orbit, given the tops and lengths of the
// longest zero runs (t0, l0) and, with spin inversion, one runs (t1, l1).
__host__ __device__ inline OrbitResult orbit_ring_tops(const RingOrbitArgs &a, uint64_t x,
                                                       uint64_t t0, int l0, uint64_t t1, int l1) {
  const uint64_t xn = ~x & a.mask;
  const int lm = l0 > l1 ? l0 : l1;
  uint64_t best = ~0ull;
  int32_t ph = 0, cnt = 0;
  bool ok = true;
  if (l0 == lm) ring_candidates(a, x, t0, l0, 0, best, ph, cnt, ok);
  if (a.has_flip && l1 == lm) ring_candidates(a, xn, t1, l1, a.flip_phase, best, ph, cnt, ok);
  return OrbitResult{best, ph, cnt, ok};
}

__host__ __device__ inline OrbitResult orbit_ring(const RingOrbitArgs &a, uint64_t x) {
  const int n = a.n;
  const uint64_t xn = ~x & a.mask;
  int l0 = 0, l1 = -1;
  const uint64_t t0 = longest_run_tops(xn, n, a.mask, l0); // zero runs of x
  uint64_t t1 = 0;
  if (a.has_flip) t1 = longest_run_tops(x, n, a.mask, l1); // one runs of x
  return orbit_ring_tops(a, x, t0, l0, t1, l1);
}

// ---------------------------------------------------------------------------
// Ring orbits of hop states.  In the row gather of a nearest-neighbour
// exchange Hamiltonian, every off-diagonal state is the row's representative
// r with one particle moved across a bond (i, i+1 mod n).  A move changes
// only the two runs next to the bond, so the longest-run tops of the moved
// state follow from the row's run data: the tops of r's runs of exact length
// L (the longest) and L-1.  Every other run is unchanged.
// ---------------------------------------------------------------------------
struct RunInfo {
  uint64_t EL;  // tops of runs of ones of length exactly L
  uint64_t EL1; // tops of runs of ones of length exactly L - 1 (L >= 2)
  int32_t L;    // longest run length
  bool usable;  // 0 < popcount < n (else the moved state takes the full scan)
};

// Tops of windows of k >= 1 ones (q with ones at q, q-1, .., q-k+1), from
// the doubling levels lv[s] = windows of 2^s ones.
__host__ __device__ inline uint64_t window_tops(const uint64_t (&lv)[6], int k, int n,
                                                uint64_t mask) {
  uint64_t cur = mask;
  int len = 0;
#pragma unroll
  for (int s = 5; s >= 0; --s) {
    if (k & (1 << s)) {
      cur &= rotl_n(lv[s], len, n, mask);
      len += 1 << s;
    }
  }
  return cur;
}

__host__ __device__ inline RunInfo run_info(uint64_t z, int n, uint64_t mask) {
  RunInfo ri{0, 0, 0, false};
  if (z == 0 || z == mask) return ri;
  int L = 0;
  ri.EL = longest_run_tops_bf(z, n, mask, L);
  ri.L = L;
  ri.usable = true;
  if (L >= 2) {
    uint64_t lv[6];
    lv[0] = z;
#pragma unroll
    for (int s = 1; s < 6; ++s)
      lv[s] = (1 << s) <= n - 1 ? lv[s - 1] & rotl_n(lv[s - 1], 1 << (s - 1), n, mask) : 0;
    const int k = L - 1;
    // exact length k: window of k ones ending at q, zero above q and below q - k + 1
    ri.EL1 = window_tops(lv, k, n, mask) & ~rotr_n(z, 1, n, mask) & ~rotl_n(z, k, n, mask);
  }
  return ri;
}

__host__ __device__ inline uint64_t bit_at(int q) { return 1ull << q; }

// Tops and length of the longest runs of ones of a word after a move at the
// bond (i, j = i + 1 mod n), from the word's run data ri and the two runs
// next to the bond: run R1 (length a, top topR1) loses a one, which joins
// run R2 (length b >= 0, top topR2); newTop* are their tops afterwards.
// Every other run is unchanged.
__host__ __device__ inline void hop_tops_finish(const RunInfo &ri, int a, int b, int topR1,
                                                int newTopR1, int topR2, int newTopR2,
                                                uint64_t &tops, int &len) {
  const int L = ri.L;
  uint64_t el = ri.EL, el1 = ri.EL1;
  el &= a == L ? ~bit_at(topR1) : ~0ull;
  el1 &= a == L - 1 ? ~bit_at(topR1) : ~0ull;
  el &= (b > 0 && b == L) ? ~bit_at(topR2) : ~0ull;
  el1 &= (b > 0 && b == L - 1) ? ~bit_at(topR2) : ~0ull;
  int lp = b + 1;
  lp = a - 1 > lp ? a - 1 : lp;
  lp = (el1 && L - 1 > lp) ? L - 1 : lp;
  lp = (el && L > lp) ? L : lp;
  uint64_t t = 0;
  t |= L == lp ? el : 0;
  t |= L - 1 == lp ? el1 : 0;
  t |= (a >= 2 && a - 1 == lp) ? bit_at(newTopR1) : 0;
  t |= b + 1 == lp ? bit_at(newTopR2) : 0;
  tops = t;
  len = lp;
}

// Longest-run tops of both words of x = r ^ (bit i | bit j), r_i != r_j:
// zero runs (t0, l0; run data ri0 of ~r) and one runs (t1, l1; ri1 of r).
// P, the word of the two that holds the one at i, sees it move up to j;
// Q = ~P sees its one at j move down to i.  The four run lengths next to
// the bond come from one rotation w of P (bit j at 0, bit i at n - 1):
//   aP = ones of P down from i     = leading ones of w,
//   bP = ones of P up from j + 1   = trailing ones of w >> 1 when aQ == 1,
//   aQ = zeros of P up from j      = trailing zeros of w,
//   bQ = zeros of P down from i-1  = leading zeros below bit n - 1 when aP == 1.
// False: the word needed (both with flip, else the zero-run word) has its
// two runs next to the bond joined around the ring, or no run data.
__host__ __device__ inline bool hop_tops_pair(const RunInfo &ri0, const RunInfo &ri1, uint64_t r,
                                              int i, int n, uint64_t mask, bool flip,
                                              uint64_t &t0, int &l0, uint64_t &t1, int &l1) {
  const int j = i + 1 == n ? 0 : i + 1;
  const int im1 = i == 0 ? n - 1 : i - 1;
  const bool ri = (r >> i) & 1;
  const uint64_t P = ri ? r : (~r & mask);
  const uint64_t w = rotr_n(P, j, n, mask);
  const int aQ = ctz64(w);
  const int aP = __builtin_clzll(~(w << (64 - n)));
  const int bP = aQ == 1 ? ctz64(~(w >> 1)) : 0;
  const uint64_t yq = w << (65 - n);
  const int bQ = aP == 1 ? (yq ? __builtin_clzll(yq) : n) : 0;
  const RunInfo &riP = ri ? ri1 : ri0;
  const RunInfo &riQ = ri ? ri0 : ri1;
  const bool okP = riP.usable && aP + bP < n - 1;
  const bool okQ = riQ.usable && aQ + bQ < n - 1;
  int tR2 = j + bP;
  tR2 = tR2 >= n ? tR2 - n : tR2;
  int tR1 = j + aQ - 1;
  tR1 = tR1 >= n ? tR1 - n : tR1;
  uint64_t tP = 0, tQ = 0;
  int lP = 0, lQ = 0;
  hop_tops_finish(riP, aP, bP, i, im1, tR2, tR2, tP, lP);
  hop_tops_finish(riQ, aQ, bQ, tR1, tR1, im1, i, tQ, lQ);
  t0 = ri ? tQ : tP;
  l0 = ri ? lQ : lP;
  t1 = ri ? tP : tQ;
  l1 = ri ? lP : lQ;
  const bool ok0 = ri ? okQ : okP;
  const bool ok1 = ri ? okP : okQ;
  if (!flip) l1 = -1;
  return ok0 && (!flip || ok1);
}

// Orbit of r with the bits at i and j = (i + 1) mod n exchanged (they
// differ), from the row's run data ri0 (zero runs of r) and ri1 (one runs).
// Same result as orbit_ring(a, r ^ (bit i | bit j)).
__host__ __device__ inline OrbitResult orbit_ring_hop(const RingOrbitArgs &a, const RunInfo &ri0,
                                                      const RunInfo &ri1, uint64_t r, int i) {
  const int n = a.n;
  const int j = i + 1 == n ? 0 : i + 1;
  const uint64_t x = r ^ (bit_at(i) | bit_at(j));
  uint64_t t0 = 0, t1 = 0;
  int l0 = 0, l1 = -1;
  const bool ok = hop_tops_pair(ri0, ri1, r, i, n, a.mask, a.has_flip, t0, l0, t1, l1);
  if (!ok) return orbit_ring(a, x);
  return orbit_ring_tops(a, x, t0, l0, t1, l1);
}

// SplitMix64 finaliser (Steele, Lea & Flood 2014): the counter hash of the
// random start vectors and the key hash of the hash index.
__host__ __device__ inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// ---------------------------------------------------------------------------
// Index of a representative.  Two forms:
//  * two-level bucket table + branchless binary search inside the bucket.
//    Representatives cluster (orbit minima start with long zero runs), so a
//    table on the top bits alone leaves tens of entries per bucket.  Level 1
//    splits the value range by its top bits into segments (v >> shift);
//    segment s owns 2^lb_s level-2 buckets, lb_s chosen from its count, and
//    dir[s] = (base_s, local shift) (dir_entry).  Level 2: buckets[b] = first i with
//    reps[i] >= start value of bucket b, in global order, so bucket b ends
//    where b + 1 starts (also across segments); buckets[nbuckets] = D.
//  * open-addressing hash table (slots != nullptr): 2^H 8-byte slots, H =
//    ceil(log2 D) + 1 (load <= 1/2), linear probing from the low H bits of
//    h = mix64(rep); a slot holds (h >> 32) << 32 | index, or kEmptySlot.  A
//    lookup reads one slot line for the fingerprint, then reps[index] to
//    confirm, which the row gather issues together with the x gather.
//  * combinatorial rank (rank_mode != 0), for complete bases whose group has
//    no permutation (the identity, with or without the global spin flip).
//    Their representatives are every state of the sector in ascending
//    order: all states (index = v); all weight-w states; or, with the flip
//    (w = n/2 or no weight), the states with site n-1 down.  Ascending order
//    of weight-w states is colexicographic order of their up-site sets, so
//    index(v) = sum_i C(p_i, i + 1) over the up sites p_0 < p_1 < ... (the
//    combinatorial number system; Knuth, TAOCP 7.2.1.3, Theorem L).  A
//    lookup is w table reads (C(p, i + 1) at rank_tab[i * 64 + p], w <= 32)
//    with no dependent loads, where the searches above chain three to five;
//    kernels that can spare 8 KB of LDS stage the table there.
// ---------------------------------------------------------------------------
constexpr uint64_t kEmptySlot = ~0ull;

// Level-1 entry: the segment's first level-2 bucket (below 2^56) and its
// local shift, packed as x = base bits 0..31, y = base bits 32..55 << 8 | lsh.
__host__ __device__ inline uint2 dir_entry(uint64_t base, uint32_t lsh) {
  uint2 d;
  d.x = static_cast<uint32_t>(base);
  d.y = static_cast<uint32_t>(base >> 32) << 8 | lsh;
  return d;
}
__host__ __device__ inline uint64_t dir_base(uint2 d) {
  return static_cast<uint64_t>(d.y >> 8) << 32 | d.x;
}
__host__ __device__ inline uint32_t dir_lsh(uint2 d) { return d.y & 0xffu; }

struct IndexArgs {
  const uint64_t *__restrict__ reps;
  const uint32_t *__restrict__ buckets; // level 2, nbuckets + 1 entries
  const uint2 *__restrict__ dir;        // level 1: (base, local shift) per segment
  uint64_t nseg;
  uint64_t segmask;                     // (1 << shift) - 1
  int32_t shift;                        // segment = v >> shift
  const uint64_t *__restrict__ slots;
  uint64_t hmask;
  // 1 when every valid orbit minimum the operator can reach is in reps (a
  // basis enumerated or verified complete, an operator that keeps the
  // Hamming weight): a bucket search then narrows to the entry without the
  // confirming load of reps.
  int32_t exact;
  // combinatorial rank: 0 off, 1 colex rank of a weight-rank_w state, 2 the
  // state itself; rank_check: reject states whose weight is not rank_w
  int32_t rank_mode;
  int32_t rank_w;
  int32_t rank_check;
  // C(p, i + 1) at [i * 64 + p], i < rank_w, clamped to 2^32 - 1: a state
  // of the sector only meets entries below its index, and indices are below
  // the dimension (< 2^32 - 1)
  const uint32_t *__restrict__ rank_tab;
  uint64_t rank_dim; // dimension: a sum at or above it is no index
};

// Largest weight of the rank index (its table: rank_w x 64 entries).
constexpr int kRankMaxW = 32;

// Up site of v with the lowest index; 63 once v is empty, so that a state
// of the wrong weight can never read outside the table.
__host__ __device__ inline int rank_site(uint64_t v) { return ctz64(v | (1ull << 63)); }

// Indices of va and vb in a rank-indexed basis (-1: not a representative,
// or use_x false).  The two sums run in one loop, so the table reads of
// both terms are in flight together.
__host__ __device__ inline void rank_of2(const IndexArgs &t, uint64_t va, bool use_a, uint64_t vb,
                                         bool use_b, int64_t &ka, int64_t &kb) {
  if (t.rank_mode == 2) {
    ka = use_a && va < t.rank_dim ? static_cast<int64_t>(va) : -1;
    kb = use_b && vb < t.rank_dim ? static_cast<int64_t>(vb) : -1;
    return;
  }
  if (t.rank_check) {
    use_a = use_a && __builtin_popcountll(va) == t.rank_w;
    use_b = use_b && __builtin_popcountll(vb) == t.rank_w;
  }
  uint64_t ra = 0, rb = 0;
  for (int i = 0; i < t.rank_w; ++i) {
    ra += t.rank_tab[i * 64 + rank_site(va)];
    rb += t.rank_tab[i * 64 + rank_site(vb)];
    va &= va - 1;
    vb &= vb - 1;
  }
  ka = use_a && ra < t.rank_dim ? static_cast<int64_t>(ra) : -1;
  kb = use_b && rb < t.rank_dim ? static_cast<int64_t>(rb) : -1;
}

// The rank index serves only bases without permutations (orbit method 0):
// kernels of the ring method leave its code out.
template <int METHOD> __host__ __device__ inline IndexArgs index_for_method(IndexArgs t) {
  if (METHOD != 0) t.rank_mode = 0;
  return t;
}

__host__ __device__ inline int64_t rank_of(const IndexArgs &t, uint64_t v) {
  int64_t k = -1, unused = -1;
  rank_of2(t, v, true, v, false, k, unused);
  return k;
}

// Level-2 bucket of v, or false when v lies beyond the last segment.
__host__ __device__ inline bool bucket_of(const IndexArgs &t, uint64_t v, uint64_t &b) {
  const uint64_t s = v >> t.shift;
  if (s >= t.nseg) return false;
  const uint2 d = t.dir[s];
  b = dir_base(d) + ((v & t.segmask) >> dir_lsh(d));
  return true;
}

// First slot index whose fingerprint matches v (its stored index is a
// candidate, to be confirmed with reps), or -1 when the probe meets an empty
// slot.  *pos is the slot where the probe stopped.
__host__ __device__ inline int64_t hash_probe_from(const IndexArgs &t, uint64_t v, uint64_t s,
                                                   uint64_t *pos) {
  const uint32_t fp = static_cast<uint32_t>(mix64(v) >> 32);
  for (uint64_t step = 0; step <= t.hmask; ++step) {
    const uint64_t e = t.slots[s];
    if (e == kEmptySlot) break;
    if (static_cast<uint32_t>(e >> 32) == fp) {
      *pos = s;
      return static_cast<int64_t>(e & 0xffffffffull);
    }
    s = (s + 1) & t.hmask;
  }
  *pos = s;
  return -1;
}

__host__ __device__ inline int64_t hash_probe(const IndexArgs &t, uint64_t v, uint64_t *pos) {
  return hash_probe_from(t, v, mix64(v) & t.hmask, pos);
}

// Exact lookup: continue past fingerprint matches that are not v.
__host__ __device__ inline int64_t hash_find(const IndexArgs &t, uint64_t v) {
  uint64_t pos = 0;
  int64_t k = hash_probe(t, v, &pos);
  while (k >= 0 && t.reps[k] != v) k = hash_probe_from(t, v, (pos + 1) & t.hmask, &pos);
  return k;
}

__host__ __device__ inline int64_t find_index(const IndexArgs &t, uint64_t v) {
  if (t.rank_mode) return rank_of(t, v);
  if (t.slots) return hash_find(t, v);
  uint64_t b = 0;
  if (!bucket_of(t, v, b)) return -1;
  uint32_t lo = t.buckets[b];
  uint32_t n = t.buckets[b + 1] - lo;
  if (n == 0) return -1;
  while (n > 1) {
    const uint32_t half = n >> 1;
    lo = (t.reps[lo + half] <= v) ? lo + half : lo;
    n -= half;
  }
  if (t.exact) return static_cast<int64_t>(lo);
  return t.reps[lo] == v ? static_cast<int64_t>(lo) : -1;
}

// Two independent searches advanced in lockstep, so a lane keeps two
// dependent load chains in flight instead of one.  use_a / use_b switch a
// search off (its result is then -1).  Exact for every index form.
__host__ __device__ inline void find_index2(const IndexArgs &t, uint64_t va, bool use_a,
                                            uint64_t vb, bool use_b, int64_t &ka, int64_t &kb) {
  if (t.rank_mode) {
    rank_of2(t, va, use_a, vb, use_b, ka, kb);
    return;
  }
  if (t.slots) {
    uint64_t pa = 0, pb = 0;
    ka = use_a ? hash_probe(t, va, &pa) : -1;
    kb = use_b ? hash_probe(t, vb, &pb) : -1;
    const bool bad_a = ka >= 0 && t.reps[ka] != va;
    const bool bad_b = kb >= 0 && t.reps[kb] != vb;
    if (bad_a) ka = hash_find(t, va);
    if (bad_b) kb = hash_find(t, vb);
    return;
  }
  uint64_t ba = 0, bb = 0;
  use_a = use_a && bucket_of(t, va, ba);
  use_b = use_b && bucket_of(t, vb, bb);
  uint32_t loa = 0, na = 0, lob = 0, nb = 0;
  if (use_a) {
    loa = t.buckets[ba];
    na = t.buckets[ba + 1] - loa;
  }
  if (use_b) {
    lob = t.buckets[bb];
    nb = t.buckets[bb + 1] - lob;
  }
  while (na > 1 || nb > 1) {
    if (na > 1) {
      const uint32_t h = na >> 1;
      loa = (t.reps[loa + h] <= va) ? loa + h : loa;
      na -= h;
    }
    if (nb > 1) {
      const uint32_t h = nb >> 1;
      lob = (t.reps[lob + h] <= vb) ? lob + h : lob;
      nb -= h;
    }
  }
  if (t.exact) {
    ka = na == 1 ? static_cast<int64_t>(loa) : -1;
    kb = nb == 1 ? static_cast<int64_t>(lob) : -1;
    return;
  }
  ka = (na == 1 && t.reps[loa] == va) ? static_cast<int64_t>(loa) : -1;
  kb = (nb == 1 && t.reps[lob] == vb) ? static_cast<int64_t>(lob) : -1;
}

// Counter-hash start vector: v[r] = U(-1, 1) from (seed, r); identical on
// host and device.
__host__ __device__ inline double hash_uniform(uint64_t seed, uint64_t r) {
  const uint64_t h = mix64(seed ^ mix64(r));
  return 2.0 * (static_cast<double>(h >> 11) * 0x1.0p-53) - 1.0;
}

} // namespace spined
