// matvec.hip -- matrix-free H*x row gather (ls_operator_matmat,
// src/SpinED/Internal.hs:377-378, :411-429).
//
// For representative r (row i) with |Stab r| = s_r:
//   y_i = D(r) x_i + sum_sigma conj(c_sigma) chi(h) sqrt(|Stab r'| / s_r) x_{idx(r')}
// where H|r> = sum_sigma c_sigma |sigma>, r' = h.sigma = min orbit(sigma),
// terms whose orbit has inconsistent phases (no symmetric state) or whose
// representative is not in the basis (another magnetisation sector) are
// dropped.  One lane owns one row and writes it once: no atomics.
//
// Per row, pass A walks the interaction tuples once: it accumulates the
// diagonal and marks the off-diagonal templates whose local state matches
// (a 64-bit word per 64 templates).  Pass B pops the marked templates in
// index order -- (tuple, a, b) order, the oracle's summation order -- and
// for each does orbit minimum, bucket + binary search and one fma gather.
// Lanes of a wave run pass B for max(#active) iterations; the orbit method
// (generic programs or ring run pruning) is a template parameter.
#include <algorithm>
#include <cmath>
#include <vector>

#include "engine.hpp"
#include "rowgather.hpp"
#include "scan.hpp"

namespace spined {

namespace {

constexpr int kBlock = 256;

// ---- real rows -------------------------------------------------------------
// Row of slot i: identity, or the tile's schedule (rows of equal work side
// by side in a wave).
__device__ inline uint64_t scheduled_row(const uint16_t *__restrict__ sched, uint64_t i) {
  return sched ? (i & ~static_cast<uint64_t>(kScheduleTile - 1)) + sched[i] : i;
}

// One wave per workgroup: rows take unequal time, and a one-wave group frees
// its slot as soon as its own rows are done, which keeps more waves resident
// than four-wave groups that hold their slots until the slowest wave ends.
constexpr int kRowBlock = 64;

template <int METHOD, int NW, bool BONDS, class XT, class YT>
__global__ void __launch_bounds__(kRowBlock)
    matvec_real_kernel(OpArgs o, BondArgs bd, RatioArgs rt, GenericOrbitArgs ga,
                       RingOrbitArgs ra, IndexArgs ix, const uint32_t *__restrict__ stab,
                       uint64_t D, const XT *__restrict__ x, YT *__restrict__ y,
                       const uint16_t *__restrict__ sched) {
  // bond coefficients in LDS: a per-term global load would have to wait for
  // every gather issued before it (vector loads complete in order)
  __shared__ double so[BONDS ? 64 * NW : 1];
  if constexpr (BONDS) {
    for (int k = threadIdx.x; k < bd.nbonds; k += kRowBlock) so[k] = bd.o[k];
    __syncthreads();
  }
  const uint64_t lb = xcd_block(blockIdx.x, gridDim.x);
  const uint64_t slot = lb * kRowBlock + threadIdx.x;
  if (slot >= D) return;
  const uint64_t i = scheduled_row(sched, slot);
  st_re(y, i,
        gather_row_real<METHOD, NW, BONDS>(o, bd, rt, ga, ra, ix, stab, i, x, BONDS ? so : nullptr));
}

// matvec_real_kernel with L lanes per row (64 / L rows per wave): the same
// sums, the row's templates spread over its lanes (gather_row_real_lanes).
template <int METHOD, int NW, int L, class XT, class YT>
__global__ void __launch_bounds__(kRowBlock)
    matvec_real_lanes_kernel(OpArgs o, BondArgs bd, RatioArgs rt, GenericOrbitArgs ga,
                             RingOrbitArgs ra, IndexArgs ix, const uint32_t *__restrict__ stab,
                             uint64_t D, const XT *__restrict__ x, YT *__restrict__ y,
                             const uint16_t *__restrict__ sched) {
  __shared__ double so[64 * NW];
  for (int k = threadIdx.x; k < bd.nbonds; k += kRowBlock) so[k] = bd.o[k];
  __syncthreads();
  const uint64_t lb = xcd_block(blockIdx.x, gridDim.x);
  // a row's L lanes share the exit: the group's shuffles stay within it
  const uint64_t slot = lb * (kRowBlock / L) + threadIdx.x / L;
  if (slot >= D) return;
  const int sub = static_cast<int>(threadIdx.x % L);
  const uint64_t i = scheduled_row(sched, slot);
  const double v =
      gather_row_real_lanes<METHOD, NW, true, L>(o, bd, rt, ga, ra, ix, stab, i, x, so, sub);
  if (sub == 0) st_re(y, i, v);
}

// ---- gather cache -----------------------------------------------------------
constexpr uint32_t kCacheIndex = 0x3FFFFFFFu; // index bits; all ones: no term
constexpr uint32_t kCacheNeg = 1u << 30;      // negative coefficient
constexpr uint32_t kCacheRatio = 1u << 31;    // sqrt(|Stab target| / |Stab row|) != 1

// Terms of the busiest row of each wave (one wave = 64 slots).
template <int NW, bool BONDS>
__global__ void __launch_bounds__(kRowBlock)
    cache_count_kernel(OpArgs o, BondArgs bd, const uint64_t *__restrict__ reps, uint64_t D,
                       const uint16_t *__restrict__ sched, uint32_t *__restrict__ wave_terms) {
  const uint64_t slot = static_cast<uint64_t>(blockIdx.x) * kRowBlock + threadIdx.x;
  uint32_t c = 0;
  if (slot < D) {
    uint64_t act[NW];
    (void)row_pass_a<NW, BONDS>(o, bd, reps[scheduled_row(sched, slot)], act);
#pragma unroll
    for (int w = 0; w < NW; ++w) c += static_cast<uint32_t>(__popcll(act[w]));
  }
  for (int off = 32; off > 0; off >>= 1) c = max(c, static_cast<uint32_t>(__shfl_xor(c, off, 64)));
  if (threadIdx.x == 0) wave_terms[blockIdx.x] = c;
}

// Entries of every term, in the order and with the orbits and lookups of
// gather_row_real.
template <int METHOD, int NW, bool BONDS>
__global__ void __launch_bounds__(kRowBlock)
    cache_fill_kernel(OpArgs o, BondArgs bd, GenericOrbitArgs ga, RingOrbitArgs ra, IndexArgs ix,
                      const uint32_t *__restrict__ stab, uint64_t D,
                      const uint16_t *__restrict__ sched, const uint64_t *__restrict__ woff,
                      uint32_t *__restrict__ entries) {
  ix = index_for_method<METHOD>(ix);
  const uint64_t slot = static_cast<uint64_t>(blockIdx.x) * kRowBlock + threadIdx.x;
  if (slot >= D) return;
  const uint64_t i = scheduled_row(sched, slot);
  const uint64_t r = ix.reps[i];
  uint64_t act[NW];
  (void)row_pass_a<NW, BONDS>(o, bd, r, act);
  const RowOrbit<METHOD, BONDS> rob(ra, bd, r);
  const int32_t si = static_cast<int32_t>(stab[i]);
  uint32_t *out = entries + woff[blockIdx.x] * 64 + threadIdx.x;
  // generic groups with trivial characters: the orbit minimum alone (no
  // phase can differ), |Stab| of the target from the basis' entry
  const bool min_only = METHOD == 0 && ga.trivial;
  auto entry = [&](int64_t k, const OrbitResult &orb) {
    if (k < 0) return kCacheIndex;
    const int32_t cnt = min_only ? static_cast<int32_t>(stab[k]) : orb.count;
    return static_cast<uint32_t>(k) | (orb.phase != 0 ? kCacheNeg : 0u) |
           (cnt != si ? kCacheRatio : 0u);
  };
  auto orbit_of = [&](int j) {
    if (min_only)
      return OrbitResult{orbit_min_generic(ga, r ^ (BONDS ? bd.flip[j] : o.flip[j])), 0, 0, true};
    return rob.orbit(o, bd, ga, ra, r, j);
  };
  int t = 0;
  while (any_set<NW>(act)) {
    const int ja = pop_first<NW>(act);
    const int jb = pop_first<NW>(act);
    const OrbitResult oa = orbit_of(ja);
    OrbitResult ob{0, 0, 0, false};
    if (jb >= 0) ob = orbit_of(jb);
    int64_t ka, kb;
    find_index2(ix, oa.rep, oa.valid, ob.rep, ob.valid, ka, kb);
    out[64 * static_cast<uint64_t>(t++)] = entry(ka, oa);
    if (jb >= 0) out[64 * static_cast<uint64_t>(t++)] = entry(kb, ob);
  }
}

// Real matvec from the gather cache: the sums of matvec_real_kernel (same
// terms, coefficients and order) without orbits or index searches.
template <int NW, bool BONDS, class XT, class YT>
__global__ void __launch_bounds__(kRowBlock)
    matvec_cached_kernel(OpArgs o, BondArgs bd, RatioArgs rt, const uint64_t *__restrict__ reps,
                         const uint32_t *__restrict__ stab, uint64_t D,
                         const XT *__restrict__ x, YT *__restrict__ y,
                         const uint16_t *__restrict__ sched, const uint64_t *__restrict__ woff,
                         const uint32_t *__restrict__ entries) {
  __shared__ double so[BONDS ? 64 * NW : 1];
  if constexpr (BONDS) {
    for (int k = threadIdx.x; k < bd.nbonds; k += kRowBlock) so[k] = bd.o[k];
    __syncthreads();
  }
  // Workgroup order: with FP64 x each XCD takes a contiguous eighth of the
  // rows (xcd_block); with FP32 x the dispatch order itself (consecutive
  // waves on different XCDs) measured faster in back-to-back launches at
  // D = 8.6e8 (85 vs 89 ms, against 98 vs 102 ms the other way round for
  // FP64 x).
  const uint64_t lb = sizeof(XT) == 8 ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
  const uint64_t slot = lb * kRowBlock + threadIdx.x;
  if (slot >= D) return;
  const uint64_t i = scheduled_row(sched, slot);
  uint64_t act[NW];
  const double diag = row_pass_a<NW, BONDS>(o, bd, reps[i], act);
  double acc = diag * ld_re(x, i);
  const uint32_t si = stab[i];
  const double sr = static_cast<double>(si);
  const uint32_t *in = entries + woff[lb] * 64 + threadIdx.x;
  // four terms per step: four x gathers in flight together with the next
  // step's four entry loads, then the four terms added in template order
  constexpr int U = 4;
  int t = 0;
  int j[U];
  uint32_t e[U];
  bool more = any_set<NW>(act);
#pragma unroll
  for (int q = 0; q < U; ++q) {
    j[q] = pop_first<NW>(act);
    e[q] = j[q] >= 0 ? in[64 * static_cast<uint64_t>(t + q)] : kCacheIndex;
  }
  t += U;
  while (more) {
    double xv[U];
#pragma unroll
    for (int q = 0; q < U; ++q) {
      const uint32_t k = e[q] & kCacheIndex;
      xv[q] = k != kCacheIndex ? ld_re(x, k) : 0.0;
    }
    more = any_set<NW>(act);
    int jn[U];
    uint32_t en[U];
#pragma unroll
    for (int q = 0; q < U; ++q) {
      jn[q] = pop_first<NW>(act);
      en[q] = jn[q] >= 0 ? in[64 * static_cast<uint64_t>(t + q)] : kCacheIndex;
    }
    t += U;
#pragma unroll
    for (int q = 0; q < U; ++q) {
      const uint32_t k = e[q] & kCacheIndex;
      if (k == kCacheIndex) continue;
      double f = 1.0;
      if (e[q] & kCacheRatio) {
        const uint32_t ck = stab[k];
        f = rt.nd ? rt.ratio[rt.div_idx[ck] * rt.nd + rt.div_idx[si]]
                  : sqrt(static_cast<double>(ck) / sr);
      }
      double c = (BONDS ? so[j[q]] : o.coef_re[j[q]]) * f;
      c = (e[q] & kCacheNeg) ? -c : c;
      acc = fma(c, xv[q], acc);
    }
#pragma unroll
    for (int q = 0; q < U; ++q) {
      j[q] = jn[q];
      e[q] = en[q];
    }
  }
  st_re(y, i, acc);
}

// Row-interleaved block of K vectors from the gather cache: per column the
// sums of matvec_block_kernel (and so of matvec_real_kernel).
template <int NW, bool BONDS, int K, class XT, class YT>
__global__ void __launch_bounds__(kRowBlock)
    matvec_cached_block_kernel(OpArgs o, BondArgs bd, RatioArgs rt,
                               const uint64_t *__restrict__ reps, const uint32_t *__restrict__ stab,
                               uint64_t D, const XT *__restrict__ x, YT *__restrict__ y,
                               const uint16_t *__restrict__ sched, const uint64_t *__restrict__ woff,
                               const uint32_t *__restrict__ entries) {
  __shared__ double so[BONDS ? 64 * NW : 1];
  if constexpr (BONDS) {
    for (int k = threadIdx.x; k < bd.nbonds; k += kRowBlock) so[k] = bd.o[k];
    __syncthreads();
  }
  using Row = RowK<XT, K>;
  const Row *__restrict__ xr = reinterpret_cast<const Row *>(x);
  const uint64_t lb = xcd_block(blockIdx.x, gridDim.x);
  const uint64_t slot = lb * kRowBlock + threadIdx.x;
  if (slot >= D) return;
  const uint64_t i = scheduled_row(sched, slot);
  uint64_t act[NW];
  const double diag = row_pass_a<NW, BONDS>(o, bd, reps[i], act);
  double acc[K];
  {
    const Row xi = xr[i];
#pragma unroll
    for (int c = 0; c < K; ++c) acc[c] = diag * static_cast<double>(xi.v[c]);
  }
  const uint32_t si = stab[i];
  const double sr = static_cast<double>(si);
  const uint32_t *in = entries + woff[lb] * 64 + threadIdx.x;
  int t = 0;
  while (any_set<NW>(act)) {
    const int j = pop_first<NW>(act);
    const uint32_t e = in[64 * static_cast<uint64_t>(t++)];
    const uint32_t k = e & kCacheIndex;
    if (k == kCacheIndex) continue;
    const Row xk = xr[k];
    double f = 1.0;
    if (e & kCacheRatio) {
      const uint32_t ck = stab[k];
      f = rt.nd ? rt.ratio[rt.div_idx[ck] * rt.nd + rt.div_idx[si]]
                : sqrt(static_cast<double>(ck) / sr);
    }
    double cf = (BONDS ? so[j] : o.coef_re[j]) * f;
    cf = (e & kCacheNeg) ? -cf : cf;
#pragma unroll
    for (int c = 0; c < K; ++c) acc[c] = fma(cf, static_cast<double>(xk.v[c]), acc[c]);
  }
  RowK<YT, K> out;
#pragma unroll
  for (int c = 0; c < K; ++c) out.v[c] = static_cast<YT>(acc[c]);
  reinterpret_cast<RowK<YT, K> *>(y)[i] = out;
}

// Row-interleaved block of K real vectors (x[row * K + c]); per column the
// same sums as matvec_real_kernel.
template <int METHOD, int NW, bool BONDS, int K, class XT, class YT>
__global__ void __launch_bounds__(kBlock)
    matvec_block_kernel(OpArgs o, BondArgs bd, RatioArgs rt, GenericOrbitArgs ga,
                        RingOrbitArgs ra, IndexArgs ix, const uint32_t *__restrict__ stab,
                        uint64_t D, const XT *__restrict__ x, YT *__restrict__ y,
                        const uint16_t *__restrict__ sched) {
  const uint64_t lb = xcd_block(blockIdx.x, gridDim.x);
  const uint64_t slot = lb * kBlock + threadIdx.x;
  if (slot >= D) return;
  const uint64_t i = scheduled_row(sched, slot);
  double acc[K];
  gather_row_real_block<METHOD, NW, BONDS, K>(o, bd, rt, ga, ra, ix, stab, i, x, acc);
  RowK<YT, K> out;
#pragma unroll
  for (int c = 0; c < K; ++c) out.v[c] = static_cast<YT>(acc[c]);
  reinterpret_cast<RowK<YT, K> *>(y)[i] = out;
}

// Column-major block (column c at x + c * stride) <-> row-interleaved block
// (xr[row * K + c]); one thread per row.
template <class A, class B>
__global__ void __launch_bounds__(kBlock)
    cols_to_rows_kernel(const A *__restrict__ x, uint64_t stride, uint64_t D, int K,
                        B *__restrict__ xr) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= D) return;
  for (int c = 0; c < K; ++c) xr[i * K + c] = static_cast<B>(x[c * stride + i]);
}

template <class A, class B>
__global__ void __launch_bounds__(kBlock)
    rows_to_cols_kernel(const A *__restrict__ yr, uint64_t D, int K, B *__restrict__ y,
                        uint64_t stride) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= D) return;
  for (int c = 0; c < K; ++c) y[c * stride + i] = static_cast<B>(yr[i * K + c]);
}

// Row schedule of one tile (one workgroup per tile): count the active
// templates of each row (pass A), then a counting sort by count, stable in
// the row order (thread 0 scatters, so the schedule is deterministic).
template <int NW, bool BONDS>
__global__ void __launch_bounds__(kBlock)
    schedule_kernel(OpArgs o, BondArgs bd, const uint64_t *__restrict__ reps, uint64_t D,
                    uint16_t *__restrict__ sched) {
  constexpr int kBins = 64 * NW + 1;
  __shared__ uint16_t count_of[kScheduleTile];
  __shared__ uint32_t offs[kBins];
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kScheduleTile;
  const int rows = static_cast<int>(D - base < kScheduleTile ? D - base : kScheduleTile);
  for (int t = threadIdx.x; t < kBins; t += kBlock) offs[t] = 0;
  __syncthreads();
  for (int k = threadIdx.x; k < rows; k += kBlock) {
    uint64_t act[NW];
    (void)row_pass_a<NW, BONDS>(o, bd, reps[base + k], act);
    int c = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) c += __popcll(act[w]);
    count_of[k] = static_cast<uint16_t>(c);
    atomicAdd(&offs[c], 1u);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = 0;
    for (int c = 0; c < kBins; ++c) {
      const uint32_t h = offs[c];
      offs[c] = run;
      run += h;
    }
    for (int k = 0; k < rows; ++k) sched[base + offs[count_of[k]]++] = static_cast<uint16_t>(k);
  }
}

// Rows with more than 256 off-diagonal templates: per-tuple loop.
template <int METHOD, class XT, class YT>
__global__ void __launch_bounds__(kBlock)
    matvec_real_direct_kernel(OpArgs o, GenericOrbitArgs ga, RingOrbitArgs ra, IndexArgs ix,
                              const uint32_t *__restrict__ stab, uint64_t D,
                              const XT *__restrict__ x, YT *__restrict__ y) {
  ix = index_for_method<METHOD>(ix);
  const uint64_t lb = xcd_block(blockIdx.x, gridDim.x);
  const uint64_t i = lb * kBlock + threadIdx.x;
  if (i >= D) return;
  const uint64_t r = ix.reps[i];
  double diag = 0.0;
  for (int t = 0; t < o.ntuples; ++t) {
    const int a = local_state(o.tuple_sites[t], static_cast<int>(o.tuple_n[t]), r);
    diag += o.diag_re[o.tuple_off[t] + a];
  }
  double acc = diag * ld_re(x, i);
  const double sr = static_cast<double>(stab[i]);
  for (int t = 0; t < o.ntuples; ++t) {
    const uint32_t off = o.tuple_off[t];
    const int a = local_state(o.tuple_sites[t], static_cast<int>(o.tuple_n[t]), r);
    const uint32_t b = o.range[off + t + a], e = o.range[off + t + a + 1];
    for (uint32_t j = b; j < e; ++j) {
      const OrbitResult orb = do_orbit<METHOD>(ga, ra, r ^ o.flip[j]);
      if (!orb.valid) continue;
      const int64_t k = find_index(ix, orb.rep);
      if (k < 0) continue;
      double coef = o.coef_re[j] * sqrt(static_cast<double>(orb.count) / sr);
      if (orb.phase != 0) coef = -coef;
      acc = fma(coef, ld_re(x, static_cast<uint64_t>(k)), acc);
    }
  }
  st_re(y, i, acc);
}

// ---- complex rows ------------------------------------------------------------
template <int METHOD, class XT, class YT>
__global__ void __launch_bounds__(kBlock)
    matvec_complex_kernel(OpArgs o, GenericOrbitArgs ga, RingOrbitArgs ra, IndexArgs ix,
                          const uint32_t *__restrict__ stab, uint64_t D,
                          const XT *__restrict__ x, YT *__restrict__ y) {
  ix = index_for_method<METHOD>(ix);
  const uint64_t lb = xcd_block(blockIdx.x, gridDim.x);
  const uint64_t i = lb * kBlock + threadIdx.x;
  if (i >= D) return;
  const uint64_t r = ix.reps[i];
  double dre = 0.0, dim = 0.0;
  for (int t = 0; t < o.ntuples; ++t) {
    const int a = local_state(o.tuple_sites[t], static_cast<int>(o.tuple_n[t]), r);
    dre += o.diag_re[o.tuple_off[t] + a];
    dim += o.diag_im[o.tuple_off[t] + a];
  }
  const double2 xi = ld_c(x, i);
  double are = fma(dre, xi.x, -(dim * xi.y));
  double aim = fma(dre, xi.y, dim * xi.x);
  const double sr = static_cast<double>(stab[i]);
  for (int t = 0; t < o.ntuples; ++t) {
    const uint32_t off = o.tuple_off[t];
    const int a = local_state(o.tuple_sites[t], static_cast<int>(o.tuple_n[t]), r);
    const uint32_t b = o.range[off + t + a], e = o.range[off + t + a + 1];
    for (uint32_t j = b; j < e; ++j) {
      const OrbitResult orb = do_orbit<METHOD>(ga, ra, r ^ o.flip[j]);
      if (!orb.valid) continue;
      const int64_t k = find_index(ix, orb.rep);
      if (k < 0) continue;
      const double f = sqrt(static_cast<double>(orb.count) / sr);
      const double cre = o.phase_cos[orb.phase], cim = o.phase_sin[orb.phase];
      const double mre = o.coef_re[j], mim = o.coef_im[j];
      const double kre = fma(mre, cre, -(mim * cim)) * f;
      const double kim = fma(mre, cim, mim * cre) * f;
      const double2 xk = ld_c(x, static_cast<uint64_t>(k));
      are = fma(kre, xk.x, are);
      are = fma(-kim, xk.y, are);
      aim = fma(kre, xk.y, aim);
      aim = fma(kim, xk.x, aim);
    }
  }
  st_c(y, i, make_double2(are, aim));
}

// ---- counting (measurement) --------------------------------------------------
template <int METHOD>
__global__ void __launch_bounds__(kBlock)
    count_offdiag_kernel(OpArgs o, GenericOrbitArgs ga, RingOrbitArgs ra, IndexArgs ix,
                         uint64_t D, unsigned long long *__restrict__ partial) {
  ix = index_for_method<METHOD>(ix);
  __shared__ unsigned long long lds[kBlock / 64];
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
  unsigned long long c = 0;
  if (i < D) {
    const uint64_t r = ix.reps[i];
    for (int t = 0; t < o.ntuples; ++t) {
      const uint32_t off = o.tuple_off[t];
      const int a = local_state(o.tuple_sites[t], static_cast<int>(o.tuple_n[t]), r);
      const uint32_t b = o.range[off + t + a], e = o.range[off + t + a + 1];
      for (uint32_t j = b; j < e; ++j) {
        const OrbitResult orb = do_orbit<METHOD>(ga, ra, r ^ o.flip[j]);
        if (orb.valid && find_index(ix, orb.rep) >= 0) ++c;
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off, 64);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long s = 0;
    for (int w = 0; w < kBlock / 64; ++w) s += lds[w];
    partial[blockIdx.x] = s;
  }
}

struct PhaseTable {
  DeviceBuffer<double> cos_t, sin_t;
  int L = 0;
};

// chi(phase) = exp(2 pi i phase / L), exact at quarter turns (same values
// as the oracle's phase_to_complex).
void phase_values(int64_t p, int64_t L, double &c, double &s) {
  p = ((p % L) + L) % L;
  if (p == 0) { c = 1.0; s = 0.0; }
  else if (2 * p == L) { c = -1.0; s = 0.0; }
  else if (4 * p == L) { c = 0.0; s = 1.0; }
  else if (4 * p == 3 * L) { c = 0.0; s = -1.0; }
  else {
    const double ang = 2.0 * M_PI * static_cast<double>(p) / static_cast<double>(L);
    c = std::cos(ang);
    s = std::sin(ang);
  }
}

OpArgs op_args(const Operator &op, const double *pc, const double *ps) {
  const auto &t = op.tables;
  OpArgs a{};
  a.tuple_sites = t.tuple_sites.ptr;
  a.tuple_n = t.tuple_n.ptr;
  a.tuple_off = t.tuple_off.ptr;
  a.range = t.range.ptr;
  a.diag_re = t.diag_re.ptr;
  a.diag_im = t.diag_im.ptr;
  a.flip = t.flip.ptr;
  a.coef_re = t.coef_re.ptr;
  a.coef_im = t.coef_im.ptr;
  a.phase_cos = pc;
  a.phase_sin = ps;
  a.ntuples = t.ntuples;
  a.ntemplates = t.ntemplates;
  return a;
}

unsigned grid_for(uint64_t n) { return static_cast<unsigned>((n + kBlock - 1) / kBlock); }

// Everything a launch needs, gathered once per call.
struct Ctx {
  OpArgs o;
  BondArgs bd;
  RatioArgs rt;
  GenericOrbitArgs ga;
  RingOrbitArgs ra;
  IndexArgs ix;
  const uint32_t *stab;
  uint64_t D;
  bool bonds;
  const uint16_t *sched;
  int lanes; // lanes per row of the exchange-bond matrix-free kernel (1, 2, 4)
};

template <int METHOD, int NW, bool BONDS, class XT, class YT>
void launch_real_nw(const Ctx &c, const XT *x, YT *y, hipStream_t s) {
  if constexpr (BONDS) {
    if (c.lanes == 2 || c.lanes == 4) {
      const uint64_t rows = kRowBlock / c.lanes;
      const unsigned g = static_cast<unsigned>((c.D + rows - 1) / rows);
      if (c.lanes == 2)
        matvec_real_lanes_kernel<METHOD, NW, 2, XT, YT><<<g, kRowBlock, 0, s>>>(
            c.o, c.bd, c.rt, c.ga, c.ra, c.ix, c.stab, c.D, x, y, c.sched);
      else
        matvec_real_lanes_kernel<METHOD, NW, 4, XT, YT><<<g, kRowBlock, 0, s>>>(
            c.o, c.bd, c.rt, c.ga, c.ra, c.ix, c.stab, c.D, x, y, c.sched);
      return;
    }
  }
  const unsigned g = static_cast<unsigned>((c.D + kRowBlock - 1) / kRowBlock);
  matvec_real_kernel<METHOD, NW, BONDS, XT, YT><<<g, kRowBlock, 0, s>>>(
      c.o, c.bd, c.rt, c.ga, c.ra, c.ix, c.stab, c.D, x, y, c.sched);
}

template <int METHOD, class XT, class YT>
void launch_real(const Ctx &c, const XT *x, YT *y, hipStream_t s) {
  const int units = c.bonds ? c.bd.nbonds : c.o.ntemplates;
  if (c.bonds && units <= 64)
    launch_real_nw<METHOD, 1, true>(c, x, y, s);
  else if (c.bonds && units <= 128)
    launch_real_nw<METHOD, 2, true>(c, x, y, s);
  else if (!c.bonds && units <= 64)
    launch_real_nw<METHOD, 1, false>(c, x, y, s);
  else if (!c.bonds && units <= 128)
    launch_real_nw<METHOD, 2, false>(c, x, y, s);
  else if (!c.bonds && units <= 256)
    launch_real_nw<METHOD, 4, false>(c, x, y, s);
  else
    matvec_real_direct_kernel<METHOD, XT, YT><<<grid_for(c.D), kBlock, 0, s>>>(
        c.o, c.ga, c.ra, c.ix, c.stab, c.D, x, y);
  SPINED_CHECK_LAUNCH();
}

template <int METHOD, class XT, class YT>
void launch_complex(const Ctx &c, const XT *x, YT *y, hipStream_t s) {
  matvec_complex_kernel<METHOD, XT, YT><<<grid_for(c.D), kBlock, 0, s>>>(c.o, c.ga, c.ra, c.ix,
                                                                        c.stab, c.D, x, y);
  SPINED_CHECK_LAUNCH();
}

template <int METHOD> void dispatch(const MatvecCall &m, const Ctx &c) {
  hipStream_t s = m.stream;
  using VT = VecType;
  if (m.xt == VT::F64 && m.yt == VT::F64)
    launch_real<METHOD>(c, static_cast<const double *>(m.x), static_cast<double *>(m.y), s);
  else if (m.xt == VT::F32 && m.yt == VT::F32)
    launch_real<METHOD>(c, static_cast<const float *>(m.x), static_cast<float *>(m.y), s);
  else if (m.xt == VT::F32 && m.yt == VT::F64)
    launch_real<METHOD>(c, static_cast<const float *>(m.x), static_cast<double *>(m.y), s);
  else if (m.xt == VT::C128 && m.yt == VT::C128)
    launch_complex<METHOD>(c, static_cast<const double2 *>(m.x), static_cast<double2 *>(m.y), s);
  else if (m.xt == VT::C64 && m.yt == VT::C64)
    launch_complex<METHOD>(c, static_cast<const float2 *>(m.x), static_cast<float2 *>(m.y), s);
  else
    throw Error(LS_INVALID_DATATYPE, "unsupported matvec type combination");
}

// Calls f(integral_constant<NW>, bool_constant<BONDS>) for the row-gather
// kernels' word count; false when the operator has too many templates.
template <class F> bool with_nw(const Ctx &c, F &&f) {
  const int units = c.bonds ? c.bd.nbonds : c.o.ntemplates;
  if (c.bonds && units <= 64)
    f(std::integral_constant<int, 1>{}, std::true_type{});
  else if (c.bonds && units <= 128)
    f(std::integral_constant<int, 2>{}, std::true_type{});
  else if (!c.bonds && units <= 64)
    f(std::integral_constant<int, 1>{}, std::false_type{});
  else if (!c.bonds && units <= 128)
    f(std::integral_constant<int, 2>{}, std::false_type{});
  else if (!c.bonds && units <= 256)
    f(std::integral_constant<int, 4>{}, std::false_type{});
  else
    return false;
  return true;
}

// Free memory that an automatically built gather cache leaves untouched,
// and the smallest basis worth one (the build costs about one matvec).
constexpr uint64_t kCacheReserve = uint64_t(4) << 30;
constexpr uint64_t kCacheMinRows = uint64_t(1) << 20;

// Builds the gather cache when it is requested (1), or when it is automatic
// (-1) and fits in free memory with kCacheReserve to spare.  Caller holds
// gc.mu.  True when the cache is ready.
template <int METHOD>
bool ensure_cache(const Operator &op, const Ctx &c, GatherCache &gc, hipStream_t s,
                  uint64_t reserve = kCacheReserve, uint64_t reusable = 0) {
  if (gc.built) return true;
  const Basis &b = *op.basis;
  if (gc.request == 0 || !op.tables.is_real || b.D >= kCacheIndex) return false;
  struct Building {
    std::atomic<bool> &f;
    explicit Building(std::atomic<bool> &x) : f(x) { f.store(true); }
    ~Building() { f.store(false); }
  } building(gc.building);
  if (gc.request < 0 && b.D < kCacheMinRows) return false;
  const uint64_t waves = (b.D + kRowBlock - 1) / kRowBlock;
  const unsigned g = static_cast<unsigned>(waves);
  if (!gc.counted) {
    DeviceBuffer<uint32_t> wt(waves);
    DeviceBuffer<uint64_t> tot(1);
    if (!with_nw(c, [&](auto nw, auto bonds) {
          cache_count_kernel<decltype(nw)::value, decltype(bonds)::value>
              <<<g, kRowBlock, 0, s>>>(c.o, c.bd, b.reps.ptr, b.D, c.sched, wt.ptr);
        }))
      return false;
    SPINED_CHECK_LAUNCH();
    gc.woff.allocate(waves);
    scan_counts(wt.ptr, waves, gc.woff.ptr, tot.ptr, s);
    uint64_t t = 0;
    SPINED_HIP(hipMemcpyAsync(&t, tot.ptr, 8, hipMemcpyDeviceToHost, s));
    SPINED_HIP(hipStreamSynchronize(s));
    gc.total = t * kRowBlock;
    gc.counted = true;
  }
  if (gc.request < 0) {
    size_t free_b = 0, total_b = 0;
    SPINED_HIP(hipMemGetInfo(&free_b, &total_b));
    free_b += pool_bytes(gc.total * 4) > 0 ? gc.total * 4 : 0; // a pooled block of this size
    if (free_b < gc.total * 4 + kCacheReserve || free_b + reusable < gc.total * 4 + reserve)
      return false;
  }
  try {
    // pooled: a rebuild of the same operator's cache takes the block back
    // without a free and a fresh multi-GB mapping
    gc.entries.allocate_pooled(std::max<uint64_t>(gc.total, 1));
  } catch (const Error &) {
    if (gc.request > 0) throw;
    return false;
  }
  with_nw(c, [&](auto nw, auto bonds) {
    cache_fill_kernel<METHOD, decltype(nw)::value, decltype(bonds)::value><<<g, kRowBlock, 0, s>>>(
        c.o, c.bd, c.ga, c.ra, c.ix, c.stab, c.D, c.sched, gc.woff.ptr, gc.entries.ptr);
  });
  SPINED_CHECK_LAUNCH();
  gc.built = true;
  return true;
}

template <class XT, class YT>
void launch_cached(const Ctx &c, const GatherCache &gc, const XT *x, YT *y, hipStream_t s) {
  const unsigned g = static_cast<unsigned>((c.D + kRowBlock - 1) / kRowBlock);
  with_nw(c, [&](auto nw, auto bonds) {
    matvec_cached_kernel<decltype(nw)::value, decltype(bonds)::value, XT, YT>
        <<<g, kRowBlock, 0, s>>>(c.o, c.bd, c.rt, c.ix.reps, c.stab, c.D, x, y, c.sched,
                                 gc.woff.ptr, gc.entries.ptr);
  });
  SPINED_CHECK_LAUNCH();
}

// Real matvec from the gather cache when there is (or can be) one.
template <int METHOD> bool try_cached(const MatvecCall &m, const Ctx &c) {
  GatherCache &gc = *m.op->cache;
  std::lock_guard<std::mutex> lk(gc.mu);
  if (m.no_alloc ? !gc.built : !ensure_cache<METHOD>(*m.op, c, gc, m.stream)) return false;
  using VT = VecType;
  if (m.xt == VT::F64 && m.yt == VT::F64)
    launch_cached(c, gc, static_cast<const double *>(m.x), static_cast<double *>(m.y), m.stream);
  else if (m.xt == VT::F32 && m.yt == VT::F32)
    launch_cached(c, gc, static_cast<const float *>(m.x), static_cast<float *>(m.y), m.stream);
  else if (m.xt == VT::F32 && m.yt == VT::F64)
    launch_cached(c, gc, static_cast<const float *>(m.x), static_cast<double *>(m.y), m.stream);
  else
    return false;
  return true;
}

template <int METHOD, int NW, bool BONDS, int K, class XT, class YT>
void launch_block_nw(const Ctx &c, const XT *x, YT *y, hipStream_t s) {
  matvec_block_kernel<METHOD, NW, BONDS, K, XT, YT><<<grid_for(c.D), kBlock, 0, s>>>(
      c.o, c.bd, c.rt, c.ga, c.ra, c.ix, c.stab, c.D, x, y, c.sched);
}

template <int METHOD, int K, class XT, class YT>
void launch_block_k(const Ctx &c, const XT *x, YT *y, hipStream_t s) {
  const int units = c.bonds ? c.bd.nbonds : c.o.ntemplates;
  if (c.bonds && units <= 64)
    launch_block_nw<METHOD, 1, true, K>(c, x, y, s);
  else if (c.bonds && units <= 128)
    launch_block_nw<METHOD, 2, true, K>(c, x, y, s);
  else if (!c.bonds && units <= 64)
    launch_block_nw<METHOD, 1, false, K>(c, x, y, s);
  else if (!c.bonds && units <= 128)
    launch_block_nw<METHOD, 2, false, K>(c, x, y, s);
  else if (!c.bonds && units <= 256)
    launch_block_nw<METHOD, 4, false, K>(c, x, y, s);
  else
    throw Error(LS_NOT_IMPLEMENTED, "block matvec with more than 256 off-diagonal templates");
  SPINED_CHECK_LAUNCH();
}

template <int METHOD, class XT, class YT>
void launch_block(const Ctx &c, int K, const XT *x, YT *y, hipStream_t s) {
  switch (K) {
  case 1: launch_real<METHOD>(c, x, y, s); break;
  case 2: launch_block_k<METHOD, 2>(c, x, y, s); break;
  case 4: launch_block_k<METHOD, 4>(c, x, y, s); break;
  case 8: launch_block_k<METHOD, 8>(c, x, y, s); break;
  default: throw Error(LS_INVALID_ARGUMENT, "block width must be 1, 2, 4 or 8");
  }
}

template <int K, class XT, class YT>
void launch_cached_block(const Ctx &c, const GatherCache &gc, const XT *x, YT *y, hipStream_t s) {
  const unsigned g = static_cast<unsigned>((c.D + kRowBlock - 1) / kRowBlock);
  with_nw(c, [&](auto nw, auto bonds) {
    matvec_cached_block_kernel<decltype(nw)::value, decltype(bonds)::value, K, XT, YT>
        <<<g, kRowBlock, 0, s>>>(c.o, c.bd, c.rt, c.ix.reps, c.stab, c.D, x, y, c.sched,
                                 gc.woff.ptr, gc.entries.ptr);
  });
  SPINED_CHECK_LAUNCH();
}

template <class XT, class YT>
void launch_cached_block_k(const Ctx &c, const GatherCache &gc, int K, const XT *x, YT *y,
                           hipStream_t s) {
  if (K == 8)
    launch_cached_block<8>(c, gc, x, y, s);
  else if (K == 4)
    launch_cached_block<4>(c, gc, x, y, s);
  else if (K == 2)
    launch_cached_block<2>(c, gc, x, y, s);
  else
    launch_cached_block<1>(c, gc, x, y, s);
}

// Block matvec from the gather cache when there is (or can be) one.
template <int METHOD> bool try_cached_block(const MatvecBlockCall &m, const Ctx &c) {
  GatherCache &gc = *m.op->cache;
  std::lock_guard<std::mutex> lk(gc.mu);
  if (!ensure_cache<METHOD>(*m.op, c, gc, m.stream)) return false;
  using VT = VecType;
  if (m.xt == VT::F64 && m.yt == VT::F64)
    launch_cached_block_k(c, gc, m.K, static_cast<const double *>(m.x), static_cast<double *>(m.y),
                          m.stream);
  else if (m.xt == VT::F32 && m.yt == VT::F32)
    launch_cached_block_k(c, gc, m.K, static_cast<const float *>(m.x), static_cast<float *>(m.y),
                          m.stream);
  else if (m.xt == VT::F32 && m.yt == VT::F64)
    launch_cached_block_k(c, gc, m.K, static_cast<const float *>(m.x), static_cast<double *>(m.y),
                          m.stream);
  else
    return false;
  return true;
}

template <int METHOD> void dispatch_block(const MatvecBlockCall &m, const Ctx &c) {
  hipStream_t s = m.stream;
  using VT = VecType;
  if (m.xt == VT::F64 && m.yt == VT::F64)
    launch_block<METHOD>(c, m.K, static_cast<const double *>(m.x), static_cast<double *>(m.y), s);
  else if (m.xt == VT::F32 && m.yt == VT::F32)
    launch_block<METHOD>(c, m.K, static_cast<const float *>(m.x), static_cast<float *>(m.y), s);
  else if (m.xt == VT::F32 && m.yt == VT::F64)
    launch_block<METHOD>(c, m.K, static_cast<const float *>(m.x), static_cast<double *>(m.y), s);
  else
    throw Error(LS_INVALID_DATATYPE, "block matvec supports real float32/float64 vectors");
}

Ctx make_ctx(const Operator &op, const double *pc, const double *ps) {
  const Basis &b = *op.basis;
  const auto &t = op.tables;
  Ctx c{};
  c.o = op_args(op, pc, ps);
  c.bonds = t.bonds;
  c.bd = BondArgs{t.bond_sites.ptr, t.bond_dp.ptr, t.bond_da.ptr, t.bond_o.ptr, t.bond_flip.ptr,
                  static_cast<int32_t>(t.h_bond_sites.size()),
                  t.nn_ring && b.tables.ring_available ? t.bond_lo.ptr : nullptr,
                  t.nn_identity ? 1 : 0};
  c.rt = RatioArgs{b.tables.div_idx.ptr, b.tables.ratio.ptr, b.tables.nd};
  c.ga = b.tables.generic_args(true);
  c.ra = b.tables.ring_args(true);
  c.ix = b.index_args(t.keeps_weight);
  c.stab = b.stab.ptr;
  c.D = b.D;
  c.sched = op.schedule_built ? op.sched.ptr : nullptr;
  c.lanes = op.lanes_per_row;
  return c;
}

std::mutex g_phase_mu;

const PhaseTable &phase_table(const Basis &b, bool no_alloc = false) {
  // one table per (device, L); small and kept for the process lifetime
  static std::vector<std::pair<std::pair<int, int>, std::unique_ptr<PhaseTable>>> cache;
  std::lock_guard<std::mutex> lk(g_phase_mu);
  const int dev = current_device();
  for (auto &kv : cache)
    if (kv.first.first == dev && kv.first.second == b.tables.L) return *kv.second;
  if (no_alloc)
    throw Error(LS_INVALID_ARGUMENT, "complex phase table not on the device yet: apply the "
                                     "operator once (ls_operator_matmat) before a call that "
                                     "allocates nothing");
  auto t = std::make_unique<PhaseTable>();
  const int L = b.tables.L;
  if (L > (1 << 20)) throw Error(LS_NOT_IMPLEMENTED, "complex sectors with huge periodicity");
  std::vector<double> c(L), s(L);
  for (int p = 0; p < L; ++p) phase_values(p, L, c[p], s[p]);
  t->cos_t.allocate(L);
  t->sin_t.allocate(L);
  SPINED_HIP(hipMemcpy(t->cos_t.ptr, c.data(), 8 * L, hipMemcpyHostToDevice));
  SPINED_HIP(hipMemcpy(t->sin_t.ptr, s.data(), 8 * L, hipMemcpyHostToDevice));
  t->L = L;
  cache.emplace_back(std::make_pair(dev, L), std::move(t));
  return *cache.back().second;
}

} // namespace

void launch_matvec(const MatvecCall &m) {
  const Operator &op = *m.op;
  const Basis &b = *op.basis;
  if (m.size != b.D) throw Error(LS_DIMENSION_MISMATCH, "vector length differs from the basis dimension");
  if (b.D == 0) return;
  const bool cplx = m.xt == VecType::C64 || m.xt == VecType::C128;
  if (!cplx && !op.tables.is_real)
    throw Error(LS_INVALID_DATATYPE, "operator is complex; use a complex datatype");
  const double *pc = nullptr, *ps = nullptr;
  if (cplx) {
    const auto &pt = phase_table(b, m.no_alloc);
    pc = pt.cos_t.ptr;
    ps = pt.sin_t.ptr;
  }
  const Ctx c = make_ctx(op, pc, ps);
  if (!cplx) {
    const bool done = op.orbit_method() == 1 ? try_cached<1>(m, c) : try_cached<0>(m, c);
    if (done) return;
  }
  if (op.orbit_method() == 1)
    dispatch<1>(m, c);
  else
    dispatch<0>(m, c);
}

namespace {
std::mutex g_ops_mu;
std::vector<Operator *> g_ops; // live operators (their gather caches can be reclaimed)
} // namespace

bool prepare_gather_cache(Operator &op, uint64_t reserve, uint64_t block_bytes) {
  if (!op.tables.is_real || op.basis->D == 0) return false;
  const Ctx c = make_ctx(op, nullptr, nullptr);
  GatherCache &gc = *op.cache;
  std::lock_guard<std::mutex> lk(gc.mu);
  const uint64_t reusable = pool_bytes(block_bytes);
  return op.orbit_method() == 1 ? ensure_cache<1>(op, c, gc, stream(), reserve, reusable)
                                : ensure_cache<0>(op, c, gc, stream(), reserve, reusable);
}

void drop_gather_cache_if_built(Operator &op) {
  std::lock_guard<std::mutex> lk(op.cache->mu);
  if (op.cache->built) {
    op.cache->built = false;
    op.cache->entries.release();
  }
}

Operator::Operator() {
  std::lock_guard<std::mutex> lk(g_ops_mu);
  g_ops.push_back(this);
}

Operator::~Operator() {
  std::lock_guard<std::mutex> lk(g_ops_mu);
  g_ops.erase(std::remove(g_ops.begin(), g_ops.end(), this), g_ops.end());
}

void drop_gather_cache(Operator &op) {
  std::lock_guard<std::mutex> lk(op.cache->mu);
  op.cache->drop();
}

// Caches in use (their lock held) are left alone.
void reclaim_gather_caches(int device) {
  std::lock_guard<std::mutex> lk(g_ops_mu);
  for (Operator *op : g_ops) {
    if (!op->basis || op->basis->device != device) continue;
    GatherCache &gc = *op->cache;
    if (gc.building.load() || !gc.mu.try_lock()) continue;
    if (gc.built) {
      gc.built = false;
      gc.entries.release();
    }
    gc.mu.unlock();
  }
}

bool matvec_block_supported(const Operator &op) {
  const auto &t = op.tables;
  if (!t.is_real) return false;
  const int units = t.bonds ? static_cast<int>(t.h_bond_sites.size()) : t.ntemplates;
  return units <= (t.bonds ? 128 : 256);
}

void launch_matvec_block(const MatvecBlockCall &m) {
  const Operator &op = *m.op;
  const Basis &b = *op.basis;
  if (m.size != b.D) throw Error(LS_DIMENSION_MISMATCH, "vector length differs from the basis dimension");
  if (b.D == 0) return;
  if (!matvec_block_supported(op))
    throw Error(LS_NOT_IMPLEMENTED, "block matvec needs a real operator with <= 256 templates");
  const Ctx c = make_ctx(op, nullptr, nullptr);
  const bool done = op.orbit_method() == 1 ? try_cached_block<1>(m, c) : try_cached_block<0>(m, c);
  if (done) return;
  if (op.orbit_method() == 1)
    dispatch_block<1>(m, c);
  else
    dispatch_block<0>(m, c);
}

void launch_cols_to_rows(VecType t, const void *x, uint64_t stride, uint64_t D, int K, void *xr,
                         hipStream_t s) {
  if (D == 0 || K == 0) return;
  if (t == VecType::F64)
    cols_to_rows_kernel<<<grid_for(D), kBlock, 0, s>>>(static_cast<const double *>(x), stride, D,
                                                       K, static_cast<double *>(xr));
  else if (t == VecType::F32)
    cols_to_rows_kernel<<<grid_for(D), kBlock, 0, s>>>(static_cast<const float *>(x), stride, D,
                                                       K, static_cast<float *>(xr));
  else
    throw Error(LS_INVALID_DATATYPE, "real vectors expected");
  SPINED_CHECK_LAUNCH();
}

void launch_rows_to_cols(VecType t, const void *yr, uint64_t D, int K, void *y, uint64_t stride,
                         hipStream_t s) {
  if (D == 0 || K == 0) return;
  if (t == VecType::F64)
    rows_to_cols_kernel<<<grid_for(D), kBlock, 0, s>>>(static_cast<const double *>(yr), D, K,
                                                       static_cast<double *>(y), stride);
  else if (t == VecType::F32)
    rows_to_cols_kernel<<<grid_for(D), kBlock, 0, s>>>(static_cast<const float *>(yr), D, K,
                                                       static_cast<float *>(y), stride);
  else
    throw Error(LS_INVALID_DATATYPE, "real vectors expected");
  SPINED_CHECK_LAUNCH();
}

int block_chunk(uint64_t remaining) { return remaining >= 8 ? 8 : remaining >= 4 ? 4 : remaining >= 2 ? 2 : 1; }

void ensure_schedule(Operator &op) {
  std::lock_guard<std::mutex> lk(op.mu);
  const Basis &b = *op.basis;
  const auto &t = op.tables;
  const int units = t.bonds ? static_cast<int>(t.h_bond_sites.size()) : t.ntemplates;
  // automatic: bases of at least one tile with the real row-gather kernels
  const bool want = op.schedule_request == 1 ||
                    (op.schedule_request < 0 && b.D >= static_cast<uint64_t>(kScheduleTile));
  if (!want || !t.is_real || units > (t.bonds ? 128 : 256) || b.D == 0) {
    if (op.schedule_built || op.sched.ptr) drop_gather_cache(op); // cache entries follow the slots
    op.schedule_built = false;
    op.sched.release();
    return;
  }
  if (op.schedule_built) return;
  drop_gather_cache(op);
  hipStream_t s = stream();
  op.sched.allocate(b.D);
  const Ctx c = make_ctx(op, nullptr, nullptr);
  const unsigned tiles = static_cast<unsigned>((b.D + kScheduleTile - 1) / kScheduleTile);
  if (c.bonds && units <= 64)
    schedule_kernel<1, true><<<tiles, kBlock, 0, s>>>(c.o, c.bd, b.reps.ptr, b.D, op.sched.ptr);
  else if (c.bonds)
    schedule_kernel<2, true><<<tiles, kBlock, 0, s>>>(c.o, c.bd, b.reps.ptr, b.D, op.sched.ptr);
  else if (units <= 64)
    schedule_kernel<1, false><<<tiles, kBlock, 0, s>>>(c.o, c.bd, b.reps.ptr, b.D, op.sched.ptr);
  else if (units <= 128)
    schedule_kernel<2, false><<<tiles, kBlock, 0, s>>>(c.o, c.bd, b.reps.ptr, b.D, op.sched.ptr);
  else
    schedule_kernel<4, false><<<tiles, kBlock, 0, s>>>(c.o, c.bd, b.reps.ptr, b.D, op.sched.ptr);
  SPINED_CHECK_LAUNCH();
  SPINED_HIP(hipStreamSynchronize(s));
  op.schedule_built = true;
}

namespace {
// <x|y> = sum_r conj(x_r) y_r in FP64 (real types: im = 0).  One partial
// (re, im) per workgroup, then one workgroup sums them in a fixed order.
__device__ inline void load_pair(const float *v, uint64_t r, double &re, double &im) {
  re = v[r];
  im = 0.0;
}
__device__ inline void load_pair(const double *v, uint64_t r, double &re, double &im) {
  re = v[r];
  im = 0.0;
}
__device__ inline void load_pair(const float2 *v, uint64_t r, double &re, double &im) {
  const float2 t = v[r];
  re = t.x;
  im = t.y;
}
__device__ inline void load_pair(const double2 *v, uint64_t r, double &re, double &im) {
  const double2 t = v[r];
  re = t.x;
  im = t.y;
}

template <class T>
__global__ void __launch_bounds__(kBlock)
    vdot_partial_kernel(const T *__restrict__ x, const T *__restrict__ y, uint64_t n,
                        double *__restrict__ partial) {
  __shared__ double lds[2][kBlock / 64];
  double re = 0.0, im = 0.0;
  for (uint64_t r = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x; r < n;
       r += static_cast<uint64_t>(gridDim.x) * kBlock) {
    double xr, xi, yr, yi;
    load_pair(x, r, xr, xi);
    load_pair(y, r, yr, yi);
    re = fma(xr, yr, fma(xi, yi, re));
    im = fma(xr, yi, fma(-xi, yr, im));
  }
  for (int off = 32; off > 0; off >>= 1) {
    re += __shfl_down(re, off, 64);
    im += __shfl_down(im, off, 64);
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
    lds[0][wid] = re;
    lds[1][wid] = im;
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    double t = 0.0;
    for (int k = 0; k < kBlock / 64; ++k) t += lds[threadIdx.x][k];
    partial[2 * static_cast<uint64_t>(blockIdx.x) + threadIdx.x] = t;
  }
}

__global__ void __launch_bounds__(kBlock)
    vdot_reduce_kernel(const double *__restrict__ partial, unsigned nb, double *__restrict__ out) {
  __shared__ double lds[2][kBlock];
  double re = 0.0, im = 0.0;
  for (unsigned b = threadIdx.x; b < nb; b += kBlock) {
    re += partial[2 * b];
    im += partial[2 * b + 1];
  }
  lds[0][threadIdx.x] = re;
  lds[1][threadIdx.x] = im;
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if (threadIdx.x < static_cast<unsigned>(off)) {
      lds[0][threadIdx.x] += lds[0][threadIdx.x + off];
      lds[1][threadIdx.x] += lds[1][threadIdx.x + off];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = lds[0][0];
    out[1] = lds[1][0];
  }
}
} // namespace

void launch_vdot(VecType t, const void *x, const void *y, uint64_t n, double *partial,
                 unsigned nb, double *out, hipStream_t s) {
  switch (t) {
  case VecType::F32:
    vdot_partial_kernel<<<nb, kBlock, 0, s>>>(static_cast<const float *>(x),
                                              static_cast<const float *>(y), n, partial);
    break;
  case VecType::F64:
    vdot_partial_kernel<<<nb, kBlock, 0, s>>>(static_cast<const double *>(x),
                                              static_cast<const double *>(y), n, partial);
    break;
  case VecType::C64:
    vdot_partial_kernel<<<nb, kBlock, 0, s>>>(static_cast<const float2 *>(x),
                                              static_cast<const float2 *>(y), n, partial);
    break;
  default:
    vdot_partial_kernel<<<nb, kBlock, 0, s>>>(static_cast<const double2 *>(x),
                                              static_cast<const double2 *>(y), n, partial);
  }
  SPINED_CHECK_LAUNCH();
  vdot_reduce_kernel<<<1, kBlock, 0, s>>>(partial, nb, out);
  SPINED_CHECK_LAUNCH();
}

unsigned vdot_blocks(uint64_t n) { return std::max(1u, std::min(grid_for(n), kVdotMaxBlocks)); }

uint64_t count_offdiag_device(const Operator &op) {
  const Basis &b = *op.basis;
  if (b.D == 0) return 0;
  hipStream_t s = stream();
  const OpArgs o = op_args(op, nullptr, nullptr);
  const auto ga = b.tables.generic_args(true);
  const auto ra = b.tables.ring_args(true);
  const IndexArgs ix = b.index_args(op.tables.keeps_weight);
  const unsigned g = grid_for(b.D);
  DeviceBuffer<unsigned long long> part(g);
  if (op.orbit_method() == 1)
    count_offdiag_kernel<1><<<g, kBlock, 0, s>>>(o, ga, ra, ix, b.D, part.ptr);
  else
    count_offdiag_kernel<0><<<g, kBlock, 0, s>>>(o, ga, ra, ix, b.D, part.ptr);
  SPINED_CHECK_LAUNCH();
  std::vector<unsigned long long> h(g);
  SPINED_HIP(hipMemcpyAsync(h.data(), part.ptr, 8ull * g, hipMemcpyDeviceToHost, s));
  SPINED_HIP(hipStreamSynchronize(s));
  uint64_t tot = 0;
  for (auto v : h) tot += v;
  return tot;
}

} // namespace spined
