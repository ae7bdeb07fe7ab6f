// basis.hip -- symmetry-adapted basis construction on the GPU (ls_build,
// src/SpinED/Internal.hs:178-182, :230-236; caller SpinED.hs:315-337).
//
// Pipeline, per segment of the candidate range:
//   1. enumerate: thread t owns 64 consecutive candidates (colex rank order,
//      which is ascending integer order; Gosper's successor steps within the
//      word), tests each for "orbit minimum with a symmetric state" and
//      writes a 64-bit hit mask; each block writes its hit count;
//   2. scan: one block turns the block counts into exclusive offsets;
//   3. emit: each thread re-walks its word and writes its hits at the block
//      offset plus its in-block prefix.
// Because enumeration is monotone and the compaction is order preserving,
// the representatives come out globally sorted: no sort pass is needed.
// Stabiliser sizes and the bucket index follow.
#include <algorithm>
#include <cstring>
#include <chrono>
#include <cmath>
#include <vector>

#include "engine.hpp"
#include "scan.hpp"

namespace spined {

namespace {

constexpr int kBlock = 256;
constexpr int kSegmentBits = 16; // level-1 segments of the bucket index

struct EnumArgs {
  const uint64_t *__restrict__ binom; // 65 x 65
  int n;
  int w;              // -1: all 2^n states
  uint64_t total;     // number of candidates
  // candidate = (y << shift) | low for the enumerated weight-w states y of
  // n bits (a monotone map: the candidates stay in ascending order)
  int shift;
  uint64_t low;
};

__device__ inline uint64_t candidate(const EnumArgs &e, uint64_t y) { return (y << e.shift) | e.low; }

__device__ inline uint64_t unrank_colex(const EnumArgs &e, uint64_t rank) {
  if (e.w < 0) return rank;
  uint64_t x = 0;
  int c = e.n - 1;
  for (int k = e.w; k >= 1; --k) {
    while (e.binom[c * 65 + k] > rank) --c;
    x |= 1ull << c;
    rank -= e.binom[c * 65 + k];
    --c;
  }
  return x;
}

__device__ inline uint64_t next_state(const EnumArgs &e, uint64_t v) {
  if (e.w < 0) return v + 1;
  // Gosper's hack: next larger integer with the same popcount
  const uint64_t t = v | (v - 1);
  return (t + 1) | (((~t & (0 - ~t)) - 1) >> (__builtin_ctzll(v) + 1));
}

template <int METHOD>
__device__ inline bool is_basis_state(const GenericOrbitArgs &ga, const RingOrbitArgs &ra,
                                      uint64_t x) {
  if constexpr (METHOD == 1) {
    const OrbitResult r = orbit_ring(ra, x);
    return r.rep == x && r.valid;
  } else {
    return rep_test_generic(ga, x) != 0;
  }
}

__device__ inline uint32_t block_sum(uint32_t v, uint32_t *lds) {
  // wave reduction then across the 4 waves of the block
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) lds[wid] = v;
  __syncthreads();
  uint32_t s = 0;
  if (threadIdx.x == 0)
    for (int i = 0; i < kBlock / 64; ++i) s += lds[i];
  return s;
}

template <int METHOD>
__global__ void __launch_bounds__(kBlock)
    enumerate_kernel(EnumArgs e, GenericOrbitArgs ga, RingOrbitArgs ra, uint64_t word0,
                     uint64_t nwords, uint64_t *__restrict__ masks,
                     uint32_t *__restrict__ block_counts) {
  __shared__ uint32_t lds[kBlock / 64];
  const uint64_t t = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
  uint64_t bits = 0;
  if (t < nwords) {
    const uint64_t r0 = (word0 + t) * 64;
    const uint64_t rem = e.total - r0;
    const int cnt = rem < 64 ? static_cast<int>(rem) : 64;
    uint64_t y = unrank_colex(e, r0);
    for (int i = 0; i < cnt; ++i) {
      if (is_basis_state<METHOD>(ga, ra, candidate(e, y))) bits |= 1ull << i;
      if (i + 1 < cnt) y = next_state(e, y);
    }
    masks[t] = bits;
  }
  const uint32_t s = block_sum(static_cast<uint32_t>(__popcll(bits)), lds);
  if (threadIdx.x == 0) block_counts[blockIdx.x] = s;
}

__global__ void __launch_bounds__(kBlock)
    emit_kernel(EnumArgs e, uint64_t word0, uint64_t nwords, const uint64_t *__restrict__ masks,
                const uint64_t *__restrict__ block_offsets, uint64_t base,
                uint64_t *__restrict__ reps) {
  __shared__ uint32_t lds[kBlock];
  const uint64_t t = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
  const uint64_t bits = t < nwords ? masks[t] : 0;
  const uint32_t c = static_cast<uint32_t>(__popcll(bits));
  lds[threadIdx.x] = c;
  __syncthreads();
  for (int off = 1; off < kBlock; off <<= 1) {
    const uint32_t v = threadIdx.x >= static_cast<unsigned>(off) ? lds[threadIdx.x - off] : 0;
    __syncthreads();
    lds[threadIdx.x] += v;
    __syncthreads();
  }
  if (bits == 0) return;
  uint64_t pos = base + block_offsets[blockIdx.x] + lds[threadIdx.x] - c;
  const uint64_t r0 = (word0 + t) * 64;
  const int last = 63 - __clzll(bits);
  uint64_t y = unrank_colex(e, r0);
  for (int i = 0; i <= last; ++i) {
    if ((bits >> i) & 1) reps[pos++] = candidate(e, y);
    if (i < last) y = next_state(e, y);
  }
}

// |Stab r| per representative; *bad becomes nonzero when an entry is not a
// valid orbit minimum of the sector's Hamming weight (w < 0: any weight).
template <int METHOD>
__global__ void __launch_bounds__(kBlock)
    stabilizer_kernel(GenericOrbitArgs ga, RingOrbitArgs ra, const uint64_t *__restrict__ reps,
                      uint64_t D, int w, uint32_t *__restrict__ stab, uint32_t *bad) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= D) return;
  const uint64_t v = reps[i];
  OrbitResult r;
  if constexpr (METHOD == 1)
    r = orbit_ring(ra, v);
  else
    r = orbit_generic(ga, v);
  stab[i] = static_cast<uint32_t>(r.count);
  if (!r.valid || r.rep != v || (w >= 0 && __popcll(v) != w)) atomicOr(bad, 1u);
}

// seg[s] = lower_bound(reps, s << shift) for s < nseg, D at s = nseg
// (grid-stride: nseg + 1 may exceed one launch's work-item range).
__global__ void __launch_bounds__(kBlock)
    segment_start_kernel(const uint64_t *__restrict__ reps, uint64_t D, int shift, uint64_t nseg,
                         uint32_t *__restrict__ seg) {
  for (uint64_t s = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; s <= nseg;
       s += static_cast<uint64_t>(gridDim.x) * kBlock) {
    if (s == nseg) {
      seg[s] = static_cast<uint32_t>(D);
      continue;
    }
    const uint64_t key = s << shift;
    uint64_t lo = 0, n = D;
    while (n > 0) {
      const uint64_t half = n >> 1;
      if (reps[lo + half] < key) {
        lo += half + 1;
        n -= half + 1;
      } else {
        n = half;
      }
    }
    seg[s] = static_cast<uint32_t>(lo);
  }
}

// Level-2 table: buckets[b] = first i in segment s with reps[i] >= the
// bucket's start value (s << shift) + (local << lsh_s); D at b = nbuckets.
// The segment of b is the last one whose base is <= b (bases ascend).
__global__ void __launch_bounds__(kBlock)
    level2_kernel(const uint64_t *__restrict__ reps, uint64_t D, const uint2 *__restrict__ dir,
                  const uint32_t *__restrict__ seg, uint64_t nseg, int shift, uint64_t nbuckets,
                  uint32_t *__restrict__ buckets) {
  for (uint64_t b = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; b <= nbuckets;
       b += static_cast<uint64_t>(gridDim.x) * kBlock) {
    if (b == nbuckets) {
      buckets[b] = static_cast<uint32_t>(D);
      continue;
    }
    uint64_t lo_s = 0, n_s = nseg; // upper_bound of b over bases, minus one
    while (n_s > 0) {
      const uint64_t half = n_s >> 1;
      if (dir_base(dir[lo_s + half]) <= b) {
        lo_s += half + 1;
        n_s -= half + 1;
      } else {
        n_s = half;
      }
    }
    const uint64_t sg = lo_s - 1;
    const uint2 d = dir[sg];
    const uint64_t key = (sg << shift) + ((b - dir_base(d)) << dir_lsh(d));
    uint64_t lo = seg[sg], n = seg[sg + 1] - seg[sg];
    while (n > 0) {
      const uint64_t half = n >> 1;
      if (reps[lo + half] < key) {
        lo += half + 1;
        n -= half + 1;
      } else {
        n = half;
      }
    }
    buckets[b] = static_cast<uint32_t>(lo);
  }
}

__global__ void __launch_bounds__(kBlock) fill_slots_kernel(uint64_t *__restrict__ p, uint64_t n) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * kBlock)
    p[i] = kEmptySlot;
}

// Linear-probing insert of every representative (64-bit compare-and-swap on
// the slot word).  The load factor is <= 1/2, so every probe meets a free
// slot; the slot order depends on the CAS order, the lookups do not.
__global__ void __launch_bounds__(kBlock)
    hash_insert_kernel(const uint64_t *__restrict__ reps, uint64_t D, uint64_t *slots,
                       uint64_t mask) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= D) return;
  const uint64_t h = mix64(reps[i]);
  const unsigned long long e = ((h >> 32) << 32) | i;
  uint64_t s = h & mask;
  for (uint64_t step = 0; step <= mask; ++step) {
    const unsigned long long old =
        atomicCAS(reinterpret_cast<unsigned long long *>(slots + s), kEmptySlot, e);
    if (old == kEmptySlot) return;
    s = (s + 1) & mask;
  }
}

uint64_t binomial(int n, int k) {
  if (k < 0 || k > n) return 0;
  static uint64_t table[65][65];
  static bool ready = false;
  if (!ready) {
    for (int a = 0; a <= 64; ++a) {
      table[a][0] = 1;
      for (int b = 1; b <= a; ++b) table[a][b] = table[a - 1][b - 1] + (b < a ? table[a - 1][b] : 0);
    }
    ready = true;
  }
  return table[n][k];
}

unsigned grid_for(uint64_t n) { return static_cast<unsigned>((n + kBlock - 1) / kBlock); }

// Fills b.stab; returns true when every entry is a valid orbit minimum of
// the sector (synchronises the stream).
bool compute_stabilizers(Basis &b, hipStream_t s) {
  if (b.D == 0) return true;
  const auto ga = b.tables.generic_args(true);
  const auto ra = b.tables.ring_args(true);
  DeviceBuffer<uint32_t> bad(1);
  SPINED_HIP(hipMemsetAsync(bad.ptr, 0, 4, s));
  if (b.tables.ring_available)
    stabilizer_kernel<1><<<grid_for(b.D), kBlock, 0, s>>>(ga, ra, b.reps.ptr, b.D, b.w, b.stab.ptr,
                                                        bad.ptr);
  else
    stabilizer_kernel<0><<<grid_for(b.D), kBlock, 0, s>>>(ga, ra, b.reps.ptr, b.D, b.w, b.stab.ptr,
                                                        bad.ptr);
  SPINED_CHECK_LAUNCH();
  uint32_t h = 1;
  SPINED_HIP(hipMemcpyAsync(&h, bad.ptr, 4, hipMemcpyDeviceToHost, s));
  SPINED_HIP(hipStreamSynchronize(s));
  return h == 0;
}

} // namespace

// Open-addressing table with 2^H slots, H = ceil(log2 D) + 1.
void build_hash_index(Basis &b, hipStream_t s) {
  int H = 1;
  while ((1ull << H) < 2 * b.D) ++H;
  const uint64_t nslots = 1ull << H;
  b.slots.release();
  b.slots.allocate(nslots);
  b.hmask = nslots - 1;
  fill_slots_kernel<<<static_cast<unsigned>(std::min<uint64_t>(grid_for(nslots), 1u << 16)), kBlock,
                      0, s>>>(b.slots.ptr, nslots);
  SPINED_CHECK_LAUNCH();
  if (b.D) {
    hash_insert_kernel<<<grid_for(b.D), kBlock, 0, s>>>(b.reps.ptr, b.D, b.slots.ptr, b.hmask);
    SPINED_CHECK_LAUNCH();
  }
  SPINED_HIP(hipStreamSynchronize(s));
}

// Rank mode of the combinatorial-rank index (IndexArgs): 1 (colex rank of
// weight-w states) or 2 (index = state) for a complete basis whose group
// has no permutation but the identity, with or without the spin flip; 0
// when the rank does not give the basis order.  The representatives are
// then every state of the sector: all 2^n or C(n, w) states, or with the
// flip the orbit minima min(v, ~v), i.e. the states with site n - 1 down
// (2^(n-1), or C(n - 1, w) for w = n / 2; other weights have no flip
// sector).  The dimension is checked against that count.  Weights above
// kRankMaxW keep the bucket index.
int rank_index_mode(const Basis &b) {
  if (!b.complete || b.tables.nperm != 1 || b.n < 1 || b.n > 63) return 0;
  const bool flip = b.tables.has_flip != 0;
  if (b.w < 0) return b.D == (1ull << (flip ? b.n - 1 : b.n)) ? 2 : 0;
  if ((flip && 2 * b.w != b.n) || b.w > kRankMaxW) return 0;
  return b.D == (flip ? binomial(b.n - 1, b.w) : binomial(b.n, b.w)) ? 1 : 0;
}

static void build_rank_index(Basis &b, int mode, hipStream_t s) {
  b.rank_mode = mode;
  b.rank_tab.release();
  const int w = mode == 1 ? b.w : 0;
  std::vector<uint32_t> tab(static_cast<size_t>(std::max(w, 1)) * 64, 0);
  for (int i = 0; i < w; ++i)
    for (int p = 0; p < 64; ++p)
      tab[i * 64 + p] = static_cast<uint32_t>(std::min<uint64_t>(binomial(p, i + 1), 0xffffffffull));
  b.rank_tab.allocate(tab.size());
  SPINED_HIP(hipMemcpyAsync(b.rank_tab.ptr, tab.data(), tab.size() * 4, hipMemcpyHostToDevice, s));
  SPINED_HIP(hipStreamSynchronize(s));
}

void build_index(Basis &b) {
  hipStream_t s = stream();
  if (b.D >= 0xffffffffull)
    throw Error(LS_NOT_IMPLEMENTED, "dimensions of 2^32 - 1 or more are not supported");
  const int rmode = rank_index_mode(b);
  const int kind = b.index_kind_request >= 0 ? b.index_kind_request : (rmode ? 2 : 0);
  if (kind == 2 && !rmode)
    throw Error(LS_INVALID_ARGUMENT, "the rank index needs a complete basis whose symmetry group "
                                     "has no permutation but the identity");
  b.index_kind = kind;
  if (kind != 2) {
    b.rank_mode = 0;
    b.rank_tab.release();
  }
  if (b.index_kind == 2) {
    b.slots.release();
    b.hmask = 0;
    b.buckets.release();
    b.dir.release();
    b.nbuckets = 0;
    b.nseg = 0;
    b.shift = 0;
    b.bucket_bits = 0;
    b.h_segstart.clear();
    build_rank_index(b, rmode, s);
    return;
  }
  if (b.index_kind == 1) {
    b.buckets.release();
    b.dir.release();
    b.nbuckets = 0;
    b.nseg = 0;
    b.shift = 0;
    b.bucket_bits = 0;
    build_hash_index(b, s);
    return;
  }
  b.slots.release();
  b.hmask = 0;
  b.buckets.release();
  b.dir.release();
  // level-2 size: about 2^bits buckets, bits = ceil(log2 D) + 3 by default
  // (eight to sixteen buckets per representative: most lookups land in a
  // bucket that holds only their target)
  int lg = 0;
  while ((1ull << lg) < b.D) ++lg;
  int bits = b.bucket_bits_request >= 0 ? b.bucket_bits_request : lg + 3;
  bits = std::max(1, std::min(kMaxBucketBits, bits));
  size_t free_bytes = 0, total_bytes = 0;
  SPINED_HIP(hipMemGetInfo(&free_bytes, &total_bytes));
  uint64_t last = 0;
  if (b.D) SPINED_HIP(hipMemcpy(&last, b.reps.ptr + (b.D - 1), 8, hipMemcpyDeviceToHost));
  const int bitlen = last ? 64 - __builtin_clzll(last) : 1;
  const int T = std::min(kSegmentBits, bitlen);
  b.shift = bitlen - T;
  b.nseg = b.D ? (last >> b.shift) + 1 : 0;
  const auto cap = [](uint64_t n) {
    return static_cast<unsigned>(std::max<uint64_t>(1, std::min<uint64_t>(grid_for(n), 1u << 20)));
  };
  DeviceBuffer<uint32_t> seg(b.nseg + 1);
  segment_start_kernel<<<cap(b.nseg + 1), kBlock, 0, s>>>(b.reps.ptr, b.D, b.shift, b.nseg, seg.ptr);
  SPINED_CHECK_LAUNCH();
  b.h_segstart.assign(b.nseg + 1, 0);
  SPINED_HIP(hipMemcpyAsync(b.h_segstart.data(), seg.ptr, 4 * (b.nseg + 1), hipMemcpyDeviceToHost,
                            s));
  SPINED_HIP(hipStreamSynchronize(s));
  // the table may take at most half the free memory
  std::vector<uint2> hdir(std::max<uint64_t>(b.nseg, 1));
  uint64_t total = level2_size(b, bits, hdir.data());
  while (4 * (total + 1) > free_bytes / 2 && bits > 1) total = level2_size(b, --bits, hdir.data());
  b.bucket_bits = bits;
  b.nbuckets = total;
  b.dir.allocate(hdir.size());
  SPINED_HIP(hipMemcpyAsync(b.dir.ptr, hdir.data(), sizeof(uint2) * hdir.size(),
                            hipMemcpyHostToDevice, s));
  b.buckets.allocate(b.nbuckets + 1);
  level2_kernel<<<cap(b.nbuckets + 1), kBlock, 0, s>>>(b.reps.ptr, b.D, b.dir.ptr, seg.ptr, b.nseg,
                                                      b.shift, b.nbuckets, b.buckets.ptr);
  SPINED_CHECK_LAUNCH();
  SPINED_HIP(hipStreamSynchronize(s));
}

// Level-2 buckets for about 2^bits in total: segment s gets 2^lb_s,
// lb_s = ceil(log2(count_s * 2^bits / D)) clamped to [0, shift].  Uses the
// segment starts of the last bucket-table build; fills dir when given.
uint64_t level2_size(const Basis &b, int bits, uint2 *dir) {
  if (b.nseg == 0 || b.h_segstart.size() != b.nseg + 1) return 0;
  const double ratio = b.D ? std::ldexp(1.0, bits) / static_cast<double>(b.D) : 1.0;
  uint64_t total = 0;
  for (uint64_t sg = 0; sg < b.nseg; ++sg) {
    const double want = static_cast<double>(b.h_segstart[sg + 1] - b.h_segstart[sg]) * ratio;
    int lb = 0;
    while (lb < b.shift && std::ldexp(1.0, lb) < want) ++lb;
    if (dir) dir[sg] = dir_entry(total, static_cast<uint32_t>(b.shift - lb));
    total += 1ull << lb;
  }
  return total;
}

// True when the permutations move site n - 1 to every site (the group acts
// transitively on the sites, as translations of a periodic lattice do).
bool site_transitive(const Group *g, int n) {
  if (!g || n <= 0) return n == 1;
  std::vector<char> seen(n, 0);
  int count = 0;
  for (const auto &p : g->perms) {
    if (static_cast<int>(p.size()) != n) return false;
    const int s = p[n - 1];
    if (s < n && !seen[s]) {
      seen[s] = 1;
      ++count;
    }
  }
  return count == n;
}

void build_basis(Basis &b) {
  const auto t0 = std::chrono::steady_clock::now();
  hipStream_t s = stream();
  const int n = b.n;
  uint64_t total;
  int en = n, ew = b.w, shift = 0;
  uint64_t low = 0;
  if (b.w < 0) {
    if (n > 48) throw Error(LS_NOT_IMPLEMENTED, "no Hamming weight restriction above 48 spins");
    total = 1ull << n;
  } else if (b.w >= 2 && b.w < n && b.tables.ring_available) {
    // all n rotations: an orbit minimum x has site n - 1 empty (some
    // rotation empties it) and site 0 occupied (else x >> 1, a rotation of
    // x, is smaller), so x = (y << 1) | 1 for y of weight w - 1 on n - 2 sites
    en = n - 2;
    ew = b.w - 1;
    shift = 1;
    low = 1;
    total = binomial(en, ew);
  } else if (b.w < n && site_transitive(b.group.get(), n)) {
    // every state with a zero has an image with site n - 1 empty, so its
    // orbit minimum lies below 2^(n-1): the candidates are the weight-w
    // states of sites 0 .. n - 2, the first C(n - 1, w) in colex order
    en = n - 1;
    total = binomial(en, ew);
  } else {
    total = binomial(n, b.w);
  }
  double imag = 0.0;
  const double expect_d = sector_dimension(b, &imag);
  const uint64_t capacity = static_cast<uint64_t>(std::llround(std::max(expect_d, 0.0)));
  if (capacity >= 0xffffffffull)
    throw Error(LS_NOT_IMPLEMENTED, "dimensions of 2^32 - 1 or more are not supported");

  std::vector<uint64_t> hb(65 * 65);
  for (int a = 0; a <= 64; ++a)
    for (int c = 0; c <= 64; ++c) hb[a * 65 + c] = binomial(a, c);
  DeviceBuffer<uint64_t> binom(hb.size());
  SPINED_HIP(hipMemcpyAsync(binom.ptr, hb.data(), hb.size() * 8, hipMemcpyHostToDevice, s));

  b.reps.allocate(std::max<uint64_t>(capacity, 1));
  const uint64_t words_total = (total + 63) / 64;
  const uint64_t seg_words = std::min<uint64_t>(words_total, 1ull << 22);
  const uint64_t seg_blocks = (seg_words + kBlock - 1) / kBlock;
  DeviceBuffer<uint64_t> masks(seg_words);
  DeviceBuffer<uint32_t> counts(seg_blocks);
  DeviceBuffer<uint64_t> offsets(seg_blocks + 1);
  DeviceBuffer<uint64_t> seg_total(1);
  EnumArgs e{binom.ptr, en, ew, total, shift, low};
  const auto ga = b.tables.generic_args(true);
  const auto ra = b.tables.ring_args(true);
  uint64_t emitted = 0;
  for (uint64_t w0 = 0; w0 < words_total; w0 += seg_words) {
    const uint64_t nw = std::min(seg_words, words_total - w0);
    const unsigned nb = grid_for(nw);
    if (b.tables.ring_available)
      enumerate_kernel<1><<<nb, kBlock, 0, s>>>(e, ga, ra, w0, nw, masks.ptr, counts.ptr);
    else
      enumerate_kernel<0><<<nb, kBlock, 0, s>>>(e, ga, ra, w0, nw, masks.ptr, counts.ptr);
    SPINED_CHECK_LAUNCH();
    scan_counts(counts.ptr, nb, offsets.ptr, seg_total.ptr, s);
    uint64_t segc = 0;
    SPINED_HIP(hipMemcpyAsync(&segc, seg_total.ptr, 8, hipMemcpyDeviceToHost, s));
    SPINED_HIP(hipStreamSynchronize(s));
    if (emitted + segc > capacity)
      throw Error(LS_INTERNAL_ERROR, "basis build found more states than the sector dimension");
    if (segc) {
      emit_kernel<<<nb, kBlock, 0, s>>>(e, w0, nw, masks.ptr, offsets.ptr, emitted, b.reps.ptr);
      SPINED_CHECK_LAUNCH();
    }
    emitted += segc;
  }
  if (emitted != capacity)
    throw Error(LS_INTERNAL_ERROR, "basis build found " + std::to_string(emitted) +
                                       " states, character formula gives " +
                                       std::to_string(capacity));
  b.D = emitted;
  b.stab.allocate(std::max<uint64_t>(b.D, 1));
  b.complete = compute_stabilizers(b, s);
  build_index(b);
  b.host_reps.reset();
  b.built = true;
  b.seconds_build =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// The index (bucket table + binary search) needs strictly ascending,
// in-range states; anything else is rejected before it reaches the GPU.
void validate_representatives(const Basis &b, uint64_t n, const uint64_t *reps) {
  const uint64_t mask = site_mask_of(b.n);
  for (uint64_t i = 0; i < n; ++i) {
    if ((reps[i] & ~mask) != 0)
      throw Error(LS_INVALID_ARGUMENT, "representative has bits beyond number_spins");
    if (i && reps[i] <= reps[i - 1])
      throw Error(LS_INVALID_ARGUMENT, "representatives must be strictly ascending");
  }
}

void adopt_basis(Basis &b, uint64_t n, const uint64_t *reps) {
  hipStream_t s = stream();
  if (n >= 0xffffffffull)
    throw Error(LS_NOT_IMPLEMENTED, "dimensions of 2^32 - 1 or more are not supported");
  validate_representatives(b, n, reps);
  b.reps.allocate(std::max<uint64_t>(n, 1));
  if (n) SPINED_HIP(hipMemcpy(b.reps.ptr, reps, n * 8, hipMemcpyHostToDevice));
  b.D = n;
  b.stab.allocate(std::max<uint64_t>(n, 1));
  // the list is the whole sector when each entry is a valid orbit minimum
  // (ascending, so distinct) and their number is the sector dimension
  const bool minima = compute_stabilizers(b, s);
  double im = 0.0;
  const double dim = sector_dimension(b, &im);
  b.complete = minima && std::fabs(im) < 0.5 && std::llround(dim) == static_cast<long long>(n);
  build_index(b);
  auto h = std::make_shared<HostStates>(n);
  if (n) std::memcpy(h->data.get(), reps, n * 8);
  b.host_reps = h;
  b.built = true;
}

} // namespace spined
