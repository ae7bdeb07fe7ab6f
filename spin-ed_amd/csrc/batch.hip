// batch.hip -- many independent, uneven ground-state solves in ONE persistent
// launch (spined_batch_eigsh).
//
// Work queue.  The instances, largest dimension first, are expanded into
// slots (instance, member): one slot per workgroup of the instance's team,
// whose size grows with the dimension (one workgroup for the small rings, up
// to kDefaultTeamCap for the largest).  Every workgroup of the launch takes the next
// slot from a device counter, solves its share of that instance together
// with its teammates, and takes the next slot when the instance is done; the
// launch ends when the slots run out.  Nothing returns to the host between
// the launch and the end of the last instance.
//
// A team runs the one-vector thick-restart Lanczos of its instance in FP64
// (the recurrence of the single solver: classical Gram-Schmidt against every
// basis vector, a second pass only when the first loses more than two digits
// (Daniel, Gragg, Kaufman & Stewart), Jacobi on the projected matrix, restart
// on the lowest Ritz vectors).  Member k owns row tiles k, k + t, k + 2t, ...
// of every vector: dot partials, updates and restarts touch only its own
// rows.  The Lanczos matvec's rows are handed out in chunks from a counter
// instead (a member that runs faster takes more; a team barrier follows),
// and every matvec gathers every row of its input.  Each member
// writes one partial per dot; after a team barrier every member sums the t
// partials in member order and runs the same Jacobi sweep on the same
// projected matrix in its own LDS, so all members take the same decisions
// without further messages, and the sums do not depend on timing.  The
// explicit residual |H y - lambda y| / |y| ends the solve; when it is above
// tol the iteration continues with a tighter estimate target.
//
// Barriers are per instance (arrival counter and generation) with release /
// acquire at agent scope, so rows written on one XCD are read correctly on
// another.  Deadlock freedom: a member waits only for teammates whose slots
// come later in the queue; they are taken as soon as any earlier instance
// finishes, and the first team is complete because the largest team fits in
// the resident workgroups (checked on the host).  Every wait is bounded: a
// member that waits longer than kWaitLimitTicks sets the abort flag, every
// waiting member leaves on it, and the call reports a device error.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <numeric>
#include <vector>

#include "engine.hpp"
#include "handles.hpp"
#include "jacobi.hpp"
#include "rowgather.hpp"

namespace spined {

namespace {

constexpr int kTile = 256;      // rows per tile = threads per workgroup
constexpr int kMaxM = 24;       // basis size limit of the batched solver
constexpr int kPS = kMaxM + 4;  // partial slots per member and reduction
constexpr int kChunk = 8;       // dots per pass
constexpr int kMaxKeep = kMaxM / 2; // kept Ritz vectors (keep = m / 2)
constexpr int kMaxTeam = 128;     // team capacity (partial slots)
constexpr int kDefaultTeamCap = 32;
constexpr uint64_t kRowsPerMember = 32768;
constexpr uint32_t kChunkTiles = 4; // matvec rows handed out per grab: 1024
constexpr int kBarWords = 3;        // per instance: arrivals, generation, chunk counter
// s_memrealtime runs at 100 MHz: 60 s
constexpr uint64_t kWaitLimitTicks = 6000000000ull;
constexpr double kEta = 1e-2; // second Gram-Schmidt pass below this |w'| / |w|

struct PInst {
  OpArgs o;
  BondArgs bd;
  RatioArgs rt;
  GenericOrbitArgs ga;
  RingOrbitArgs ra;
  IndexArgs ix;
  const uint32_t *stab;
  uint64_t D;
  double *V;      // (m + 2) vectors of D doubles: basis 0..m, slot m + 1 = H y
  double *w;      // D
  double *pm;     // 2 x kMaxTeam x kPS member partials (double buffered)
  unsigned *bar;  // [0] arrivals, [1] generation, [2] matvec chunk counter
  double *out;    // [0] lambda [1] residual [2] cycles [3] matvecs [4] finish ticks [5] status
  uint32_t team;
  uint32_t ntiles;
};

struct Slot {
  uint32_t inst;
  uint32_t member;
};

__device__ inline uint64_t now_ticks() { return wall_clock64(); }

// Team barrier: everything this workgroup wrote before it is visible to the
// teammates after it, on any XCD.  False when the wait was abandoned (abort).
// The workgroup barrier orders every wave's stores before thread 0's
// agent-scope release (which writes the XCD's L2 back), and thread 0's
// agent-scope acquire (which invalidates the CU's and XCD's caches) before
// every wave's later loads; one fence pair per workgroup suffices.
__device__ bool team_barrier(unsigned *bar, uint32_t t, int *abort) {
  __shared__ int s_ok;
  __syncthreads();
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    int ok = 1;
    if (t > 1) {
      const unsigned g = __hip_atomic_load(bar + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const unsigned a = __hip_atomic_fetch_add(bar, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
      if (a == t - 1) {
        __hip_atomic_store(bar, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(bar + 1, g + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      } else {
        const uint64_t t0 = now_ticks();
        while (__hip_atomic_load(bar + 1, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == g) {
          if (__hip_atomic_load(abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
            ok = 0;
            break;
          }
          if (now_ticks() - t0 > kWaitLimitTicks) {
            __hip_atomic_store(abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            ok = 0;
            break;
          }
          __builtin_amdgcn_s_sleep(8);
        }
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    s_ok = ok;
  }
  __syncthreads();
  return s_ok != 0;
}

// Workgroup sums of acc[0..nq) (wave shuffles, then the 4 waves in order)
// into dst[q] (thread q writes).
template <int NQ>
__device__ inline void block_sums(const double (&acc)[NQ], int nq, double *__restrict__ dst) {
  __shared__ double lds[NQ][kTile / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    if (q < nq) {
      double v = acc[q];
      for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
      if (lane == 0) lds[q][wid] = v;
    }
  }
  __syncthreads();
  if (static_cast<int>(threadIdx.x) < nq) {
    double s = 0.0;
    for (int k = 0; k < kTile / 64; ++k) s += lds[threadIdx.x][k];
    dst[threadIdx.x] = s;
  }
  __syncthreads();
}

struct Team {
  const PInst &I;
  uint32_t k, t;
  int *abort;
  int red = 0; // which half of pm the next reduction writes
  __device__ Team(const PInst &inst, uint32_t member, int *ab)
      : I(inst), k(member), t(inst.team), abort(ab) {}
  __device__ double *pm_mine() { return I.pm + (static_cast<uint64_t>(red) * kMaxTeam + k) * kPS; }
  // after the member partials are written: barrier, then out[q] = sum over
  // members in order (every member the same); false on abort
  __device__ bool reduce(int nq, double *out) {
    if (!team_barrier(I.bar, t, abort)) return false;
    const double *base = I.pm + static_cast<uint64_t>(red) * kMaxTeam * kPS;
    if (static_cast<int>(threadIdx.x) < nq) {
      double s = 0.0;
      for (uint32_t mm = 0; mm < t; ++mm) s += base[mm * kPS + threadIdx.x];
      out[threadIdx.x] = s;
    }
    red ^= 1;
    __syncthreads();
    return true;
  }
  __device__ bool sync() { return team_barrier(I.bar, t, abort); }
};

template <int METHOD, int NW>
__device__ inline double matvec_row(const PInst &I, const IndexArgs &ix, uint64_t r,
                                    const double *x) {
  return gather_row_real<METHOD, NW, true>(I.o, I.bd, I.rt, I.ga, I.ra, ix, I.stab, r, x);
}

// partials of V_q . w for q in [0, nv) and (with ww) w . w in slot nv, over
// the member's rows, into the member's partial slots
__device__ void dots_pass(Team &tm, int nv, bool ww) {
  const PInst &I = tm.I;
  double *dst = tm.pm_mine();
  const int total = nv + (ww ? 1 : 0);
  for (int q0 = 0; q0 < total; q0 += kChunk) {
    double acc[kChunk];
#pragma unroll
    for (int c = 0; c < kChunk; ++c) acc[c] = 0.0;
    for (uint32_t tile = tm.k; tile < I.ntiles; tile += tm.t) {
      const uint64_t r = static_cast<uint64_t>(tile) * kTile + threadIdx.x;
      if (r >= I.D) continue;
      const double wr = I.w[r];
#pragma unroll
      for (int c = 0; c < kChunk; ++c) {
        const int q = q0 + c;
        if (q < nv)
          acc[c] = fma(I.V[static_cast<uint64_t>(q) * I.D + r], wr, acc[c]);
        else if (q == nv && ww)
          acc[c] = fma(wr, wr, acc[c]);
      }
    }
    block_sums<kChunk>(acc, std::min(kChunk, total - q0), dst + q0);
  }
}

// w -= sum_{q < nv} h[q] V_q over the member's rows; then the partials of
// V_q . w (q < nv) when dots, else of w . w in slot 0
__device__ void update_pass(Team &tm, int nv, const double *h, bool dots) {
  const PInst &I = tm.I;
  double acc[1] = {0.0};
  for (uint32_t tile = tm.k; tile < I.ntiles; tile += tm.t) {
    const uint64_t r = static_cast<uint64_t>(tile) * kTile + threadIdx.x;
    if (r >= I.D) continue;
    double s = 0.0;
    for (int q = 0; q < nv; ++q) s = fma(h[q], I.V[static_cast<uint64_t>(q) * I.D + r], s);
    const double wr = I.w[r] - s;
    I.w[r] = wr;
    acc[0] = fma(wr, wr, acc[0]);
  }
  if (dots)
    dots_pass(tm, nv, false);
  else
    block_sums<1>(acc, 1, tm.pm_mine());
}

// v[r] = (w[r] - sum_{q < nv} h[q] V_q[r]) * inv over the member's rows
// (the update and the store of the next basis vector in one pass)
__device__ void update_store_pass(Team &tm, int nv, const double *h, double inv, double *v) {
  const PInst &I = tm.I;
  for (uint32_t tile = tm.k; tile < I.ntiles; tile += tm.t) {
    const uint64_t r = static_cast<uint64_t>(tile) * kTile + threadIdx.x;
    if (r >= I.D) continue;
    double s = 0.0;
    for (int q = 0; q < nv; ++q) s = fma(h[q], I.V[static_cast<uint64_t>(q) * I.D + r], s);
    v[r] = (I.w[r] - s) * inv;
  }
}

template <int METHOD, int NW>
__device__ void solve_instance(const PInst &I, uint32_t inst_id, uint32_t member, int m, int keep,
                               double tol, int max_cycles, uint64_t seed, int *abort,
                               const uint64_t *t_start) {
  __shared__ double sT[kMaxM * kMaxM];
  __shared__ double sS[kMaxM * kMaxM];
  __shared__ double stheta[kMaxM];
  __shared__ double sbeta[kMaxM + 1];
  __shared__ double sh1[kPS], sh2[kPS], sred[kPS];
  __shared__ double sC[kMaxM * kMaxKeep]; // restart coefficients [q][i]
  // the rank index (bases without permutations, orbit method 0): its
  // binomial table is read from LDS by every lookup
  __shared__ uint32_t sRank[METHOD == 0 ? kRankMaxW * 64 : 1];
  IndexArgs ix = I.ix;
  if constexpr (METHOD == 0) {
    if (ix.rank_mode == 1) {
      for (int e = threadIdx.x; e < ix.rank_w * 64; e += kTile) sRank[e] = ix.rank_tab[e];
      __syncthreads();
    }
    ix.rank_tab = sRank;
  }
  auto mv = [&](uint64_t r, const double *x) {
    if constexpr (METHOD == 0) return matvec_row<METHOD, NW>(I, ix, r, x);
    else return matvec_row<METHOD, NW>(I, I.ix, r, x);
  };
  Team tm(I, member, abort);
  const uint64_t D = I.D;
  const int mm = static_cast<int>(std::min<uint64_t>(static_cast<uint64_t>(m), D));
  const int kp = std::max(1, std::min({keep, mm - 1, kMaxKeep}));
  double *V = I.V;
  auto vec = [&](int q) { return V + static_cast<uint64_t>(q) * D; };
  int64_t matvecs = 0;
  // w = H x over all rows, in chunks of kChunkTiles tiles handed out by the
  // instance's counter: members that run faster (next to lighter work on
  // their CUs) take more chunks, where a fixed split would make the team
  // wait for its slowest member.  Every member makes exactly one failing
  // grab per matvec, so the counter advances by nch + t per matvec and a
  // grab's chunk is its value minus that base; the next grab is issued
  // before the current chunk runs.  A row's sum does not depend on who
  // computes it.  Ends with a team barrier (the Gram-Schmidt passes read
  // their own rows of w); false on abort.
  __shared__ uint32_t s_chunk;
  uint32_t mv_round = 0;
  auto matvec_chunks = [&](const double *x) {
    const uint32_t nch = (I.ntiles + kChunkTiles - 1) / kChunkTiles;
    const uint32_t base = mv_round++ * (nch + tm.t);
    unsigned *ctr = I.bar + 2;
    if (threadIdx.x == 0) s_chunk = atomicAdd(ctr, 1u) - base;
    __syncthreads();
    uint32_t g = s_chunk;
    __syncthreads();
    while (g < nch) {
      if (threadIdx.x == 0) s_chunk = atomicAdd(ctr, 1u) - base;
      const uint32_t t_end = std::min((g + 1) * kChunkTiles, I.ntiles);
      for (uint32_t tile = g * kChunkTiles; tile < t_end; ++tile) {
        const uint64_t r = static_cast<uint64_t>(tile) * kTile + threadIdx.x;
        if (r < D) I.w[r] = mv(r, x);
      }
      __syncthreads();
      g = s_chunk;
      __syncthreads();
    }
    return tm.sync();
  };

  // start vector: w = hash(seed + instance, r), v_0 = w / |w|
  {
    double acc[1] = {0.0};
    for (uint32_t tile = member; tile < I.ntiles; tile += tm.t) {
      const uint64_t r = static_cast<uint64_t>(tile) * kTile + threadIdx.x;
      if (r >= D) continue;
      const double v = hash_uniform(seed + inst_id, r);
      I.w[r] = v;
      acc[0] = fma(v, v, acc[0]);
    }
    block_sums<1>(acc, 1, tm.pm_mine());
    if (!tm.reduce(1, sred)) return;
    const double inv = 1.0 / sqrt(sred[0]);
    for (uint32_t tile = member; tile < I.ntiles; tile += tm.t) {
      const uint64_t r = static_cast<uint64_t>(tile) * kTile + threadIdx.x;
      if (r < D) vec(0)[r] = I.w[r] * inv;
    }
    if (!tm.sync()) return;
  }
  for (int e = threadIdx.x; e < kMaxM * kMaxM; e += kTile) sT[e] = 0.0;
  __syncthreads();

  int j0 = 0, cycle = 0, msz = mm;
  double est_target = tol;
  double lambda = 0.0, resid = 0.0;
  bool done = false;
  while (!done) {
    ++cycle;
    msz = mm;
    for (int j = j0; j < mm; ++j) {
      // w = H v_j (own rows); dots with V_0..V_j and w.w
      const double *x = vec(j);
      if (tm.t > 1) {
        if (!matvec_chunks(x)) return;
      } else {
        for (uint32_t tile = 0; tile < I.ntiles; ++tile) {
          const uint64_t r = static_cast<uint64_t>(tile) * kTile + threadIdx.x;
          if (r < D) I.w[r] = mv(r, x);
        }
      }
      ++matvecs;
      dots_pass(tm, j + 1, true);
      if (!tm.reduce(j + 2, sh1)) return;
      double ss = 0.0;
      for (int q = 0; q <= j; ++q) ss = fma(sh1[q], sh1[q], ss);
      const double ww = sh1[j + 1];
      // one Gram-Schmidt pass suffices when it keeps at least kEta of |w|;
      // then |w'|^2 = w.w - |h|^2 (Pythagoras, relative error below
      // eps / kEta^2) and the update stores v_{j+1} directly
      const bool skip = ww > 0.0 && ww - ss >= kEta * kEta * ww;
      double bj = 0.0;
      if (skip) {
        bj = sqrt(ww - ss);
        if (threadIdx.x <= static_cast<unsigned>(j)) sh2[threadIdx.x] = 0.0;
      } else {
        update_pass(tm, j + 1, sh1, true);
        if (!tm.reduce(j + 1, sh2)) return;
        update_pass(tm, j + 1, sh2, false);
        if (!tm.reduce(1, sred)) return;
        bj = sqrt(std::max(sred[0], 0.0));
      }
      if (threadIdx.x <= static_cast<unsigned>(j)) {
        const double v = sh1[threadIdx.x] + sh2[threadIdx.x];
        sT[threadIdx.x * kMaxM + j] = v;
        sT[j * kMaxM + threadIdx.x] = v;
      }
      if (threadIdx.x == 0) sbeta[j] = bj;
      __syncthreads();
      // invariant subspace: the cycle ends here (exact Ritz values)
      double scale = 0.0;
      for (int q = 0; q <= j; ++q) scale = std::max(scale, fabs(sT[q * kMaxM + q]));
      if (bj <= 1e-14 * std::max(scale, 1.0)) {
        msz = j + 1;
        if (threadIdx.x == 0) sbeta[j] = 0.0;
        __syncthreads();
        break;
      }
      const double inv = 1.0 / bj;
      double *vn = vec(j + 1);
      if (skip) {
        update_store_pass(tm, j + 1, sh1, inv, vn);
      } else {
        for (uint32_t tile = member; tile < I.ntiles; tile += tm.t) {
          const uint64_t r = static_cast<uint64_t>(tile) * kTile + threadIdx.x;
          if (r < D) vn[r] = I.w[r] * inv;
        }
      }
      if (!tm.sync()) return;
    }
    // projected problem (every member the same)
    jacobi_eigh_device<kMaxM>(sT, kMaxM, msz, stheta, sS, 40);
    __syncthreads();
    const double est = fabs(sbeta[msz - 1] * sS[msz - 1]);
    const bool exhausted = msz < mm || static_cast<uint64_t>(mm) >= D;
    if (est <= est_target || exhausted || cycle >= max_cycles) {
      // Ritz vector y = V S[:, 0] into w, H y into V_{m+1}
      for (uint32_t tile = member; tile < I.ntiles; tile += tm.t) {
        const uint64_t r = static_cast<uint64_t>(tile) * kTile + threadIdx.x;
        if (r >= D) continue;
        double acc = 0.0;
        for (int q = 0; q < msz; ++q) acc = fma(sS[q], vec(q)[r], acc);
        I.w[r] = acc;
      }
      if (!tm.sync()) return;
      double *hy = vec(m + 1);
      double acc2[2] = {0.0, 0.0};
      for (uint32_t tile = member; tile < I.ntiles; tile += tm.t) {
        const uint64_t r = static_cast<uint64_t>(tile) * kTile + threadIdx.x;
        if (r >= D) continue;
        const double h = mv(r, I.w);
        hy[r] = h;
        const double y = I.w[r];
        acc2[0] = fma(y, h, acc2[0]);
        acc2[1] = fma(y, y, acc2[1]);
      }
      ++matvecs;
      block_sums<2>(acc2, 2, tm.pm_mine());
      if (!tm.reduce(2, sred)) return;
      const double yy = sred[1];
      lambda = sred[0] / yy;
      double acc[1] = {0.0};
      for (uint32_t tile = member; tile < I.ntiles; tile += tm.t) {
        const uint64_t r = static_cast<uint64_t>(tile) * kTile + threadIdx.x;
        if (r >= D) continue;
        const double d = fma(-lambda, I.w[r], hy[r]);
        acc[0] = fma(d, d, acc[0]);
      }
      block_sums<1>(acc, 1, tm.pm_mine());
      if (!tm.reduce(1, sred)) return;
      resid = sqrt(std::max(sred[0], 0.0) / yy);
      if (resid <= tol || exhausted || cycle >= max_cycles) {
        done = true;
        break;
      }
      // the estimate was optimistic: aim lower and go on
      est_target = std::max(0.25 * est_target, 1e-3 * tol);
    }
    // thick restart: V_i <- sum_q S[i][q] V_q (i < kp), V_kp <- V_msz
    for (int e = threadIdx.x; e < msz * kMaxKeep; e += kTile) {
      const int q = e / kMaxKeep, i = e % kMaxKeep;
      sC[e] = i < kp ? sS[i * msz + q] : 0.0;
    }
    __syncthreads();
    for (uint32_t tile = member; tile < I.ntiles; tile += tm.t) {
      const uint64_t r = static_cast<uint64_t>(tile) * kTile + threadIdx.x;
      if (r >= D) continue;
      // each input read once; the kept outputs in registers, written after
      // the whole row is read (in place)
      double out[kMaxKeep];
#pragma unroll
      for (int i = 0; i < kMaxKeep; ++i) out[i] = 0.0;
      for (int q = 0; q < msz; ++q) {
        const double x = vec(q)[r];
#pragma unroll
        for (int i = 0; i < kMaxKeep; ++i) out[i] = fma(sC[q * kMaxKeep + i], x, out[i]);
      }
      const double last = vec(msz)[r];
#pragma unroll
      for (int i = 0; i < kMaxKeep; ++i)
        if (i < kp) vec(i)[r] = out[i];
      vec(kp)[r] = last;
    }
    for (int e = threadIdx.x; e < kMaxM * kMaxM; e += kTile) {
      const int i = e / kMaxM, jj = e % kMaxM;
      sT[e] = (i == jj && i < kp) ? stheta[i] : 0.0;
    }
    if (!tm.sync()) return;
    j0 = kp;
  }
  if (member == 0 && threadIdx.x == 0) {
    I.out[0] = lambda;
    I.out[1] = resid;
    I.out[2] = cycle;
    I.out[3] = static_cast<double>(matvecs);
    I.out[4] = static_cast<double>(now_ticks() - *t_start);
    I.out[5] = resid <= tol ? 1.0 : 2.0;
  }
}

template <int METHOD, int NW>
__global__ void __launch_bounds__(kTile)
    batch_persistent_kernel(const PInst *__restrict__ insts, const Slot *__restrict__ slots,
                            uint32_t nslots, unsigned *ticket, int m, int keep, double tol,
                            int max_cycles, uint64_t seed, int *abort, uint64_t *t_start) {
  __shared__ uint32_t s_slot;
  if (threadIdx.x == 0) atomicMin(reinterpret_cast<unsigned long long *>(t_start),
                                  static_cast<unsigned long long>(now_ticks()));
  for (;;) {
    if (threadIdx.x == 0)
      s_slot = __hip_atomic_load(abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                   ? nslots
                   : atomicAdd(ticket, 1u);
    __syncthreads();
    // wave-uniform (scalar) slot, so the instance's arguments load into
    // scalar registers
    const uint32_t sl = __builtin_amdgcn_readfirstlane(s_slot);
    __syncthreads();
    if (sl >= nslots) return;
    const uint32_t inst = __builtin_amdgcn_readfirstlane(slots[sl].inst);
    const uint32_t member = __builtin_amdgcn_readfirstlane(slots[sl].member);
    solve_instance<METHOD, NW>(insts[inst], inst, member, m, keep, tol, max_cycles, seed, abort,
                               t_start);
  }
}

__global__ void batch_init_kernel(unsigned *ticket, int *abort, uint64_t *t_start, unsigned *bars,
                                  uint64_t nbar) {
  const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
  if (i == 0) {
    *ticket = 0;
    *abort = 0;
    *t_start = ~0ull;
  }
  if (i < nbar) bars[i] = 0;
}

template <int METHOD, int NW>
int resident_blocks() {
  int per_cu = 0;
  SPINED_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &per_cu, reinterpret_cast<const void *>(&batch_persistent_kernel<METHOD, NW>), kTile, 0));
  int dev = 0, cus = 0;
  SPINED_HIP(hipGetDevice(&dev));
  SPINED_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  return std::max(1, per_cu) * std::max(1, cus);
}

} // namespace

// One batched solve; see the file comment.
int batch_eigsh(const std::vector<Operator *> &ops, double tol, int m_req, uint64_t seed,
                int max_cycles, double *evals, double *rnorms, int32_t *cycles_out,
                double *seconds_out, int64_t *matvecs_out, spined_batch_stats *stats) {
  const auto t_begin = std::chrono::steady_clock::now();
  const int n_inst = static_cast<int>(ops.size());
  if (n_inst == 0) return LS_SUCCESS;
  hipStream_t s = stream();
  const int method = ops[0]->orbit_method();
  int nw = 1;
  for (auto *op : ops) {
    if (!op->tables.bonds || !op->tables.is_real)
      throw Error(LS_NOT_IMPLEMENTED, "batched solver needs real exchange-bond operators");
    if (op->orbit_method() != method)
      throw Error(LS_NOT_IMPLEMENTED, "batched instances must share the orbit method");
    const int nb = static_cast<int>(op->tables.h_bond_sites.size());
    if (nb > 128) throw Error(LS_NOT_IMPLEMENTED, "at most 128 bonds per batched instance");
    nw = std::max(nw, nb > 64 ? 2 : 1);
  }
  // 16 vectors: the 512-ring batch launch is about 4% shorter than with 20
  // (fewer Gram-Schmidt and restart passes per step outweigh the extra
  // matvecs); 12 is as slow as 20, 10 slower
  const int m = std::max(4, std::min(m_req > 0 ? m_req : 16, kMaxM));
  const int keep = m / 2;
  const int resident = method == 1 ? (nw == 1 ? resident_blocks<1, 1>() : resident_blocks<1, 2>())
                                   : (nw == 1 ? resident_blocks<0, 1>() : resident_blocks<0, 2>());
  // team sizing: about rows_per_member rows per workgroup, at most max_team
  // workgroups (SPINED_BATCH_ROWS_PER_MEMBER / SPINED_BATCH_MAX_TEAM override
  // them, for tuning)
  uint64_t rows_per_member = kRowsPerMember;
  int team_cap = kDefaultTeamCap;
  if (const char *e = std::getenv("SPINED_BATCH_ROWS_PER_MEMBER"))
    rows_per_member = std::max<uint64_t>(kTile, std::strtoull(e, nullptr, 10));
  if (const char *e = std::getenv("SPINED_BATCH_MAX_TEAM"))
    team_cap = std::max(1, std::min(kMaxTeam, std::atoi(e)));
  const uint32_t max_team = static_cast<uint32_t>(std::max(1, std::min(team_cap, resident)));

  // per-instance storage
  std::vector<uint64_t> voff(n_inst);
  uint64_t total_rows = 0, total_vec = 0;
  for (int i = 0; i < n_inst; ++i) {
    const uint64_t D = ops[i]->basis->D;
    if (D == 0) throw Error(LS_INVALID_ARGUMENT, "empty basis in batch");
    voff[i] = total_vec;
    total_vec += D * static_cast<uint64_t>(m + 3); // m + 2 basis slots and w
    total_rows += D;
  }
  DeviceBuffer<double> vpool(total_vec);
  const uint64_t pm_per = 2ull * kMaxTeam * kPS;
  DeviceBuffer<double> pm(pm_per * n_inst);
  DeviceBuffer<double> outs(8ull * n_inst);
  DeviceBuffer<unsigned> bars(static_cast<uint64_t>(kBarWords) * n_inst);
  std::vector<PInst> hinst(n_inst);
  for (int i = 0; i < n_inst; ++i) {
    Operator &op = *ops[i];
    Basis &b = *op.basis;
    const auto &t = op.tables;
    PInst &I = hinst[i];
    I.o = OpArgs{};
    I.bd = BondArgs{t.bond_sites.ptr, t.bond_dp.ptr, t.bond_da.ptr, t.bond_o.ptr, t.bond_flip.ptr,
                    static_cast<int32_t>(t.h_bond_sites.size()), nullptr, 0};
    I.rt = RatioArgs{b.tables.div_idx.ptr, b.tables.ratio.ptr, b.tables.nd};
    I.ga = b.tables.generic_args(true);
    I.ra = b.tables.ring_args(true);
    I.ix = b.index_args(t.keeps_weight);
    I.stab = b.stab.ptr;
    I.D = b.D;
    I.V = vpool.ptr + voff[i];
    I.w = I.V + b.D * static_cast<uint64_t>(m + 2);
    I.pm = pm.ptr + pm_per * i;
    I.bar = bars.ptr + static_cast<uint64_t>(kBarWords) * i;
    I.out = outs.ptr + 8ull * i;
    I.ntiles = static_cast<uint32_t>((b.D + kTile - 1) / kTile);
    const uint64_t want = (b.D + rows_per_member - 1) / rows_per_member;
    I.team = static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>({want, max_team, I.ntiles})));
  }
  // queue: largest instances first (their teams form first), members in order
  std::vector<int> order(n_inst);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return hinst[a].D > hinst[b].D; });
  std::vector<Slot> hslots;
  for (int i : order)
    for (uint32_t k = 0; k < hinst[i].team; ++k) hslots.push_back(Slot{static_cast<uint32_t>(i), k});
  const uint32_t nslots = static_cast<uint32_t>(hslots.size());
  DeviceBuffer<PInst> dinst(n_inst);
  DeviceBuffer<Slot> dslots(nslots);
  SPINED_HIP(hipMemcpyAsync(dinst.ptr, hinst.data(), sizeof(PInst) * n_inst, hipMemcpyHostToDevice, s));
  SPINED_HIP(hipMemcpyAsync(dslots.ptr, hslots.data(), sizeof(Slot) * nslots, hipMemcpyHostToDevice, s));
  DeviceBuffer<unsigned> ticket(1);
  DeviceBuffer<int> abort_flag(1);
  DeviceBuffer<uint64_t> t_start(1);
  const uint64_t nbar = static_cast<uint64_t>(kBarWords) * n_inst;
  batch_init_kernel<<<static_cast<unsigned>((nbar + 255) / 256), 256, 0, s>>>(
      ticket.ptr, abort_flag.ptr, t_start.ptr, bars.ptr, nbar);
  SPINED_CHECK_LAUNCH();
  // one workgroup per slot at most, no more than fit on the device at once
  const unsigned grid = static_cast<unsigned>(std::min<uint64_t>(nslots, static_cast<uint64_t>(resident)));
  SPINED_HIP(hipStreamSynchronize(s));
  hipEvent_t ev0, ev1;
  SPINED_HIP(hipEventCreate(&ev0));
  SPINED_HIP(hipEventCreate(&ev1));
  SPINED_HIP(hipEventRecord(ev0, s));
  auto launch = [&](auto meth, auto nwc) {
    constexpr int M = decltype(meth)::value, W = decltype(nwc)::value;
    batch_persistent_kernel<M, W><<<grid, kTile, 0, s>>>(dinst.ptr, dslots.ptr, nslots, ticket.ptr,
                                                         m, keep, tol, max_cycles, seed,
                                                         abort_flag.ptr, t_start.ptr);
  };
  if (method == 1) {
    if (nw == 1) launch(std::integral_constant<int, 1>{}, std::integral_constant<int, 1>{});
    else launch(std::integral_constant<int, 1>{}, std::integral_constant<int, 2>{});
  } else {
    if (nw == 1) launch(std::integral_constant<int, 0>{}, std::integral_constant<int, 1>{});
    else launch(std::integral_constant<int, 0>{}, std::integral_constant<int, 2>{});
  }
  SPINED_CHECK_LAUNCH();
  SPINED_HIP(hipEventRecord(ev1, s));
  SPINED_HIP(hipStreamSynchronize(s));
  float ms = 0.f;
  SPINED_HIP(hipEventElapsedTime(&ms, ev0, ev1));
  (void)hipEventDestroy(ev0);
  (void)hipEventDestroy(ev1);
  const double t_solve = ms * 1e-3;
  int aborted = 0;
  std::vector<double> hout(8ull * n_inst);
  SPINED_HIP(hipMemcpy(&aborted, abort_flag.ptr, sizeof(int), hipMemcpyDeviceToHost));
  SPINED_HIP(hipMemcpy(hout.data(), outs.ptr, 8ull * n_inst * sizeof(double), hipMemcpyDeviceToHost));
  if (aborted)
    throw Error(LS_DEVICE_ERROR, "batched solve abandoned: a team barrier waited longer than its limit");
  int dev = 0, khz = 0;
  SPINED_HIP(hipGetDevice(&dev));
  SPINED_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
  const double tick_s = khz > 0 ? 1.0 / (1e3 * khz) : 1e-8;
  int nconv = 0, max_cycles_seen = 0;
  int64_t max_mv = 0;
  for (int i = 0; i < n_inst; ++i) {
    const double *o = hout.data() + 8ull * i;
    evals[i] = o[0];
    rnorms[i] = o[1];
    const int cyc = static_cast<int>(o[2]);
    const bool ok = o[5] == 1.0 && o[1] <= tol;
    if (ok) ++nconv;
    if (cycles_out) cycles_out[i] = ok ? cyc : -1;
    if (seconds_out) seconds_out[i] = o[4] * tick_s;
    if (matvecs_out) matvecs_out[i] = static_cast<int64_t>(o[3]);
    max_cycles_seen = std::max(max_cycles_seen, cyc);
    max_mv = std::max<int64_t>(max_mv, static_cast<int64_t>(o[3]));
  }
  if (logging_enabled())
    log_line("batch: %d instances in %u slots on %u workgroups (largest team %u): %.3f s",
             n_inst, nslots, grid, max_team, t_solve);
  if (stats) {
    stats->cycles = max_cycles_seen;
    stats->converged = nconv;
    stats->seconds_total =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t_begin).count();
    stats->seconds_cycles = t_solve;
    stats->total_rows = total_rows;
    stats->matvecs_per_instance_max = max_mv;
  }
  return nconv == n_inst ? LS_SUCCESS : LS_NOT_CONVERGED;
}

} // namespace spined
