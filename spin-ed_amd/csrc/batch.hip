// batch.hip -- many independent, uneven ground-state solves at once.
//
// Every instance (an ls_operator on a built basis; instances may share a
// basis) gets its own thick-restart Lanczos state on the device.  All
// instances advance in lockstep through Lanczos steps; each step is a fixed
// sequence of batched kernels whose grid covers (instance, 256-row tile)
// pairs of ALL instances, so one launch serves every Hamiltonian regardless
// of its dimension.  One restart cycle (steps, projected Jacobi solves,
// convergence test, restart) is captured once as a hipGraph and replayed
// until every instance has converged; converged instances drop out (their
// tiles return immediately).  The host reads one counter per cycle.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "engine.hpp"
#include "handles.hpp"
#include "jacobi.hpp"
#include "rowgather.hpp"

namespace spined {

namespace {

constexpr int kTile = 256;
constexpr int kMaxM = 32;       // basis size limit of the batched solver
constexpr int kPS = 40;         // partial-sum slots per tile (>= kMaxM + 1)
constexpr int kChunk = 8;

struct BatchInst {
  // operator / basis (read uniformly by a workgroup)
  OpArgs o;
  BondArgs bd;
  RatioArgs rt;
  GenericOrbitArgs ga;
  RingOrbitArgs ra;
  IndexArgs ix;
  const uint32_t *stab;
  uint64_t D;
  // solver state
  double *V;     // (m + 1) vectors of D doubles, contiguous
  double *w;     // D
  double *T;     // kMaxM x kMaxM
  double *theta; // kMaxM
  double *S;     // kMaxM x kMaxM
  double *beta;  // kMaxM + 1 (the last slot normalises the start)
  double *h1;    // kMaxM + 1
  double *h2;    // kMaxM + 1
  double *scal;  // 4 scalars: lambda, |y|^2, residual^2, spare
  int *flag;     // 0 active, > 0: cycle at which it converged
  uint32_t tile_begin, tile_end;
};

struct Tile {
  uint32_t inst;
  uint32_t pad;
  uint64_t row0;
};

__device__ inline bool tile_row(const BatchInst *insts, const Tile *tiles, const BatchInst *&I,
                                uint64_t &r) {
  const Tile t = tiles[blockIdx.x];
  I = insts + t.inst;
  r = t.row0 + threadIdx.x;
  return *I->flag == 0 && r < I->D;
}

template <int METHOD, int NW>
__global__ void __launch_bounds__(kTile)
    batch_matvec_kernel(const BatchInst *insts, const Tile *tiles, int src, int dst, int final_pass) {
  const BatchInst *I;
  uint64_t r;
  const Tile t = tiles[blockIdx.x];
  I = insts + t.inst;
  r = t.row0 + threadIdx.x;
  if (r >= I->D) return;
  const int f = *I->flag;
  if (final_pass ? f == 0 : f != 0) return;
  // src / dst: vector slot (0..m) or -1 for w
  const double *x = src < 0 ? I->w : I->V + static_cast<uint64_t>(src) * I->D;
  double *y = dst < 0 ? I->w : I->V + static_cast<uint64_t>(dst) * I->D;
  y[r] = gather_row_real<METHOD, NW, true>(I->o, I->bd, I->rt, I->ga, I->ra, I->ix, I->stab, r, x);
}

// partial[tile * kPS + q] = sum over the tile's rows of V_q . w (q < nvec),
// or, with mode 1, w . w in slot 0.
__global__ void __launch_bounds__(kTile)
    batch_multidot_kernel(const BatchInst *insts, const Tile *tiles, int nvec, int mode,
                          double *partial) {
  __shared__ double lds[kChunk][kTile / 64];
  const Tile t = tiles[blockIdx.x];
  const BatchInst *I = insts + t.inst;
  const int c = blockIdx.y;
  const int q0 = c * kChunk;
  const uint64_t r = t.row0 + threadIdx.x;
  const bool active = *I->flag == 0;
  double acc[kChunk];
#pragma unroll
  for (int q = 0; q < kChunk; ++q) acc[q] = 0.0;
  if (active && r < I->D) {
    const double wr = I->w[r];
    if (mode == 1) {
      acc[0] = wr * wr;
    } else {
#pragma unroll
      for (int q = 0; q < kChunk; ++q)
        if (q0 + q < nvec) acc[q] = I->V[static_cast<uint64_t>(q0 + q) * I->D + r] * wr;
    }
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < kChunk; ++q) {
    double v = acc[q];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if (lane == 0) lds[q][wid] = v;
  }
  __syncthreads();
  if (threadIdx.x < kChunk && q0 + static_cast<int>(threadIdx.x) < (mode == 1 ? 1 : nvec)) {
    double s = 0.0;
    for (int k = 0; k < kTile / 64; ++k) s += lds[threadIdx.x][k];
    partial[static_cast<uint64_t>(blockIdx.x) * kPS + q0 + threadIdx.x] = s;
  }
}

// Per instance (one block each): out[q] = sum of its tiles' partials
// (fixed order).  target 0: h1, 1: h2, 2: beta[j] = sqrt(sum).
__global__ void __launch_bounds__(256)
    batch_reduce_kernel(const BatchInst *insts, const double *partial, int nvec, int target, int j) {
  const BatchInst *I = insts + blockIdx.x;
  if (*I->flag != 0) return;
  __shared__ double lds[256];
  for (int q = 0; q < nvec; ++q) {
    double s = 0.0;
    for (uint32_t t = I->tile_begin + threadIdx.x; t < I->tile_end; t += 256)
      s += partial[static_cast<uint64_t>(t) * kPS + q];
    lds[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
      if (threadIdx.x < static_cast<unsigned>(off)) lds[threadIdx.x] += lds[threadIdx.x + off];
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      if (target == 0) I->h1[q] = lds[0];
      else if (target == 1) I->h2[q] = lds[0];
      else I->beta[j] = sqrt(lds[0]);
    }
    __syncthreads();
  }
}

// w -= sum_q h V_q  (h = h1 or h2)
__global__ void __launch_bounds__(kTile)
    batch_gemv_sub_kernel(const BatchInst *insts, const Tile *tiles, int nvec, int which) {
  const BatchInst *I;
  uint64_t r;
  if (!tile_row(insts, tiles, I, r)) return;
  const double *h = which == 0 ? I->h1 : I->h2;
  double acc = 0.0;
  for (int q = 0; q < nvec; ++q) acc = fma(h[q], I->V[static_cast<uint64_t>(q) * I->D + r], acc);
  I->w[r] -= acc;
}

__global__ void batch_store_column_kernel(const BatchInst *insts, int j) {
  const BatchInst *I = insts + blockIdx.x;
  if (*I->flag != 0) return;
  const int i = threadIdx.x;
  if (i <= j) {
    const double v = I->h1[i] + I->h2[i];
    I->T[i * kMaxM + j] = v;
    I->T[j * kMaxM + i] = v;
  }
}

// V_dst = w / beta[bidx]
__global__ void __launch_bounds__(kTile)
    batch_scale_store_kernel(const BatchInst *insts, const Tile *tiles, int bidx, int dst) {
  const BatchInst *I;
  uint64_t r;
  if (!tile_row(insts, tiles, I, r)) return;
  const double b = I->beta[bidx];
  const double inv = b > 1e-300 ? 1.0 / b : 0.0;
  I->V[static_cast<uint64_t>(dst) * I->D + r] = I->w[r] * inv;
}

__global__ void __launch_bounds__(256) batch_jacobi_kernel(const BatchInst *insts, int m) {
  const BatchInst *I = insts + blockIdx.x;
  if (*I->flag != 0) return;
  const int mm = static_cast<uint64_t>(m) < I->D ? m : static_cast<int>(I->D);
  jacobi_eigh_device(I->T, kMaxM, mm, I->theta, I->S, 40);
}

// Converged when |beta_{m-1} S[m-1, 0]| <= tol (or the Krylov space is
// exhausted); the flag records the cycle number.
__global__ void batch_check_kernel(const BatchInst *insts, int n_inst, int m, double tol,
                                   const int *cycle_counter, int *remaining) {
  const int cycle = *cycle_counter;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_inst) return;
  const BatchInst *I = insts + i;
  if (*I->flag != 0) return;
  const int mm = static_cast<uint64_t>(m) < I->D ? m : static_cast<int>(I->D);
  const double est = fabs(I->beta[mm - 1] * I->S[mm - 1]);
  if (est <= tol || mm < m) {
    *I->flag = cycle;
  } else {
    atomicAdd(remaining, 1);
  }
}

// in place, row-local: V_i <- sum_q S[i m + q] V_q for i < keep; V_keep <- V_m
__global__ void __launch_bounds__(kTile)
    batch_restart_kernel(const BatchInst *insts, const Tile *tiles, int m, int keep) {
  const BatchInst *I;
  uint64_t r;
  if (!tile_row(insts, tiles, I, r)) return;
  double row[kMaxM];
  for (int q = 0; q < m; ++q) row[q] = I->V[static_cast<uint64_t>(q) * I->D + r];
  for (int i = 0; i < keep; ++i) {
    double acc = 0.0;
    for (int q = 0; q < m; ++q) acc = fma(I->S[i * m + q], row[q], acc);
    I->V[static_cast<uint64_t>(i) * I->D + r] = acc;
  }
  I->V[static_cast<uint64_t>(keep) * I->D + r] = I->V[static_cast<uint64_t>(m) * I->D + r];
}

__global__ void batch_reset_kernel(const BatchInst *insts, int keep) {
  const BatchInst *I = insts + blockIdx.x;
  if (*I->flag != 0) return;
  for (int e = threadIdx.x; e < kMaxM * kMaxM; e += blockDim.x) {
    const int i = e / kMaxM, j = e % kMaxM;
    I->T[e] = (i == j && i < keep) ? I->theta[i] : 0.0;
  }
}

__global__ void batch_zero_kernel(int *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// start of a cycle (inside the captured graph): count it, clear `remaining`
__global__ void batch_cycle_begin_kernel(int *cycle_counter, int *remaining) {
  *cycle_counter += 1;
  *remaining = 0;
}

// instances still active at the end: report their current lowest Ritz pair
// (after the last restart it is V_0, i.e. S = e_0)
__global__ void batch_mark_unconverged_kernel(const BatchInst *insts, int n_inst) {
  const int i = blockIdx.x;
  if (i >= n_inst) return;
  const BatchInst *I = insts + i;
  if (*I->flag != 0) return;
  __syncthreads();
  for (int q = threadIdx.x; q < kMaxM; q += blockDim.x) I->S[q] = q == 0 ? 1.0 : 0.0;
  __syncthreads();
  if (threadIdx.x == 0) *I->flag = -1;
}

// ---- finalisation (converged instances: flag > 0) ----------------------------
// w = Ritz vector of the lowest Ritz value (FP64)
__global__ void __launch_bounds__(kTile)
    batch_ritz_kernel(const BatchInst *insts, const Tile *tiles, int m) {
  const Tile t = tiles[blockIdx.x];
  const BatchInst *I = insts + t.inst;
  const uint64_t r = t.row0 + threadIdx.x;
  if (*I->flag == 0 || r >= I->D) return;
  const int mm = static_cast<uint64_t>(m) < I->D ? m : static_cast<int>(I->D);
  double acc = 0.0;
  for (int q = 0; q < mm; ++q) acc = fma(I->S[q], I->V[static_cast<uint64_t>(q) * I->D + r], acc);
  I->w[r] = acc;
}

// partials of (y.Hy, y.y) with y = w, Hy = V_m; or (mode 1) |Hy - lambda y|^2
__global__ void __launch_bounds__(kTile)
    batch_final_dots_kernel(const BatchInst *insts, const Tile *tiles, int m, int mode,
                            double *partial) {
  __shared__ double lds[2][kTile / 64];
  const Tile t = tiles[blockIdx.x];
  const BatchInst *I = insts + t.inst;
  const uint64_t r = t.row0 + threadIdx.x;
  double a = 0.0, b = 0.0;
  if (*I->flag != 0 && r < I->D) {
    const double y = I->w[r], hy = I->V[static_cast<uint64_t>(m) * I->D + r];
    if (mode == 0) {
      a = y * hy;
      b = y * y;
    } else {
      const double d = fma(-I->scal[0], y, hy);
      a = d * d;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_down(a, off, 64);
    b += __shfl_down(b, off, 64);
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
    lds[0][wid] = a;
    lds[1][wid] = b;
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    double s = 0.0;
    for (int k = 0; k < kTile / 64; ++k) s += lds[threadIdx.x][k];
    partial[static_cast<uint64_t>(blockIdx.x) * kPS + threadIdx.x] = s;
  }
}

__global__ void __launch_bounds__(256)
    batch_final_reduce_kernel(const BatchInst *insts, const double *partial, int mode) {
  const BatchInst *I = insts + blockIdx.x;
  if (*I->flag == 0) return;
  __shared__ double lds[2][256];
  double a = 0.0, b = 0.0;
  for (uint32_t t = I->tile_begin + threadIdx.x; t < I->tile_end; t += 256) {
    a += partial[static_cast<uint64_t>(t) * kPS];
    b += partial[static_cast<uint64_t>(t) * kPS + 1];
  }
  lds[0][threadIdx.x] = a;
  lds[1][threadIdx.x] = b;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < static_cast<unsigned>(off)) {
      lds[0][threadIdx.x] += lds[0][threadIdx.x + off];
      lds[1][threadIdx.x] += lds[1][threadIdx.x + off];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (mode == 0) {
      I->scal[0] = lds[0][0] / lds[1][0]; // Rayleigh quotient
      I->scal[1] = lds[1][0];             // |y|^2
    } else {
      I->scal[2] = sqrt(lds[0][0] / I->scal[1]); // |Hy - lambda y| / |y|
    }
  }
}

__global__ void __launch_bounds__(kTile)
    batch_random_kernel(const BatchInst *insts, const Tile *tiles, uint64_t seed) {
  const Tile t = tiles[blockIdx.x];
  const BatchInst *I = insts + t.inst;
  const uint64_t r = t.row0 + threadIdx.x;
  if (r >= I->D) return;
  I->w[r] = hash_uniform(seed + t.inst, r);
}

unsigned blocks_for(uint64_t n, unsigned b) { return static_cast<unsigned>((n + b - 1) / b); }

} // namespace

// One batched solve; see the file comment.
int batch_eigsh(const std::vector<Operator *> &ops, double tol, int m_req, uint64_t seed,
                int max_cycles, double *evals, double *rnorms, int32_t *cycles_out,
                double *seconds_out, spined_batch_stats *stats) {
  const auto t_start = std::chrono::steady_clock::now();
  const int n_inst = static_cast<int>(ops.size());
  if (n_inst == 0) return LS_SUCCESS;
  hipStream_t s = stream();
  // all instances must share the kernel variant (orbit method, bond form)
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
  const int m = std::max(4, std::min(m_req > 0 ? m_req : 20, kMaxM));
  const int keep = m / 2;
  // tiles and per-instance storage
  std::vector<Tile> htiles;
  std::vector<uint64_t> voff(n_inst);
  uint64_t total_rows = 0;
  for (int i = 0; i < n_inst; ++i) {
    const uint64_t D = ops[i]->basis->D;
    if (D == 0) throw Error(LS_INVALID_ARGUMENT, "empty basis in batch");
    voff[i] = total_rows;
    total_rows += D;
  }
  DeviceBuffer<double> Vpool(total_rows * static_cast<uint64_t>(m + 1));
  DeviceBuffer<double> wpool(total_rows);
  const size_t per_small = kMaxM * kMaxM * 2 + kMaxM + 3 * (kMaxM + 1) + 4;
  DeviceBuffer<double> small(per_small * n_inst);
  DeviceBuffer<int> flags(n_inst + 1);
  std::vector<BatchInst> hinst(n_inst);
  for (int i = 0; i < n_inst; ++i) {
    Operator &op = *ops[i];
    Basis &b = *op.basis;
    const auto &t = op.tables;
    BatchInst &I = hinst[i];
    I.o = OpArgs{};
    I.bd = BondArgs{t.bond_sites.ptr, t.bond_dp.ptr, t.bond_da.ptr, t.bond_o.ptr, t.bond_flip.ptr,
                    static_cast<int32_t>(t.h_bond_sites.size()), nullptr, 0};
    I.rt = RatioArgs{b.tables.div_idx.ptr, b.tables.ratio.ptr, b.tables.nd};
    I.ga = b.tables.generic_args(true);
    I.ra = b.tables.ring_args(true);
    I.ix = b.index_args(t.keeps_weight);
    I.stab = b.stab.ptr;
    I.D = b.D;
    I.V = Vpool.ptr + voff[i] * static_cast<uint64_t>(m + 1);
    I.w = wpool.ptr + voff[i];
    double *sm = small.ptr + per_small * i;
    I.T = sm;
    I.S = sm + kMaxM * kMaxM;
    I.theta = sm + 2 * kMaxM * kMaxM;
    I.beta = I.theta + kMaxM;
    I.h1 = I.beta + kMaxM + 1;
    I.h2 = I.h1 + kMaxM + 1;
    I.scal = I.h2 + kMaxM + 1;
    I.flag = flags.ptr + i;
    I.tile_begin = static_cast<uint32_t>(htiles.size());
    for (uint64_t r0 = 0; r0 < b.D; r0 += kTile)
      htiles.push_back(Tile{static_cast<uint32_t>(i), 0u, r0});
    I.tile_end = static_cast<uint32_t>(htiles.size());
  }
  // instances whose dimension is at most m would need m > D vectors; they
  // run with m' = D (exhausted Krylov space converges in the first cycle)
  const unsigned ntiles = static_cast<unsigned>(htiles.size());
  DeviceBuffer<BatchInst> dinst(n_inst);
  DeviceBuffer<Tile> dtiles(ntiles);
  SPINED_HIP(hipMemcpy(dinst.ptr, hinst.data(), sizeof(BatchInst) * n_inst, hipMemcpyHostToDevice));
  SPINED_HIP(hipMemcpy(dtiles.ptr, htiles.data(), sizeof(Tile) * ntiles, hipMemcpyHostToDevice));
  SPINED_HIP(hipMemsetAsync(small.ptr, 0, small.bytes(), s));
  DeviceBuffer<double> partial(static_cast<uint64_t>(ntiles) * kPS);
  batch_zero_kernel<<<blocks_for(n_inst + 1, 256), 256, 0, s>>>(flags.ptr, n_inst + 1);
  SPINED_CHECK_LAUNCH();

  const unsigned ib = static_cast<unsigned>(n_inst);
  auto matvec = [&](int src, int dst, int final_pass) {
    if (method == 1) {
      if (nw == 1) batch_matvec_kernel<1, 1><<<ntiles, kTile, 0, s>>>(dinst.ptr, dtiles.ptr, src, dst, final_pass);
      else batch_matvec_kernel<1, 2><<<ntiles, kTile, 0, s>>>(dinst.ptr, dtiles.ptr, src, dst, final_pass);
    } else {
      if (nw == 1) batch_matvec_kernel<0, 1><<<ntiles, kTile, 0, s>>>(dinst.ptr, dtiles.ptr, src, dst, final_pass);
      else batch_matvec_kernel<0, 2><<<ntiles, kTile, 0, s>>>(dinst.ptr, dtiles.ptr, src, dst, final_pass);
    }
    SPINED_CHECK_LAUNCH();
  };
  auto multidot = [&](int nvec, int mode, int target, int j) {
    const unsigned chunks = mode == 1 ? 1u : static_cast<unsigned>((nvec + kChunk - 1) / kChunk);
    batch_multidot_kernel<<<dim3(ntiles, chunks), kTile, 0, s>>>(dinst.ptr, dtiles.ptr, nvec, mode,
                                                                 partial.ptr);
    SPINED_CHECK_LAUNCH();
    batch_reduce_kernel<<<ib, 256, 0, s>>>(dinst.ptr, partial.ptr, mode == 1 ? 1 : nvec, target, j);
    SPINED_CHECK_LAUNCH();
  };
  auto step = [&](int j) {
    matvec(j, -1, 0);
    multidot(j + 1, 0, 0, 0);
    batch_gemv_sub_kernel<<<ntiles, kTile, 0, s>>>(dinst.ptr, dtiles.ptr, j + 1, 0);
    SPINED_CHECK_LAUNCH();
    multidot(j + 1, 0, 1, 0);
    batch_gemv_sub_kernel<<<ntiles, kTile, 0, s>>>(dinst.ptr, dtiles.ptr, j + 1, 1);
    SPINED_CHECK_LAUNCH();
    batch_store_column_kernel<<<ib, kMaxM, 0, s>>>(dinst.ptr, j);
    SPINED_CHECK_LAUNCH();
    multidot(1, 1, 2, j);
    batch_scale_store_kernel<<<ntiles, kTile, 0, s>>>(dinst.ptr, dtiles.ptr, j, j + 1);
    SPINED_CHECK_LAUNCH();
  };
  // one restart cycle from step j0 (captured as a graph)
  DeviceBuffer<int> counters(2); // [0] cycle number, [1] instances still active
  auto cycle = [&](int j0) {
    batch_cycle_begin_kernel<<<1, 1, 0, s>>>(counters.ptr, counters.ptr + 1);
    SPINED_CHECK_LAUNCH();
    for (int j = j0; j < m; ++j) step(j);
    batch_jacobi_kernel<<<ib, 256, 0, s>>>(dinst.ptr, m);
    SPINED_CHECK_LAUNCH();
    batch_check_kernel<<<blocks_for(n_inst, 256), 256, 0, s>>>(dinst.ptr, n_inst, m, tol,
                                                               counters.ptr, counters.ptr + 1);
    SPINED_CHECK_LAUNCH();
    batch_restart_kernel<<<ntiles, kTile, 0, s>>>(dinst.ptr, dtiles.ptr, m, keep);
    SPINED_CHECK_LAUNCH();
    batch_reset_kernel<<<ib, 256, 0, s>>>(dinst.ptr, keep);
    SPINED_CHECK_LAUNCH();
  };
  auto capture = [&](int j0) {
    hipGraph_t g = nullptr;
    hipGraphExec_t e = nullptr;
    SPINED_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    cycle(j0);
    SPINED_HIP(hipStreamEndCapture(s, &g));
    SPINED_HIP(hipGraphInstantiate(&e, g, nullptr, nullptr, 0));
    SPINED_HIP(hipGraphDestroy(g));
    return e;
  };

  // start vectors: w = hash(seed + instance, r), v_0 = w / |w|
  batch_zero_kernel<<<1, 2, 0, s>>>(counters.ptr, 2);
  SPINED_CHECK_LAUNCH();
  batch_random_kernel<<<ntiles, kTile, 0, s>>>(dinst.ptr, dtiles.ptr, seed);
  SPINED_CHECK_LAUNCH();
  multidot(1, 1, 2, kMaxM);
  batch_scale_store_kernel<<<ntiles, kTile, 0, s>>>(dinst.ptr, dtiles.ptr, kMaxM, 0);
  SPINED_CHECK_LAUNCH();
  SPINED_HIP(hipStreamSynchronize(s));

  hipGraphExec_t first = capture(0), steady = capture(keep);
  std::vector<double> t_cycle;
  int done_cycles = 0;
  int active = n_inst;
  const auto t0 = std::chrono::steady_clock::now();
  for (int c = 0; c < max_cycles && active > 0; ++c) {
    SPINED_HIP(hipGraphLaunch(c == 0 ? first : steady, s));
    SPINED_HIP(hipMemcpyAsync(&active, counters.ptr + 1, sizeof(int), hipMemcpyDeviceToHost, s));
    SPINED_HIP(hipStreamSynchronize(s));
    t_cycle.push_back(std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    ++done_cycles;
    if (logging_enabled()) log_line("batch cycle %d: %d instances active", done_cycles, active);
  }
  (void)hipGraphExecDestroy(first);
  (void)hipGraphExecDestroy(steady);

  // finalisation: Ritz vector, H y, Rayleigh quotient, explicit residual
  batch_mark_unconverged_kernel<<<ib, 64, 0, s>>>(dinst.ptr, n_inst);
  SPINED_CHECK_LAUNCH();
  batch_ritz_kernel<<<ntiles, kTile, 0, s>>>(dinst.ptr, dtiles.ptr, m);
  SPINED_CHECK_LAUNCH();
  matvec(-1, m, 1);
  batch_final_dots_kernel<<<ntiles, kTile, 0, s>>>(dinst.ptr, dtiles.ptr, m, 0, partial.ptr);
  SPINED_CHECK_LAUNCH();
  batch_final_reduce_kernel<<<ib, 256, 0, s>>>(dinst.ptr, partial.ptr, 0);
  SPINED_CHECK_LAUNCH();
  batch_final_dots_kernel<<<ntiles, kTile, 0, s>>>(dinst.ptr, dtiles.ptr, m, 1, partial.ptr);
  SPINED_CHECK_LAUNCH();
  batch_final_reduce_kernel<<<ib, 256, 0, s>>>(dinst.ptr, partial.ptr, 1);
  SPINED_CHECK_LAUNCH();
  std::vector<int> hflags(n_inst);
  std::vector<double> hsmall(per_small * n_inst);
  SPINED_HIP(hipMemcpyAsync(hflags.data(), flags.ptr, sizeof(int) * n_inst, hipMemcpyDeviceToHost, s));
  SPINED_HIP(hipMemcpyAsync(hsmall.data(), small.ptr, small.bytes(), hipMemcpyDeviceToHost, s));
  SPINED_HIP(hipStreamSynchronize(s));
  const double total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
  int nconv = 0;
  for (int i = 0; i < n_inst; ++i) {
    const double *sc = hsmall.data() + per_small * i + (hinst[i].scal - hinst[i].T);
    evals[i] = sc[0];
    rnorms[i] = sc[2];
    const int f = hflags[i];
    if (cycles_out) cycles_out[i] = f;
    if (seconds_out) seconds_out[i] = f > 0 ? t_cycle[f - 1] : (t_cycle.empty() ? 0.0 : t_cycle.back());
    if (f > 0 && rnorms[i] <= tol) ++nconv;
  }
  if (stats) {
    stats->cycles = done_cycles;
    stats->converged = nconv;
    stats->seconds_total = total;
    stats->seconds_cycles = t_cycle.empty() ? 0.0 : t_cycle.back();
    stats->total_rows = total_rows;
    stats->matvecs_per_instance_max = done_cycles ? m + (done_cycles - 1) * (m - keep) + 1 : 0;
  }
  return nconv == n_inst ? LS_SUCCESS : LS_NOT_CONVERGED;
}

} // namespace spined
