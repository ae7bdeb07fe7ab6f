// blocklanczos.hpp -- block thick-restart Lanczos over row-interleaved
// blocks of K vectors (included by solver.hip only; uses its helpers).
//
// Every matvec of this phase is one block matvec (launch_matvec_block): the
// orbit scan, the index search and the gathered row line of each term serve
// K vectors at once, which is where the phase gains over the one-vector
// recurrence.  Step j of a cycle:
//   W = H V_j                                   (block matvec, FP64 rows)
//   C = V^T W;  W -= V C        (twice)         (FP64 dots, all blocks so far)
//   W = Q B                     (SVQB twice)    (Stathopoulos & Wu, SISC 23 (2002))
//   T[:, block j] = C1 + C2,  T[block j+1, block j] = B,  V_{j+1} = Q
// T is the dense projected matrix (<= 64), solved by the one-workgroup
// Jacobi; the residual estimate of Ritz pair i is |B s_i[last block]|.  The
// thick restart keeps `keepb` blocks of the lowest Ritz vectors (row-local,
// in place) followed by the last Q, exactly as the one-vector phase keeps
// Ritz vectors and v_m (Wu & Simon); the coupling rows of T are refilled by
// the next step's reorthogonalisation coefficients.
#pragma once

namespace spined {
namespace {

constexpr int kMaxKeep = 32;         // kept Ritz vectors per restart (registers)
constexpr int kMaxBlockCycles = 400; // the one-vector phase takes over after this

// partial[(b*K*K + a*K + c) * G + g] = sum over the rows of chunk g of
// V_b[r][a] * W[r][c] (FP64), for blocks b = blockIdx.x and chunks g =
// blockIdx.y (G = gridDim.y): the workgroups of one chunk are dispatched
// side by side, so W's rows, read once per block b, come from the Infinity
// Cache after the first read.
template <class VT, int K>
__global__ void __launch_bounds__(kBlock)
    block_dot_kernel(const VT *const *__restrict__ V, const double *__restrict__ W, uint64_t n,
                     double *__restrict__ partial) {
  __shared__ double lds[K * K][kBlock / 64];
  const int b = blockIdx.x;
  const uint64_t g = blockIdx.y, ng = gridDim.y;
  const RowK<VT, K> *vb = reinterpret_cast<const RowK<VT, K> *>(V[b]);
  const RowK<double, K> *wr = reinterpret_cast<const RowK<double, K> *>(W);
  double acc[K][K];
#pragma unroll
  for (int a = 0; a < K; ++a)
#pragma unroll
    for (int c = 0; c < K; ++c) acc[a][c] = 0.0;
  for (uint64_t r = g * kBlock + threadIdx.x; r < n; r += ng * kBlock) {
    const RowK<VT, K> v = vb[r];
    const RowK<double, K> w = wr[r];
#pragma unroll
    for (int a = 0; a < K; ++a)
#pragma unroll
      for (int c = 0; c < K; ++c) acc[a][c] = fma(to_d(v.v[a]), w.v[c], acc[a][c]);
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int a = 0; a < K; ++a)
#pragma unroll
    for (int c = 0; c < K; ++c) {
      double s = acc[a][c];
      for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
      if (lane == 0) lds[a * K + c][wid] = s;
    }
  __syncthreads();
  if (threadIdx.x < K * K) {
    double s = 0.0;
    for (int q = 0; q < kBlock / 64; ++q) s += lds[threadIdx.x][q];
    partial[(static_cast<uint64_t>(b) * K * K + threadIdx.x) * ng + g] = s;
  }
}

// W[r][c] -= sum_{b < nblk, a < K} V_b[r][a] C[(b*K + a)*K + c]
template <class VT, int K>
__global__ void __launch_bounds__(kBlock)
    block_sub_kernel(const VT *const *__restrict__ V, int nblk, const double *__restrict__ C,
                     double *__restrict__ W, uint64_t n) {
  RowK<double, K> *wr = reinterpret_cast<RowK<double, K> *>(W);
  for (uint64_t r = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x; r < n;
       r += static_cast<uint64_t>(gridDim.x) * kBlock) {
    double acc[K];
#pragma unroll
    for (int c = 0; c < K; ++c) acc[c] = 0.0;
    for (int b = 0; b < nblk; ++b) {
      const RowK<VT, K> v = reinterpret_cast<const RowK<VT, K> *>(V[b])[r];
      const double *cb = C + static_cast<size_t>(b) * K * K;
#pragma unroll
      for (int a = 0; a < K; ++a)
#pragma unroll
        for (int c = 0; c < K; ++c) acc[c] = fma(to_d(v.v[a]), cb[a * K + c], acc[c]);
    }
    RowK<double, K> w = wr[r];
#pragma unroll
    for (int c = 0; c < K; ++c) w.v[c] -= acc[c];
    wr[r] = w;
  }
}

// out[r][c] = sum_a W[r][a] X[a*K + c]   (in place allowed: row-local)
template <int K, class OT>
__global__ void __launch_bounds__(kBlock)
    block_transform_kernel(const double *W, const double *__restrict__ X, OT *out, uint64_t n) {
  const RowK<double, K> *wr = reinterpret_cast<const RowK<double, K> *>(W);
  RowK<OT, K> *orow = reinterpret_cast<RowK<OT, K> *>(out);
  for (uint64_t r = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x; r < n;
       r += static_cast<uint64_t>(gridDim.x) * kBlock) {
    const RowK<double, K> w = wr[r];
    RowK<OT, K> o;
#pragma unroll
    for (int c = 0; c < K; ++c) {
      double s = 0.0;
#pragma unroll
      for (int a = 0; a < K; ++a) s = fma(w.v[a], X[a * K + c], s);
      o.v[c] = static_cast<OT>(s);
    }
    orow[r] = o;
  }
}

// Scaled SVQB factors of the K x K Gram matrix G = W^T W (one thread; every
// loop over K is unrolled, so the matrices stay in registers):
//   G~ = D^-1/2 G D^-1/2 = U L U^T,  X = D^-1/2 U L^-1/2,  F = L^1/2 U^T D^1/2
// so that Q = W X is orthonormal and W = Q F.  Bout = F * Bin (Bin NULL: F).
// status[0] = 1 when the block is (numerically) rank deficient.
template <int K>
__global__ void __launch_bounds__(64)
    svqb_kernel(const double *__restrict__ G, double *__restrict__ X,
                const double *__restrict__ Bin, double *__restrict__ Bout, int *__restrict__ status) {
  if (threadIdx.x != 0) return;
  double A[K][K], U[K][K], d[K];
  bool deficient = false;
#pragma unroll
  for (int a = 0; a < K; ++a) {
    d[a] = G[a * K + a] > 0.0 ? sqrt(G[a * K + a]) : 0.0;
    if (!(d[a] > 0.0)) deficient = true;
  }
#pragma unroll
  for (int a = 0; a < K; ++a)
#pragma unroll
    for (int c = 0; c < K; ++c) {
      A[a][c] = deficient ? 0.0 : G[a * K + c] / (d[a] * d[c]);
      U[a][c] = a == c ? 1.0 : 0.0;
    }
  // cyclic Jacobi on the K x K matrix
  for (int sweep = 0; sweep < 30 && !deficient; ++sweep) {
    double off = 0.0;
#pragma unroll
    for (int p = 0; p < K; ++p)
#pragma unroll
      for (int q = p + 1; q < K; ++q) off += A[p][q] * A[p][q];
    if (off < 1e-32) break;
#pragma unroll
    for (int p = 0; p < K; ++p)
#pragma unroll
      for (int q = p + 1; q < K; ++q) {
        if (A[p][q] == 0.0) continue;
        const double tau = (A[q][q] - A[p][p]) / (2.0 * A[p][q]);
        const double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
        const double cs = 1.0 / sqrt(1.0 + t * t), sn = t * cs;
#pragma unroll
        for (int r = 0; r < K; ++r) {
          const double ap = A[r][p], aq = A[r][q];
          A[r][p] = cs * ap - sn * aq;
          A[r][q] = sn * ap + cs * aq;
        }
#pragma unroll
        for (int r = 0; r < K; ++r) {
          const double ap = A[p][r], aq = A[q][r];
          A[p][r] = cs * ap - sn * aq;
          A[q][r] = sn * ap + cs * aq;
        }
#pragma unroll
        for (int r = 0; r < K; ++r) {
          const double up = U[r][p], uq = U[r][q];
          U[r][p] = cs * up - sn * uq;
          U[r][q] = sn * up + cs * uq;
        }
      }
  }
  double lmax = 0.0;
#pragma unroll
  for (int a = 0; a < K; ++a) lmax = fmax(lmax, A[a][a]);
#pragma unroll
  for (int a = 0; a < K; ++a)
    if (!(A[a][a] > 1e-24 * lmax)) deficient = true;
  status[0] = deficient ? 1 : 0;
  double F[K][K];
#pragma unroll
  for (int a = 0; a < K; ++a)
#pragma unroll
    for (int c = 0; c < K; ++c) {
      const double l = deficient ? 1.0 : A[c][c];
      const double da = d[a] > 0.0 ? d[a] : 1.0;
      X[a * K + c] = U[a][c] / (da * sqrt(l));
      F[c][a] = sqrt(l) * U[a][c] * da; // F = L^1/2 U^T D^1/2
    }
#pragma unroll
  for (int a = 0; a < K; ++a)
#pragma unroll
    for (int c = 0; c < K; ++c) {
      double s = 0.0;
      if (Bin) {
#pragma unroll
        for (int q = 0; q < K; ++q) s += F[a][q] * Bin[q * K + c];
      } else {
        s = F[a][c];
      }
      Bout[a * K + c] = s;
    }
}

// Gram-Schmidt test after one projection pass (C1 = V^T W over `rows`
// basis vectors, G = W'^T W' of the projected block): column c keeps
// |w'_c|^2 = G_cc of its |w_c|^2 = G_cc + sum_q C1[q][c]^2 (V orthonormal).
// When every column keeps at least 1e-4 of it (1% of the norm: at most two
// digits lost to cancellation, the criterion of the one-vector phase) one
// pass suffices and *need = 0.  C2, the second pass's coefficients, is
// zeroed either way (a second pass overwrites it).
template <int K>
__global__ void second_pass_test_kernel(const double *__restrict__ C1, int rows,
                                        const double *__restrict__ G, double *__restrict__ C2,
                                        int *__restrict__ need) {
  for (int e = threadIdx.x; e < rows * K; e += blockDim.x) C2[e] = 0.0;
  if (threadIdx.x == 0) {
    int any = 0;
    for (int c = 0; c < K; ++c) {
      double cn = 0.0;
      for (int q = 0; q < rows; ++q) cn = fma(C1[q * K + c], C1[q * K + c], cn);
      const double wn = G[c * K + c];
      if (!(wn >= 1e-4 * (wn + cn))) any = 1; // NaN: take the second pass
    }
    *need = any;
  }
}

// T entries of block column j: rows q < (j+1)K from C1 + C2 (the block
// diagonal symmetrised), rows of block j+1 from B; mirrored.
template <int K>
__global__ void store_block_column_kernel(const double *__restrict__ C1,
                                          const double *__restrict__ C2,
                                          const double *__restrict__ B, int j, int last,
                                          double *__restrict__ T, int ldt) {
  const int rows = (j + 1) * K;
  for (int e = threadIdx.x; e < (rows + K) * K; e += blockDim.x) {
    const int q = e / K, c = e % K;
    const int col = j * K + c;
    double v;
    if (q < j * K) {
      v = C1[q * K + c] + C2[q * K + c];
    } else if (q < rows) {
      const int a = q - j * K;
      const double x = C1[q * K + c] + C2[q * K + c];
      const double y = C1[(j * K + c) * K + a] + C2[(j * K + c) * K + a];
      v = 0.5 * (x + y);
    } else {
      if (last) continue; // the coupling to the block after the last one
      v = B[(q - rows) * K + c];
    }
    T[q * ldt + col] = v;
    T[col * ldt + q] = v;
  }
}

// In place, row-local restart: the first p = keepb*K vectors become
// Y_i = sum_{q < M} S[i*M + q] V_q (vector q = row element q%K of block q/K),
// p <= KP (8, 16 or 32).  The coefficients are staged once per workgroup in
// LDS as [q][i] (zero for i >= p); each output sums in q order.
template <class VT, int K, int KP>
__global__ void __launch_bounds__(kBlock)
    block_restart_kernel(VT *const *__restrict__ V, int nblk, int keepb,
                         const double *__restrict__ S, uint64_t n) {
  const int M = nblk * K, p = keepb * K;
  __shared__ double sS[KP * kMaxBasis];
  for (int t = threadIdx.x; t < KP * M; t += kBlock) {
    const int q = t / KP, i = t % KP;
    sS[t] = i < p ? S[i * M + q] : 0.0;
  }
  __syncthreads();
  for (uint64_t r = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x; r < n;
       r += static_cast<uint64_t>(gridDim.x) * kBlock) {
    double out[KP];
#pragma unroll
    for (int i = 0; i < KP; ++i) out[i] = 0.0;
    for (int b = 0; b < nblk; ++b) {
      const RowK<VT, K> v = reinterpret_cast<const RowK<VT, K> *>(V[b])[r];
#pragma unroll
      for (int a = 0; a < K; ++a) {
        const double x = to_d(v.v[a]);
        const double *c = sS + (b * K + a) * KP;
#pragma unroll
        for (int i = 0; i < KP; ++i) out[i] = fma(c[i], x, out[i]);
      }
    }
#pragma unroll
    for (int ob = 0; ob < KP / K; ++ob) {
      if (ob < keepb) {
        RowK<VT, K> o;
#pragma unroll
        for (int a = 0; a < K; ++a) o.v[a] = static_cast<VT>(out[ob * K + a]);
        reinterpret_cast<RowK<VT, K> *>(V[ob])[r] = o;
      }
    }
  }
}

template <class VT, int K>
void launch_block_restart(VT *const *V, int nblk, int keepb, const double *S, uint64_t n,
                          unsigned G, hipStream_t s) {
  const int p = keepb * K;
  if (p <= 8)
    block_restart_kernel<VT, K, (K > 8 ? K : 8)><<<G, kBlock, 0, s>>>(V, nblk, keepb, S, n);
  else if (p <= 16)
    block_restart_kernel<VT, K, (K > 16 ? K : 16)><<<G, kBlock, 0, s>>>(V, nblk, keepb, S, n);
  else
    block_restart_kernel<VT, K, kMaxKeep><<<G, kBlock, 0, s>>>(V, nblk, keepb, S, n);
  SPINED_CHECK_LAUNCH();
}

// out[r] = sum_q s[q] V_q[r] over M = nblk*K block vectors (FP64)
template <class VT, int K>
__global__ void __launch_bounds__(kBlock)
    block_ritz_kernel(const VT *const *__restrict__ V, int nblk, const double *__restrict__ s,
                      double *__restrict__ out, uint64_t n) {
  for (uint64_t r = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x; r < n;
       r += static_cast<uint64_t>(gridDim.x) * kBlock) {
    double acc = 0.0;
    for (int b = 0; b < nblk; ++b) {
      const RowK<VT, K> v = reinterpret_cast<const RowK<VT, K> *>(V[b])[r];
#pragma unroll
      for (int a = 0; a < K; ++a) acc = fma(s[b * K + a], to_d(v.v[a]), acc);
    }
    out[r] = acc;
  }
}

// Start block: column 0 = start vector, columns 1.. = counter-hash uniforms.
template <int K>
__global__ void __launch_bounds__(kBlock)
    block_start_kernel(const double *__restrict__ start, uint64_t seed, double *__restrict__ W,
                       uint64_t n) {
  RowK<double, K> *wr = reinterpret_cast<RowK<double, K> *>(W);
  for (uint64_t r = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x; r < n;
       r += static_cast<uint64_t>(gridDim.x) * kBlock) {
    RowK<double, K> w;
    w.v[0] = start[r];
#pragma unroll
    for (int c = 1; c < K; ++c) w.v[c] = hash_uniform(seed + 0x9E3779B97F4A7C15ull * c, r);
    wr[r] = w;
  }
}

template <class VT, int K> struct BlockPhase {
  Operator &op;
  uint64_t n;
  int mb, k, keepb;
  hipStream_t s;
  std::vector<DeviceBuffer<VT>> blocks; // mb + 1 blocks of n*K
  std::vector<VT *> hptr;               // current order of the blocks
  DeviceBuffer<VT *> vptr;
  DeviceBuffer<const double *> wptr;
  DeviceBuffer<double> W, partial, C, G, X, B, Bt, T, theta, S;
  DeviceBuffer<int> status;
  std::vector<hipEvent_t> ev;
  int64_t matvecs = 0; // vector applications (K per block matvec)
  int64_t second_passes = 0; // steps whose Gram-Schmidt took a second pass
  double matvec_seconds = 0.0;
  unsigned Gd;
  int mlast = 0;

  BlockPhase(Operator &o, uint64_t n_, int mb_, int k_, int keepb_, hipStream_t s_)
      : op(o), n(n_), mb(mb_), k(k_), keepb(keepb_), s(s_) {
    blocks.resize(mb + 1);
    hptr.resize(mb + 1);
    for (int i = 0; i <= mb; ++i) {
      blocks[i].allocate(n * K);
      hptr[i] = blocks[i].ptr;
    }
    vptr.allocate(mb + 1);
    upload_pointers();
    W.allocate(n * K);
    const double *wp = W.ptr;
    wptr.allocate(1);
    SPINED_HIP(hipMemcpy(wptr.ptr, &wp, sizeof(wp), hipMemcpyHostToDevice));
    Gd = grid_for(n, 1024);
    partial.allocate(static_cast<size_t>(mb + 1) * K * K * Gd);
    C.allocate(2 * static_cast<size_t>(mb + 1) * K * K);
    G.allocate(K * K);
    X.allocate(K * K);
    B.allocate(K * K);
    Bt.allocate(K * K);
    T.allocate(kMaxBasis * kMaxBasis);
    theta.allocate(kMaxBasis);
    S.allocate(kMaxBasis * kMaxBasis);
    status.allocate(3);
    reset_projected_kernel<<<1, 256, 0, s>>>(T.ptr, kMaxBasis, theta.ptr, 0);
    SPINED_CHECK_LAUNCH();
  }
  ~BlockPhase() {
    for (auto e : ev) (void)hipEventDestroy(e);
  }

  void upload_pointers() {
    SPINED_HIP(hipMemcpyAsync(vptr.ptr, hptr.data(), sizeof(VT *) * (mb + 1),
                              hipMemcpyHostToDevice, s));
    SPINED_HIP(hipStreamSynchronize(s));
  }

  void harvest_events() {
    SPINED_HIP(hipStreamSynchronize(s));
    for (size_t i = 0; i + 1 < ev.size(); i += 2) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, ev[i], ev[i + 1]) == hipSuccess) matvec_seconds += ms * 1e-3;
      (void)hipEventDestroy(ev[i]);
      (void)hipEventDestroy(ev[i + 1]);
    }
    ev.clear();
  }

  void block_matvec(const VT *x, double *y) {
    hipEvent_t a, b;
    SPINED_HIP(hipEventCreate(&a));
    SPINED_HIP(hipEventCreate(&b));
    SPINED_HIP(hipEventRecord(a, s));
    const VecType xt = std::is_same<VT, float>::value ? VecType::F32 : VecType::F64;
    launch_matvec_block(MatvecBlockCall{&op, xt, VecType::F64, x, y, n, K, s});
    SPINED_HIP(hipEventRecord(b, s));
    ev.push_back(a);
    ev.push_back(b);
    matvecs += K;
  }

  // out[(b*K + a)*K + c] = sum_r V_b[r][a] W[r][c] for b < nblk
  template <class PT> void dots(const PT *const *V, int nblk, double *out) {
    block_dot_kernel<PT, K><<<dim3(nblk, Gd), kBlock, 0, s>>>(V, W.ptr, n, partial.ptr);
    SPINED_CHECK_LAUNCH();
    reduce_partials_kernel<<<nblk * K * K, kBlock, 0, s>>>(partial.ptr, nblk * K * K, Gd, out, 0);
    SPINED_CHECK_LAUNCH();
  }

  // Classical Gram-Schmidt of W against the first nblk blocks: one pass,
  // and a second when second_pass_test_kernel asks for it.  Returns true
  // when G already holds W^T W of the projected block (one pass: the test's
  // Gram matrix is the first SVQB pass's).
  bool reorth(int nblk, double *c1, double *c2) {
    dots<VT>(vptr.ptr, nblk, c1);
    block_sub_kernel<VT, K><<<Gd, kBlock, 0, s>>>(vptr.ptr, nblk, c1, W.ptr, n);
    SPINED_CHECK_LAUNCH();
    dots<double>(wptr.ptr, 1, G.ptr);
    second_pass_test_kernel<K><<<1, 256, 0, s>>>(c1, nblk * K, G.ptr, c2, status.ptr + 2);
    SPINED_CHECK_LAUNCH();
    int need = 1;
    SPINED_HIP(hipMemcpyAsync(&need, status.ptr + 2, sizeof(int), hipMemcpyDeviceToHost, s));
    SPINED_HIP(hipStreamSynchronize(s));
    if (!need) return true;
    ++second_passes;
    dots<VT>(vptr.ptr, nblk, c2);
    block_sub_kernel<VT, K><<<Gd, kBlock, 0, s>>>(vptr.ptr, nblk, c2, W.ptr, n);
    SPINED_CHECK_LAUNCH();
    return false;
  }

  // W = Q B with Q orthonormal (two SVQB passes) stored into block `dst`;
  // returns false when W is rank deficient.  gram_ready: G holds W^T W.
  bool orthonormalize_into(int dst, bool gram_ready = false) {
    if (!gram_ready) dots<double>(wptr.ptr, 1, G.ptr);
    svqb_kernel<K><<<1, 64, 0, s>>>(G.ptr, X.ptr, nullptr, Bt.ptr, status.ptr);
    SPINED_CHECK_LAUNCH();
    block_transform_kernel<K, double><<<Gd, kBlock, 0, s>>>(W.ptr, X.ptr, W.ptr, n);
    SPINED_CHECK_LAUNCH();
    dots<double>(wptr.ptr, 1, G.ptr);
    svqb_kernel<K><<<1, 64, 0, s>>>(G.ptr, X.ptr, Bt.ptr, B.ptr, status.ptr + 1);
    SPINED_CHECK_LAUNCH();
    block_transform_kernel<K, VT><<<Gd, kBlock, 0, s>>>(W.ptr, X.ptr, hptr[dst], n);
    SPINED_CHECK_LAUNCH();
    int st[2] = {0, 0};
    SPINED_HIP(hipMemcpyAsync(st, status.ptr, sizeof(st), hipMemcpyDeviceToHost, s));
    SPINED_HIP(hipStreamSynchronize(s));
    return st[0] == 0 && st[1] == 0;
  }

  // V_0 from an FP64 start vector (column 0) plus random columns
  bool set_start(const double *start, uint64_t seed) {
    block_start_kernel<K><<<Gd, kBlock, 0, s>>>(start, seed, W.ptr, n);
    SPINED_CHECK_LAUNCH();
    return orthonormalize_into(0);
  }

  bool step(int j) {
    block_matvec(hptr[j], W.ptr);
    double *c1 = C.ptr, *c2 = C.ptr + static_cast<size_t>(mb + 1) * K * K;
    const bool gram = reorth(j + 1, c1, c2);
    const bool ok = orthonormalize_into(j + 1, gram);
    store_block_column_kernel<K><<<1, 256, 0, s>>>(c1, c2, B.ptr, j, j + 1 == mb ? 1 : 0, T.ptr,
                                                   kMaxBasis);
    SPINED_CHECK_LAUNCH();
    return ok;
  }

  PhaseResult run(double tol, int64_t max_matvecs, bool stall_detect, int phase_id,
                  spined_monitor_fn monitor, void *ctx, bool verbose, bool &deficient) {
    PhaseResult res;
    deficient = false;
    int j0 = 0;
    double best = 1e300;
    int stalled = 0;
    const int M = mb * K;
    std::vector<double> th(M), Sh(static_cast<size_t>(M) * M), Bh(K * K);
    auto project = [&](int nb, double &worst, int &nconv) {
      const int msz = nb * K;
      jacobi_eigh_kernel<<<1, 256, 0, s>>>(T.ptr, kMaxBasis, msz, theta.ptr, S.ptr, 40);
      SPINED_CHECK_LAUNCH();
      SPINED_HIP(hipMemcpyAsync(th.data(), theta.ptr, 8 * msz, hipMemcpyDeviceToHost, s));
      SPINED_HIP(hipMemcpyAsync(Sh.data(), S.ptr, 8ull * msz * msz, hipMemcpyDeviceToHost, s));
      SPINED_HIP(hipMemcpyAsync(Bh.data(), B.ptr, 8 * K * K, hipMemcpyDeviceToHost, s));
      SPINED_HIP(hipStreamSynchronize(s));
      res.theta.assign(th.begin(), th.begin() + k);
      res.est.resize(k);
      worst = 0.0;
      nconv = 0;
      for (int i = 0; i < k; ++i) {
        const double *si = Sh.data() + static_cast<size_t>(i) * msz + (msz - K);
        double r2 = 0.0;
        for (int a = 0; a < K; ++a) {
          double t = 0.0;
          for (int c = 0; c < K; ++c) t += Bh[a * K + c] * si[c];
          r2 += t * t;
        }
        res.est[i] = std::sqrt(r2);
        worst = std::max(worst, res.est[i]);
        if (res.est[i] <= tol) ++nconv;
      }
      mlast = msz;
    };
    for (;;) {
      double worst = 0.0;
      int nconv = 0;
      bool early = false;
      for (int j = j0; j < mb; ++j) {
        if (!step(j)) {
          harvest_events();
          deficient = true;
          return res;
        }
        const int nb = j + 1;
        if (nb < mb && nb * K >= k) {
          project(nb, worst, nconv);
          if (nconv == k) {
            early = true;
            break;
          }
        }
      }
      if (!early) project(mb, worst, nconv);
      harvest_events();
      ++res.cycles;
      if (verbose)
        log_line("block phase %d (K=%d) cycle %d matvecs %lld theta0 %.15g est %.3e "
                 "second Gram-Schmidt passes %lld",
                 phase_id, K, res.cycles, static_cast<long long>(matvecs), th[0], worst,
                 static_cast<long long>(second_passes));
      int stop = 0;
      if (monitor) {
        spined_monitor_info info{matvecs, res.cycles, phase_id, nconv, k, res.theta.data(),
                                 res.est.data()};
        stop = monitor(&info, ctx);
      }
      if (nconv == k) {
        res.converged = true;
        return res;
      }
      res.stopped = stop != 0;
      if (stop || (max_matvecs > 0 && matvecs >= max_matvecs) || res.cycles >= kMaxBlockCycles)
        return res;
      if (stall_detect) {
        if (worst < 0.5 * best) {
          best = worst;
          stalled = 0;
        } else if (++stalled >= 2) {
          return res;
        }
      }
      // thick restart: keep the lowest keepb*K Ritz vectors, then the last Q
      launch_block_restart<VT, K>(vptr.ptr, mb, keepb, S.ptr, n, Gd, s);
      std::swap(hptr[keepb], hptr[mb]);
      upload_pointers();
      reset_projected_kernel<<<1, 256, 0, s>>>(T.ptr, kMaxBasis, theta.ptr, keepb * K);
      SPINED_CHECK_LAUNCH();
      j0 = keepb;
    }
  }

  // FP64 Ritz vector i of the last projected problem into out
  void ritz(int i, double *out) {
    block_ritz_kernel<VT, K><<<Gd, kBlock, 0, s>>>(vptr.ptr, mlast / K,
                                                   S.ptr + static_cast<size_t>(i) * mlast, out, n);
    SPINED_CHECK_LAUNCH();
  }
};

} // namespace
} // namespace spined
