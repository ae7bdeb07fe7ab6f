// solver.hip -- device-resident eigensolver (replaces PRIMME's eigh at
// src/SpinED.hs:383-404, which calls back into ls_operator_matmat through
// inplaceApply, src/SpinED/Internal.hs:411-429, on every iteration).
//
// Thick-restart Lanczos (Wu & Simon, SIAM J. Matrix Anal. Appl. 22 (2000))
// with full (twice classical Gram-Schmidt) reorthogonalisation:
//   step j:  w = H v_j;  h = V_j^T w;  w -= V_j h;  (twice, FP64 dots)
//            T[:j, j] = h;  beta_j = |w|;  v_{j+1} = w / beta_j
//   cycle:   T = S diag(theta) S^T (one-workgroup Jacobi, FP64, on device);
//            residual estimates |beta_{m-1} S[m-1, i]|;
//            restart keeps the lowest Ritz vectors V S[:, :keep] (row-local,
//            in place) and v_m.
// Everything stays on the GPU; the host reads m + k scalars per cycle.
//
// Precision modes: FP64 stores V in double.  MIXED stores V in float while
// the matvec accumulates rows in FP64 into a double w, and the dots, the
// recurrence, the projected solve and the Ritz assembly are FP64; once the
// FP32-storage phase stalls at its floor (~1e-6 |H|) the Ritz vector is
// assembled in FP64 and an FP64 phase restarts from it to reach tol.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "engine.hpp"
#include "jacobi.hpp"
#include "rowgather.hpp"

namespace spined {

namespace {

constexpr int kBlock = 256;
constexpr int kMaxBasis = 64;
constexpr int kCacheMinBasis = 16; // smallest automatic basis that keeps a gather cache
constexpr int kDotChunk = 8;

unsigned grid_for(uint64_t n, uint64_t cap = 8192) {
  uint64_t g = (n + kBlock - 1) / kBlock;
  return static_cast<unsigned>(std::max<uint64_t>(1, std::min(g, cap)));
}

__device__ inline double to_d(float v) { return static_cast<double>(v); }
__device__ inline double to_d(double v) { return v; }

__global__ void random_fill_kernel(double *v, uint64_t n, uint64_t seed) {
  for (uint64_t r = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; r < n;
       r += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    v[r] = hash_uniform(seed, r);
}

// partial[c*G + b] for vectors i in [8c, 8c+8): sum over this block's rows of
// V_i[r] * w[r] (FP64).  V vectors are given by a device array of pointers.
template <class VT>
__global__ void __launch_bounds__(kBlock)
    multidot_kernel(const VT *const *__restrict__ V, int nvec, const double *__restrict__ w,
                    uint64_t n, double *__restrict__ partial) {
  __shared__ double lds[kDotChunk][kBlock / 64];
  const int c = blockIdx.y;
  const int i0 = c * kDotChunk;
  double acc[kDotChunk];
#pragma unroll
  for (int q = 0; q < kDotChunk; ++q) acc[q] = 0.0;
  const VT *vp[kDotChunk];
#pragma unroll
  for (int q = 0; q < kDotChunk; ++q) vp[q] = (i0 + q < nvec) ? V[i0 + q] : nullptr;
  for (uint64_t r = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x; r < n;
       r += static_cast<uint64_t>(gridDim.x) * kBlock) {
    const double wr = w[r];
#pragma unroll
    for (int q = 0; q < kDotChunk; ++q)
      if (i0 + q < nvec) acc[q] = fma(to_d(vp[q][r]), wr, acc[q]);
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < kDotChunk; ++q) {
    double v = acc[q];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if (lane == 0) lds[q][wid] = v;
  }
  __syncthreads();
  if (threadIdx.x < kDotChunk && i0 + static_cast<int>(threadIdx.x) < nvec) {
    double s = 0.0;
    for (int k = 0; k < kBlock / 64; ++k) s += lds[threadIdx.x][k];
    partial[static_cast<uint64_t>(i0 + threadIdx.x) * gridDim.x + blockIdx.x] = s;
  }
}

// h[i] (+)= sum_b partial[i*G + b], fixed order (deterministic).  Does
// nothing when skip is given and *skip != 0.
__global__ void reduce_partials_kernel(const double *__restrict__ partial, int nvec, int G,
                                       double *__restrict__ h, int accumulate,
                                       const int *__restrict__ skip = nullptr) {
  if (skip && *skip) return;
  const int i = blockIdx.x;
  __shared__ double lds[kBlock];
  double s = 0.0;
  for (int b = threadIdx.x; b < G; b += kBlock) s += partial[static_cast<uint64_t>(i) * G + b];
  lds[threadIdx.x] = s;
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if (threadIdx.x < static_cast<unsigned>(off)) lds[threadIdx.x] += lds[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0 && i < nvec) h[i] = accumulate ? h[i] + lds[0] : lds[0];
}

// w[r] -= sum_{i < nvec} h[i] V_i[r]
template <class VT>
__global__ void __launch_bounds__(kBlock)
    gemv_sub_kernel(const VT *const *__restrict__ V, int nvec, const double *__restrict__ h,
                    double *__restrict__ w, uint64_t n) {
  for (uint64_t r = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x; r < n;
       r += static_cast<uint64_t>(gridDim.x) * kBlock) {
    double acc = 0.0;
    for (int i = 0; i < nvec; ++i) acc = fma(h[i], to_d(V[i][r]), acc);
    w[r] -= acc;
  }
}

// Complex Hermitian Lanczos (vectors of double2, addressed as 2n doubles by
// the componentwise kernels): only the reorthogonalisation needs the
// conjugated dots and the complex update.
// partial[(2 (8c + q) + {0: re, 1: im}) * G + b] = sum over this block's rows
// of conj(V_i[r]) w[r].
__global__ void __launch_bounds__(kBlock)
    cmultidot_kernel(const double2 *const *__restrict__ V, int nvec,
                     const double2 *__restrict__ w, uint64_t n, double *__restrict__ partial) {
  __shared__ double lds[2 * kDotChunk][kBlock / 64];
  const int c = blockIdx.y;
  const int i0 = c * kDotChunk;
  double re[kDotChunk], im[kDotChunk];
#pragma unroll
  for (int q = 0; q < kDotChunk; ++q) re[q] = im[q] = 0.0;
  const double2 *vp[kDotChunk];
#pragma unroll
  for (int q = 0; q < kDotChunk; ++q) vp[q] = (i0 + q < nvec) ? V[i0 + q] : nullptr;
  for (uint64_t r = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x; r < n;
       r += static_cast<uint64_t>(gridDim.x) * kBlock) {
    const double2 wr = w[r];
#pragma unroll
    for (int q = 0; q < kDotChunk; ++q)
      if (i0 + q < nvec) {
        const double2 v = vp[q][r];
        re[q] = fma(v.y, wr.y, fma(v.x, wr.x, re[q]));
        im[q] = fma(-v.y, wr.x, fma(v.x, wr.y, im[q]));
      }
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < kDotChunk; ++q) {
    double a = re[q], b = im[q];
    for (int off = 32; off > 0; off >>= 1) {
      a += __shfl_down(a, off, 64);
      b += __shfl_down(b, off, 64);
    }
    if (lane == 0) {
      lds[2 * q][wid] = a;
      lds[2 * q + 1][wid] = b;
    }
  }
  __syncthreads();
  if (threadIdx.x < 2 * kDotChunk && i0 + static_cast<int>(threadIdx.x / 2) < nvec) {
    double sum = 0.0;
    for (int k = 0; k < kBlock / 64; ++k) sum += lds[threadIdx.x][k];
    partial[(2 * static_cast<uint64_t>(i0) + threadIdx.x) * gridDim.x + blockIdx.x] = sum;
  }
}

// w[r] -= sum_i h_i V_i[r], h interleaved (re, im)
__global__ void __launch_bounds__(kBlock)
    cgemv_sub_kernel(const double2 *const *__restrict__ V, int nvec, const double *__restrict__ h,
                     double2 *__restrict__ w, uint64_t n) {
  for (uint64_t r = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x; r < n;
       r += static_cast<uint64_t>(gridDim.x) * kBlock) {
    double ar = 0.0, ai = 0.0;
    for (int i = 0; i < nvec; ++i) {
      const double2 v = V[i][r];
      const double hr = h[2 * i], hi = h[2 * i + 1];
      ar = fma(-hi, v.y, fma(hr, v.x, ar));
      ai = fma(hi, v.x, fma(hr, v.y, ai));
    }
    double2 x = w[r];
    x.x -= ar;
    x.y -= ai;
    w[r] = x;
  }
}

constexpr int kFuseMax = 32; // vectors held in registers by the fused passes
// The second Gram-Schmidt pass runs only when the first update keeps less
// than this fraction of |w| (second_pass_test_kernel).
constexpr double kSecondPassEta = 1e-2;
constexpr int64_t kAutoMaxMatvecs = 20000; // spined_eigsh_options.max_matvecs = 0

// Block reduction of acc[q] (q < nq) into partial[q * gridDim.x + blockIdx.x]
// with the multidot kernel's order (wave shuffles, then the 4 waves in
// order), so fused and unfused passes give bit-identical sums.
template <int NQ>
__device__ inline void block_partials(const double (&acc)[NQ], int nq, double *__restrict__ partial) {
  __shared__ double lds[NQ][kBlock / 64];
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
    double sum = 0.0;
    for (int k = 0; k < kBlock / 64; ++k) sum += lds[threadIdx.x][k];
    partial[static_cast<uint64_t>(threadIdx.x) * gridDim.x + blockIdx.x] = sum;
  }
}

// Second Gram-Schmidt pass fused with the first update:
//   w[r] -= sum_i h1[i] V_i[r];  partial dots of V_i with the new w.
// Per row the arithmetic equals gemv_sub_kernel followed by multidot_kernel.
// Does nothing when skip is given and *skip != 0.
template <class VT, int NQ>
__global__ void __launch_bounds__(kBlock)
    sub_dot_kernel(const VT *const *__restrict__ V, int nvec, const double *__restrict__ h1,
                   double *__restrict__ w, uint64_t n, double *__restrict__ partial,
                   const int *__restrict__ skip = nullptr) {
  if (skip && *skip) return;
  // pointers hoisted out of the row loop; slots past nvec read V[0] (a valid
  // address) so every load is unconditional and the loads of a row are all
  // in flight together; their values are masked out of the arithmetic
  const VT *vp[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) vp[q] = V[q < nvec ? q : 0];
  double acc[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) acc[q] = 0.0;
  for (uint64_t r = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x; r < n;
       r += static_cast<uint64_t>(gridDim.x) * kBlock) {
    double vals[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) vals[q] = to_d(vp[q][r]);
    double sub = 0.0;
#pragma unroll
    for (int q = 0; q < NQ; ++q) sub = q < nvec ? fma(h1[q], vals[q], sub) : sub;
    const double wr = w[r] - sub;
    w[r] = wr;
#pragma unroll
    for (int q = 0; q < NQ; ++q) acc[q] = q < nvec ? fma(vals[q], wr, acc[q]) : acc[q];
  }
  block_partials<NQ>(acc, nvec, partial);
}

// Last update fused with the norm: w[r] -= sum_i h2[i] V_i[r]; partial
// sums of w[r]^2 (the norm_w multidot's arithmetic).  When *skip != 0 (the
// first pass was the only one) the update uses h1 and stores the next basis
// vector directly, vnext[r] = (w[r] - sum_i h1[i] V_i[r]) / *beta, with the
// Pythagorean beta of second_pass_test_kernel (w and the partials are then
// left alone).
template <class VT, int NQ>
__global__ void __launch_bounds__(kBlock)
    sub_norm_kernel(const VT *const *__restrict__ V, int nvec, const double *__restrict__ h2,
                    double *__restrict__ w, uint64_t n, double *__restrict__ partial,
                    const double *__restrict__ h1, const int *__restrict__ skip,
                    VT *__restrict__ vnext, const double *__restrict__ beta) {
  const bool store = *skip != 0;
  const double *c = store ? h1 : h2;
  const double b = store ? *beta : 0.0;
  const double inv = b > 0.0 ? 1.0 / b : 0.0;
  const VT *vp[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) vp[q] = V[q < nvec ? q : 0];
  double acc[1] = {0.0};
  for (uint64_t r = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x; r < n;
       r += static_cast<uint64_t>(gridDim.x) * kBlock) {
    double vals[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) vals[q] = to_d(vp[q][r]);
    double sub = 0.0;
#pragma unroll
    for (int q = 0; q < NQ; ++q) sub = q < nvec ? fma(c[q], vals[q], sub) : sub;
    const double wr = w[r] - sub;
    if (store) {
      vnext[r] = static_cast<VT>(wr * inv);
    } else {
      w[r] = wr;
      acc[0] = fma(wr, wr, acc[0]);
    }
  }
  if (!store) block_partials<1>(acc, 1, partial);
}

// First Gram-Schmidt pass with the squared norm of w: partial dots of V_i
// (i < nvec <= kFuseMax) with w in slots 0..nvec-1 and w.w in slot nvec.
template <class VT, int NQ>
__global__ void __launch_bounds__(kBlock)
    dot_ww_kernel(const VT *const *__restrict__ V, int nvec, const double *__restrict__ w,
                  uint64_t n, double *__restrict__ partial) {
  const VT *vp[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) vp[q] = V[q < nvec ? q : 0];
  double acc[NQ + 1];
#pragma unroll
  for (int q = 0; q <= NQ; ++q) acc[q] = 0.0;
  for (uint64_t r = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x; r < n;
       r += static_cast<uint64_t>(gridDim.x) * kBlock) {
    double vals[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) vals[q] = to_d(vp[q][r]);
    const double wr = w[r];
#pragma unroll
    for (int q = 0; q < NQ; ++q) acc[q] = q < nvec ? fma(vals[q], wr, acc[q]) : acc[q];
    acc[NQ] = fma(wr, wr, acc[NQ]);
  }
  // w.w moves to slot nvec
  double out[NQ + 1];
#pragma unroll
  for (int q = 0; q <= NQ; ++q) out[q] = q < nvec ? acc[q] : (q == nvec ? acc[NQ] : 0.0);
  block_partials<NQ + 1>(out, nvec + 1, partial);
}

// Whether the second Gram-Schmidt pass can be left out: h[0..nvec) are the
// first-pass coefficients and h[nvec] = w.w.  After the first update
// |w'|^2 = w.w - |h|^2; when |w'| >= eta |w| the first pass lost at most a
// factor 1/eta of accuracy to cancellation and its result is kept (the
// criterion of Daniel, Gragg, Kaufman & Stewart, Math. Comp. 30 (1976) 772,
// with eta below their 1/sqrt(2)).  Sets *skip and, when set, h2 = 0 and
// *beta = |w'| = sqrt(w.w - |h|^2) (relative error below eps / eta^2).
__global__ void second_pass_test_kernel(const double *__restrict__ h, int nvec, double eta,
                                        double *__restrict__ h2, int *__restrict__ skip,
                                        double *__restrict__ beta) {
  __shared__ int flag;
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < nvec; ++i) s = fma(h[i], h[i], s);
    const double ww = h[nvec];
    flag = (ww > 0.0 && ww - s >= eta * eta * ww) ? 1 : 0;
    *skip = flag;
    if (flag) *beta = sqrt(ww - s);
  }
  __syncthreads();
  if (flag)
    for (int i = threadIdx.x; i < nvec; i += blockDim.x) h2[i] = 0.0;
}

// v[r] = w[r] / (*beta) (beta == 0 -> zero vector); nothing when skip is
// given and set (sub_norm_kernel stored v already)
template <class VT>
__global__ void __launch_bounds__(kBlock)
    scale_store_kernel(const double *__restrict__ w, const double *__restrict__ beta,
                       VT *__restrict__ v, uint64_t n, const int *__restrict__ skip = nullptr) {
  if (skip && *skip) return;
  const double b = *beta;
  const double inv = b > 0.0 ? 1.0 / b : 0.0;
  for (uint64_t r = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x; r < n;
       r += static_cast<uint64_t>(gridDim.x) * kBlock)
    v[r] = static_cast<VT>(w[r] * inv);
}

__global__ void sqrt_inplace_kernel(double *v, const int *__restrict__ skip = nullptr) {
  if (skip && *skip) return;
  v[0] = sqrt(v[0]);
}

// T[i*ldt + j] = T[j*ldt + i] = h[i], i <= j; beta store
__global__ void store_column_kernel(const double *__restrict__ h, int j, double *__restrict__ T,
                                    int ldt) {
  const int i = threadIdx.x;
  if (i <= j) {
    T[i * ldt + j] = h[i];
    T[j * ldt + i] = h[i];
  }
}

__global__ void __launch_bounds__(256)
    jacobi_eigh_kernel(const double *__restrict__ T, int ldt, int m, double *__restrict__ evals,
                       double *__restrict__ S, int max_sweeps) {
  jacobi_eigh_device(T, ldt, m, evals, S, max_sweeps);
}

constexpr int kMaxKeepOne = 32; // kept Ritz vectors of the one-vector restart

// In place, row-local: V_i[r] <- sum_q S[i*m + q] V_q[r] for i < keep <= KB
// (KB = 8, 16 or 32).  The coefficients are staged once per workgroup in LDS
// as [q][i] (zero for i >= keep), so each input value meets its KB
// coefficients in broadcast LDS reads; the outputs live in registers and the
// inputs stream once.  Each output sums in q order, as a plain loop would.
template <class VT, int KB>
__global__ void __launch_bounds__(kBlock)
    restart_kernel(VT *const *__restrict__ V, int m, int keep, const double *__restrict__ S,
                   uint64_t n) {
  __shared__ double sS[KB * kMaxBasis];
  for (int t = threadIdx.x; t < KB * m; t += kBlock) {
    const int q = t / KB, i = t % KB;
    sS[t] = i < keep ? S[i * m + q] : 0.0;
  }
  __syncthreads();
  for (uint64_t r = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x; r < n;
       r += static_cast<uint64_t>(gridDim.x) * kBlock) {
    double out[KB];
#pragma unroll
    for (int i = 0; i < KB; ++i) out[i] = 0.0;
    for (int q = 0; q < m; ++q) {
      const double x = to_d(V[q][r]);
      const double *c = sS + q * KB;
#pragma unroll
      for (int i = 0; i < KB; ++i) out[i] = fma(c[i], x, out[i]);
    }
#pragma unroll
    for (int i = 0; i < KB; ++i)
      if (i < keep) V[i][r] = static_cast<VT>(out[i]);
  }
}

template <class VT>
void launch_restart(VT *const *V, int m, int keep, const double *S, uint64_t n, unsigned G,
                    hipStream_t s) {
  if (keep <= 8)
    restart_kernel<VT, 8><<<G, kBlock, 0, s>>>(V, m, keep, S, n);
  else if (keep <= 16)
    restart_kernel<VT, 16><<<G, kBlock, 0, s>>>(V, m, keep, S, n);
  else
    restart_kernel<VT, 32><<<G, kBlock, 0, s>>>(V, m, keep, S, n);
  SPINED_CHECK_LAUNCH();
}

// out[r] = sum_q S[col*m + q] V_q[r] (FP64 Ritz vector)
template <class VT>
__global__ void __launch_bounds__(kBlock)
    ritz_kernel(const VT *const *__restrict__ V, int m, const double *__restrict__ s,
                double *__restrict__ out, uint64_t n) {
  for (uint64_t r = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x; r < n;
       r += static_cast<uint64_t>(gridDim.x) * kBlock) {
    double acc = 0.0;
    for (int q = 0; q < m; ++q) acc = fma(s[q], to_d(V[q][r]), acc);
    out[r] = acc;
  }
}

// w[r] = hy[r] - theta * y[r]
__global__ void __launch_bounds__(kBlock)
    residual_kernel(const double *__restrict__ hy, const double *__restrict__ y, double theta,
                    double *__restrict__ w, uint64_t n) {
  for (uint64_t r = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x; r < n;
       r += static_cast<uint64_t>(gridDim.x) * kBlock)
    w[r] = fma(-theta, y[r], hy[r]);
}

// b[r] = B(a[r] * s); b may be a (in place)
template <class B>
__global__ void __launch_bounds__(kBlock) scale_convert_kernel(const double *a, double s, B *b,
                                                               uint64_t n) {
  for (uint64_t r = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x; r < n;
       r += static_cast<uint64_t>(gridDim.x) * kBlock)
    b[r] = static_cast<B>(a[r] * s);
}

template <class A, class B>
__global__ void __launch_bounds__(kBlock) convert_kernel(const A *__restrict__ a, B *__restrict__ b,
                                                         uint64_t n) {
  for (uint64_t r = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x; r < n;
       r += static_cast<uint64_t>(gridDim.x) * kBlock)
    b[r] = static_cast<B>(a[r]);
}


// stride 2: h1, h2 hold complex coefficients (re, im); T takes the real
// parts (the projected matrix of a Hermitian operator is real symmetric in
// this basis; the imaginary parts are rounding).
__global__ void store_column2_kernel(const double *__restrict__ h1, const double *__restrict__ h2,
                                     int j, double *__restrict__ T, int ldt, int stride) {
  const int i = threadIdx.x;
  if (i <= j) {
    const double v = h1[stride * i] + h2[stride * i];
    T[i * ldt + j] = v;
    T[j * ldt + i] = v;
  }
}

// T <- diag(theta[0..keep)), zero elsewhere
__global__ void reset_projected_kernel(double *__restrict__ T, int ldt,
                                       const double *__restrict__ theta, int keep) {
  for (int e = threadIdx.x; e < ldt * ldt; e += blockDim.x) {
    const int i = e / ldt, j = e % ldt;
    T[e] = (i == j && i < keep) ? theta[i] : 0.0;
  }
}

double seconds_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

struct PhaseResult {
  std::vector<double> theta;  // lowest k Ritz values
  std::vector<double> est;    // residual estimates
  bool converged = false;
  bool stopped = false;       // the monitor asked to stop
  int cycles = 0;
};

// One thick-restart Lanczos phase with vector storage type VT.
// C = true: complex Hermitian operator; VT = double and every vector holds
// n double2 values, i.e. ne = 2n doubles for the componentwise kernels.
template <class VT, bool C = false> struct Phase {
  Operator &op;
  uint64_t n;
  uint64_t ne; // doubles (or floats) per vector
  int m, k, keep;
  hipStream_t s;
  std::vector<DeviceBuffer<VT>> vecs;
  DeviceBuffer<VT *> vptr;
  DeviceBuffer<const double *> wptr;
  DeviceBuffer<double> w, partial, h, T, theta, S, beta;
  DeviceBuffer<int> skip; // second Gram-Schmidt pass left out (FP64 phase)
  std::vector<hipEvent_t> ev;
  int64_t matvecs = 0;
  double matvec_seconds = 0.0;
  unsigned G;
  bool small;
  int mlast = 0; // size of the last projected problem (Ritz assembly)
  // checkpoint after every cp_every-th restart (spined_eigsh_options)
  spined_checkpoint_fn cp_fn = nullptr;
  void *cp_ctx = nullptr;
  int cp_every = 0;
  void set_checkpoint(const spined_eigsh_options &o) {
    cp_fn = o.checkpoint;
    cp_ctx = o.checkpoint_ctx;
    cp_every = o.checkpoint_every;
  }

  Phase(Operator &o, uint64_t n_, int m_, int k_, int keep_, hipStream_t s_)
      : op(o), n(n_), ne(C ? 2 * n_ : n_), m(m_), k(k_), keep(keep_), s(s_) {
    static_assert(!C || std::is_same<VT, double>::value, "complex vectors are stored in FP64");
    // FP64-sized pooled blocks for either storage type: the FP32 phase of a
    // mixed solve hands them on to the FP64 phase without a free/allocate
    vecs.resize(m + 1);
    for (auto &v : vecs) v.allocate_pooled(ne, ne * sizeof(double));
    std::vector<VT *> hp(m + 1);
    for (int i = 0; i <= m; ++i) hp[i] = vecs[i].ptr;
    vptr.allocate(m + 1);
    SPINED_HIP(hipMemcpy(vptr.ptr, hp.data(), sizeof(VT *) * (m + 1), hipMemcpyHostToDevice));
    w.allocate_pooled(ne);
    const double *wp = w.ptr;
    wptr.allocate(1);
    SPINED_HIP(hipMemcpy(wptr.ptr, &wp, sizeof(wp), hipMemcpyHostToDevice));
    G = grid_for(ne, 2048);
    const int chunks = (m + 1 + kDotChunk - 1) / kDotChunk;
    partial.allocate(static_cast<size_t>(chunks) * kDotChunk * G * (C ? 2 : 1));
    h.allocate(4 * kMaxBasis + 4);
    T.allocate(kMaxBasis * kMaxBasis);
    theta.allocate(kMaxBasis);
    S.allocate(kMaxBasis * kMaxBasis);
    beta.allocate(kMaxBasis + 1);
    skip.allocate(1);
    reset_projected_kernel<<<1, 256, 0, s>>>(T.ptr, kMaxBasis, theta.ptr, 0);
    SPINED_CHECK_LAUNCH();
    small = n <= (1u << 20);
  }
  ~Phase() {
    for (auto e : ev) (void)hipEventDestroy(e);
  }

  void matvec(const void *x, VecType xt, double *y) {
    hipEvent_t a, b;
    SPINED_HIP(hipEventCreate(&a));
    SPINED_HIP(hipEventCreate(&b));
    SPINED_HIP(hipEventRecord(a, s));
    MatvecCall c{&op, xt, xt == VecType::C128 ? VecType::C128 : VecType::F64, x, y, n, s};
    launch_matvec(c);
    SPINED_HIP(hipEventRecord(b, s));
    ev.push_back(a);
    ev.push_back(b);
    ++matvecs;
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

  // real dots over `len` elements (default: the whole vector, ne)
  template <class PT>
  void multidot(const PT *const *V, int nvec, const double *wv, double *hout, uint64_t len = 0) {
    const int chunks = (nvec + kDotChunk - 1) / kDotChunk;
    multidot_kernel<PT><<<dim3(G, chunks), kBlock, 0, s>>>(V, nvec, wv, len ? len : ne,
                                                           partial.ptr);
    SPINED_CHECK_LAUNCH();
    reduce_partials_kernel<<<nvec, kBlock, 0, s>>>(partial.ptr, nvec, G, hout, 0);
    SPINED_CHECK_LAUNCH();
  }

  void norm_w(double *out) {
    multidot<double>(wptr.ptr, 1, w.ptr, out);
    sqrt_inplace_kernel<<<1, 1, 0, s>>>(out);
    SPINED_CHECK_LAUNCH();
  }

  // complex conjugated dots h[2i], h[2i+1] = conj(V_i) . w
  void cdots(int nvec, double *hout) {
    const int chunks = (nvec + kDotChunk - 1) / kDotChunk;
    const auto *V2 = reinterpret_cast<const double2 *const *>(vptr.ptr);
    cmultidot_kernel<<<dim3(G, chunks), kBlock, 0, s>>>(V2, nvec,
                                                        reinterpret_cast<const double2 *>(w.ptr), n,
                                                        partial.ptr);
    SPINED_CHECK_LAUNCH();
    reduce_partials_kernel<<<2 * nvec, kBlock, 0, s>>>(partial.ptr, 2 * nvec, G, hout, 0);
    SPINED_CHECK_LAUNCH();
  }

  void csub(int nvec, const double *hc) {
    const auto *V2 = reinterpret_cast<const double2 *const *>(vptr.ptr);
    cgemv_sub_kernel<<<G, kBlock, 0, s>>>(V2, nvec, hc, reinterpret_cast<double2 *>(w.ptr), n);
    SPINED_CHECK_LAUNCH();
  }

  // orthogonalise w against V[0..nvec) twice; h1 + h2 are the coefficients
  void reorth(int nvec, double *h1, double *h2) {
    if constexpr (C) {
      cdots(nvec, h1);
      csub(nvec, h1);
      cdots(nvec, h2);
      csub(nvec, h2);
      return;
    }
    multidot<VT>(vptr.ptr, nvec, w.ptr, h1);
    gemv_sub_kernel<VT><<<G, kBlock, 0, s>>>(vptr.ptr, nvec, h1, w.ptr, n);
    SPINED_CHECK_LAUNCH();
    multidot<VT>(vptr.ptr, nvec, w.ptr, h2);
    gemv_sub_kernel<VT><<<G, kBlock, 0, s>>>(vptr.ptr, nvec, h2, w.ptr, n);
    SPINED_CHECK_LAUNCH();
  }

  // v_0 from an FP64 start vector in w (normalised here)
  void set_start_from_w() {
    norm_w(beta.ptr);
    scale_store_kernel<VT><<<G, kBlock, 0, s>>>(w.ptr, beta.ptr, vecs[0].ptr, ne);
    SPINED_CHECK_LAUNCH();
  }

  // Classical Gram-Schmidt against V[0..nvec) (real vectors, nvec <=
  // kFuseMax) and the norm of the result: dots with w.w; update + dots
  // (the second pass); update + norm.  The second pass is left out on the
  // device (no host round trip) when the first update loses little to
  // cancellation (second_pass_test_kernel); the norm is then w.w - |h|^2
  // and the update writes the next basis vector vnext = w' / |w'| itself
  // (the caller's scale_store_kernel, given skip, does nothing).  Complex
  // vectors take reorth() and norm_w().
  void reorth_norm_fused(int nvec, double *h1, double *h2, double *norm_out, VT *vnext) {
    // register width: the smallest multiple of 8 holding nvec vectors (the
    // slots past nvec still issue loads, of V[0])
    const int nq = (nvec + 7) / 8 * 8;
    auto passes = [&](auto width) {
      constexpr int NQ = decltype(width)::value;
      dot_ww_kernel<VT, NQ><<<G, kBlock, 0, s>>>(vptr.ptr, nvec, w.ptr, n, partial.ptr);
      SPINED_CHECK_LAUNCH();
      reduce_partials_kernel<<<nvec + 1, kBlock, 0, s>>>(partial.ptr, nvec + 1, G, h1, 0);
      SPINED_CHECK_LAUNCH();
      second_pass_test_kernel<<<1, 64, 0, s>>>(h1, nvec, kSecondPassEta, h2, skip.ptr, norm_out);
      SPINED_CHECK_LAUNCH();
      sub_dot_kernel<VT, NQ>
          <<<G, kBlock, 0, s>>>(vptr.ptr, nvec, h1, w.ptr, n, partial.ptr, skip.ptr);
      SPINED_CHECK_LAUNCH();
      reduce_partials_kernel<<<nvec, kBlock, 0, s>>>(partial.ptr, nvec, G, h2, 0, skip.ptr);
      SPINED_CHECK_LAUNCH();
      sub_norm_kernel<VT, NQ><<<G, kBlock, 0, s>>>(vptr.ptr, nvec, h2, w.ptr, n, partial.ptr,
                                                   h1, skip.ptr, vnext, norm_out);
      SPINED_CHECK_LAUNCH();
    };
    if (nq <= 8)
      passes(std::integral_constant<int, 8>{});
    else if (nq <= 16)
      passes(std::integral_constant<int, 16>{});
    else if (nq <= 24)
      passes(std::integral_constant<int, 24>{});
    else
      passes(std::integral_constant<int, kFuseMax>{});
    reduce_partials_kernel<<<1, kBlock, 0, s>>>(partial.ptr, 1, G, norm_out, 0, skip.ptr);
    SPINED_CHECK_LAUNCH();
    sqrt_inplace_kernel<<<1, 1, 0, s>>>(norm_out, skip.ptr);
    SPINED_CHECK_LAUNCH();
  }

  void step(int j, uint64_t seed_breakdown) {
    matvec(vecs[j].ptr,
           C ? VecType::C128 : (std::is_same<VT, float>::value ? VecType::F32 : VecType::F64),
           w.ptr);
    double *h1 = h.ptr, *h2 = h.ptr + 2 * kMaxBasis + 2;
    const bool fused = !C && j + 1 <= kFuseMax;
    if (fused) {
      if constexpr (!C) reorth_norm_fused(j + 1, h1, h2, beta.ptr + j, vecs[j + 1].ptr);
    } else {
      reorth(j + 1, h1, h2);
      norm_w(beta.ptr + j);
    }
    store_column2_kernel<<<1, kMaxBasis, 0, s>>>(h1, h2, j, T.ptr, kMaxBasis, C ? 2 : 1);
    SPINED_CHECK_LAUNCH();
    if (small) {
      // breakdown guard (invariant subspace): continue with a random vector
      double bj = 0.0, tj = 0.0;
      SPINED_HIP(hipMemcpyAsync(&bj, beta.ptr + j, 8, hipMemcpyDeviceToHost, s));
      SPINED_HIP(hipMemcpyAsync(&tj, T.ptr + j * kMaxBasis + j, 8, hipMemcpyDeviceToHost, s));
      SPINED_HIP(hipStreamSynchronize(s));
      if (bj <= 1e-12 * std::max(1.0, std::fabs(tj)) && j + 1 < static_cast<int>(n)) {
        launch_random_fill(w.ptr, ne, seed_breakdown + 7919u * (j + 1), s);
        double *g1 = h.ptr, *g2 = h.ptr + 2 * kMaxBasis + 2; // scratch (column already stored)
        reorth(j + 1, g1, g2);
        double nb = 0.0;
        norm_w(beta.ptr + kMaxBasis);
        SPINED_HIP(hipMemcpyAsync(&nb, beta.ptr + kMaxBasis, 8, hipMemcpyDeviceToHost, s));
        SPINED_HIP(hipStreamSynchronize(s));
        scale_store_kernel<VT><<<G, kBlock, 0, s>>>(w.ptr, beta.ptr + kMaxBasis, vecs[j + 1].ptr, ne);
        SPINED_CHECK_LAUNCH();
        const double zero = 0.0;
        SPINED_HIP(hipMemcpyAsync(beta.ptr + j, &zero, 8, hipMemcpyHostToDevice, s));
        SPINED_HIP(hipStreamSynchronize(s));
        return;
      }
    }
    scale_store_kernel<VT><<<G, kBlock, 0, s>>>(w.ptr, beta.ptr + j, vecs[j + 1].ptr, ne,
                                                fused ? skip.ptr : nullptr);
    SPINED_CHECK_LAUNCH();
  }

  // Runs cycles until the k lowest estimates are <= tol, the matvec budget is
  // spent, or (stall_detect) the estimates stop improving.
  PhaseResult run(double tol, int64_t max_matvecs, bool stall_detect, int phase_id,
                  spined_monitor_fn monitor, void *ctx, bool verbose, uint64_t seed) {
    PhaseResult res;
    int j0 = 0;
    double best = 1e300;
    int stalled = 0;
    std::vector<double> th(m), Sh(static_cast<size_t>(m) * m);
    // Convergence is also checked inside a cycle (on the leading block of T)
    // so that a good start vector is not held back to a restart boundary:
    // every step when a matvec is expensive, every 4th otherwise.
    const int check_every = n >= (1u << 20) ? 1 : 4;
    // solve the leading msz x msz projected problem; estimates of the k lowest
    auto project = [&](int msz, double &worst, int &nconv) {
      jacobi_eigh_kernel<<<1, 256, 0, s>>>(T.ptr, kMaxBasis, msz, theta.ptr, S.ptr, 40);
      SPINED_CHECK_LAUNCH();
      double bm = 0.0;
      SPINED_HIP(hipMemcpyAsync(th.data(), theta.ptr, 8 * msz, hipMemcpyDeviceToHost, s));
      SPINED_HIP(hipMemcpyAsync(Sh.data(), S.ptr, 8ull * msz * msz, hipMemcpyDeviceToHost, s));
      SPINED_HIP(hipMemcpyAsync(&bm, beta.ptr + (msz - 1), 8, hipMemcpyDeviceToHost, s));
      SPINED_HIP(hipStreamSynchronize(s));
      res.theta.assign(th.begin(), th.begin() + k);
      res.est.resize(k);
      worst = 0.0;
      nconv = 0;
      for (int i = 0; i < k; ++i) {
        res.est[i] = std::fabs(bm * Sh[static_cast<size_t>(i) * msz + (msz - 1)]);
        worst = std::max(worst, res.est[i]);
        if (res.est[i] <= tol) ++nconv;
      }
      mlast = msz;
    };
    for (;;) {
      double worst = 0.0;
      int nconv = 0;
      bool early = false;
      for (int j = j0; j < m; ++j) {
        step(j, seed);
        const int msz = j + 1;
        if (msz < m && msz > k && (msz - j0) % check_every == 0) {
          project(msz, worst, nconv);
          if (nconv == k) {
            early = true;
            break;
          }
        }
      }
      if (!early) project(m, worst, nconv);
      harvest_events();
      ++res.cycles;
      if (verbose)
        log_line("phase %d cycle %d matvecs %lld theta0 %.15g est %.3e", phase_id, res.cycles,
                 static_cast<long long>(matvecs), th[0], worst);
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
      if (stop || (max_matvecs > 0 && matvecs >= max_matvecs) || m >= static_cast<int>(n))
        return res;
      if (stall_detect) {
        if (worst < 0.5 * best) {
          best = worst;
          stalled = 0;
        } else if (++stalled >= 2) {
          return res;
        }
      }
      // thick restart: keep the lowest `keep` Ritz vectors and v_m
      launch_restart<VT>(vptr.ptr, m, keep, S.ptr, ne, G, s);
      SPINED_HIP(hipMemcpyAsync(vecs[keep].ptr, vecs[m].ptr, ne * sizeof(VT),
                                hipMemcpyDeviceToDevice, s));
      reset_projected_kernel<<<1, 256, 0, s>>>(T.ptr, kMaxBasis, theta.ptr, keep);
      SPINED_CHECK_LAUNCH();
      j0 = keep;
      // V_0 now holds the lowest Ritz vector
      if (cp_fn && cp_every > 0 && res.cycles % cp_every == 0) {
        SPINED_HIP(hipStreamSynchronize(s));
        spined_monitor_info info{matvecs, res.cycles, phase_id, nconv, k, res.theta.data(),
                                 res.est.data()};
        cp_fn(vecs[0].ptr,
              C ? LS_COMPLEX128 : (std::is_same<VT, float>::value ? LS_FLOAT32 : LS_FLOAT64), n,
              &info, cp_ctx);
      }
    }
  }

  // FP64 Ritz vector i of the last projected problem into out
  void ritz(int i, double *out) {
    ritz_kernel<VT><<<G, kBlock, 0, s>>>(vptr.ptr, mlast, S.ptr + static_cast<size_t>(i) * mlast,
                                         out, ne);
    SPINED_CHECK_LAUNCH();
  }
};

} // namespace
} // namespace spined

#include "blocklanczos.hpp"

namespace spined {
namespace {

void pick_sizes(uint64_t n, uint64_t k, const spined_eigsh_options &o, int &m, int &keep) {
  int mm = o.max_basis_size > 0 ? o.max_basis_size : 0;
  if (mm <= 0) mm = static_cast<int>(std::max<uint64_t>(2 * k + 10, 24));
  mm = std::max<int>(mm, static_cast<int>(k) + 2);
  mm = std::min(mm, kMaxBasis);
  if (static_cast<uint64_t>(mm) > n) mm = static_cast<int>(n);
  m = mm;
  int kp = o.min_restart_size > 0 ? o.min_restart_size : (m + static_cast<int>(k)) / 2;
  kp = std::max(kp, static_cast<int>(k));
  kp = std::min(kp, m - 1);
  kp = std::min(kp, kMaxKeepOne);
  keep = std::max(kp, 0);
}

// Block width from max_block_size (PRIMME's maxBlockSize): 1 (one-vector
// recurrence), 2, 4 or 8.
int block_width(int32_t req) { return req >= 8 ? 8 : req >= 4 ? 4 : req >= 2 ? 2 : 1; }

// Blocks per cycle and kept blocks: M = mb*K <= 64 vectors, about half kept
// (at least ceil(k/K) blocks, at most 32 vectors).
void pick_block_sizes(uint64_t k, int K, const spined_eigsh_options &o, int &mb, int &keepb) {
  int M = o.max_basis_size > 0 ? o.max_basis_size : 4 * K;
  M = std::min(std::max(M, 3 * K), kMaxBasis);
  mb = std::max(3, M / K);
  while (mb * K > kMaxBasis) --mb;
  const int need = static_cast<int>((k + K - 1) / K);
  keepb = o.min_restart_size > 0 ? (o.min_restart_size + K - 1) / K : mb / 2;
  keepb = std::max(keepb, need);
  keepb = std::min(keepb, kMaxKeep / K);
  keepb = std::min(keepb, mb - 1);
}

struct BlockOutcome {
  bool ran = false;       // false: rank deficient, start left unchanged
  int64_t matvecs = 0;
  double matvec_seconds = 0.0;
  int cycles = 0;
  bool converged = false;
};

// Block phase with storage VT; on success start <- sum of the k lowest Ritz
// vectors (FP64).
template <class VT, int K>
BlockOutcome run_block_phase(Operator &op, uint64_t n, uint64_t k, double tol, bool stall_detect,
                             int phase_id, const spined_eigsh_options &opts, double *start,
                             double *y, spined_monitor_fn monitor, void *ctx, hipStream_t s) {
  BlockOutcome out;
  int mb = 0, keepb = 0;
  pick_block_sizes(k, K, opts, mb, keepb);
  BlockPhase<VT, K> p(op, n, mb, static_cast<int>(k), keepb, s);
  bool deficient = !p.set_start(start, opts.seed ^ 0xB10Cull);
  PhaseResult r;
  if (!deficient)
    r = p.run(tol, opts.max_matvecs, stall_detect, phase_id, monitor, ctx, opts.verbose,
              deficient);
  out.matvecs = p.matvecs;
  out.matvec_seconds = p.matvec_seconds;
  out.cycles = r.cycles;
  if (deficient) {
    if (opts.verbose) log_line("block phase: rank-deficient block, one-vector recurrence instead");
    return out;
  }
  SPINED_HIP(hipMemsetAsync(start, 0, n * 8, s));
  for (uint64_t i = 0; i < k; ++i) {
    p.ritz(static_cast<int>(i), y);
    residual_kernel<<<grid_for(n), kBlock, 0, s>>>(start, y, -1.0, start, n);
    SPINED_CHECK_LAUNCH();
  }
  SPINED_HIP(hipStreamSynchronize(s));
  out.ran = true;
  out.converged = r.converged;
  return out;
}

template <class VT>
BlockOutcome run_block(int K, Operator &op, uint64_t n, uint64_t k, double tol, bool stall_detect,
                       int phase_id, const spined_eigsh_options &opts, double *start, double *y,
                       spined_monitor_fn monitor, void *ctx, hipStream_t s) {
  switch (K) {
  case 2:
    return run_block_phase<VT, 2>(op, n, k, tol, stall_detect, phase_id, opts, start, y, monitor,
                                  ctx, s);
  case 4:
    return run_block_phase<VT, 4>(op, n, k, tol, stall_detect, phase_id, opts, start, y, monitor,
                                  ctx, s);
  default:
    return run_block_phase<VT, 8>(op, n, k, tol, stall_detect, phase_id, opts, start, y, monitor,
                                  ctx, s);
  }
}

} // namespace

void launch_random_fill(double *v, uint64_t n, uint64_t seed, hipStream_t s) {
  random_fill_kernel<<<grid_for(n), kBlock, 0, s>>>(v, n, seed);
  SPINED_CHECK_LAUNCH();
}

namespace {
// Plain streaming copy (16-byte elements): the calibration kernel for the
// HBM byte counters (a copy of B bytes should read B and write B).
__global__ void __launch_bounds__(kBlock)
    copy_calibration_kernel(const uint4 *__restrict__ src, uint4 *__restrict__ dst, uint64_t n) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * kBlock)
    dst[i] = src[i];
}
} // namespace

namespace {
// Random-gather calibration: each group of `words` consecutive lanes reads
// words 0..words-1 (4 bytes each) of one 128-byte line of src, the lines
// visited once each in the order i -> (i * odd) mod n_lines (n_lines a power
// of two, so every line is read exactly once and neighbours in time are
// far apart in memory).  Each wave writes one word.  With words = 1 the
// counters show what one isolated 4-byte gather costs; with words = 32 the
// whole line is read, 128 bytes per line.
__global__ void __launch_bounds__(kBlock)
    gather_calibration_kernel(const uint32_t *__restrict__ src, uint32_t *__restrict__ dst,
                              uint64_t n_lines, int words) {
  const uint64_t lanes = n_lines * static_cast<uint64_t>(words);
  uint32_t acc = 0;
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x; i < lanes;
       i += static_cast<uint64_t>(gridDim.x) * kBlock) {
    const uint64_t line = ((i / words) * 0x9E3779B97F4A7C15ull) & (n_lines - 1);
    acc += src[line * 32 + i % words];
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if ((threadIdx.x & 63) == 0)
    dst[(blockIdx.x * static_cast<uint64_t>(kBlock) + threadIdx.x) / 64] = acc;
}
} // namespace

// dst must hold one word per wave of the grid: grid_for(n_lines * words,
// 1 << 16) * 4 words at most.
uint64_t gather_calibration_out_words() { return (uint64_t(1) << 16) * (kBlock / 64); }

void launch_gather_calibration(const void *src, void *dst, uint64_t n_lines, int words,
                               hipStream_t s) {
  gather_calibration_kernel<<<grid_for(n_lines * words, 1u << 16), kBlock, 0, s>>>(
      static_cast<const uint32_t *>(src), static_cast<uint32_t *>(dst), n_lines, words);
  SPINED_CHECK_LAUNCH();
}

void launch_copy_calibration(const void *src, void *dst, uint64_t bytes, hipStream_t s) {
  const uint64_t n = bytes / 16;
  copy_calibration_kernel<<<grid_for(n, 1u << 16), kBlock, 0, s>>>(
      static_cast<const uint4 *>(src), static_cast<uint4 *>(dst), n);
  SPINED_CHECK_LAUNCH();
}

// The FP64 phase that finishes every solve (real: C = false; complex
// Hermitian: C = true, vectors of n double2), with the explicit residual
// check of the Ritz vectors and the eigenvector copy-out.  start: the FP64
// start vector (ne doubles); y, hy: scratch vectors of ne doubles.
template <bool C>
bool final_phase(Operator &op, uint64_t n, int m, int keep, uint64_t k, double tol, int dtype,
                 const spined_eigsh_options &opts, DeviceBuffer<double> &start,
                 DeviceBuffer<double> &y, DeviceBuffer<double> &hy, void *evecs, uint64_t ld,
                 spined_monitor_fn monitor, void *ctx, spined_stats &st, std::vector<double> &theta,
                 std::vector<double> &est, hipStream_t s) {
  bool converged = false;
  const uint64_t ne = C ? 2 * n : n;
  const VecType vt = C ? VecType::C128 : VecType::F64;
  Phase<double, C> p(op, n, m, static_cast<int>(k), keep, s);
  p.set_checkpoint(opts);
  SPINED_HIP(hipMemcpyAsync(p.w.ptr, start.ptr, ne * 8, hipMemcpyDeviceToDevice, s));
  p.set_start_from_w();
  start.release();
  int64_t budget = opts.max_matvecs > 0 ? std::max<int64_t>(1, opts.max_matvecs - st.matvecs) : 0;
  // the caller's tolerance judges every explicit residual; only the target
  // handed to the recurrence tightens when an estimate proves optimistic
  const double tol_accept = tol;
  double tol_run = tol;
  for (int attempt = 0; attempt < 8; ++attempt) {
    PhaseResult r = p.run(tol_run, budget, false, 0, monitor, ctx, opts.verbose, opts.seed + 1);
    st.restarts += r.cycles;
    // explicit residuals of the Ritz vectors (FP64 matvec); their sum goes
    // to w, the start of a further attempt
    bool ok = true;
    SPINED_HIP(hipMemsetAsync(p.w.ptr, 0, ne * 8, s));
    for (uint64_t i = 0; i < k; ++i) {
      p.ritz(static_cast<int>(i), y.ptr);
      residual_kernel<<<grid_for(ne), kBlock, 0, s>>>(p.w.ptr, y.ptr, -1.0, p.w.ptr, ne);
      SPINED_CHECK_LAUNCH();
      p.matvec(y.ptr, vt, hy.ptr);
      // Rayleigh quotient and residual norm
      const double *yp = y.ptr;
      DeviceBuffer<const double *> yptr(1);
      SPINED_HIP(hipMemcpyAsync(yptr.ptr, &yp, sizeof(yp), hipMemcpyHostToDevice, s));
      DeviceBuffer<double> sc(4);
      p.template multidot<double>(yptr.ptr, 1, hy.ptr, sc.ptr);     // y.Hy
      p.template multidot<double>(yptr.ptr, 1, y.ptr, sc.ptr + 1);  // y.y
      double hsc[2];
      SPINED_HIP(hipMemcpyAsync(hsc, sc.ptr, 16, hipMemcpyDeviceToHost, s));
      SPINED_HIP(hipStreamSynchronize(s));
      const double lam = hsc[0] / hsc[1];
      const double ynorm = std::sqrt(hsc[1]);
      residual_kernel<<<grid_for(ne), kBlock, 0, s>>>(hy.ptr, y.ptr, lam, hy.ptr, ne);
      SPINED_CHECK_LAUNCH();
      const double *hp = hy.ptr;
      SPINED_HIP(hipMemcpyAsync(yptr.ptr, &hp, sizeof(hp), hipMemcpyHostToDevice, s));
      p.template multidot<double>(yptr.ptr, 1, hy.ptr, sc.ptr + 2);
      double rr = 0.0;
      SPINED_HIP(hipMemcpyAsync(&rr, sc.ptr + 2, 8, hipMemcpyDeviceToHost, s));
      SPINED_HIP(hipStreamSynchronize(s));
      theta[i] = lam;
      est[i] = std::sqrt(std::max(rr, 0.0)) / ynorm;
      if (est[i] > tol_accept) ok = false;
      if (evecs) {
        // normalise (and convert) on the device, then one copy into the
        // caller's column: y is recomputed for the next i, and hy (its
        // residual norm already read) holds the FP32 values
        const double inv = 1.0 / ynorm;
        // complex: ld counts complex elements, each two reals
        const uint64_t per = C ? 2 : 1;
        if (dtype == LS_FLOAT64 || dtype == LS_COMPLEX128) {
          scale_convert_kernel<double><<<grid_for(ne), kBlock, 0, s>>>(y.ptr, inv, y.ptr, ne);
          SPINED_CHECK_LAUNCH();
          SPINED_HIP(hipMemcpyAsync(static_cast<double *>(evecs) + i * ld * per, y.ptr, ne * 8,
                                    hipMemcpyDeviceToHost, s));
        } else {
          float *tmp = reinterpret_cast<float *>(hy.ptr);
          scale_convert_kernel<float><<<grid_for(ne), kBlock, 0, s>>>(y.ptr, inv, tmp, ne);
          SPINED_CHECK_LAUNCH();
          SPINED_HIP(hipMemcpyAsync(static_cast<float *>(evecs) + i * ld * per, tmp, ne * 4,
                                    hipMemcpyDeviceToHost, s));
        }
        SPINED_HIP(hipStreamSynchronize(s));
      }
    }
    p.harvest_events();
    // a stop requested by the monitor ends the solve here, as does the budget
    if (ok || r.stopped || (budget > 0 && p.matvecs >= budget)) {
      converged = ok;
      break;
    }
    // estimates said converged but the explicit residual disagrees (the
    // basis has drifted from the recurrence): a fresh recurrence from the
    // current Ritz vectors, with a tighter target for its estimates
    p.set_start_from_w();
    tol_run *= 0.5;
  }
  st.matvecs += p.matvecs;
  st.seconds_matvec += p.matvec_seconds;
  return converged;
}

// Real operators on at most kMaxBasis states: H column by column from unit
// vectors (D matvecs), then the one-workgroup Jacobi eigensolver on the
// symmetrised dense matrix.  Exact, including degenerate levels (which a
// single-vector recurrence returns once per eigenspace).
void dense_small_eigsh(Operator &op, uint64_t n, uint64_t k, int dtype, double *evals,
                       void *evecs, uint64_t ld, double *rnorms, spined_stats &st,
                       hipStream_t s) {
  std::vector<double> H(n * n), e(n, 0.0);
  DeviceBuffer<double> dx(n), dy(n);
  for (uint64_t j = 0; j < n; ++j) {
    std::fill(e.begin(), e.end(), 0.0);
    e[j] = 1.0;
    SPINED_HIP(hipMemcpyAsync(dx.ptr, e.data(), 8 * n, hipMemcpyHostToDevice, s));
    launch_matvec(MatvecCall{&op, VecType::F64, VecType::F64, dx.ptr, dy.ptr, n, s});
    SPINED_HIP(hipMemcpyAsync(H.data() + j * n, dy.ptr, 8 * n, hipMemcpyDeviceToHost, s));
    SPINED_HIP(hipStreamSynchronize(s));
  }
  st.matvecs += static_cast<int64_t>(n);
  // T = (H + H^T) / 2 in the projected-matrix layout (ld = kMaxBasis)
  std::vector<double> T(kMaxBasis * kMaxBasis, 0.0);
  for (uint64_t a = 0; a < n; ++a)
    for (uint64_t b = 0; b < n; ++b) T[a * kMaxBasis + b] = 0.5 * (H[a * n + b] + H[b * n + a]);
  DeviceBuffer<double> dT(kMaxBasis * kMaxBasis), dth(kMaxBasis), dS(kMaxBasis * kMaxBasis);
  SPINED_HIP(hipMemcpyAsync(dT.ptr, T.data(), 8 * T.size(), hipMemcpyHostToDevice, s));
  jacobi_eigh_kernel<<<1, 256, 0, s>>>(dT.ptr, kMaxBasis, static_cast<int>(n), dth.ptr, dS.ptr, 60);
  SPINED_CHECK_LAUNCH();
  std::vector<double> th(n), S(n * n);
  SPINED_HIP(hipMemcpyAsync(th.data(), dth.ptr, 8 * n, hipMemcpyDeviceToHost, s));
  SPINED_HIP(hipMemcpyAsync(S.data(), dS.ptr, 8 * n * n, hipMemcpyDeviceToHost, s));
  SPINED_HIP(hipStreamSynchronize(s));
  for (uint64_t i = 0; i < k; ++i) {
    const double *v = S.data() + i * n; // eigenvector i (column i of S)
    double r2 = 0.0;
    for (uint64_t a = 0; a < n; ++a) {
      double hv = 0.0;
      for (uint64_t b = 0; b < n; ++b) hv += H[b * n + a] * v[b];
      r2 += (hv - th[i] * v[a]) * (hv - th[i] * v[a]);
    }
    evals[i] = th[i];
    if (rnorms) rnorms[i] = std::sqrt(r2);
    if (evecs) {
      if (dtype == LS_FLOAT64) {
        double *o = static_cast<double *>(evecs) + i * ld;
        for (uint64_t a = 0; a < n; ++a) o[a] = v[a];
      } else {
        float *o = static_cast<float *>(evecs) + i * ld;
        for (uint64_t a = 0; a < n; ++a) o[a] = static_cast<float>(v[a]);
      }
    }
  }
}

int eigsh(Operator &op, int dtype, uint64_t k, double tol, const spined_eigsh_options &opts_in,
          double *evals, void *evecs, uint64_t ld, double *rnorms, spined_monitor_fn monitor,
          void *ctx) {
  const auto t0 = std::chrono::steady_clock::now();
  // max_matvecs 0 = automatic: a cap far beyond any solve that converges,
  // so that one which stops converging returns LS_NOT_CONVERGED
  spined_eigsh_options opts = opts_in;
  if (opts.max_matvecs <= 0) opts.max_matvecs = kAutoMaxMatvecs;
  Basis &b = *op.basis;
  const uint64_t n = b.D;
  if (k == 0 || k > n) throw Error(LS_INVALID_ARGUMENT, "number of eigenpairs must be in [1, dim]");
  if (k > 16) throw Error(LS_NOT_IMPLEMENTED, "at most 16 eigenpairs per solve");
  const bool cplx = !op.tables.is_real;
  if (!cplx && dtype != LS_FLOAT64 && dtype != LS_FLOAT32)
    throw Error(LS_INVALID_DATATYPE, "a real operator takes float32/float64 eigenvectors");
  if (cplx && dtype != LS_COMPLEX128 && dtype != LS_COMPLEX64)
    throw Error(LS_INVALID_DATATYPE, "a complex operator takes complex64/complex128 eigenvectors");
  const uint64_t ne = cplx ? 2 * n : n;
  if (evecs && ld < n) throw Error(LS_INVALID_ARGUMENT, "leading dimension smaller than dim");
  if (!(tol > 0.0)) tol = 1e-10; // PRIMME's eps = 0 means "its default"; ours is 1e-10
  hipStream_t s = stream();
  int m = 0, keep = 0;
  pick_sizes(n, k, opts, m, keep);
  spined_stats st{};
  if (!cplx && n <= static_cast<uint64_t>(kMaxBasis)) {
    dense_small_eigsh(op, n, k, dtype, evals, evecs, ld, rnorms, st, s);
    st.seconds_total = seconds_since(t0);
    std::lock_guard<std::mutex> lk(op.mu);
    op.last_stats = st;
    return LS_SUCCESS;
  }

  // Gather cache (real one-vector solves): a cached matvec costs about a
  // third of the orbit-and-index one, so with an automatic basis size the
  // cache is kept at the price of a smaller Krylov basis, down to
  // kCacheMinBasis vectors; otherwise it gives its memory to the vectors.
  if (!cplx && block_width(opts.max_block_size) == 1) {
    const uint64_t blk = n * 8;
    const int extra = 5; // w, start, y, hy and one block of slack
    const int mmin = opts.max_basis_size > 0 ? m : std::min(m, kCacheMinBasis);
    if (prepare_gather_cache(op, static_cast<uint64_t>(mmin + 1 + extra) * blk, blk)) {
      size_t free_b = 0, total_b = 0;
      SPINED_HIP(hipMemGetInfo(&free_b, &total_b));
      const int fit = static_cast<int>((free_b + pool_bytes(blk)) / blk) - 1 - extra;
      if (fit < mmin) {
        drop_gather_cache_if_built(op);
      } else if (fit < m) {
        spined_eigsh_options o2 = opts;
        o2.max_basis_size = fit;
        pick_sizes(n, k, o2, m, keep);
      }
    }
  }

  // ---- start vector (FP64, in a scratch buffer) --------------------------
  DeviceBuffer<double> start;
  start.allocate_pooled(ne);

  const auto ts = std::chrono::steady_clock::now();
  std::vector<double> host_start;
  if (opts.start == SPINED_START_VECTOR) {
    if (!opts.start_vector) throw Error(LS_INVALID_ARGUMENT, "start_vector is NULL");
    if (cplx) {
      // real start vector as the real parts of the complex one
      host_start.assign(ne, 0.0);
      for (uint64_t q = 0; q < n; ++q) host_start[2 * q] = opts.start_vector[q];
      SPINED_HIP(hipMemcpyAsync(start.ptr, host_start.data(), ne * 8, hipMemcpyHostToDevice, s));
    } else {
      SPINED_HIP(hipMemcpyAsync(start.ptr, opts.start_vector, n * 8, hipMemcpyHostToDevice, s));
    }
  } else if (opts.start == SPINED_START_LEARNED) {
    if (!opts.model) throw Error(LS_INVALID_ARGUMENT, "learned start needs a model");
    if (cplx) throw Error(LS_INVALID_ARGUMENT, "the learned start covers real sectors only");
    cnn_start_vector(static_cast<const spined_cnn *>(opts.model), b, opts.start_sign_mask,
                     start.ptr);
  } else {
    launch_random_fill(start.ptr, ne, opts.seed, s);
  }
  SPINED_HIP(hipStreamSynchronize(s));
  st.seconds_start = seconds_since(ts);
  if (opts.verbose) log_line("start vector %.3f s (set-up before it %.3f s)", st.seconds_start,
                             std::chrono::duration<double>(ts - t0).count());

  std::vector<double> theta(k), est(k);
  DeviceBuffer<double> y, hy;
  y.allocate_pooled(ne);
  hy.allocate_pooled(ne);

  // Block phase (max_block_size >= 2): FP32 storage in mixed mode, FP64
  // otherwise; the one-vector FP64 phase below then finishes from the sum of
  // its Ritz vectors (and owns the explicit residual check).
  const int K = block_width(opts.max_block_size);
  bool block_done = false;
  if (!cplx && K > 1 && matvec_block_supported(op) && n >= (1u << 14)) {
    const bool mixed = opts.precision == SPINED_PRECISION_MIXED;
    const BlockOutcome bo =
        mixed ? run_block<float>(K, op, n, k, std::max(tol, 1e-5), true, 1, opts, start.ptr,
                                 y.ptr, monitor, ctx, s)
              : run_block<double>(K, op, n, k, tol, true, 0, opts, start.ptr, y.ptr, monitor,
                                  ctx, s);
    st.matvecs += bo.matvecs;
    if (mixed) {
      st.matvecs_mixed += bo.matvecs;
      st.seconds_matvec_mixed += bo.matvec_seconds;
    }
    st.seconds_matvec += bo.matvec_seconds;
    st.restarts += bo.cycles;
    block_done = bo.ran;
  }

  // complex sectors: FP64 throughout (the mixed phase stores real vectors)
  if (!cplx && !block_done && opts.precision == SPINED_PRECISION_MIXED &&
      n > static_cast<uint64_t>(m)) {
    // FP32-storage phase down to its floor, then hand the FP64 Ritz vector
    // (sum of the k lowest) to the FP64 phase.
    const auto ta = std::chrono::steady_clock::now();
    Phase<float> p(op, n, m, static_cast<int>(k), keep, s);
    p.set_checkpoint(opts);
    if (opts.verbose)
      log_line("FP32 phase: %d vectors allocated in %.3f s (since the start vector %.3f s)", m + 1,
               seconds_since(ta), seconds_since(ts) - st.seconds_start);
    SPINED_HIP(hipMemcpyAsync(p.w.ptr, start.ptr, n * 8, hipMemcpyDeviceToDevice, s));
    p.set_start_from_w();
    const double tol32 = std::max(tol, 1e-5);
    PhaseResult r = p.run(tol32, opts.max_matvecs, true, 1, monitor, ctx, opts.verbose, opts.seed);
    double scale = 0.0;
    for (double v : r.theta) scale = std::max(scale, std::fabs(v));
    (void)scale;
    // start = sum_i Ritz_i (FP64 assembly)
    SPINED_HIP(hipMemsetAsync(start.ptr, 0, n * 8, s));
    for (uint64_t i = 0; i < k; ++i) {
      p.ritz(static_cast<int>(i), y.ptr);
      residual_kernel<<<grid_for(n), kBlock, 0, s>>>(start.ptr, y.ptr, -1.0, start.ptr, n);
      SPINED_CHECK_LAUNCH();
    }
    SPINED_HIP(hipStreamSynchronize(s));
    st.matvecs += p.matvecs;
    st.matvecs_mixed += p.matvecs;
    st.seconds_matvec += p.matvec_seconds;
    st.seconds_matvec_mixed += p.matvec_seconds;
    st.restarts += r.cycles;
  }

  const bool converged =
      cplx ? final_phase<true>(op, n, m, keep, k, tol, dtype, opts, start, y, hy, evecs, ld,
                               monitor, ctx, st, theta, est, s)
           : final_phase<false>(op, n, m, keep, k, tol, dtype, opts, start, y, hy, evecs, ld,
                                monitor, ctx, st, theta, est, s);
  for (uint64_t i = 0; i < k; ++i) {
    evals[i] = theta[i];
    if (rnorms) rnorms[i] = est[i];
  }
  st.seconds_total = seconds_since(t0);
  {
    std::lock_guard<std::mutex> lk(op.mu);
    op.last_stats = st;
  }
  return converged ? LS_SUCCESS : LS_NOT_CONVERGED;
}

} // namespace spined
