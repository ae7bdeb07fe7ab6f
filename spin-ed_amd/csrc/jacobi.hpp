// jacobi.hpp -- one-workgroup parallel cyclic Jacobi eigensolver for the
// small symmetric projected matrices of the Lanczos solvers (FP64, in LDS).
// Call from a kernel launched with 256 threads per block.
#pragma once

#include <hip/hip_runtime.h>

namespace spined {

constexpr int kJacobiMax = 64;

// Parallel cyclic Jacobi (round-robin pairing) for the symmetric m x m
// projected matrix, m <= MAXM (even, <= 64), in LDS.  evals ascending; S
// column-major.  MAXM sizes the LDS workspace (2 MAXM (MAXM + 1) doubles).
template <int MAXM = kJacobiMax>
__device__ inline void jacobi_eigh_device(const double *__restrict__ T, int ldt, int m,
                                          double *__restrict__ evals, double *__restrict__ S,
                                          int max_sweeps) {
  static_assert(MAXM % 2 == 0 && MAXM <= kJacobiMax, "even size up to kJacobiMax");
  __shared__ double A[MAXM][MAXM + 1];
  __shared__ double V[MAXM][MAXM + 1];
  __shared__ double cs[MAXM / 2], sn[MAXM / 2];
  __shared__ int pp[MAXM / 2], qq[MAXM / 2];
  __shared__ double red[2][4]; // per-wave sums of (off, dia)
  __shared__ int done;
  const int tid = threadIdx.x;
  const int mp = m + (m & 1);
  for (int e = tid; e < mp * mp; e += blockDim.x) {
    const int i = e / mp, j = e % mp;
    double a = 0.0;
    if (i < m && j < m) a = 0.5 * (T[i * ldt + j] + T[j * ldt + i]);
    A[i][j] = a;
    V[i][j] = i == j ? 1.0 : 0.0;
  }
  __syncthreads();
  for (int sweep = 0; sweep < max_sweeps; ++sweep) {
    // convergence: off-diagonal mass vs diagonal mass
    double off = 0.0, dia = 0.0;
    for (int e = tid; e < mp * mp; e += blockDim.x) {
      const int i = e / mp, j = e % mp;
      const double a = A[i][j];
      if (i == j)
        dia += a * a;
      else
        off += a * a;
    }
    // wave sums, then the four waves in order (256 threads)
    for (int off2 = 32; off2 > 0; off2 >>= 1) {
      off += __shfl_down(off, off2, 64);
      dia += __shfl_down(dia, off2, 64);
    }
    if ((tid & 63) == 0) {
      red[0][tid >> 6] = off;
      red[1][tid >> 6] = dia;
    }
    __syncthreads();
    const double off_tot = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
    const double dia_tot = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
    if (tid == 0) done = off_tot <= 1e-30 * (dia_tot + off_tot) || off_tot == 0.0;
    __syncthreads();
    if (done) break;
    for (int step = 0; step < mp - 1; ++step) {
      const int half = mp / 2;
      if (tid < half) {
        // round robin: position 0 holds player 0, others rotate
        auto player = [&](int pos) { return pos == 0 ? 0 : 1 + ((pos - 1 + step) % (mp - 1)); };
        int p = player(tid), q = player(mp - 1 - tid);
        if (p > q) {
          const int t = p;
          p = q;
          q = t;
        }
        pp[tid] = p;
        qq[tid] = q;
        const double apq = A[p][q];
        double c = 1.0, s = 0.0;
        if (apq != 0.0) {
          const double app = A[p][p], aqq = A[q][q];
          const double tau = (aqq - app) / (2.0 * apq);
          const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
          c = 1.0 / sqrt(1.0 + t * t);
          s = t * c;
        }
        cs[tid] = c;
        sn[tid] = s;
      }
      __syncthreads();
      // A <- A J (columns p, q) and V <- V J
      for (int e = tid; e < half * mp; e += blockDim.x) {
        const int k = e / mp, i = e % mp;
        const int p = pp[k], q = qq[k];
        const double c = cs[k], s = sn[k];
        const double aip = A[i][p], aiq = A[i][q];
        A[i][p] = c * aip - s * aiq;
        A[i][q] = s * aip + c * aiq;
        const double vip = V[i][p], viq = V[i][q];
        V[i][p] = c * vip - s * viq;
        V[i][q] = s * vip + c * viq;
      }
      __syncthreads();
      // A <- J^T A (rows p, q)
      for (int e = tid; e < half * mp; e += blockDim.x) {
        const int k = e / mp, j = e % mp;
        const int p = pp[k], q = qq[k];
        const double c = cs[k], s = sn[k];
        const double apj = A[p][j], aqj = A[q][j];
        A[p][j] = c * apj - s * aqj;
        A[q][j] = s * apj + c * aqj;
      }
      __syncthreads();
    }
  }
  __syncthreads();
  __shared__ int order[MAXM];
  if (tid == 0) {
    // selection sort of the m real eigenpairs (the padding index is dropped)
    for (int i = 0; i < m; ++i) order[i] = i;
    for (int i = 0; i < m; ++i) {
      int best = i;
      for (int j = i + 1; j < m; ++j)
        if (A[order[j]][order[j]] < A[order[best]][order[best]]) best = j;
      const int t = order[i];
      order[i] = order[best];
      order[best] = t;
    }
    for (int i = 0; i < m; ++i) {
      const int c = order[i];
      evals[i] = A[c][c];
      for (int r = 0; r < m; ++r) S[i * m + r] = V[r][c];
    }
  }
}

} // namespace spined
