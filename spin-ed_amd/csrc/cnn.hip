// cnn.hip -- the learned start vector: a periodic 1D CNN over +-1 spins with
// hand-written BF16 MFMA forward and backward kernels (FP32 accumulate),
// Adam training on the GPU and evaluation over all representatives.
//
// Network (2,722 parameters; the oracle's cnn_oracle.c is its FP64 twin):
//   h0[i] = +1 if bit i is set else -1
//   z_l[c][i] = b_l[c] + sum_{c', d<5} W_l[c][c'][d] h_{l-1}[c'][(i+d-2) mod N]
//   h_l = GELU_tanh(z_l), l = 1..3, 16 channels;  f[c] = sum_i h3[c][i]
//   a = u0 + u.f  (log|psi|),  t = v0 + v.f  (sign logit)
//
// MFMA mapping (v_mfma_f32_16x16x32_bf16, one wave = 16 samples of equal N):
//   D[co][s] += A[co][k] B[k][s], k = 16 d + ci (80 used of 96 = 3 steps)
//   A = weights (registers), B = activations from LDS laid out [s][pos][c]
//   (one 16-byte load per lane per step), D lands as 4 channels of one
//   sample per lane (one 8-byte LDS store).
// Fragment layout (checked on the GPU by spined_debug_mfma_probe):
//   A: lane l holds row l%16, k = 8(l/16) .. +7;  B: column l%16, same k;
//   D: column l%16, rows 4(l/16) .. +3.
// Backward: the data gradient is the same conv with W'[ci][co][d] =
// W[co][ci][4-d]; weight gradients are MFMAs with K running over 32
// (sample, position) slots.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "handles.hpp"

namespace spined {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int CH = 16;
constexpr int KW = 5;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + CH * KW, OFF_W2 = OFF_B1 + CH,
              OFF_B2 = OFF_W2 + CH * CH * KW, OFF_W3 = OFF_B2 + CH,
              OFF_B3 = OFF_W3 + CH * CH * KW, OFF_U = OFF_B3 + CH, OFF_U0 = OFF_U + CH,
              OFF_V = OFF_U0 + 1, OFF_V0 = OFF_V + CH, NPARAMS = OFF_V0 + 1;
constexpr int kMaxN = 64;
constexpr int kInPlaceMinN = 8; // layer 2 in place for n >= 8 (see cnn_forward_kernel)

constexpr float GELU_A = 0.7978845608028654f;
constexpr float GELU_B = 0.044715f;

// GELU in its tanh form, 0.5 x (1 + tanh(u)) with u = A (x + B x^3), is
// x s(2u) with the logistic s(y) = 1 / (1 + exp(-y)): one native base-2
// exponential and one native reciprocal (no division sequence).  s
// saturates to 0 / 1 through the exponential's overflow / underflow; the
// error is a few FP32 ulp (the network's operands are BF16).  Training and
// the layer-1 table use this form.
constexpr float LOG2E = 1.4426950408889634f;
__device__ inline float logistic2(float u) {
  return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-2.0f * LOG2E * u));
}

__device__ inline float gelu_f(float x) {
  return x * logistic2(GELU_A * (x + GELU_B * x * x * x));
}
// d/dx [x s(2u)] = s + 2 x s (1 - s) A (1 + 3 B x^2)
__device__ inline float gelu_grad_f(float x) {
  const float sg = logistic2(GELU_A * (x + GELU_B * x * x * x));
  return sg + 2.0f * x * sg * (1.0f - sg) * GELU_A * (1.0f + 3.0f * GELU_B * x * x);
}

// The same tanh-form GELU without a transcendental, for the inference
// kernel: x G(x) with the gate G(x) = clamp(1/2 + x P(x^2), 0, 1), where P
// has degree 4 and a positive leading coefficient, fitted (linear program,
// minimax) to (GELU(x) - x/2) / x^2 on [0, 3.4] under the constraint
// x P(x^2) >= 1/2 on [3.4, 40]: beyond 3.4 the clamp gives x (and 0 for
// x <= -3.4), so y = x^2 needs no clamp of its own.  Largest error over all
// x, evaluated in FP32: 1.06e-3 absolute, below the BF16 resolution of the
// layer-2 activations it produces (2^-8 relative) wherever they exceed
// 0.27; the network's outputs move by less than their BF16 rounding does
// (both measured against the FP64 network).  Seven VALU operations with
// the product (the last fma clamps with its output modifier) instead of an
// exponential and a reciprocal, which run at a quarter of the VALU rate.
__device__ inline float gelu_gate4(float x) {
  const float y = x * x;
  float p = 1.063038355e-05f;
  p = fmaf(p, y, -4.166025794e-04f);
  p = fmaf(p, y, 6.714980851e-03f);
  p = fmaf(p, y, -5.973969748e-02f);
  p = fmaf(p, y, 3.940491249e-01f);
  return __builtin_amdgcn_fmed3f(fmaf(x, p, 0.5f), 0.0f, 1.0f);
}

__device__ inline floatx4 mfma(bf16x8 a, bf16x8 b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ inline bf16x8 zero8() {
  bf16x8 z;
#pragma unroll
  for (int j = 0; j < 8; ++j) z[j] = static_cast<__bf16>(0.0f);
  return z;
}

// A fragments of a 16->16 conv: A[co][k] = W[co][ci][d], k = 16 d + ci.
// transpose = true builds the data-gradient weights W'[ci][co][d] =
// W[co][ci][4-d] (rows indexed by ci, k = 16 d + co).
__device__ inline void load_conv_weights(const float *W, bool transpose, bf16x8 (&a)[3]) {
  const int lane = threadIdx.x & 63, row = lane & 15, g = lane >> 4;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int k = 32 * q + 8 * g + j;
      float v = 0.0f;
      if (k < 16 * KW) {
        const int d = k >> 4, c = k & 15;
        v = transpose ? W[(c * CH + row) * KW + (KW - 1 - d)] : W[(row * CH + c) * KW + d];
      }
      a[q][j] = static_cast<__bf16>(v);
    }
  }
}

// A fragment of layer 1: A[co][k] = W1[co][0][k] for k < 5.
__device__ inline bf16x8 load_first_weights(const float *W1) {
  const int lane = threadIdx.x & 63, row = lane & 15, g = lane >> 4;
  bf16x8 a;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int k = 8 * g + j;
    a[j] = static_cast<__bf16>(k < KW ? W1[row * KW + k] : 0.0f);
  }
  return a;
}

// B fragment for layer 1 at position pos: spins of sample s at pos + k - 2.
__device__ inline bf16x8 spin_fragment(uint64_t state, int n, int pos) {
  const int g = (threadIdx.x & 63) >> 4;
  bf16x8 b = zero8();
  if (g == 0) {
#pragma unroll
    for (int k = 0; k < KW; ++k) {
      int p = pos + k - 2;
      p = p < 0 ? p + n : (p >= n ? p - n : p);
      b[k] = static_cast<__bf16>(((state >> p) & 1) ? 1.0f : -1.0f);
    }
  }
  return b;
}

// act: [16][n][16] bf16 in LDS; B fragment of step q at position pos.
__device__ inline bf16x8 act_fragment(const __bf16 *act, int n, int pos, int q) {
  const int lane = threadIdx.x & 63, s = lane & 15, g = lane >> 4;
  const int k0 = 32 * q + 8 * g;
  const int d = k0 >> 4, c0 = k0 & 15;
  if (d >= KW) return zero8();
  int p = pos + d - 2;
  p = p < 0 ? p + n : (p >= n ? p - n : p);
  return *reinterpret_cast<const bf16x8 *>(act + (static_cast<size_t>(s) * n + p) * CH + c0);
}

__device__ inline floatx4 conv_at(const bf16x8 (&a)[3], const __bf16 *act, int n, int pos) {
  floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int q = 0; q < 3; ++q) acc = mfma(a[q], act_fragment(act, n, pos, q), acc);
  return acc;
}

// store 4 channels (rows 4g..4g+3) of sample s at pos
__device__ inline void store4(__bf16 *act, int n, int pos, floatx4 v) {
  const int lane = threadIdx.x & 63, s = lane & 15, g = lane >> 4;
  bf16x4 o;
#pragma unroll
  for (int r = 0; r < 4; ++r) o[r] = static_cast<__bf16>(v[r]);
  *reinterpret_cast<bf16x4 *>(act + (static_cast<size_t>(s) * n + pos) * CH + 4 * g) = o;
}

__device__ inline void store4_f32(float *z, int n, int pos, floatx4 v) {
  // z: [16][n][16] fp32 (global), same indexing as the LDS activations
  const int lane = threadIdx.x & 63, s = lane & 15, g = lane >> 4;
  *reinterpret_cast<floatx4 *>(z + (static_cast<size_t>(s) * n + pos) * CH + 4 * g) = v;
}

__device__ inline floatx4 load4_f32(const float *z, int n, int pos) {
  const int lane = threadIdx.x & 63, s = lane & 15, g = lane >> 4;
  return *reinterpret_cast<const floatx4 *>(z + (static_cast<size_t>(s) * n + pos) * CH + 4 * g);
}

struct CnnArgs {
  const float *params; // FP32 master parameters
  int n;
  uint64_t count;      // samples
};

// Forward: one wave per 16 samples; writes a, t per sample and optionally
// the pre-activations z1..z3 ([sample][pos][c], fp32) for the backward pass.
__global__ void __launch_bounds__(64)
    cnn_forward_kernel(CnnArgs c, const uint64_t *__restrict__ states, float *__restrict__ out_a,
                       float *__restrict__ out_t, float *__restrict__ z_store) {
  extern __shared__ __attribute__((aligned(16))) __bf16 lds[];
  const int n = c.n;
  // n >= kInPlaceMinN: layer 2 overwrites layer 1's activations in place
  // (one 16 x n x 16 buffer; twice the resident waves per CU)
  const bool in_place = n >= kInPlaceMinN;
  __bf16 *h1 = lds;
  __bf16 *h2 = in_place ? lds : lds + 16 * n * CH;
  const int lane = threadIdx.x, s = lane & 15, g = lane >> 4;
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * 16;
  const uint64_t sample = base + s;
  const uint64_t state = sample < c.count ? states[sample] : 0;
  const float *P = c.params;
  float *z1 = nullptr, *z2 = nullptr, *z3 = nullptr;
  if (z_store) {
    const size_t per = static_cast<size_t>(n) * CH;
    z1 = z_store + base * per;
    z2 = z_store + (c.count + 15) / 16 * 16 * per + base * per;
    z3 = z_store + 2 * ((c.count + 15) / 16 * 16) * per + base * per;
  }
  float bias[4];
  // layer 1
  if (!z1) {
    // inference: layer 1 at a position sees only the five spins around it,
    // so its 32 possible outputs (x 16 channels) are evaluated once per
    // workgroup (BF16 weights, FP32 sums, as the MFMA path) and looked up
    __bf16 *tab = lds + (in_place ? 1 : 2) * 16 * n * CH; // [32][16]
    for (int e = lane; e < 32 * CH; e += 64) {
      const int pat = e / CH, ch = e % CH;
      float v = 0.0f;
      for (int k = 0; k < KW; ++k) {
        const float wk = static_cast<float>(static_cast<__bf16>(P[OFF_W1 + ch * KW + k]));
        v += ((pat >> k) & 1) ? wk : -wk;
      }
      tab[e] = static_cast<__bf16>(gelu_f(v + P[OFF_B1 + ch]));
    }
    __syncthreads();
    for (int pos = 0; pos < n; ++pos) {
      // bit k of the pattern: spin at pos + k - 2 (cyclic)
      const int start = pos >= 2 ? pos - 2 : pos - 2 + n;
      const uint64_t rot = start == 0 ? state : ((state >> start) | (state << (n - start)));
      const int pat = static_cast<int>(rot & 31u);
      *reinterpret_cast<bf16x4 *>(h1 + (static_cast<size_t>(s) * n + pos) * CH + 4 * g) =
          *reinterpret_cast<const bf16x4 *>(tab + pat * CH + 4 * g);
    }
  } else {
    const bf16x8 a = load_first_weights(P + OFF_W1);
#pragma unroll
    for (int r = 0; r < 4; ++r) bias[r] = P[OFF_B1 + 4 * g + r];
    for (int pos = 0; pos < n; ++pos) {
      floatx4 acc = {0.f, 0.f, 0.f, 0.f};
      acc = mfma(a, spin_fragment(state, n, pos), acc);
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[r] += bias[r];
      if (z1) store4_f32(z1, n, pos, acc);
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[r] = gelu_f(acc[r]);
      store4(h1, n, pos, acc);
    }
  }
  __syncthreads();
  // layer 2
  {
    bf16x8 a[3];
    load_conv_weights(P + OFF_W2, false, a);
#pragma unroll
    for (int r = 0; r < 4; ++r) bias[r] = P[OFF_B2 + 4 * g + r];
    if (in_place) {
      // slot s of h1 is last read by output s + 2, except slots n-2, n-1, 0
      // and 1, which the wrap-around of the last outputs still reads: output
      // pos is written when output pos + 2 is done; outputs 0, 1, n-2, n-1
      // are written after the loop.  (One wave per workgroup: its LDS
      // accesses stay in program order.)
      floatx4 o0 = {0.f, 0.f, 0.f, 0.f}, o1 = o0, prev1 = o0, prev2 = o0;
      for (int pos = 0; pos < n; ++pos) {
        floatx4 acc = conv_at(a, h1, n, pos);
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[r] += bias[r];
        if (z2) store4_f32(z2, n, pos, acc);
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[r] = gelu_f(acc[r]);
        if (pos >= 4) store4(h1, n, pos - 2, prev2);
        if (pos == 0) o0 = acc;
        if (pos == 1) o1 = acc;
        prev2 = prev1;
        prev1 = acc;
      }
      store4(h1, n, n - 2, prev2);
      store4(h1, n, n - 1, prev1);
      store4(h1, n, 0, o0);
      store4(h1, n, 1, o1);
    } else {
      for (int pos = 0; pos < n; ++pos) {
        floatx4 acc = conv_at(a, h1, n, pos);
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[r] += bias[r];
        if (z2) store4_f32(z2, n, pos, acc);
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[r] = gelu_f(acc[r]);
        store4(h2, n, pos, acc);
      }
    }
  }
  __syncthreads();
  // layer 3 + global sum
  float f[4] = {0.f, 0.f, 0.f, 0.f};
  {
    bf16x8 a[3];
    load_conv_weights(P + OFF_W3, false, a);
#pragma unroll
    for (int r = 0; r < 4; ++r) bias[r] = P[OFF_B3 + 4 * g + r];
    for (int pos = 0; pos < n; ++pos) {
      floatx4 acc = conv_at(a, h2, n, pos);
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[r] += bias[r];
      if (z3) store4_f32(z3, n, pos, acc);
#pragma unroll
      for (int r = 0; r < 4; ++r) f[r] += gelu_f(acc[r]);
    }
  }
  // heads: reduce over the 4 lane groups holding channels 4g..4g+3
  float pa = 0.f, pt = 0.f;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    pa += P[OFF_U + 4 * g + r] * f[r];
    pt += P[OFF_V + 4 * g + r] * f[r];
  }
  pa += __shfl_xor(pa, 16, 64);
  pa += __shfl_xor(pa, 32, 64);
  pt += __shfl_xor(pt, 16, 64);
  pt += __shfl_xor(pt, 32, 64);
  if (g == 0 && sample < c.count) {
    out_a[sample] = pa + P[OFF_U0];
    out_t[sample] = pt + P[OFF_V0];
  }
}

// Streaming inference (no stored pre-activations, n >= kInferMinN): the
// forward kernel's network (same BF16 operands; the polynomial GELU gates
// in layers 2 and 3; MFMA sums that start from the bias instead of adding
// it last) from 7.5 KB of LDS per wave instead of 21 KB, so about three
// times the waves stay resident.
//  * layer 1 is the table of its 32 input patterns; layer 2 reads its B
//    fragments from the table directly, so layer 1 is never stored;
//  * layer 2's outputs go to a ring of the last 8 positions (slot pos mod 8,
//    slot 0 mirrored in slot 8, plus a copy of positions 0..3 for the
//    wrap-around), and layer 3 at position p - 2 runs as soon as layer 2 at
//    p is done;
//  * the position loop is unrolled over the ring, so every ring address is
//    a per-lane base plus a constant: no per-position address arithmetic
//    beyond the two table lookups' pattern extraction.
constexpr int kRing = 8;
constexpr int kInferMinN = 9;
constexpr int kSlotBytes = 16 * CH * 2; // one position of 16 samples x 16 channels

// What every wave of the inference kernel needs, once per evaluation
// instead of once per wave (64 x 7 16-byte words):
//  * words 0..63, the layer-1 table tab[h][pattern][8 channels]: channels
//    8h .. 8h + 7 of GELU(b1 + W1 h0) for each 5-spin pattern (bit k of the
//    pattern = spin at offset k - 2), 16 bytes per pattern;
//  * words 64 + 6 lane + q, lane's BF16 A fragments of layers 2 (q < 3) and
//    3 (q >= 3), as load_conv_weights builds them.
constexpr int kInferPrepWords = 64 * 7;
__global__ void __launch_bounds__(64) cnn_table_kernel(const float *__restrict__ P,
                                                       uint4 *__restrict__ prep) {
  __bf16 *tab = reinterpret_cast<__bf16 *>(prep);
  bf16x8 a[3];
  load_conv_weights(P + OFF_W2, false, a);
#pragma unroll
  for (int q = 0; q < 3; ++q) reinterpret_cast<bf16x8 *>(prep + 64 + 6 * threadIdx.x)[q] = a[q];
  load_conv_weights(P + OFF_W3, false, a);
#pragma unroll
  for (int q = 0; q < 3; ++q) reinterpret_cast<bf16x8 *>(prep + 64 + 6 * threadIdx.x)[3 + q] = a[q];
  for (int e = threadIdx.x; e < 32 * CH; e += 64) {
    const int pat = e / CH, ch = e % CH;
    float v = 0.0f;
    for (int k = 0; k < KW; ++k) {
      const float wk = static_cast<float>(static_cast<__bf16>(P[OFF_W1 + ch * KW + k]));
      v += ((pat >> k) & 1) ? wk : -wk;
    }
    tab[(ch >> 3) * 256 + pat * 8 + (ch & 7)] = static_cast<__bf16>(gelu_f(v + P[OFF_B1 + ch]));
  }
}

__global__ void __launch_bounds__(64)
    cnn_infer_kernel(CnnArgs c, const uint4 *__restrict__ table, const uint64_t *__restrict__ states,
                     float *__restrict__ out_a, float *__restrict__ out_t) {
  // the layer-1 table (cnn_table_kernel's layout), one 16-byte load per lane
  __shared__ __attribute__((aligned(16))) __bf16 tab[2 * 32 * 8];
  __shared__ __attribute__((aligned(16))) __bf16 ring[(kRing + 1) * 16 * CH];
  __shared__ __attribute__((aligned(16))) __bf16 head[4 * 16 * CH];
  const int n = c.n;
  const int lane = threadIdx.x, s = lane & 15, g = lane >> 4;
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * 16;
  const uint64_t sample = base + s;
  const uint64_t state = sample < c.count ? states[sample] : 0;
  const float *P = c.params;
  reinterpret_cast<uint4 *>(tab)[lane] = table[lane];
  __syncthreads();
  bf16x8 a2[3], a3[3];
  const bf16x8 *fr = reinterpret_cast<const bf16x8 *>(table + 64 + 6 * lane);
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    a2[q] = fr[q];
    a3[q] = fr[3 + q];
  }
  float b2[4], b3[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    b2[r] = P[OFF_B2 + 4 * g + r];
    b3[r] = P[OFF_B3 + 4 * g + r];
  }
  // B fragment q holds rows k = 32 q + 8 g .. + 7: tap offset d = k >> 4,
  // channels 8h .. 8h + 7.  Rows k >= 80 (offset 5) meet zero weights
  // (load_conv_weights), so their lanes read the offset-4 data instead of
  // building zeros, and no fragment load needs a branch: d = 2q + hi for
  // q < 2, d = 4 for q = 2.  (A 16x16x16 step for tap 4 alone would need
  // its accumulator from a 16x16x32 step, and such a mixed-shape dependent
  // pair gives wrong sums as the compiler issues it, with no wait states.)
  const int hi = g >> 1, h = g & 1;
  const int dq[3] = {hi, 2 + hi, 4};
  // Ring and head rows (32 bytes per sample) hold channel c of sample s at
  // c ^ 8 fs(s): the 16-byte halves swap for fs = 1, so the lanes of one
  // LDS phase (8 or 16 consecutive samples reading the same half) spread
  // over all banks instead of meeting every 4th / 8th sample.
  const int fs = ((s >> 2) ^ (s >> 3)) & 1;
  const int c0 = 8 * (h ^ fs);                             // this lane's read half
  const char *tabc = reinterpret_cast<const char *>(tab);
  char *ringc = reinterpret_cast<char *>(ring);
  char *headc = reinterpret_cast<char *>(head);
  const uint32_t hoff = 512u * h;                          // table half, bytes
  const uint32_t rl = static_cast<uint32_t>(s * CH + c0) * 2; // ring read base
  const uint32_t rlh = rl + hi * kSlotBytes;               // ... one slot up for d odd
  // ring write base: channels 4g .. 4g + 3 of sample s
  const uint32_t wl = static_cast<uint32_t>(s * CH + 4 * (g ^ (2 * fs))) * 2;
  // ext bit k = state bit (k - 4) mod n for k < 2n (n >= kInferMinN makes
  // 2n >= n + 9): layer 2 at pos reads the layer-1 table at the 5-spin
  // pattern around pos + d - 2, i.e. bits d .. d + 4 of win = ext >> pos.
  const uint64_t nmask = n == 64 ? ~0ull : ((1ull << n) - 1);
  const uint64_t rot = ((state >> (n - 4)) | (state << 4)) & nmask;
  const uint64_t ext_lo = n == 64 ? rot : (rot | (rot << n));
  const uint64_t ext_hi = n == 64 ? rot : (rot >> (64 - n));
  const uint32_t e0 = static_cast<uint32_t>(ext_lo), e1 = static_cast<uint32_t>(ext_lo >> 32),
                 e2 = static_cast<uint32_t>(ext_hi);
  auto ld = [](const char *p) { return *reinterpret_cast<const bf16x8 *>(p); };
  // the last four layer-3 positions wrap around to the head copy of 0..3
  auto tail_src = [&](int p) {
    p = p < 0 ? p + n : (p >= n ? p - n : p);
    return p < 4 ? headc + p * kSlotBytes : ringc + (p & (kRing - 1)) * kSlotBytes;
  };
  // the biases start the MFMA accumulators
  float f[4] = {0.f, 0.f, 0.f, 0.f};
  auto layer3_out = [&](floatx4 acc) {
#pragma unroll
    for (int r = 0; r < 4; ++r) f[r] = fmaf(acc[r], gelu_gate4(acc[r]), f[r]);
  };
  // (one wave per workgroup: its LDS accesses stay in program order)
  for (int pos0 = 0; pos0 < n; pos0 += kRing) {
    const uint32_t wlo = pos0 < 32 ? e0 : e1, whi = pos0 < 32 ? e1 : e2;
    const int sh0 = pos0 & 31;
#pragma unroll
    for (int u = 0; u < kRing; ++u) {
      const int pos = pos0 + u;
      if (pos >= n) break;
      const uint32_t win = __builtin_amdgcn_alignbit(whi, wlo, sh0 + u);
      floatx4 acc = {b2[0], b2[1], b2[2], b2[3]};
      acc = mfma(a2[0], ld(tabc + ((__builtin_amdgcn_ubfe(win, dq[0], 5) << 4) | hoff)), acc);
      acc = mfma(a2[1], ld(tabc + ((__builtin_amdgcn_ubfe(win, dq[1], 5) << 4) | hoff)), acc);
      acc = mfma(a2[2], ld(tabc + ((win & 0x1f0u) | hoff)), acc);
      bf16x4 o;
#pragma unroll
      for (int r = 0; r < 4; ++r) o[r] = static_cast<__bf16>(acc[r] * gelu_gate4(acc[r]));
      *reinterpret_cast<bf16x4 *>(ringc + wl + u * kSlotBytes) = o;
      if (u == 0) *reinterpret_cast<bf16x4 *>(ringc + wl + kRing * kSlotBytes) = o;
      if (pos < 4) *reinterpret_cast<bf16x4 *>(headc + wl + pos * kSlotBytes) = o;
      if (pos >= 4) {
        // layer 3 at pos - 2 reads layer 2 at pos - 4 + d, slot (u - 4 + d)
        // mod 8: d = 2q + hi for q < 2 (slot 8 stands for slot 0), d = 4
        // (slot u) for q = 2
        floatx4 acc3 = {b3[0], b3[1], b3[2], b3[3]};
        acc3 = mfma(a3[0], ld(ringc + rlh + ((u - 4) & (kRing - 1)) * kSlotBytes), acc3);
        acc3 = mfma(a3[1], ld(ringc + rlh + ((u - 2) & (kRing - 1)) * kSlotBytes), acc3);
        acc3 = mfma(a3[2], ld(ringc + rl + u * kSlotBytes), acc3);
        layer3_out(acc3);
      }
    }
  }
  const int tail_pos[4] = {n - 2, n - 1, 0, 1};
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int tp = tail_pos[t];
    floatx4 acc3 = {b3[0], b3[1], b3[2], b3[3]};
#pragma unroll
    for (int q = 0; q < 3; ++q) acc3 = mfma(a3[q], ld(tail_src(tp + dq[q] - 2) + rl), acc3);
    layer3_out(acc3);
  }
  float pa = 0.f, pt = 0.f;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    pa += P[OFF_U + 4 * g + r] * f[r];
    pt += P[OFF_V + 4 * g + r] * f[r];
  }
  pa += __shfl_xor(pa, 16, 64);
  pa += __shfl_xor(pa, 32, 64);
  pt += __shfl_xor(pt, 16, 64);
  pt += __shfl_xor(pt, 32, 64);
  if (g == 0 && sample < c.count) {
    out_a[sample] = pa + P[OFF_U0];
    out_t[sample] = pt + P[OFF_V0];
  }
}

// Weight gradient of one conv layer: C_d[co][ci] += sum_slots dz[slot][co] *
// h[slot shifted by d][ci]; 32 (sample, position) slots per MFMA.
// Layer 1 (first = true): C[co][d] with h = spins.
__device__ inline void weight_grad(const __bf16 *dz, const __bf16 *h, const uint64_t *st, int n,
                                   bool first, floatx4 (&cacc)[KW], float &dbias) {
  const int lane = threadIdx.x & 63, col = lane & 15, g = lane >> 4;
  const int nslots = 16 * n;
  for (int c0 = 0; c0 < nslots; c0 += 32) {
    bf16x8 a;
    int ss[8], pp[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int slot = c0 + 8 * g + j;
      const bool ok = slot < nslots;
      ss[j] = ok ? slot / n : 0;
      pp[j] = ok ? slot % n : 0;
      const float v = ok ? static_cast<float>(dz[(static_cast<size_t>(ss[j]) * n + pp[j]) * CH + col]) : 0.0f;
      a[j] = static_cast<__bf16>(v);
      dbias += v;
      if (!ok) ss[j] = -1;
    }
    if (first) {
      bf16x8 b;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        float v = 0.0f;
        if (ss[j] >= 0 && col < KW) {
          int p = pp[j] + col - 2;
          p = p < 0 ? p + n : (p >= n ? p - n : p);
          v = ((st[ss[j]] >> p) & 1) ? 1.0f : -1.0f;
        }
        b[j] = static_cast<__bf16>(v);
      }
      cacc[0] = mfma(a, b, cacc[0]);
    } else {
#pragma unroll
      for (int d = 0; d < KW; ++d) {
        bf16x8 b;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          float v = 0.0f;
          if (ss[j] >= 0) {
            int p = pp[j] + d - 2;
            p = p < 0 ? p + n : (p >= n ? p - n : p);
            v = static_cast<float>(h[(static_cast<size_t>(ss[j]) * n + p) * CH + col]);
          }
          b[j] = static_cast<__bf16>(v);
        }
        cacc[d] = mfma(a, b, cacc[d]);
      }
    }
  }
}

// Backward for one wave of 16 samples: per-wave partial gradients
// (NPARAMS floats) of sum_s ga[s] a_s + gt[s] t_s.
__global__ void __launch_bounds__(64)
    cnn_backward_kernel(CnnArgs c, const uint64_t *__restrict__ states,
                        const float *__restrict__ z_store, const float *__restrict__ ga,
                        const float *__restrict__ gt, float *__restrict__ partial) {
  extern __shared__ __attribute__((aligned(16))) __bf16 lds[];
  const int n = c.n;
  const size_t act = static_cast<size_t>(16) * n * CH;
  __bf16 *bufA = lds, *bufB = lds + act, *bufC = lds + 2 * act;
  __shared__ uint64_t st[16];
  __shared__ float gsh[2][16];
  const int lane = threadIdx.x, s = lane & 15, g = lane >> 4;
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * 16;
  const uint64_t sample = base + s;
  const bool valid = sample < c.count;
  if (g == 0) {
    st[s] = valid ? states[sample] : 0;
    gsh[0][s] = valid ? ga[sample] : 0.0f;
    gsh[1][s] = valid ? gt[sample] : 0.0f;
  }
  __syncthreads();
  const float *P = c.params;
  const size_t per = static_cast<size_t>(n) * CH;
  const uint64_t padded = (c.count + 15) / 16 * 16;
  const float *z1 = z_store + base * per;
  const float *z2 = z_store + padded * per + base * per;
  const float *z3 = z_store + 2 * padded * per + base * per;
  float *out = partial + static_cast<size_t>(blockIdx.x) * NPARAMS;
  const float gas = gsh[0][s], gts = gsh[1][s];

  // heads and d3 = (ga u + gt v) gelu'(z3)  -> bufA; h2 = gelu(z2) -> bufB
  float f[4] = {0.f, 0.f, 0.f, 0.f};
  float coef[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) coef[r] = gas * P[OFF_U + 4 * g + r] + gts * P[OFF_V + 4 * g + r];
  for (int pos = 0; pos < n; ++pos) {
    floatx4 z = load4_f32(z3, n, pos), d;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      f[r] += gelu_f(z[r]);
      d[r] = valid ? coef[r] * gelu_grad_f(z[r]) : 0.0f;
    }
    store4(bufA, n, pos, d);
    floatx4 z2v = load4_f32(z2, n, pos), h;
#pragma unroll
    for (int r = 0; r < 4; ++r) h[r] = gelu_f(z2v[r]);
    store4(bufB, n, pos, h);
  }
  // du[co] = sum_s ga[s] f[s][co]: reduce over the 16 samples of a lane group
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    float du = valid ? gas * f[r] : 0.0f, dv = valid ? gts * f[r] : 0.0f;
    for (int off = 8; off > 0; off >>= 1) {
      du += __shfl_xor(du, off, 64);
      dv += __shfl_xor(dv, off, 64);
    }
    if (s == 0) {
      out[OFF_U + 4 * g + r] = du;
      out[OFF_V + 4 * g + r] = dv;
    }
  }
  {
    float a0 = gas, t0 = gts;
    for (int off = 8; off > 0; off >>= 1) {
      a0 += __shfl_xor(a0, off, 64);
      t0 += __shfl_xor(t0, off, 64);
    }
    if (lane == 0) {
      out[OFF_U0] = a0;
      out[OFF_V0] = t0;
    }
  }
  __syncthreads();

  auto write_wgrad = [&](int off_w, int off_b, floatx4 (&cacc)[KW], float dbias, bool first) {
    // cacc[d]: rows co = 4g + r, column = ci (or d for the first layer)
    const int col = lane & 15;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int co = 4 * g + r;
      if (first) {
        if (col < KW) out[off_w + co * KW + col] = cacc[0][r];
      } else {
#pragma unroll
        for (int d = 0; d < KW; ++d) out[off_w + (co * CH + col) * KW + d] = cacc[d][r];
      }
    }
    // bias: dbias holds this lane's share for channel `col` (A rows are co = col)
    float b = dbias;
    b += __shfl_xor(b, 16, 64);
    b += __shfl_xor(b, 32, 64);
    if (g == 0) out[off_b + col] = b;
  };

  // layer 3 weight gradient (dz = bufA, h = bufB)
  {
    floatx4 cacc[KW];
#pragma unroll
    for (int d = 0; d < KW; ++d) cacc[d] = floatx4{0.f, 0.f, 0.f, 0.f};
    float db = 0.0f;
    weight_grad(bufA, bufB, st, n, false, cacc, db);
    write_wgrad(OFF_W3, OFF_B3, cacc, db, false);
  }
  __syncthreads();
  // d2 = conv^T(d3) * gelu'(z2) -> bufC ; h1 = gelu(z1) -> bufB
  {
    bf16x8 a[3];
    load_conv_weights(P + OFF_W3, true, a);
    for (int pos = 0; pos < n; ++pos) {
      floatx4 acc = conv_at(a, bufA, n, pos);
      floatx4 z = load4_f32(z2, n, pos);
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[r] *= gelu_grad_f(z[r]);
      store4(bufC, n, pos, acc);
    }
  }
  __syncthreads();
  for (int pos = 0; pos < n; ++pos) {
    floatx4 z = load4_f32(z1, n, pos), h;
#pragma unroll
    for (int r = 0; r < 4; ++r) h[r] = gelu_f(z[r]);
    store4(bufB, n, pos, h);
  }
  __syncthreads();
  // layer 2 weight gradient (dz = bufC, h = bufB)
  {
    floatx4 cacc[KW];
#pragma unroll
    for (int d = 0; d < KW; ++d) cacc[d] = floatx4{0.f, 0.f, 0.f, 0.f};
    float db = 0.0f;
    weight_grad(bufC, bufB, st, n, false, cacc, db);
    write_wgrad(OFF_W2, OFF_B2, cacc, db, false);
  }
  __syncthreads();
  // d1 = conv^T(d2) * gelu'(z1) -> bufA
  {
    bf16x8 a[3];
    load_conv_weights(P + OFF_W2, true, a);
    for (int pos = 0; pos < n; ++pos) {
      floatx4 acc = conv_at(a, bufC, n, pos);
      floatx4 z = load4_f32(z1, n, pos);
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[r] *= gelu_grad_f(z[r]);
      store4(bufA, n, pos, acc);
    }
  }
  __syncthreads();
  // layer 1 weight gradient (dz = bufA, input spins)
  {
    floatx4 cacc[KW];
#pragma unroll
    for (int d = 0; d < KW; ++d) cacc[d] = floatx4{0.f, 0.f, 0.f, 0.f};
    float db = 0.0f;
    weight_grad(bufA, nullptr, st, n, true, cacc, db);
    write_wgrad(OFF_W1, OFF_B1, cacc, db, true);
  }
}

// grad[p] = sum_w partial[w][p] (fixed order)
__global__ void reduce_grad_kernel(const float *__restrict__ partial, int nwaves,
                                   float *__restrict__ grad) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= NPARAMS) return;
  float s = 0.0f;
  for (int w = 0; w < nwaves; ++w) s += partial[static_cast<size_t>(w) * NPARAMS + p];
  grad[p] = s;
}

// Adam (Kingma & Ba 2015) on the FP32 master parameters.
__global__ void adam_kernel(float *__restrict__ p, const float *__restrict__ g,
                            float *__restrict__ m, float *__restrict__ v, float lr, float b1,
                            float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= NPARAMS) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// start vector c_r = marshall(r) * sign(t_r) * exp(a_r - amax) / N_r
__global__ void start_vector_kernel(const float *__restrict__ a, const float *__restrict__ t,
                                    const uint64_t *__restrict__ reps,
                                    const uint32_t *__restrict__ stab, uint64_t D, float amax,
                                    uint64_t marshall_mask, double inv_group, double *__restrict__ v) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < D;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    double amp = exp(static_cast<double>(a[i] - amax));
    if (t[i] < 0.0f) amp = -amp;
    if (__popcll(reps[i] & marshall_mask) & 1) amp = -amp;
    v[i] = amp / sqrt(static_cast<double>(stab[i]) * inv_group);
  }
}

__global__ void max_partial_kernel(const float *__restrict__ a, uint64_t n, float *__restrict__ out) {
  __shared__ float lds[256];
  float m = -3.0e38f;
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    m = fmaxf(m, a[i]);
  lds[threadIdx.x] = m;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < static_cast<unsigned>(off)) lds[threadIdx.x] = fmaxf(lds[threadIdx.x], lds[threadIdx.x + off]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = lds[0];
}

// One MFMA with known operands, for the lane-layout check.
__global__ void mfma_probe_kernel(float *out) {
  const int l = threadIdx.x, row = l & 15, g = l >> 4;
  bf16x8 a, b;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int k = 8 * g + j;
    a[j] = static_cast<__bf16>(static_cast<float>((row * 3 + k) % 7 - 3));     // A[row][k]
    b[j] = static_cast<__bf16>(static_cast<float>((k * 3 + row * 7) % 5 - 2)); // B[k][col=row]
  }
  floatx4 c = {0.f, 0.f, 0.f, 0.f};
  c = mfma(a, b, c);
#pragma unroll
  for (int r = 0; r < 4; ++r) out[(4 * g + r) * 16 + row] = c[r]; // D[4g+r][col]
}

// activations (one or two 16 x n x 16 BF16 buffers) + the layer-1 table
size_t lds_bytes_forward(int n) {
  return static_cast<size_t>(n >= kInPlaceMinN ? 1 : 2) * 16 * n * CH * 2 + 32 * CH * 2;
}
size_t lds_bytes_backward(int n) { return static_cast<size_t>(3) * 16 * n * CH * 2; }

} // namespace

struct CnnModel {
  DeviceBuffer<float> params, m, v, grad;
  DeviceBuffer<uint4> table; // the inference kernel's table and weight fragments
  int64_t step = 0;
};

// the table from the current parameters, then the inference kernel
inline void launch_infer(const CnnModel &m, const CnnArgs &a, unsigned blocks,
                         const uint64_t *states, float *out_a, float *out_t, hipStream_t s) {
  cnn_table_kernel<<<1, 64, 0, s>>>(a.params, m.table.ptr);
  cnn_infer_kernel<<<blocks, 64, 0, s>>>(a, m.table.ptr, states, out_a, out_t);
}

} // namespace spined

using spined::CnnModel;
using spined::DeviceBuffer;
using spined::Error;

struct spined_cnn {
  CnnModel m;
};

extern "C" {

int spined_cnn_num_params(void) { return spined::NPARAMS; }

int spined_cnn_create(spined_cnn **out, const float *init_params) {
  return spined::guarded([&] {
    if (!out || !init_params) throw Error(LS_INVALID_ARGUMENT, "NULL argument");
    auto *c = new spined_cnn();
    c->m.params.allocate(spined::NPARAMS);
    c->m.m.allocate(spined::NPARAMS);
    c->m.v.allocate(spined::NPARAMS);
    c->m.grad.allocate(spined::NPARAMS);
    c->m.table.allocate(spined::kInferPrepWords);
    SPINED_HIP(hipMemcpy(c->m.params.ptr, init_params, 4 * spined::NPARAMS, hipMemcpyHostToDevice));
    SPINED_HIP(hipMemset(c->m.m.ptr, 0, 4 * spined::NPARAMS));
    SPINED_HIP(hipMemset(c->m.v.ptr, 0, 4 * spined::NPARAMS));
    *out = c;
  });
}

void spined_cnn_destroy(spined_cnn *c) { delete c; }

int spined_cnn_get_params(const spined_cnn *c, float *out) {
  return spined::guarded([&] {
    if (!c || !out) throw Error(LS_INVALID_ARGUMENT, "NULL argument");
    SPINED_HIP(hipMemcpy(out, c->m.params.ptr, 4 * spined::NPARAMS, hipMemcpyDeviceToHost));
  });
}

int spined_cnn_set_params(spined_cnn *c, const float *in) {
  return spined::guarded([&] {
    if (!c || !in) throw Error(LS_INVALID_ARGUMENT, "NULL argument");
    SPINED_HIP(hipMemcpy(c->m.params.ptr, in, 4 * spined::NPARAMS, hipMemcpyHostToDevice));
  });
}

// Forward on device buffers: states (count), out_a / out_t (count floats);
// z_store (3 * ceil16(count) * n * 16 floats) may be NULL.
int spined_cnn_forward_device(const spined_cnn *c, int n, uint64_t count, const uint64_t *states,
                              float *out_a, float *out_t, float *z_store) {
  return spined::guarded([&] {
    if (!c || n < 5 || n > spined::kMaxN) throw Error(LS_INVALID_ARGUMENT, "n must be in [5, 64]");
    if (count == 0) return;
    spined::CnnArgs a{c->m.params.ptr, n, count};
    const unsigned blocks = static_cast<unsigned>((count + 15) / 16);
    if (!z_store && n >= spined::kInferMinN)
      spined::launch_infer(c->m, a, blocks, states, out_a, out_t, spined::stream());
    else
      hipLaunchKernelGGL(spined::cnn_forward_kernel, dim3(blocks), dim3(64),
                         spined::lds_bytes_forward(n), spined::stream(), a, states, out_a, out_t,
                         z_store);
    SPINED_CHECK_LAUNCH();
  });
}

// Gradient (device buffer grad_out, NPARAMS floats) of sum ga*a + gt*t over
// `count` samples whose pre-activations were stored by a forward call.
int spined_cnn_backward_device(spined_cnn *c, int n, uint64_t count, const uint64_t *states,
                               const float *z_store, const float *ga, const float *gt,
                               float *grad_out) {
  return spined::guarded([&] {
    if (!c || n < 5 || n > spined::kMaxN) throw Error(LS_INVALID_ARGUMENT, "n must be in [5, 64]");
    if (count == 0) return;
    spined::CnnArgs a{c->m.params.ptr, n, count};
    const unsigned blocks = static_cast<unsigned>((count + 15) / 16);
    DeviceBuffer<float> partial(static_cast<size_t>(blocks) * spined::NPARAMS);
    hipStream_t s = spined::stream();
    hipLaunchKernelGGL(spined::cnn_backward_kernel, dim3(blocks), dim3(64),
                       spined::lds_bytes_backward(n), s, a, states, z_store, ga, gt, partial.ptr);
    SPINED_CHECK_LAUNCH();
    float *dst = grad_out ? grad_out : c->m.grad.ptr;
    hipLaunchKernelGGL(spined::reduce_grad_kernel, dim3((spined::NPARAMS + 255) / 256), dim3(256),
                       0, s, partial.ptr, static_cast<int>(blocks), dst);
    SPINED_CHECK_LAUNCH();
    SPINED_HIP(hipStreamSynchronize(s)); // `partial` is released at scope exit
  });
}

// One Adam step with the gradient left in the model by backward (grad_out = NULL).
int spined_cnn_adam_step(spined_cnn *c, float lr) {
  return spined::guarded([&] {
    if (!c) throw Error(LS_INVALID_ARGUMENT, "NULL model");
    const float b1 = 0.9f, b2 = 0.999f;
    c->m.step += 1;
    const float c1 = 1.0f - std::pow(b1, static_cast<float>(c->m.step));
    const float c2 = 1.0f - std::pow(b2, static_cast<float>(c->m.step));
    hipLaunchKernelGGL(spined::adam_kernel, dim3((spined::NPARAMS + 255) / 256), dim3(256), 0,
                       spined::stream(), c->m.params.ptr, c->m.grad.ptr, c->m.m.ptr, c->m.v.ptr,
                       lr, b1, b2, 1e-8f, c1, c2);
    SPINED_CHECK_LAUNCH();
  });
}

} // extern "C"

extern "C" {

// Learned start vector over the whole basis (device, FP64, length D):
// c_r = marshall(r) sign(t_r) exp(a_r - max a) / N_r.
int spined_cnn_start_vector(const spined_cnn *c, const ls_spin_basis *basis,
                            uint64_t marshall_mask, double *v_device) {
  return spined::guarded([&] {
    if (!c || !basis || !v_device) throw Error(LS_INVALID_ARGUMENT, "NULL argument");
    spined::Basis &b = *basis->p;
    if (!b.built) throw Error(LS_CACHE_NOT_BUILT, "basis has not been built");
    spined::cnn_start_vector(c, b, marshall_mask, v_device);
  });
}

int spined_debug_mfma_probe(float *out_host) {
  return spined::guarded([&] {
    DeviceBuffer<float> d(256);
    hipLaunchKernelGGL(spined::mfma_probe_kernel, dim3(1), dim3(64), 0, spined::stream(), d.ptr);
    SPINED_CHECK_LAUNCH();
    SPINED_HIP(hipMemcpyAsync(out_host, d.ptr, 1024, hipMemcpyDeviceToHost, spined::stream()));
    SPINED_HIP(hipStreamSynchronize(spined::stream()));
  });
}

} // extern "C"

namespace spined {
// launched from api.cpp (needs the basis internals)
void cnn_start_vector(const spined_cnn *c, const Basis &b, uint64_t marshall_mask, double *v) {
  hipStream_t s = stream();
  const uint64_t D = b.D;
  if (D == 0) return;
  // network outputs in FP64-vector-sized pooled blocks: a solve that follows
  // takes them over as Krylov vectors without a free and a fresh allocation
  DeviceBuffer<float> a, t;
  a.allocate_pooled(D, D * sizeof(double));
  t.allocate_pooled(D, D * sizeof(double));
  CnnArgs args{c->m.params.ptr, b.n, D};
  const unsigned blocks = static_cast<unsigned>((D + 15) / 16);
  if (b.n >= kInferMinN)
    launch_infer(c->m, args, blocks, b.reps.ptr, a.ptr, t.ptr, s);
  else
    hipLaunchKernelGGL(cnn_forward_kernel, dim3(blocks), dim3(64), lds_bytes_forward(b.n), s,
                       args, b.reps.ptr, a.ptr, t.ptr, static_cast<float *>(nullptr));
  SPINED_CHECK_LAUNCH();
  const unsigned mb = 1024;
  DeviceBuffer<float> part(mb);
  max_partial_kernel<<<mb, 256, 0, s>>>(a.ptr, D, part.ptr);
  SPINED_CHECK_LAUNCH();
  std::vector<float> hp(mb);
  SPINED_HIP(hipMemcpyAsync(hp.data(), part.ptr, 4 * mb, hipMemcpyDeviceToHost, s));
  SPINED_HIP(hipStreamSynchronize(s));
  const float amax = *std::max_element(hp.begin(), hp.end());
  const double inv_group = 1.0 / static_cast<double>(b.tables.nperm * (b.tables.has_flip ? 2 : 1));
  start_vector_kernel<<<4096, 256, 0, s>>>(a.ptr, t.ptr, b.reps.ptr, b.stab.ptr, D, amax,
                                           marshall_mask, inv_group, v);
  SPINED_CHECK_LAUNCH();
  SPINED_HIP(hipStreamSynchronize(s));
}
} // namespace spined
