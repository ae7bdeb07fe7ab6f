/*
 * cnn_oracle.c -- FP64 CPU forward/backward of the learned-start network
 * (TEST INFRASTRUCTURE ONLY: the checker of the engine's BF16-MFMA kernels).
 *
 * The network has no counterpart in the reference; it is specified by the
 * engine's design (periodic 1D CNN over +-1 spins):
 *   h0[0][i] = +1 if bit i of the state is set, else -1      (1 channel)
 *   z_l[c][i] = b_l[c] + sum_{c', d<5} W_l[c][c'][d] h_{l-1}[c'][(i + d - 2) mod N]
 *   h_l = GELU(z_l)            l = 1, 2, 3; 16 channels; GELU in tanh form
 *   f[c] = sum_i h3[c][i]
 *   a = u0 + sum_c u[c] f[c]   (log |psi|)
 *   t = v0 + sum_c v[c] f[c]   (sign logit)
 * Parameter vector layout (2,722 doubles):
 *   W1[16][1][5], b1[16], W2[16][16][5], b2[16], W3[16][16][5], b3[16],
 *   u[16], u0, v[16], v0.
 * Pinned by finite differences (tests/test_cnn_oracle.py).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define C 16
#define K 5

enum {
  OFF_W1 = 0,
  OFF_B1 = OFF_W1 + C * 1 * K,
  OFF_W2 = OFF_B1 + C,
  OFF_B2 = OFF_W2 + C * C * K,
  OFF_W3 = OFF_B2 + C,
  OFF_B3 = OFF_W3 + C * C * K,
  OFF_U = OFF_B3 + C,
  OFF_U0 = OFF_U + C,
  OFF_V = OFF_U0 + 1,
  OFF_V0 = OFF_V + C,
  NPARAMS = OFF_V0 + 1,
};

int cnn_num_params(void) { return NPARAMS; }

/* GELU in its tanh form (Hendrycks & Gimpel, arXiv:1606.08415):
 * 0.5 x (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3))) */
#define GELU_A 0.7978845608028654 /* sqrt(2/pi) */
#define GELU_B 0.044715
static double gelu(double x) {
  return 0.5 * x * (1.0 + tanh(GELU_A * (x + GELU_B * x * x * x)));
}
static double gelu_grad(double x) {
  const double u = GELU_A * (x + GELU_B * x * x * x);
  const double t = tanh(u);
  return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * GELU_A * (1.0 + 3.0 * GELU_B * x * x);
}

/* z[co][i] = b[co] + sum W[co][ci][d] h[ci][(i+d-2) mod n] */
static void conv(int n, int cin, const double *W, const double *b, const double *h, double *z) {
  for (int co = 0; co < C; ++co)
    for (int i = 0; i < n; ++i) {
      double acc = b[co];
      for (int ci = 0; ci < cin; ++ci)
        for (int d = 0; d < K; ++d) {
          int j = (i + d - 2) % n;
          if (j < 0) j += n;
          acc += W[(co * cin + ci) * K + d] * h[ci * n + j];
        }
      z[co * n + i] = acc;
    }
}

/* Forward for `count` states: out[2*s] = a (log|psi|), out[2*s+1] = t. */
void cnn_forward(const double *params, int n, int64_t count, const uint64_t *states,
                 double *out) {
  double *h0 = (double *)malloc(sizeof(double) * n);
  double *z = (double *)malloc(sizeof(double) * C * n);
  double *h1 = (double *)malloc(sizeof(double) * C * n);
  double *h2 = (double *)malloc(sizeof(double) * C * n);
  for (int64_t s = 0; s < count; ++s) {
    for (int i = 0; i < n; ++i) h0[i] = ((states[s] >> i) & 1) ? 1.0 : -1.0;
    conv(n, 1, params + OFF_W1, params + OFF_B1, h0, z);
    for (int q = 0; q < C * n; ++q) h1[q] = gelu(z[q]);
    conv(n, C, params + OFF_W2, params + OFF_B2, h1, z);
    for (int q = 0; q < C * n; ++q) h2[q] = gelu(z[q]);
    conv(n, C, params + OFF_W3, params + OFF_B3, h2, z);
    double a = params[OFF_U0], t = params[OFF_V0];
    for (int c = 0; c < C; ++c) {
      double f = 0.0;
      for (int i = 0; i < n; ++i) f += gelu(z[c * n + i]);
      a += params[OFF_U + c] * f;
      t += params[OFF_V + c] * f;
    }
    out[2 * s] = a;
    out[2 * s + 1] = t;
  }
  free(h0);
  free(z);
  free(h1);
  free(h2);
}

/* dL/dz_in from dL/dh_out through conv (transpose), and parameter grads. */
static void conv_backward(int n, int cin, const double *W, const double *hin, const double *dz,
                          double *dW, double *db, double *dhin) {
  for (int co = 0; co < C; ++co)
    for (int i = 0; i < n; ++i) {
      const double g = dz[co * n + i];
      db[co] += g;
      for (int ci = 0; ci < cin; ++ci)
        for (int d = 0; d < K; ++d) {
          int j = (i + d - 2) % n;
          if (j < 0) j += n;
          dW[(co * cin + ci) * K + d] += g * hin[ci * n + j];
          if (dhin) dhin[ci * n + j] += g * W[(co * cin + ci) * K + d];
        }
    }
}

/* Accumulates into grad (NPARAMS) the gradient of sum_s (ga[s] a_s + gt[s] t_s). */
void cnn_backward(const double *params, int n, int64_t count, const uint64_t *states,
                  const double *ga, const double *gt, double *grad) {
  double *h0 = (double *)malloc(sizeof(double) * n);
  double *z1 = (double *)malloc(sizeof(double) * C * n);
  double *z2 = (double *)malloc(sizeof(double) * C * n);
  double *z3 = (double *)malloc(sizeof(double) * C * n);
  double *h1 = (double *)malloc(sizeof(double) * C * n);
  double *h2 = (double *)malloc(sizeof(double) * C * n);
  double *d3 = (double *)malloc(sizeof(double) * C * n);
  double *d2 = (double *)malloc(sizeof(double) * C * n);
  double *d1 = (double *)malloc(sizeof(double) * C * n);
  for (int64_t s = 0; s < count; ++s) {
    for (int i = 0; i < n; ++i) h0[i] = ((states[s] >> i) & 1) ? 1.0 : -1.0;
    conv(n, 1, params + OFF_W1, params + OFF_B1, h0, z1);
    for (int q = 0; q < C * n; ++q) h1[q] = gelu(z1[q]);
    conv(n, C, params + OFF_W2, params + OFF_B2, h1, z2);
    for (int q = 0; q < C * n; ++q) h2[q] = gelu(z2[q]);
    conv(n, C, params + OFF_W3, params + OFF_B3, h2, z3);
    /* heads */
    for (int c = 0; c < C; ++c) {
      double f = 0.0;
      for (int i = 0; i < n; ++i) f += gelu(z3[c * n + i]);
      grad[OFF_U + c] += ga[s] * f;
      grad[OFF_V + c] += gt[s] * f;
      const double df = ga[s] * params[OFF_U + c] + gt[s] * params[OFF_V + c];
      for (int i = 0; i < n; ++i) d3[c * n + i] = df * gelu_grad(z3[c * n + i]);
    }
    grad[OFF_U0] += ga[s];
    grad[OFF_V0] += gt[s];
    memset(d2, 0, sizeof(double) * C * n);
    conv_backward(n, C, params + OFF_W3, h2, d3, grad + OFF_W3, grad + OFF_B3, d2);
    for (int q = 0; q < C * n; ++q) d2[q] *= gelu_grad(z2[q]);
    memset(d1, 0, sizeof(double) * C * n);
    conv_backward(n, C, params + OFF_W2, h1, d2, grad + OFF_W2, grad + OFF_B2, d1);
    for (int q = 0; q < C * n; ++q) d1[q] *= gelu_grad(z1[q]);
    conv_backward(n, 1, params + OFF_W1, h0, d1, grad + OFF_W1, grad + OFF_B1, NULL);
  }
  free(h0);
  free(z1);
  free(z2);
  free(z3);
  free(h1);
  free(h2);
  free(d3);
  free(d2);
  free(d1);
}
