/*
 * spined.h -- entry points of libspined beyond the ls_* ABI.
 *
 * SpinED's ground-state solve calls PRIMME's eigh with a Haskell callback
 * that crosses into ls_operator_matmat with host vectors on every iteration
 * (src/SpinED.hs:383-404, src/SpinED/Internal.hs:411-429, :457-458).  The
 * device-resident replacement is spined_eigsh: the whole Lanczos solve runs
 * on the GPU and only eigenvalues, residuals and (optionally) eigenvectors
 * come back.  It is the one line SpinED's `diagonalize` would change
 * (SpinED.hs:404).  The rest of this header is the device-pointer matvec
 * used by tests and benchmarks, the batched solver and the learned start.
 */
#ifndef SPINED_H
#define SPINED_H

#include <stddef.h>
#include <stdint.h>

#include "lattice_symmetries.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Storage / arithmetic mode of the eigensolver. */
typedef enum {
  SPINED_PRECISION_FP64 = 0,  /* vectors stored and accumulated in FP64 */
  SPINED_PRECISION_MIXED = 1, /* FP32 vector storage, FP64 row accumulation and
                                 FP64 recurrence / dots / projected solve, then
                                 an FP64 refinement phase to the tolerance */
} spined_precision;

typedef enum {
  SPINED_START_RANDOM = 0,  /* v0[r] = U(-1, 1) from a counter hash of (seed, r) */
  SPINED_START_VECTOR = 1,  /* host vector in start_vector (length = dimension) */
  SPINED_START_LEARNED = 2, /* amplitudes from the periodic CNN in `model` */
} spined_start;

typedef struct spined_monitor_info {
  int64_t matvecs;
  int32_t restarts;
  int32_t phase;            /* 0 = FP64, 1 = FP32-storage phase */
  int32_t nconverged;
  int32_t k;
  const double *ritz_values;   /* k values */
  const double *residuals;     /* k residual estimates */
} spined_monitor_info;

/* Called once per restart cycle (mirrors PrimmeMonitor, SpinED.hs:379-382).
 * Return non-zero to stop the current phase at this cycle: the FP32-storage
 * (or block) phase hands over to the FP64 phase; the FP64 phase returns its
 * current Ritz values with their explicit residuals (LS_NOT_CONVERGED when
 * they miss tol) and makes no further attempt. */
typedef int (*spined_monitor_fn)(const spined_monitor_info *info, void *ctx);

/* Checkpoint (no counterpart in the reference, which creates /_workspace
 * and leaves it empty): after every checkpoint_every-th thick restart the
 * solver passes the lowest Ritz vector of its current basis, on the device
 * and of unit norm (valid during the call only): dtype LS_FLOAT32 in the
 * FP32-storage phase, else LS_FLOAT64 (LS_COMPLEX128 for complex sectors),
 * n elements.  Handing it back as start_vector resumes the solve. */
typedef void (*spined_checkpoint_fn)(const void *ritz_device, int dtype, uint64_t n,
                                     const spined_monitor_info *info, void *ctx);

typedef struct spined_eigsh_options {
  int32_t precision;        /* spined_precision */
  int32_t start;            /* spined_start */
  uint64_t seed;
  int32_t max_basis_size;   /* Lanczos basis m (max_primme_basis_size); 0 = auto:
                               max(2k + 10, 24), lowered (down to 16) when that
                               keeps the operator's gather cache in memory */
  int32_t min_restart_size; /* vectors kept on restart (min_primme_restart_size); 0 = auto */
  int32_t max_block_size;   /* max_primme_block_size: 0/1 = one-vector recurrence;
                               2, 4, 8 = block Lanczos phase on row-interleaved
                               blocks of that width (block matvec), then the
                               one-vector FP64 phase finishes */
  int32_t verbose;          /* 1 = one log line per restart on stderr */
  int64_t max_matvecs;      /* 0 = auto: 20,000 (LS_NOT_CONVERGED beyond) */
  const double *start_vector;
  const void *model;        /* spined_cnn* for SPINED_START_LEARNED */
  uint64_t start_sign_mask; /* SPINED_START_LEARNED: Marshall sublattice mask */
  spined_checkpoint_fn checkpoint; /* NULL = none */
  void *checkpoint_ctx;
  int32_t checkpoint_every; /* restarts between checkpoints; 0 = none */
} spined_eigsh_options;

typedef struct spined_stats {
  int64_t matvecs;          /* total H*v applications */
  int64_t matvecs_mixed;    /* of which in the FP32-storage phase */
  int32_t restarts;
  double seconds_total;
  double seconds_matvec;    /* wall time inside matvec kernels (events) */
  double seconds_start;     /* start-vector construction */
  double seconds_matvec_mixed; /* of seconds_matvec: the FP32-x matvecs (FP32-storage phase) */
} spined_stats;

void spined_eigsh_options_init(spined_eigsh_options *opts);

/* Lowest k eigenpairs of the operator to residual ||H v - lambda v|| <= tol
 * (absolute, unit v).  evecs: host, column-major size x k with leading
 * dimension ld (may be NULL); dtype LS_FLOAT64 or LS_FLOAT32 selects its
 * element type (datatype of SpinED.hs:142-156).  Replaces eigh at
 * SpinED.hs:404. */
int spined_eigsh(const ls_operator *op, int dtype, uint64_t k, double tol,
                 const spined_eigsh_options *opts, double *evals, void *evecs, uint64_t ld,
                 double *rnorms, spined_monitor_fn monitor, void *monitor_ctx);
/* Statistics of the last spined_eigsh on this operator. */
int spined_get_last_stats(const ls_operator *op, spined_stats *out);

/* ---- device-resident helpers ------------------------------------------- */
/* The library's HIP stream for the current device (hipStream_t). */
void *spined_stream(void);
/* Waits for the library stream. */
int spined_synchronize(void);
/* Waits for the whole device. */
int spined_device_synchronize(void);
/* Device memory owned by the caller (hipMalloc/hipFree of this library's
 * runtime), and copies / fills on the library stream (synchronous). */
int spined_device_malloc(void **out, uint64_t bytes);
int spined_device_free(void *ptr);
/* The eigensolver keeps its vector blocks (those of 256 MiB or more) after a
 * solve and reuses them for the next solve or phase of the same size on the
 * device, and so does a dropped gather cache for its next build; this returns
 * them to the driver (hipFree).  The library also does
 * so by itself before any allocation the driver would refuse.  Call it
 * before allocating large device memory outside this library. */
int spined_device_release_pool(void);
/* kind: 0 host->device, 1 device->host, 2 device->device */
int spined_memcpy(void *dst, const void *src, uint64_t bytes, int kind);
int spined_memset(void *dst, int value, uint64_t bytes);
/* Device pointer to the basis' representatives (ascending) and stabiliser
 * sizes; valid while the basis lives and is built. */
int spined_basis_device_states(const ls_spin_basis *basis, const uint64_t **reps,
                               const uint32_t **stabilizers, uint64_t *size);
/* Host copy of the stabiliser sizes |Stab r| (norm N_r = sqrt(|Stab r| / |G|)). */
int spined_basis_get_stabilizers(const ls_spin_basis *basis, uint32_t *out);
/* y = H x with DEVICE pointers, enqueued on spined_stream(); dtype as in
 * ls_operator_matmat plus mixed mode: x float32 / y float64 when
 * dtype = LS_FLOAT32 and y_is_fp64 = 1. */
int spined_operator_matmat_device(const ls_operator *op, int dtype, uint64_t size,
                                  uint64_t block_size, const void *x, uint64_t x_stride,
                                  void *y, uint64_t y_stride, int y_is_fp64);
/* <x_c|H|x_c> for the block_size columns of x (DEVICE, column-major with
 * leading dimension x_stride), not divided by <x_c|x_c>: the device form of
 * ls_operator_expectation (src/SpinED/Internal.hs:437-452, used by
 * src/SpinED.hs:296-313).  y_scratch: one DEVICE vector of `size` elements
 * of dtype.  out: HOST, 2 * block_size doubles (re, im per column). */
int spined_operator_expectation_device(const ls_operator *op, int dtype, uint64_t size,
                                       uint64_t block_size, const void *x, uint64_t x_stride,
                                       void *y_scratch, double *out);
/* The same as spined_operator_expectation_device, but allocating no device
 * memory: no counterpart in the reference; for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The dot's partial sums live in a fixed device array of
 * 2 * 2048 + 2 * 64 doubles (33 KB) that each GPU holds until the process
 * ends and every call on that GPU shares: calls that overlap on one GPU race.
 * At most 64 columns per call (more: LS_INVALID_ARGUMENT).  The operator must
 * have been applied once on this device (its tables, and for a complex
 * operator its phase table, must be there already; else LS_INVALID_ARGUMENT);
 * an existing gather cache is used, none is built. */
int spined_operator_expectation_device_noalloc(const ls_operator *op, int dtype, uint64_t size,
                                               uint64_t block_size, const void *x,
                                               uint64_t x_stride, void *y_scratch, double *out);
/* Y = H X for a row-interleaved block of block_size in {1, 2, 4, 8} real
 * vectors (element (row, c) at row * block_size + c; DEVICE pointers): one
 * orbit scan and index search per term serve the whole block, and every
 * gather reads one row of the block.  Per column the sums are those of
 * spined_operator_matmat_device.  ls_operator_matmat and
 * spined_operator_matmat_device use it for real blocks (chunks of <= 8
 * columns), the block form of ls_operator_matmat's block_size
 * (src/SpinED/Internal.hs:411-429). */
int spined_operator_matmat_rows_device(const ls_operator *op, int dtype, uint64_t size,
                                       uint64_t block_size, const void *x, void *y,
                                       int y_is_fp64);
/* Number of off-diagonal terms (sigma states) H generates from the basis:
 * the per-matvec unit count behind the algorithmic byte figure. */
int spined_operator_count_offdiag(const ls_operator *op, uint64_t *out);
/* Orbit method used by the matvec: 0 = generic per-element programs,
 * 1 = ring (rotation/reflection run pruning). */
int spined_operator_orbit_method(const ls_operator *op);
/* Force an orbit method (for parity tests): -1 = automatic. */
int spined_operator_set_orbit_method(ls_operator *op, int method);
/* Row schedule of the real matvec kernels: within tiles of 4096 rows the
 * rows are processed in the order of their number of active off-diagonal
 * terms, so the 64 lanes of a wave carry equal work (the sums per row are
 * unchanged).  mode: -1 automatic (on for bases of >= 4096 rows), 0 off,
 * 1 on; built at the next matvec.  get: 1 if a schedule is in use. */
int spined_operator_set_schedule(ls_operator *op, int mode);
int spined_operator_get_schedule(const ls_operator *op);
/* Rows per wave of the matrix-free real matvec of exchange-bond operators:
 * 64 (one lane per row), 32 or 16 (the row's templates spread over 2 or 4
 * lanes, its terms summed in the same order: the sums are unchanged); -1 =
 * automatic (64).  The autotuner times all three.  Other operators and the
 * gather-cache matvec keep one lane per row. */
int spined_operator_set_rows_per_wave(ls_operator *op, int rows);
int spined_operator_get_rows_per_wave(const ls_operator *op);
/* Gather cache of a real operator: the target of every term of every row
 * (4 bytes per term, padded per wave: ~74 GB at D = 8.6e8), written once by the orbit and
 * index kernels so that later real single-vector matvecs read it instead of
 * recomputing orbits and searching the index; the sums are identical.
 * mode: -1 automatic (built at a real matvec of a basis of at least 2^20
 * rows when free device memory exceeds it by 4 GiB, and given back whenever
 * a device allocation of the library needs the memory), 0 off, 1 on (built
 * at the next real matvec; an error if it does not fit).  get: 1 while a
 * cache is built.  Changing the row schedule drops the cache. */
int spined_operator_set_gather_cache(ls_operator *op, int mode);
int spined_operator_get_gather_cache(const ls_operator *op);
/* Bucket bits of the bucket index (about 2^bits level-2 buckets of 4 bytes,
 * 1..36; -1 = automatic: ceil(log2 D) + 3).  The build lowers it until the
 * table takes at most half the free device memory.  Setting it on a built
 * basis rebuilds the index. */
int spined_basis_set_bucket_bits(ls_spin_basis *basis, int bits);
int spined_basis_get_bucket_bits(const ls_spin_basis *basis);
/* Index of representatives used by every lookup: 0 = bucket table on the
 * top bits + binary search; 1 = open-addressing hash table (2^H slots of
 * 8 bytes, H = ceil(log2 D) + 1, one slot line + one confirming load per
 * lookup); 2 = combinatorial rank, for complete bases whose symmetry group
 * has no permutation but the identity (with or without spin inversion):
 * index = sum of C(p_i, i + 1) over the up sites p_0 < p_1 < ..., or the
 * state itself without a Hamming weight, no table beyond w x 64 binomials;
 * -1 = automatic (2 where it applies, else 0).  Setting it on a built basis
 * rebuilds the index; 2 on a basis it does not apply to is an error. */
int spined_basis_set_index_kind(ls_spin_basis *basis, int kind);
int spined_basis_get_index_kind(const ls_spin_basis *basis);
/* 1 when the built basis holds every valid orbit minimum of its sector (an
 * ls_build basis, or an ls_build_unsafe list of valid orbit minima as long as
 * the sector dimension), 0 otherwise, -1 before a build.  On a complete
 * basis with an operator that keeps the Hamming weight the bucket search
 * skips its confirming load of the representative: every valid image is
 * known to be present. */
int spined_basis_is_complete(const ls_spin_basis *basis);
/* Autotuner: times the operator's real matvec on its own basis for every
 * candidate setting and keeps the fastest.  Candidates: bucket-table sizes
 * (bits ceil(log2 D) + 1 .. + 4) with the row schedule, the default size
 * without it, the hash table when the bucket search must confirm its
 * hits (spined_basis_is_complete), and the rank index where it applies;
 * then, on the chosen index, 64, 32 and 16 rows per wave for exchange-bond
 * operators (spined_operator_set_rows_per_wave); then the per-vector time
 * of the block matvec for K = 1, 2, 4, 8.  Index tables larger than a quarter of the
 * free device memory are skipped, leaving room for the eigensolver's
 * vectors.  Each timing is the mean of `repeats` launches after one warm-up.
 * The chosen setting stays on the basis / operator. */
typedef struct spined_autotune_report {
  int32_t index_kind;         /* chosen: 0 bucket table, 1 hash, 2 rank */
  int32_t bucket_bits;        /* chosen (bucket table) */
  int32_t schedule;           /* chosen: 1 row schedule on */
  int32_t candidates;         /* settings timed */
  double ms_per_matvec;       /* chosen setting, FP64 x */
  double ms_per_matvec_first; /* the setting in force before tuning */
  double ms_per_vector_block[4]; /* block matvec, K = 1, 2, 4, 8 (0: not timed) */
  int32_t best_block;         /* K with the lowest per-vector time */
  int32_t rows_per_wave;      /* chosen: 64, 32 or 16 (exchange-bond operators) */
  double ms_per_rows_per_wave[3]; /* FP64 matvec at 64, 32, 16 rows per wave (0: not timed) */
} spined_autotune_report;
int spined_operator_autotune(ls_operator *op, int32_t repeats, spined_autotune_report *report);
/* Exact sector dimension by character counting (before ls_build). */
int spined_basis_sector_dimension(const ls_spin_basis *basis, double *out);

/* ---- batched solver (no counterpart in the reference) ------------------ */
/* Lowest eigenpair of `count` independent real exchange-bond operators
 * (instances may have different sizes and may share a basis), solved in ONE
 * persistent launch: the instances, largest first, are split into team
 * slots (one workgroup each, team size growing with the dimension) that the
 * launch's workgroups take from a device counter; each team runs its
 * instance's thick-restart Lanczos (FP64) with per-instance barriers, and no
 * host round trip happens until every instance is done.  Replaces one SpinED
 * run per Hamiltonian. */
typedef struct spined_batch_stats {
  int32_t cycles;                   /* restart cycles of the slowest instance */
  int32_t converged;                /* instances with residual <= tol */
  double seconds_total;             /* setup + solve + read-back */
  double seconds_cycles;            /* the persistent launch (HIP events) */
  uint64_t total_rows;              /* sum of the instances' dimensions */
  int64_t matvecs_per_instance_max; /* matvecs of the slowest instance */
} spined_batch_stats;
/* evals, rnorms: [count]; cycles[i]: restart cycles instance i ran (-1: not
 * converged); seconds[i]: time from the launch to the end of instance i;
 * matvecs[i]: its matvecs (cycles, seconds, matvecs may be NULL).
 * max_basis_size <= 24 (0: 16). */
int spined_batch_eigsh(const ls_operator *const ops[], uint64_t count, double tol,
                       int32_t max_basis_size, uint64_t seed, int32_t max_cycles, double *evals,
                       double *rnorms, int32_t *cycles, double *seconds, int64_t *matvecs,
                       spined_batch_stats *stats);

/* ---- learned start vector (no counterpart in the reference) ------------ */
/* Periodic 1D CNN over +-1 spins: 3 circular conv layers (16 channels,
 * kernel 5, tanh-GELU), global sum, heads log|psi| and sign logit; 2,722
 * FP32 master parameters, BF16 MFMA kernels with FP32 accumulation. */
typedef struct spined_cnn spined_cnn;
int spined_cnn_num_params(void);
int spined_cnn_create(spined_cnn **out, const float *init_params);
void spined_cnn_destroy(spined_cnn *model);
int spined_cnn_get_params(const spined_cnn *model, float *out);
int spined_cnn_set_params(spined_cnn *model, const float *in);
/* Device buffers: states[count]; out_a/out_t[count]; z_store (may be NULL)
 * holds 3 * ceil(count/16)*16 * n * 16 floats of pre-activations. */
int spined_cnn_forward_device(const spined_cnn *model, int n, uint64_t count,
                              const uint64_t *states, float *out_a, float *out_t, float *z_store);
/* Gradient of sum_s ga[s] a_s + gt[s] t_s (device buffers) into grad_out
 * (device, num_params floats) or, if NULL, into the model for adam_step. */
int spined_cnn_backward_device(spined_cnn *model, int n, uint64_t count, const uint64_t *states,
                               const float *z_store, const float *ga, const float *gt,
                               float *grad_out);
int spined_cnn_adam_step(spined_cnn *model, float lr);
/* Start vector over the basis (device FP64, length = dimension):
 * c_r = (-1)^popcount(r & marshall_mask) sign(t_r) exp(a_r - max a) / N_r. */
int spined_cnn_start_vector(const spined_cnn *model, const ls_spin_basis *basis,
                            uint64_t marshall_mask, double *v_device);

/* ---- test hooks (no counterpart in the reference) ---------------------- */
/* One v_mfma_f32_16x16x32_bf16 on known operands (lane-layout check):
 * out[16 x 16] row-major = A B with A[i][k] = (3i + k) % 7 - 3 and
 * B[k][j] = (3k + 7j) % 5 - 2. */
int spined_debug_mfma_probe(float *out_host);
/* Runs the device orbit code (the same inline functions the kernels use) on
 * the host for `count` states: method 0 = generic (the torus factorisation
 * when the group has one), 1 = ring, 2 = generic with one program per
 * element (the factorisation off), 3 = the generic orbit minimum alone
 * (reps only; phases, counts and valid are left untouched).  For
 * CPU-side verification against the oracle; no compute path uses it. */
int spined_debug_orbit(const ls_spin_basis *basis, int method, uint64_t count,
                       const uint64_t *states, uint64_t *reps, int32_t *phases, int32_t *counts,
                       uint8_t *valid);
/* Ring orbit of states[q] with its spins at sites[q] and sites[q] + 1
 * (mod n) exchanged (they must differ), computed from the run data of
 * states[q] -- the path of the nearest-neighbour row gather -- on the host.
 * For CPU-side verification against the full orbit scan. */
int spined_debug_orbit_hop(const ls_spin_basis *basis, uint64_t count, const uint64_t *states,
                           const int32_t *sites, uint64_t *reps, int32_t *phases, int32_t *counts,
                           uint8_t *valid);
/* 1 if the basis' group is a ring (rotation/reflection) group. */
int spined_debug_ring_available(const ls_spin_basis *basis);
/* Shape l1 x l2 of the torus factorisation that the generic orbit uses (the
 * group holds every translation of a row-major l1 x l2 torus, site = l1 y +
 * x, and each element is a translation after one of a few point elements);
 * 0 x 0 when the orbit runs one program per element. */
int spined_debug_torus_shape(const ls_spin_basis *basis, int32_t *l1, int32_t *l2);
/* Plain device copy of `bytes` (multiple of 16): the calibration kernel for
 * HBM byte counters in profiles. */
int spined_debug_copy_device(void *dst, const void *src, uint64_t bytes);
/* Random-gather calibration for the same counters: reads `words` (1..32)
 * 4-byte words of each of the n_lines (a power of two) 128-byte lines of
 * src, every line once, in a scattered order; dst holds dst_words >=
 * spined_debug_gather_out_words() words of per-wave sums. */
int spined_debug_gather_device(void *dst, uint64_t dst_words, const void *src, uint64_t n_lines,
                               int32_t words);
uint64_t spined_debug_gather_out_words(void);

#ifdef __cplusplus
}
#endif
#endif
