/*
 * spined_oracle.h -- CPU oracle for the exact-diagonalization hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the GPU engine is
 * compared against; only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load it.  The engine (libspined) never links or calls it.
 *
 * It restates, in plain C (FP64, optional OpenMP), what the SpinED reference
 * obtains from lattice-symmetries and PRIMME through the FFI in
 * src/SpinED/Internal.hs:
 *   - symmetry validation and periodicity   (ls_create_symmetry, Internal.hs:69-116)
 *   - group closure with exact phases        (ls_create_group,    Internal.hs:120-150)
 *   - sector dimension by character counting (pin for ls_build)
 *   - basis build: orbit minimum + stabilizer test, ascending reps
 *                                            (ls_build, Internal.hs:178-182, :230-236)
 *   - matrix-free H*x row gather             (ls_operator_matmat, Internal.hs:377-378, :411-429)
 *   - <x|H|x> per column                     (ls_operator_expectation, Internal.hs:380-381, :437-452)
 *   - FP64 Lanczos with full reorthogonalisation and a tridiagonal
 *     implicit-QL eigensolver                (role of PRIMME eigh, SpinED.hs:383-404)
 *
 * Conventions (the reference leaves them to lattice-symmetries, so they are
 * fixed here and shared with the engine):
 *   - bit i of a uint64 state is site i; hamming weight = popcount.
 *   - a permutation p acts on states by (g.x)_i = x_{p[i]}.
 *   - a generator with periodicity n and sector k has character
 *     exp(2 pi i k / n); spin inversion F has character +1 / -1.
 *     Phases are kept exactly as integers modulo L = lcm(2, all periodicities).
 *   - the representative of x is min_g g.x (integer order).
 *   - local index of an n-site interaction: the first listed site is the most
 *     significant bit (kron order).
 *
 * "parity unpinned" items are listed at the end of the oracle's C file.
 */
#ifndef SPINED_ORACLE_H
#define SPINED_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_system orc_system;

/* Status codes shared with the engine's error table. */
enum {
  ORC_OK = 0,
  ORC_INVALID_PERMUTATION = 1,
  ORC_INVALID_SECTOR = 2,
  ORC_INCOMPATIBLE_SYMMETRIES = 3,
  ORC_INVALID_NUMBER_SPINS = 4,
  ORC_INVALID_HAMMING_WEIGHT = 5,
  ORC_INVALID_SPIN_INVERSION = 6,
  ORC_OUT_OF_MEMORY = 7,
  ORC_NOT_BUILT = 8,
  ORC_INVALID_ARGUMENT = 9,
  ORC_NOT_CONVERGED = 10,
};

/* Periodicity of a permutation (0 if it is not a permutation of 0..n-1). */
unsigned orc_periodicity(unsigned n, const unsigned *perm);

/* Build the system: group closure over the generators (+ spin inversion).
 * perms: n_gens x n_sites, sectors: n_gens.  Returns NULL and *status != 0 on
 * error. */
orc_system *orc_create(int n_sites, int hamming_weight, int spin_inversion,
                       int n_gens, const unsigned *perms, const unsigned *sectors,
                       int *status);
void orc_destroy(orc_system *s);

int orc_group_size(const orc_system *s);      /* permutations only (ls_get_group_size) */
int orc_full_group_size(const orc_system *s); /* including spin inversion */
int orc_phase_denominator(const orc_system *s);
/* Copies the full group: perms (G x n), flips (G), phase numerators (G). */
void orc_get_elements(const orc_system *s, unsigned *perms, int *flips, int *phases);
/* g.x for element index g. */
uint64_t orc_apply(const orc_system *s, int g, uint64_t x);
/* Orbit scan: rep = min_g g.x, phase of the first element reaching it,
 * count = #{g : g.x = rep}, valid = all those elements share the phase. */
void orc_orbit(const orc_system *s, uint64_t x, uint64_t *rep, int *phase,
               int *count, int *valid);

/* Sector dimension by character counting: (1/|G|) sum_g chi(g)* Fix(g).
 * Returned as a double holding an exact integer for |D| < 2^53; *imag gets
 * the imaginary residue (0 for consistent input). */
double orc_sector_dimension(const orc_system *s, double *imag);

/* Basis build (ascending representatives with their stabilizer sizes). */
int64_t orc_build(orc_system *s);
int64_t orc_build_range(orc_system *s, uint64_t first, uint64_t last); /* enumeration restricted to [first,last] */
int orc_set_basis(orc_system *s, int64_t n, const uint64_t *reps); /* like ls_build_unsafe */
/* Adopt representatives together with their stabiliser sizes (no orbit
 * recomputation; used to time the CPU matvec on a large basis). */
int orc_set_basis_with_stab(orc_system *s, int64_t n, const uint64_t *reps, const uint32_t *stab);
int64_t orc_dimension(const orc_system *s);
void orc_get_basis(const orc_system *s, uint64_t *reps, uint32_t *stab);
int64_t orc_index(const orc_system *s, uint64_t rep); /* -1 if absent */

/* Operator: add interaction terms.  matrix: 2^n x 2^n complex, row-major,
 * interleaved (re, im).  sites: n_tuples x n. */
int orc_add_interaction(orc_system *s, int n, const double *matrix, int n_tuples,
                        const unsigned *sites);
int orc_operator_is_real(const orc_system *s);
int orc_number_offdiag_templates(const orc_system *s);

/* y = H x.  Real variant (operator and sector must be real).  Rows in
 * [row_begin, row_end) only (pass 0, -1 for all). */
int orc_matvec_real(const orc_system *s, const double *x, double *y,
                    int64_t row_begin, int64_t row_end);
/* Complex variant; x, y interleaved (re, im). */
int orc_matvec_complex(const orc_system *s, const double *x, double *y,
                       int64_t row_begin, int64_t row_end);
/* <x|H|x> for one real column (not normalised). */
int orc_expectation_real(const orc_system *s, const double *x, double *out);
/* Number of off-diagonal (sigma) terms generated by rows [b, e). */
int64_t orc_count_offdiag(const orc_system *s, int64_t row_begin, int64_t row_end);

/* Counter-hash start vector: v[r] = U(-1, 1) from (seed, r). */
void orc_random_vector(uint64_t seed, int64_t n, double *v);

/* Symmetric tridiagonal eigensolver (implicit QL, Wilkinson shift).
 * alpha: n diagonal, beta: n-1 off-diagonal.  evals ascending; evecs
 * (n x n, column j = eigenvector j, column-major) may be NULL. */
int orc_tridiag_eigh(int n, const double *alpha, const double *beta,
                     double *evals, double *evecs);

/* Lanczos with full reorthogonalisation for the k lowest eigenpairs of the
 * real operator.  evecs: D x k column-major (may be NULL).  Uses start vector
 * v0 if non-NULL, else orc_random_vector(seed).  alpha/beta (may be NULL,
 * length max_iter) receive the recurrence coefficients. */
int orc_lanczos(const orc_system *s, int k, int max_iter, double tol, uint64_t seed,
                const double *v0, double *evals, double *evecs, double *rnorms,
                int *iterations, double *alpha, double *beta);

#ifdef __cplusplus
}
#endif
#endif
