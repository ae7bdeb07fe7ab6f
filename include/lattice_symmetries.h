/*
 * lattice_symmetries.h -- the ls_* C ABI exported by libspined.
 *
 * libspined is a drop-in replacement for liblattice_symmetries on SpinED's
 * ground-state path: these are exactly the 32 symbols SpinED imports with
 * `foreign import ccall` in src/SpinED/Internal.hs, with C types taken from
 * the Haskell FFI types there.  The link contract is configure:54-66
 * (extra-libraries: lattice_symmetries gomp stdc++).
 *
 * Conventions (shared with the CPU oracle):
 *   - every fallible call returns int, 0 = success; on failure an out-pointer
 *     is left untouched (Internal.hs:58-65, :358-360, :398-400);
 *   - every object comes back through an out-pointer and is released by the
 *     matching ls_destroy_*, which may run on any thread, in any order
 *     (ForeignPtr finalisers, Internal.hs:116, :150, :228, :362, :402); objects
 *     keep what they depend on alive (basis -> group, operator -> basis and
 *     interactions), and device memory is reference counted;
 *   - bit i of a uint64 state is site i; (g.x)_i = x_{perm[i]}; a generator
 *     of periodicity n in sector k has character exp(2 pi i k / n);
 *     representatives are orbit minima; the local index of an n-site
 *     interaction has the first listed site as its most significant bit;
 *   - x, y and out of ls_operator_matmat / ls_operator_expectation are HOST
 *     pointers (Haskell storable vectors), column-major blocks with the given
 *     leading dimensions; the device-resident path is spined_eigsh (spined.h).
 */
#ifndef LATTICE_SYMMETRIES_H
#define LATTICE_SYMMETRIES_H

#include <stdbool.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ls_symmetry ls_symmetry;
typedef struct ls_group ls_group;
typedef struct ls_spin_basis ls_spin_basis;
typedef struct ls_states ls_states;
typedef struct ls_interaction ls_interaction;
typedef struct ls_operator ls_operator;

/* Status codes (ls_error_to_string gives the text). */
typedef enum {
  LS_SUCCESS = 0,
  LS_INVALID_PERMUTATION = 1,
  LS_INVALID_SECTOR = 2,
  LS_INCOMPATIBLE_SYMMETRIES = 3,
  LS_INVALID_NUMBER_SPINS = 4,
  LS_INVALID_HAMMING_WEIGHT = 5,
  LS_INVALID_SPIN_INVERSION = 6,
  LS_OUT_OF_MEMORY = 7,
  LS_CACHE_NOT_BUILT = 8,
  LS_INVALID_ARGUMENT = 9,
  LS_NOT_CONVERGED = 10,
  LS_INVALID_DATATYPE = 11,
  LS_DIMENSION_MISMATCH = 12,
  LS_DEVICE_ERROR = 13,
  LS_NOT_IMPLEMENTED = 14,
  LS_INTERNAL_ERROR = 15,
} ls_error_code;

/* Element types of ls_operator_matmat / ls_operator_expectation
 * (toCdatatype, Internal.hs:404-409). */
typedef enum {
  LS_FLOAT32 = 0,
  LS_FLOAT64 = 1,
  LS_COMPLEX64 = 2,
  LS_COMPLEX128 = 3,
} ls_datatype;

/* ---- errors & logging (Internal.hs:39-45) ------------------------------ */
/* Returns a string owned by the caller, released with ls_destroy_string
 * (getErrorMessage, Internal.hs:48-54). */
const char *ls_error_to_string(int status);
void ls_destroy_string(const char *message);
void ls_enable_logging(void);
void ls_disable_logging(void);

/* ---- symmetry (Internal.hs:67-116) ------------------------------------- */
/* perm must be a permutation of 0..length-1 (length <= 64) and
 * sector < periodicity(perm); otherwise LS_INVALID_PERMUTATION /
 * LS_INVALID_SECTOR (test/Spec.hs:38-43). */
int ls_create_symmetry(ls_symmetry **out, unsigned length, const unsigned permutation[],
                       unsigned sector);
void ls_destroy_symmetry(ls_symmetry *symmetry);
unsigned ls_get_sector(const ls_symmetry *symmetry);
/* sector / periodicity, in [0, 1): the character is exp(2 pi i phase). */
double ls_get_phase(const ls_symmetry *symmetry);
unsigned ls_get_periodicity(const ls_symmetry *symmetry);

/* ---- group (Internal.hs:118-155) --------------------------------------- */
/* Closure of the generators; LS_INCOMPATIBLE_SYMMETRIES when the sectors do
 * not define a one-dimensional representation (test/Spec.hs:72-79).  The
 * group copies its generators. */
int ls_create_group(ls_group **out, unsigned size, const ls_symmetry *generators[]);
void ls_destroy_group(ls_group *group);
/* Number of permutations (spin inversion not included). */
unsigned ls_get_group_size(const ls_group *group);

/* ---- basis (Internal.hs:157-254) --------------------------------------- */
/* hamming_weight: -1 = no restriction; spin_inversion: 0 = none, +1 / -1. */
int ls_create_spin_basis(ls_spin_basis **out, const ls_group *group, unsigned number_spins,
                         int hamming_weight, int spin_inversion);
void ls_destroy_spin_basis(ls_spin_basis *basis);
/* Builds the ascending list of representatives on the GPU (safe FFI import,
 * Internal.hs:178-179). */
int ls_build(ls_spin_basis *basis);
/* Adopts representatives (e.g. reloaded from /basis/representatives) without
 * validating them (Internal.hs:181-182, SpinED.hs:319-331). */
int ls_build_unsafe(ls_spin_basis *basis, uint64_t size, const uint64_t representatives[]);
int ls_get_number_states(const ls_spin_basis *basis, uint64_t *out);
/* Host copy of the representatives; valid until ls_destroy_states, which
 * may run on another thread (basisGetStates, Internal.hs:238-244). */
int ls_get_states(ls_states **out, const ls_spin_basis *basis);
const uint64_t *ls_states_get_data(const ls_states *states);
uint64_t ls_states_get_size(const ls_states *states);
void ls_destroy_states(ls_states *states);

/* ---- interactions (Internal.hs:256-367) -------------------------------- */
/* matrix: complex double [2^n][2^n] row-major, interleaved (re, im);
 * sites: uint16 [number_sites][n] (toMatrix / unsafeMkInteraction,
 * Internal.hs:278-362). */
int ls_create_interaction1(ls_interaction **out, const void *matrix, unsigned number_sites,
                           const uint16_t *sites);
int ls_create_interaction2(ls_interaction **out, const void *matrix, unsigned number_sites,
                           const uint16_t *sites);
int ls_create_interaction3(ls_interaction **out, const void *matrix, unsigned number_sites,
                           const uint16_t *sites);
int ls_create_interaction4(ls_interaction **out, const void *matrix, unsigned number_sites,
                           const uint16_t *sites);
bool ls_interaction_is_real(const ls_interaction *interaction);
void ls_destroy_interaction(ls_interaction *interaction);

/* ---- operator (Internal.hs:369-458) ------------------------------------ */
/* The operator may be created before ls_build (SpinED.hs:243-248 then
 * :315-337); it binds the basis' device state lazily. */
int ls_create_operator(ls_operator **out, const ls_spin_basis *basis, unsigned number_terms,
                       const ls_interaction *const terms[]);
void ls_destroy_operator(ls_operator *op);
/* y = H x for a column-major block (inplaceApply, Internal.hs:411-429). */
int ls_operator_matmat(const ls_operator *op, int dtype, uint64_t size, uint64_t block_size,
                       const void *x, uint64_t x_stride, void *y, uint64_t y_stride);
/* out[j] = <x_j|H|x_j> (complex double per column; not divided by <x_j|x_j>)
 * (expectation, Internal.hs:437-452).  Each column is copied to the device
 * once; H x_j and the FP64 dot are computed there and only out returns. */
int ls_operator_expectation(const ls_operator *op, int dtype, uint64_t size, uint64_t block_size,
                            const void *x, uint64_t x_stride, void *out);
bool ls_operator_is_real(const ls_operator *op);

#ifdef __cplusplus
}
#endif
#endif
