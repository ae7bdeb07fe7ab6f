// engine.hpp -- host-side objects behind the ls_* handles and the kernel
// launchers they call.  Ownership follows the reference's ForeignPtr model
// (src/SpinED/Internal.hs:116, :150, :228, :362, :402): every handle holds a
// shared_ptr, dependants keep their dependencies alive, and the last release
// (from any thread) frees host and device memory.
#pragma once

#include <array>
#include <complex>
#include <cstdint>
#include <memory>
#include <atomic>
#include <mutex>
#include <vector>

#include "../../include/spined.h"
#include "common.hpp"
#include "orbit.hpp"

namespace spined {

using cplx = std::complex<double>;

struct Symmetry {
  std::vector<unsigned> perm;
  unsigned sector = 0;
  unsigned periodicity = 1;
};

// Closure of the generators with exact phases (integers modulo L).
struct Group {
  int n = 0; // number of sites; 0 when built from no generators
  std::vector<Symmetry> generators;
  std::vector<std::vector<uint8_t>> perms; // perms[0] is the identity
  std::vector<int64_t> phases;
  int64_t L = 2;
};

unsigned periodicity_of(const std::vector<unsigned> &perm); // 0 if not a permutation
std::shared_ptr<Group> close_group(const std::vector<Symmetry> &gens);

// Device tables describing the full symmetry group (permutations x optional
// spin inversion) of a basis.
struct OrbitTables {
  int n = 0;
  int nperm = 0;
  int has_flip = 0;
  int32_t flip_phase = 0;
  int32_t L = 2;
  bool trivial_chars = true; // every phase is 0 (incl. the flip)
  bool real_chars = true;    // every phase is 0 or L/2
  bool ring_available = false;
  DeviceBuffer<uint64_t> masks;
  DeviceBuffer<int32_t> shifts;
  DeviceBuffer<uint32_t> offsets;
  DeviceBuffer<uint32_t> flags;
  DeviceBuffer<int32_t> phases;
  DeviceBuffer<int32_t> ring_phases;
  // host mirrors (used by spined_debug_orbit)
  std::vector<uint64_t> h_masks;
  std::vector<int32_t> h_shifts;
  std::vector<uint32_t> h_offsets, h_flags;
  std::vector<int32_t> h_phases, h_ring_phases;
  int ring_has_rev = 0;
  uint32_t total_terms = 0;
  // torus factorisation of the generic orbit (GenericOrbitArgs); l1 = 0: off
  int torus_l1 = 0, torus_l2 = 0;
  std::vector<uint32_t> h_point_elem;
  std::vector<int32_t> h_tp_phase;
  std::vector<uint64_t> h_row_lo, h_row_hi;
  DeviceBuffer<uint32_t> point_elem;
  DeviceBuffer<int32_t> tp_phase;
  DeviceBuffer<uint64_t> row_lo, row_hi;
  // sqrt(c / s) for stabiliser sizes c, s (divisors of |G|): ratio[i * nd + j]
  // with i = div_idx[c], j = div_idx[s]; nd = 0 when |G| is too large.
  int nd = 0;
  std::vector<uint16_t> h_div_idx;
  std::vector<double> h_ratio;
  DeviceBuffer<uint16_t> div_idx;
  DeviceBuffer<double> ratio;
  bool on_device = false;

  GenericOrbitArgs generic_args(bool device) const;
  RingOrbitArgs ring_args(bool device) const;
};

// Host tables at object creation (no GPU needed); uploaded on first use.
void build_orbit_tables(const Group &g, int n, int spin_inversion, OrbitTables &t);
void ensure_device(OrbitTables &t);

// Largest level-2 size of the bucket index: about 2^36 buckets (256 GiB).
constexpr int kMaxBucketBits = 36;

// Host copy of the representatives (the buffer ls_get_states lends):
// allocated without initialisation and filled by one copy.
struct HostStates {
  std::unique_ptr<uint64_t[]> data;
  uint64_t size = 0;
  explicit HostStates(uint64_t n) : data(new uint64_t[n > 0 ? n : 1]), size(n) {}
};

struct Basis {
  std::shared_ptr<const Group> group;
  int n = 0, w = -1, z = 0;
  int device = 0;
  OrbitTables tables;
  std::mutex mu;
  bool built = false;
  uint64_t D = 0;
  DeviceBuffer<uint64_t> reps;
  DeviceBuffer<uint32_t> stab;
  // two-level bucket index (see IndexArgs)
  DeviceBuffer<uint32_t> buckets;
  DeviceBuffer<uint2> dir;
  int bucket_bits_request = -1; // -1: automatic
  int bucket_bits = 0;          // about 2^bits level-2 buckets in total
  int shift = 0;
  uint64_t nseg = 0;
  uint64_t nbuckets = 0;
  // index form: 0 = bucket table + binary search, 1 = hash table,
  // 2 = combinatorial rank (rank_index_mode)
  int index_kind_request = -1; // -1: automatic
  int index_kind = 0;
  DeviceBuffer<uint64_t> slots;
  uint64_t hmask = 0;
  int rank_mode = 0;                  // IndexArgs::rank_mode
  DeviceBuffer<uint32_t> rank_tab;    // w x 64 binomials (IndexArgs)
  std::shared_ptr<HostStates> host_reps; // lazily downloaded
  double seconds_build = 0.0;

  // reps holds every valid orbit minimum of the sector (set by the
  // enumerating build, or by an adopted list that passed the check)
  bool complete = false;
  // segment starts of the last bucket-table build (nseg + 1 entries)
  std::vector<uint32_t> h_segstart;

  // images_keep_weight: the operator maps weight-w states to weight-w states
  IndexArgs index_args(bool images_keep_weight) const {
    const bool exact = index_kind == 0 && complete && (w < 0 || images_keep_weight);
    const bool ranked = index_kind == 2;
    return IndexArgs{reps.ptr,
                     buckets.ptr,
                     dir.ptr,
                     nseg,
                     shift >= 64 ? ~0ull : (1ull << shift) - 1,
                     shift,
                     index_kind == 1 ? slots.ptr : nullptr,
                     hmask,
                     exact ? 1 : 0,
                     ranked ? rank_mode : 0,
                     w,
                     ranked && rank_mode == 1 && !images_keep_weight ? 1 : 0,
                     ranked ? rank_tab.ptr : nullptr,
                     D};
  }
};

double sector_dimension(const Basis &b, double *imag);
void build_basis(Basis &b);                                  // basis.hip
void adopt_basis(Basis &b, uint64_t n, const uint64_t *reps); // basis.hip (ls_build_unsafe)
void validate_representatives(const Basis &b, uint64_t n, const uint64_t *reps);
void build_index(Basis &b);                                   // basis.hip
int rank_index_mode(const Basis &b);                          // basis.hip
// level-2 bucket count for about 2^bits buckets (fills dir when given)
uint64_t level2_size(const Basis &b, int bits, uint2 *dir);   // basis.hip

struct Interaction {
  int n = 0;
  std::vector<cplx> matrix; // 2^n x 2^n row-major
  std::vector<std::array<uint16_t, 4>> sites;
  bool is_real = true;
};

// Compiled operator tables.  Off-diagonal templates are ordered by (tuple,
// local state a, target state b); the row gather visits the active ones in
// that order, which is also the oracle's summation order.
struct OperatorTables {
  int ntuples = 0;
  int ntemplates = 0;
  int max_local = 0; // max 2^n over tuples
  bool is_real = true;
  bool keeps_weight = true; // every template preserves the number of up spins
  // per tuple: packed sites (4 x 8 bits), n, offset into diag values and
  // into the (a -> template range) table
  std::vector<uint32_t> h_tuple_sites, h_tuple_n, h_tuple_off;
  std::vector<double> h_diag_re, h_diag_im;   // per tuple, 2^n entries
  std::vector<uint32_t> h_range;              // per tuple, 2^n + 1 template offsets
  std::vector<uint64_t> h_flip;               // per template: XOR mask
  std::vector<double> h_coef_re, h_coef_im;   // per template: conj(M[b][a])
  DeviceBuffer<uint32_t> tuple_sites, tuple_n, tuple_off, range;
  DeviceBuffer<double> diag_re, diag_im, coef_re, coef_im;
  DeviceBuffer<uint64_t> flip;
  // exchange-bond form (see compile_operator): one record per tuple
  bool bonds = false;
  std::vector<uint32_t> h_bond_sites;        // s0 | s1 << 8
  std::vector<double> h_bond_dp, h_bond_da, h_bond_o;
  std::vector<uint64_t> h_bond_flip;
  // every bond joins ring neighbours (i, i+1 mod n): its lower site i
  bool nn_ring = false;
  bool nn_identity = false; // bond k joins sites k and k + 1 for every k
  std::vector<uint8_t> h_bond_lo;
  DeviceBuffer<uint8_t> bond_lo;
  DeviceBuffer<uint32_t> bond_sites;
  DeviceBuffer<double> bond_dp, bond_da, bond_o;
  DeviceBuffer<uint64_t> bond_flip;
  bool on_device = false;
};

void ensure_device(OperatorTables &t);

// Gather cache of a real operator (matvec.hip): the target of every term of
// every row, written once by the orbit and index kernels, so that later real
// matvecs read it instead of recomputing orbits and searching the index.
// Rows in schedule order, one wave per 64 slots; term t of lane l of wave w
// at entries[woff[w] + 64 t + l] (coalesced).  An entry holds the target
// index (bits 0..29, all ones: no term), a negative coefficient (bit 30) and
// a stabiliser ratio other than 1 (bit 31, taken from stab[target]).
struct GatherCache {
  std::mutex mu;
  // set while the holder of mu allocates: a refused allocation reclaims
  // caches from the same thread, which must pass this one by instead of
  // locking mu a second time
  std::atomic<bool> building{false};
  int request = -1;       // -1 automatic (built when it fits in free memory), 0 off, 1 on
  bool counted = false;   // woff and total known
  bool built = false;
  uint64_t total = 0;     // entries
  DeviceBuffer<uint64_t> woff;
  DeviceBuffer<uint32_t> entries;
  void drop() {           // caller holds mu
    built = false;
    counted = false;
    total = 0;
    entries.release();
    woff.release();
  }
};

struct Operator {
  std::shared_ptr<Basis> basis;
  std::vector<std::shared_ptr<const Interaction>> terms;
  OperatorTables tables;
  int orbit_method_override = -1;
  std::mutex mu;
  spined_stats last_stats{};
  int orbit_method() const; // 0 generic, 1 ring
  // Row schedule of the real matvec kernels: within every tile of
  // kScheduleTile consecutive rows, the rows ordered by their number of
  // active off-diagonal templates, so a wave's lanes carry equal work.
  // sched[i] = offset within the tile of the row processed at slot i.
  int schedule_request = -1; // -1 automatic, 0 off, 1 on
  // Lanes per row of the matrix-free real matvec of exchange-bond
  // operators: 1 (kRowsPerWaveMax rows per wave), 2 or 4; the sums do not
  // depend on it
  int lanes_per_row = 1;
  bool schedule_built = false;
  DeviceBuffer<uint16_t> sched;
  std::unique_ptr<GatherCache> cache = std::make_unique<GatherCache>();
  Operator();
  ~Operator();
};

constexpr int kScheduleTile = 4096;
constexpr int kRowsPerWaveMax = 64; // one lane per row
void ensure_schedule(Operator &op); // matvec.hip
// Drops the operator's gather cache (the schedule or the basis changed).
void drop_gather_cache(Operator &op);  // matvec.hip
// Frees the gather caches on a device (a device allocation failed).
void reclaim_gather_caches(int device); // matvec.hip
// Builds the operator's gather cache (unless it is off) when it fits with
// reserve bytes of device memory (free or pooled in blocks of block_bytes)
// to spare; true when a cache is ready.
bool prepare_gather_cache(Operator &op, uint64_t reserve, uint64_t block_bytes); // matvec.hip
void drop_gather_cache_if_built(Operator &op);                                   // matvec.hip

void compile_operator(Operator &op);

// ---- kernel launchers (matvec.hip) ----------------------------------------
enum class VecType { F32, F64, C64, C128 };

struct MatvecCall {
  const Operator *op;
  VecType xt, yt;
  const void *x;
  void *y;
  uint64_t size;
  hipStream_t stream;
  // no device allocation: the gather cache is read when it is built but not
  // built, and a complex operator's phase table must exist already
  bool no_alloc = false;
};
void launch_matvec(const MatvecCall &c);
// out[0], out[1] (device) = re, im of sum_r conj(x_r) y_r in FP64, in a fixed
// order; partial holds 2 * nb doubles, nb = vdot_blocks(n) <= kVdotMaxBlocks.
constexpr unsigned kVdotMaxBlocks = 2048;
unsigned vdot_blocks(uint64_t n);
void launch_vdot(VecType t, const void *x, const void *y, uint64_t n, double *partial,
                 unsigned nb, double *out, hipStream_t s);
// Row-interleaved block of K in {1, 2, 4, 8} real vectors: x[row * K + c].
struct MatvecBlockCall {
  const Operator *op;
  VecType xt, yt;
  const void *x;
  void *y;
  uint64_t size;
  int K;
  hipStream_t stream;
};
bool matvec_block_supported(const Operator &op);
void launch_matvec_block(const MatvecBlockCall &c);
// Layout changes between column-major (stride in elements) and
// row-interleaved blocks of real vectors of type t.
void launch_cols_to_rows(VecType t, const void *x, uint64_t stride, uint64_t D, int K, void *xr,
                         hipStream_t s);
void launch_rows_to_cols(VecType t, const void *yr, uint64_t D, int K, void *y, uint64_t stride,
                         hipStream_t s);
// Largest supported block width (8, 4, 2, 1) not above `remaining`.
int block_chunk(uint64_t remaining);
uint64_t count_offdiag_device(const Operator &op);

// ---- eigensolver (solver.hip) ---------------------------------------------
int eigsh(Operator &op, int dtype, uint64_t k, double tol, const spined_eigsh_options &opts,
          double *evals, void *evecs, uint64_t ld, double *rnorms, spined_monitor_fn monitor,
          void *ctx);

// ---- batched solver (batch.hip) ---------------------------------------------
int batch_eigsh(const std::vector<Operator *> &ops, double tol, int m_req, uint64_t seed,
                int max_cycles, double *evals, double *rnorms, int32_t *cycles_out,
                double *seconds_out, int64_t *matvecs_out, spined_batch_stats *stats);

// ---- misc kernels (solver.hip) ----------------------------------------------
void launch_random_fill(double *v, uint64_t n, uint64_t seed, hipStream_t s);
void launch_copy_calibration(const void *src, void *dst, uint64_t bytes, hipStream_t s);
uint64_t gather_calibration_out_words();
void launch_gather_calibration(const void *src, void *dst, uint64_t n_lines, int words,
                               hipStream_t s);

// ---- learned start (cnn.hip) ------------------------------------------------
void cnn_start_vector(const spined_cnn *c, const Basis &b, uint64_t marshall_mask, double *v);

} // namespace spined
