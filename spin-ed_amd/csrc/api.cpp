// api.cpp -- the ls_* C ABI (include/lattice_symmetries.h) and the spined_*
// entry points (include/spined.h).  Each function cites the reference
// binding it serves in src/SpinED/Internal.hs.
#include <algorithm>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "handles.hpp"

using namespace spined;

namespace spined {

static std::atomic<bool> g_logging{false};

bool logging_enabled() { return g_logging.load(std::memory_order_relaxed); }

void log_line(const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  fprintf(stderr, "[spined] %s\n", buf);
  fflush(stderr);
}

[[noreturn]] void throw_hip(hipError_t e, const char *expr, const char *file, int line) {
  (void)hipGetLastError();
  throw Error(LS_DEVICE_ERROR, std::string(hipGetErrorString(e)) + " in " + expr + " at " +
                                   file + ":" + std::to_string(line));
}

int current_device() {
  int d = 0;
  SPINED_HIP(hipGetDevice(&d));
  return d;
}

namespace {
struct PoolBlock {
  int device;
  size_t bytes;
  void *ptr;
};
std::mutex g_pool_mu;
std::vector<PoolBlock> g_pool;

void release_blocks(int device) {
  std::vector<void *> drop;
  {
    std::lock_guard<std::mutex> lk(g_pool_mu);
    for (size_t i = 0; i < g_pool.size();) {
      if (g_pool[i].device == device) {
        drop.push_back(g_pool[i].ptr);
        g_pool[i] = g_pool.back();
        g_pool.pop_back();
      } else {
        ++i;
      }
    }
  }
  for (void *p : drop) (void)hipFree(p);
}
} // namespace

void *device_alloc(size_t bytes, bool pooled) {
  const int d = current_device();
  if (pooled) {
    std::lock_guard<std::mutex> lk(g_pool_mu);
    for (size_t i = 0; i < g_pool.size(); ++i) {
      if (g_pool[i].device == d && g_pool[i].bytes == bytes) {
        void *p = g_pool[i].ptr;
        g_pool[i] = g_pool.back();
        g_pool.pop_back();
        return p;
      }
    }
  }
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    reclaim_gather_caches(d); // their entries go to the pool, freed next
    release_blocks(d);
    e = hipMalloc(&p, bytes);
  }
  if (e != hipSuccess) {
    (void)hipGetLastError();
    throw Error(LS_OUT_OF_MEMORY, "hipMalloc of " + std::to_string(bytes) +
                                      " bytes failed: " + hipGetErrorString(e));
  }
  return p;
}

void device_free(void *p, size_t bytes, bool pooled, int device) {
  if (!p) return;
  if (pooled) {
    std::lock_guard<std::mutex> lk(g_pool_mu);
    g_pool.push_back(PoolBlock{device, bytes, p});
    return;
  }
  (void)hipFree(p);
}

void pool_release_device() { release_blocks(current_device()); }

size_t pool_bytes(size_t bytes) {
  const int d = current_device();
  std::lock_guard<std::mutex> lk(g_pool_mu);
  size_t s = 0;
  for (const PoolBlock &b : g_pool)
    if (b.device == d && b.bytes == bytes) s += b.bytes;
  return s;
}

hipStream_t stream() {
  static std::mutex mu;
  static std::vector<hipStream_t> streams;
  const int d = current_device();
  std::lock_guard<std::mutex> lk(mu);
  if (static_cast<int>(streams.size()) <= d) streams.resize(d + 1, nullptr);
  if (!streams[d]) SPINED_HIP(hipStreamCreateWithFlags(&streams[d], hipStreamNonBlocking));
  return streams[d];
}

} // namespace spined

namespace {

const char *message_of(int status) {
  switch (status) {
  case LS_SUCCESS: return "no error";
  case LS_INVALID_PERMUTATION: return "argument is not a valid permutation";
  case LS_INVALID_SECTOR: return "invalid symmetry sector: must be smaller than the periodicity";
  case LS_INCOMPATIBLE_SYMMETRIES: return "symmetries are incompatible";
  case LS_INVALID_NUMBER_SPINS: return "invalid number of spins";
  case LS_INVALID_HAMMING_WEIGHT: return "invalid Hamming weight";
  case LS_INVALID_SPIN_INVERSION: return "invalid spin inversion";
  case LS_OUT_OF_MEMORY: return "failed to allocate memory";
  case LS_CACHE_NOT_BUILT: return "basis has not been built; call ls_build first";
  case LS_INVALID_ARGUMENT: return "invalid argument";
  case LS_NOT_CONVERGED: return "eigensolver did not converge";
  case LS_INVALID_DATATYPE: return "invalid or unsupported datatype";
  case LS_DIMENSION_MISMATCH: return "dimensions do not match";
  case LS_DEVICE_ERROR: return "GPU (HIP) error";
  case LS_NOT_IMPLEMENTED: return "not implemented";
  case LS_INTERNAL_ERROR: return "internal error";
  default: return "unknown error";
  }
}

// Make sure the basis' group tables are on the device.
void ensure_basis_device(Basis &b) { ensure_device(b.tables); }

void require_built(const Basis &b) {
  if (!b.built) throw Error(LS_CACHE_NOT_BUILT, "basis has not been built");
}

VecType vec_type(int dtype) {
  switch (dtype) {
  case LS_FLOAT32: return VecType::F32;
  case LS_FLOAT64: return VecType::F64;
  case LS_COMPLEX64: return VecType::C64;
  case LS_COMPLEX128: return VecType::C128;
  default: throw Error(LS_INVALID_DATATYPE, "datatype must be 0..3");
  }
}

size_t elem_size(int dtype) {
  switch (dtype) {
  case LS_FLOAT32: return 4;
  case LS_FLOAT64: return 8;
  case LS_COMPLEX64: return 8;
  default: return 16;
  }
}

// Locks the operator's basis for device work: builds are serialised with
// matvecs, and device tables are uploaded on first use.
Operator &prepare_operator(const ls_operator *op) {
  if (!op || !op->p) throw Error(LS_INVALID_ARGUMENT, "operator is NULL");
  Operator &o = *op->p;
  Basis &b = *o.basis;
  require_built(b);
  ensure_basis_device(b);
  ensure_device(o.tables);
  ensure_schedule(o);
  return o;
}

int create_interaction(ls_interaction **out, const void *matrix, unsigned number_sites,
                       const uint16_t *sites, int n) {
  return guarded([&] {
    if (!out || (!matrix) || (number_sites && !sites))
      throw Error(LS_INVALID_ARGUMENT, "NULL argument");
    auto it = std::make_shared<Interaction>();
    it->n = n;
    const int dim = 1 << n;
    const double *m = static_cast<const double *>(matrix);
    it->matrix.resize(dim * dim);
    for (int i = 0; i < dim * dim; ++i) {
      it->matrix[i] = cplx(m[2 * i], m[2 * i + 1]);
      if (m[2 * i + 1] != 0.0) it->is_real = false;
    }
    for (unsigned t = 0; t < number_sites; ++t) {
      std::array<uint16_t, 4> s{0, 0, 0, 0};
      for (int k = 0; k < n; ++k) {
        s[k] = sites[t * n + k];
        if (s[k] >= 64) throw Error(LS_INVALID_ARGUMENT, "site index must be < 64");
        for (int q = 0; q < k; ++q)
          if (s[q] == s[k]) throw Error(LS_INVALID_ARGUMENT, "repeated site in an interaction");
      }
      it->sites.push_back(s);
    }
    *out = new ls_interaction{it};
  });
}

} // namespace

extern "C" {

// ---- errors & logging (Internal.hs:39-54) ----------------------------------
const char *ls_error_to_string(int status) {
  const char *m = message_of(status);
  char *s = static_cast<char *>(std::malloc(std::strlen(m) + 1));
  if (s) std::strcpy(s, m);
  return s;
}

void ls_destroy_string(const char *message) { std::free(const_cast<char *>(message)); }
void ls_enable_logging(void) { g_logging.store(true); }
void ls_disable_logging(void) { g_logging.store(false); }

// ---- symmetry (Internal.hs:69-116) -----------------------------------------
int ls_create_symmetry(ls_symmetry **out, unsigned length, const unsigned permutation[],
                       unsigned sector) {
  return guarded([&] {
    if (!out || (length && !permutation)) throw Error(LS_INVALID_ARGUMENT, "NULL argument");
    auto s = std::make_shared<Symmetry>();
    s->perm.assign(permutation, permutation + length);
    s->periodicity = periodicity_of(s->perm);
    if (s->periodicity == 0) throw Error(LS_INVALID_PERMUTATION, "not a permutation");
    if (sector >= s->periodicity) throw Error(LS_INVALID_SECTOR, "sector out of range");
    s->sector = sector;
    *out = new ls_symmetry{s};
  });
}

void ls_destroy_symmetry(ls_symmetry *symmetry) { delete symmetry; }
unsigned ls_get_sector(const ls_symmetry *s) { return s ? s->p->sector : 0; }
double ls_get_phase(const ls_symmetry *s) {
  return s ? static_cast<double>(s->p->sector) / static_cast<double>(s->p->periodicity) : 0.0;
}
unsigned ls_get_periodicity(const ls_symmetry *s) { return s ? s->p->periodicity : 0; }

// ---- group (Internal.hs:120-155) ------------------------------------------
int ls_create_group(ls_group **out, unsigned size, const ls_symmetry *generators[]) {
  return guarded([&] {
    if (!out || (size && !generators)) throw Error(LS_INVALID_ARGUMENT, "NULL argument");
    std::vector<Symmetry> gens;
    for (unsigned i = 0; i < size; ++i) {
      if (!generators[i]) throw Error(LS_INVALID_ARGUMENT, "NULL generator");
      gens.push_back(*generators[i]->p);
    }
    *out = new ls_group{close_group(gens)};
  });
}

void ls_destroy_group(ls_group *group) { delete group; }
unsigned ls_get_group_size(const ls_group *g) {
  return g ? static_cast<unsigned>(g->p->perms.size()) : 0;
}

// ---- basis (Internal.hs:172-254) -------------------------------------------
int ls_create_spin_basis(ls_spin_basis **out, const ls_group *group, unsigned number_spins,
                         int hamming_weight, int spin_inversion) {
  return guarded([&] {
    if (!out || !group) throw Error(LS_INVALID_ARGUMENT, "NULL argument");
    const Group &g = *group->p;
    if (number_spins == 0 || number_spins > 64)
      throw Error(LS_INVALID_NUMBER_SPINS, "number of spins must be in [1, 64]");
    if (g.n != 0 && static_cast<unsigned>(g.n) != number_spins)
      throw Error(LS_INVALID_NUMBER_SPINS, "number of spins differs from the symmetries' size");
    if (hamming_weight < -1 || hamming_weight > static_cast<int>(number_spins))
      throw Error(LS_INVALID_HAMMING_WEIGHT, "Hamming weight must be -1 or in [0, n]");
    if (spin_inversion < -1 || spin_inversion > 1)
      throw Error(LS_INVALID_SPIN_INVERSION, "spin inversion must be -1, 0 or +1");
    if (spin_inversion != 0 && hamming_weight >= 0 &&
        2 * hamming_weight != static_cast<int>(number_spins))
      throw Error(LS_INVALID_SPIN_INVERSION, "spin inversion needs hamming_weight = n/2");
    auto b = std::make_shared<Basis>();
    b->group = group->p;
    b->n = static_cast<int>(number_spins);
    b->w = hamming_weight;
    b->z = spin_inversion;
    build_orbit_tables(g, b->n, spin_inversion, b->tables);
    *out = new ls_spin_basis{b};
  });
}

void ls_destroy_spin_basis(ls_spin_basis *basis) { delete basis; }

int ls_build(ls_spin_basis *basis) {
  return guarded([&] {
    if (!basis) throw Error(LS_INVALID_ARGUMENT, "NULL basis");
    Basis &b = *basis->p;
    std::lock_guard<std::mutex> lk(b.mu);
    if (b.built) return;
    ensure_basis_device(b);
    build_basis(b);
  });
}

int ls_build_unsafe(ls_spin_basis *basis, uint64_t size, const uint64_t representatives[]) {
  return guarded([&] {
    if (!basis || (size && !representatives)) throw Error(LS_INVALID_ARGUMENT, "NULL argument");
    Basis &b = *basis->p;
    std::lock_guard<std::mutex> lk(b.mu);
    // operators bound to this basis keep row schedules, gather caches and
    // index arguments sized for its current list: it cannot be replaced
    if (b.built)
      throw Error(LS_INVALID_ARGUMENT, "basis is already built; its representatives cannot be replaced");
    validate_representatives(b, size, representatives);
    ensure_basis_device(b);
    adopt_basis(b, size, representatives);
  });
}

int ls_get_number_states(const ls_spin_basis *basis, uint64_t *out) {
  return guarded([&] {
    if (!basis || !out) throw Error(LS_INVALID_ARGUMENT, "NULL argument");
    const Basis &b = *basis->p;
    require_built(b);
    *out = b.D;
  });
}

int ls_get_states(ls_states **out, const ls_spin_basis *basis) {
  return guarded([&] {
    if (!basis || !out) throw Error(LS_INVALID_ARGUMENT, "NULL argument");
    Basis &b = *basis->p;
    std::lock_guard<std::mutex> lk(b.mu);
    require_built(b);
    if (!b.host_reps) {
      auto h = std::make_shared<HostStates>(b.D);
      if (b.D)
        SPINED_HIP(hipMemcpy(h->data.get(), b.reps.ptr, b.D * 8, hipMemcpyDeviceToHost));
      b.host_reps = h;
    }
    *out = new ls_states{b.host_reps};
  });
}

const uint64_t *ls_states_get_data(const ls_states *states) {
  return states && states->data ? states->data->data.get() : nullptr;
}
uint64_t ls_states_get_size(const ls_states *states) {
  return states && states->data ? states->data->size : 0;
}
void ls_destroy_states(ls_states *states) { delete states; }

// ---- interactions (Internal.hs:256-367) ------------------------------------
int ls_create_interaction1(ls_interaction **out, const void *matrix, unsigned number_sites,
                           const uint16_t *sites) {
  return create_interaction(out, matrix, number_sites, sites, 1);
}
int ls_create_interaction2(ls_interaction **out, const void *matrix, unsigned number_sites,
                           const uint16_t *sites) {
  return create_interaction(out, matrix, number_sites, sites, 2);
}
int ls_create_interaction3(ls_interaction **out, const void *matrix, unsigned number_sites,
                           const uint16_t *sites) {
  return create_interaction(out, matrix, number_sites, sites, 3);
}
int ls_create_interaction4(ls_interaction **out, const void *matrix, unsigned number_sites,
                           const uint16_t *sites) {
  return create_interaction(out, matrix, number_sites, sites, 4);
}
bool ls_interaction_is_real(const ls_interaction *it) { return it && it->p->is_real; }
void ls_destroy_interaction(ls_interaction *it) { delete it; }

// ---- operator (Internal.hs:369-458) ----------------------------------------
int ls_create_operator(ls_operator **out, const ls_spin_basis *basis, unsigned number_terms,
                       const ls_interaction *const terms[]) {
  return guarded([&] {
    if (!out || !basis || (number_terms && !terms))
      throw Error(LS_INVALID_ARGUMENT, "NULL argument");
    auto op = std::make_shared<Operator>();
    op->basis = basis->p;
    for (unsigned i = 0; i < number_terms; ++i) {
      if (!terms[i]) throw Error(LS_INVALID_ARGUMENT, "NULL interaction");
      op->terms.push_back(terms[i]->p);
    }
    compile_operator(*op);
    *out = new ls_operator{op};
  });
}

void ls_destroy_operator(ls_operator *op) { delete op; }

bool ls_operator_is_real(const ls_operator *op) { return op && op->p->tables.is_real; }

int ls_operator_matmat(const ls_operator *op, int dtype, uint64_t size, uint64_t block_size,
                       const void *x, uint64_t x_stride, void *y, uint64_t y_stride) {
  return guarded([&] {
    Operator &o = prepare_operator(op);
    const VecType vt = vec_type(dtype);
    if (size != o.basis->D) throw Error(LS_DIMENSION_MISMATCH, "size differs from dimension");
    if (block_size == 0) return;
    if (!x || !y) throw Error(LS_INVALID_ARGUMENT, "NULL vector");
    if (x_stride < size || y_stride < size)
      throw Error(LS_INVALID_ARGUMENT, "stride smaller than size");
    const size_t es = elem_size(dtype);
    hipStream_t s = stream();
    if (block_size >= 2 && (vt == VecType::F64 || vt == VecType::F32) &&
        matvec_block_supported(o)) {
      // chunks of up to 8 columns through the row-interleaved block matvec
      const uint64_t w = std::min<uint64_t>(block_size, 8);
      DeviceBuffer<char> dc(size * es * w), dxr(size * es * w), dyr(size * es * w);
      for (uint64_t c0 = 0; c0 < block_size;) {
        const int kc = block_chunk(block_size - c0);
        const char *xc = static_cast<const char *>(x) + c0 * x_stride * es;
        char *yc = static_cast<char *>(y) + c0 * y_stride * es;
        SPINED_HIP(hipMemcpy2DAsync(dc.ptr, size * es, xc, x_stride * es, size * es, kc,
                                    hipMemcpyHostToDevice, s));
        launch_cols_to_rows(vt, dc.ptr, size, size, kc, dxr.ptr, s);
        launch_matvec_block(MatvecBlockCall{&o, vt, vt, dxr.ptr, dyr.ptr, size, kc, s});
        launch_rows_to_cols(vt, dyr.ptr, size, kc, dc.ptr, size, s);
        SPINED_HIP(hipMemcpy2DAsync(yc, y_stride * es, dc.ptr, size * es, size * es, kc,
                                    hipMemcpyDeviceToHost, s));
        c0 += kc;
      }
      SPINED_HIP(hipStreamSynchronize(s));
      return;
    }
    DeviceBuffer<char> dx(size * es), dy(size * es);
    for (uint64_t c = 0; c < block_size; ++c) {
      const char *xc = static_cast<const char *>(x) + c * x_stride * es;
      char *yc = static_cast<char *>(y) + c * y_stride * es;
      SPINED_HIP(hipMemcpyAsync(dx.ptr, xc, size * es, hipMemcpyHostToDevice, s));
      launch_matvec(MatvecCall{&o, vt, vt, dx.ptr, dy.ptr, size, s});
      SPINED_HIP(hipMemcpyAsync(yc, dy.ptr, size * es, hipMemcpyDeviceToHost, s));
    }
    SPINED_HIP(hipStreamSynchronize(s));
  });
}

int ls_operator_expectation(const ls_operator *op, int dtype, uint64_t size, uint64_t block_size,
                            const void *x, uint64_t x_stride, void *out) {
  // <x_c|H|x_c> per column c (not divided by <x_c|x_c>).  Each column goes to
  // the device once; H x_c and the FP64 dot stay there, and only the two
  // doubles of the result come back.
  return guarded([&] {
    Operator &o = prepare_operator(op);
    const VecType vt = vec_type(dtype);
    if (size != o.basis->D) throw Error(LS_DIMENSION_MISMATCH, "size differs from dimension");
    if (block_size == 0) return;
    if (!x || !out) throw Error(LS_INVALID_ARGUMENT, "NULL argument");
    if (x_stride < size) throw Error(LS_INVALID_ARGUMENT, "stride smaller than size");
    const size_t es = elem_size(dtype);
    hipStream_t s = stream();
    DeviceBuffer<char> dx(std::max<uint64_t>(size, 1) * es), dy(std::max<uint64_t>(size, 1) * es);
    const unsigned nb = vdot_blocks(size);
    DeviceBuffer<double> partial(2 * static_cast<uint64_t>(nb)), res(2 * block_size);
    for (uint64_t c = 0; c < block_size; ++c) {
      const char *xc = static_cast<const char *>(x) + c * x_stride * es;
      SPINED_HIP(hipMemcpyAsync(dx.ptr, xc, size * es, hipMemcpyHostToDevice, s));
      launch_matvec(MatvecCall{&o, vt, vt, dx.ptr, dy.ptr, size, s});
      launch_vdot(vt, dx.ptr, dy.ptr, size, partial.ptr, nb, res.ptr + 2 * c, s);
    }
    SPINED_HIP(hipMemcpyAsync(out, res.ptr, 16 * block_size, hipMemcpyDeviceToHost, s));
    SPINED_HIP(hipStreamSynchronize(s));
  });
}

int spined_operator_expectation_device(const ls_operator *op, int dtype, uint64_t size,
                                       uint64_t block_size, const void *x, uint64_t x_stride,
                                       void *y_scratch, double *out) {
  return guarded([&] {
    Operator &o = prepare_operator(op);
    const VecType vt = vec_type(dtype);
    if (size != o.basis->D) throw Error(LS_DIMENSION_MISMATCH, "size differs from dimension");
    if (block_size == 0) return;
    if (!x || !out || !y_scratch) throw Error(LS_INVALID_ARGUMENT, "NULL argument");
    if (x_stride < size) throw Error(LS_INVALID_ARGUMENT, "stride smaller than size");
    const size_t es = elem_size(dtype);
    hipStream_t s = stream();
    const unsigned nb = vdot_blocks(size);
    DeviceBuffer<double> partial(2 * static_cast<uint64_t>(nb)), res(2 * block_size);
    for (uint64_t c = 0; c < block_size; ++c) {
      const char *xc = static_cast<const char *>(x) + c * x_stride * es;
      launch_matvec(MatvecCall{&o, vt, vt, xc, y_scratch, size, s});
      launch_vdot(vt, xc, y_scratch, size, partial.ptr, nb, res.ptr + 2 * c, s);
    }
    SPINED_HIP(hipMemcpyAsync(out, res.ptr, 16 * block_size, hipMemcpyDeviceToHost, s));
    SPINED_HIP(hipStreamSynchronize(s));
  });
}

// ---- spined_* ---------------------------------------------------------------
void spined_eigsh_options_init(spined_eigsh_options *o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->precision = SPINED_PRECISION_FP64;
  o->start = SPINED_START_RANDOM;
  o->seed = 0x5eed;
}

int spined_eigsh(const ls_operator *op, int dtype, uint64_t k, double tol,
                 const spined_eigsh_options *opts, double *evals, void *evecs, uint64_t ld,
                 double *rnorms, spined_monitor_fn monitor, void *monitor_ctx) {
  int status = LS_SUCCESS;
  const int g = guarded([&] {
    Operator &o = prepare_operator(op);
    if (!evals) throw Error(LS_INVALID_ARGUMENT, "evals is NULL");
    spined_eigsh_options d;
    spined_eigsh_options_init(&d);
    status = eigsh(o, dtype, k, tol, opts ? *opts : d, evals, evecs, ld, rnorms, monitor,
                   monitor_ctx);
  });
  return g ? g : status;
}

int spined_batch_eigsh(const ls_operator *const ops[], uint64_t count, double tol,
                       int32_t max_basis_size, uint64_t seed, int32_t max_cycles, double *evals,
                       double *rnorms, int32_t *cycles, double *seconds, int64_t *matvecs,
                       spined_batch_stats *stats) {
  int status = LS_SUCCESS;
  const int g = guarded([&] {
    if (count && (!ops || !evals || !rnorms)) throw Error(LS_INVALID_ARGUMENT, "NULL argument");
    if (max_basis_size < 0 || max_basis_size > 24)
      throw Error(LS_INVALID_ARGUMENT, "max_basis_size must be in [0, 24]");
    std::vector<Operator *> v;
    v.reserve(count);
    for (uint64_t i = 0; i < count; ++i) v.push_back(&prepare_operator(ops[i]));
    status = batch_eigsh(v, tol > 0.0 ? tol : 1e-9, max_basis_size, seed,
                         max_cycles > 0 ? max_cycles : 200, evals, rnorms, cycles, seconds,
                         matvecs, stats);
  });
  return g ? g : status;
}

int spined_get_last_stats(const ls_operator *op, spined_stats *out) {
  return guarded([&] {
    if (!op || !out) throw Error(LS_INVALID_ARGUMENT, "NULL argument");
    std::lock_guard<std::mutex> lk(op->p->mu);
    *out = op->p->last_stats;
  });
}

void *spined_stream(void) {
  try {
    return static_cast<void *>(stream());
  } catch (...) {
    return nullptr;
  }
}

int spined_synchronize(void) {
  return guarded([] { SPINED_HIP(hipStreamSynchronize(stream())); });
}

int spined_device_synchronize(void) {
  return guarded([] { SPINED_HIP(hipDeviceSynchronize()); });
}

int spined_device_malloc(void **out, uint64_t bytes) {
  return guarded([&] {
    if (!out) throw Error(LS_INVALID_ARGUMENT, "NULL argument");
    *out = device_alloc(bytes ? bytes : 1, false);
  });
}

int spined_device_release_pool(void) {
  return guarded([&] { pool_release_device(); });
}

int spined_device_free(void *ptr) {
  return guarded([&] {
    if (ptr) SPINED_HIP(hipFree(ptr));
  });
}

int spined_memcpy(void *dst, const void *src, uint64_t bytes, int kind) {
  return guarded([&] {
    if (bytes == 0) return;
    if (!dst || !src) throw Error(LS_INVALID_ARGUMENT, "NULL argument");
    hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice
                      : kind == 1 ? hipMemcpyDeviceToHost
                                  : hipMemcpyDeviceToDevice;
    if (kind < 0 || kind > 2) throw Error(LS_INVALID_ARGUMENT, "kind must be 0, 1 or 2");
    SPINED_HIP(hipMemcpyAsync(dst, src, bytes, k, stream()));
    SPINED_HIP(hipStreamSynchronize(stream()));
  });
}

int spined_memset(void *dst, int value, uint64_t bytes) {
  return guarded([&] {
    if (bytes == 0) return;
    if (!dst) throw Error(LS_INVALID_ARGUMENT, "NULL argument");
    SPINED_HIP(hipMemsetAsync(dst, value, bytes, stream()));
    SPINED_HIP(hipStreamSynchronize(stream()));
  });
}

int spined_basis_device_states(const ls_spin_basis *basis, const uint64_t **reps,
                               const uint32_t **stabilizers, uint64_t *size) {
  return guarded([&] {
    if (!basis) throw Error(LS_INVALID_ARGUMENT, "NULL basis");
    const Basis &b = *basis->p;
    require_built(b);
    if (reps) *reps = b.reps.ptr;
    if (stabilizers) *stabilizers = b.stab.ptr;
    if (size) *size = b.D;
  });
}

int spined_basis_get_stabilizers(const ls_spin_basis *basis, uint32_t *out) {
  return guarded([&] {
    if (!basis || !out) throw Error(LS_INVALID_ARGUMENT, "NULL argument");
    const Basis &b = *basis->p;
    require_built(b);
    if (b.D) SPINED_HIP(hipMemcpy(out, b.stab.ptr, b.D * 4, hipMemcpyDeviceToHost));
  });
}

int spined_operator_matmat_device(const ls_operator *op, int dtype, uint64_t size,
                                  uint64_t block_size, const void *x, uint64_t x_stride,
                                  void *y, uint64_t y_stride, int y_is_fp64) {
  return guarded([&] {
    Operator &o = prepare_operator(op);
    const VecType xt = vec_type(dtype);
    VecType yt = xt;
    if (y_is_fp64) {
      if (xt != VecType::F32) throw Error(LS_INVALID_DATATYPE, "y_is_fp64 needs float32 x");
      yt = VecType::F64;
    }
    if (size != o.basis->D) throw Error(LS_DIMENSION_MISMATCH, "size differs from dimension");
    if (x_stride < size || y_stride < size)
      throw Error(LS_INVALID_ARGUMENT, "stride smaller than size");
    const size_t xs = elem_size(dtype), ys = y_is_fp64 ? 8 : elem_size(dtype);
    if (block_size >= 2 && (xt == VecType::F64 || xt == VecType::F32) &&
        matvec_block_supported(o)) {
      const uint64_t w = std::min<uint64_t>(block_size, 8);
      DeviceBuffer<char> dxr(size * xs * w), dyr(size * ys * w);
      hipStream_t s = stream();
      for (uint64_t c0 = 0; c0 < block_size;) {
        const int kc = block_chunk(block_size - c0);
        launch_cols_to_rows(xt, static_cast<const char *>(x) + c0 * x_stride * xs, x_stride, size,
                            kc, dxr.ptr, s);
        launch_matvec_block(MatvecBlockCall{&o, xt, yt, dxr.ptr, dyr.ptr, size, kc, s});
        launch_rows_to_cols(yt, dyr.ptr, size, kc, static_cast<char *>(y) + c0 * y_stride * ys,
                            y_stride, s);
        c0 += kc;
      }
      // the temporaries are freed when this returns: wait for their users
      SPINED_HIP(hipStreamSynchronize(s));
      return;
    }
    for (uint64_t c = 0; c < block_size; ++c)
      launch_matvec(MatvecCall{&o, xt, yt, static_cast<const char *>(x) + c * x_stride * xs,
                               static_cast<char *>(y) + c * y_stride * ys, size, stream()});
  });
}

int spined_operator_matmat_rows_device(const ls_operator *op, int dtype, uint64_t size,
                                       uint64_t block_size, const void *x, void *y,
                                       int y_is_fp64) {
  return guarded([&] {
    Operator &o = prepare_operator(op);
    const VecType xt = vec_type(dtype);
    if (xt != VecType::F64 && xt != VecType::F32)
      throw Error(LS_INVALID_DATATYPE, "row-interleaved blocks are real");
    VecType yt = xt;
    if (y_is_fp64) {
      if (xt != VecType::F32) throw Error(LS_INVALID_DATATYPE, "y_is_fp64 needs float32 x");
      yt = VecType::F64;
    }
    if (size != o.basis->D) throw Error(LS_DIMENSION_MISMATCH, "size differs from dimension");
    if (block_size != 1 && block_size != 2 && block_size != 4 && block_size != 8)
      throw Error(LS_INVALID_ARGUMENT, "block width must be 1, 2, 4 or 8");
    if (size && (!x || !y)) throw Error(LS_INVALID_ARGUMENT, "NULL vector");
    launch_matvec_block(MatvecBlockCall{&o, xt, yt, x, y, size, static_cast<int>(block_size),
                                        stream()});
  });
}

namespace {
// Mean device time of `repeats` FP64 matvecs (block width K) after one warm-up.
double time_matvec(Operator &o, int K, int repeats, double *x, double *y) {
  hipStream_t s = stream();
  const uint64_t n = o.basis->D;
  auto launch = [&] {
    if (K == 1)
      launch_matvec(MatvecCall{&o, VecType::F64, VecType::F64, x, y, n, s});
    else
      launch_matvec_block(MatvecBlockCall{&o, VecType::F64, VecType::F64, x, y, n, K, s});
  };
  launch();
  hipEvent_t a, b;
  SPINED_HIP(hipEventCreate(&a));
  SPINED_HIP(hipEventCreate(&b));
  SPINED_HIP(hipEventRecord(a, s));
  for (int r = 0; r < repeats; ++r) launch();
  SPINED_HIP(hipEventRecord(b, s));
  SPINED_HIP(hipEventSynchronize(b));
  float ms = 0.f;
  SPINED_HIP(hipEventElapsedTime(&ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return ms / repeats;
}

uint64_t free_device_bytes() {
  size_t f = 0, t = 0;
  SPINED_HIP(hipMemGetInfo(&f, &t));
  return f;
}
} // namespace

int spined_operator_autotune(ls_operator *op, int32_t repeats, spined_autotune_report *report) {
  return guarded([&] {
    Operator &o = prepare_operator(op);
    Basis &b = *o.basis;
    if (!report) throw Error(LS_INVALID_ARGUMENT, "NULL report");
    if (!o.tables.is_real) throw Error(LS_INVALID_DATATYPE, "the autotuner times real operators");
    if (repeats < 1) repeats = 2;
    *report = spined_autotune_report{};
    const uint64_t n = b.D;
    if (n == 0) return;
    // block widths whose two FP64 blocks fit in half the free memory
    int kmax = matvec_block_supported(o) ? 8 : 1;
    while (kmax > 1 && 2 * n * kmax * 8 > free_device_bytes() / 2) kmax /= 2;
    DeviceBuffer<double> x(n * kmax), y(n * kmax);
    launch_random_fill(x.ptr, n * kmax, 0x7E57, stream());
    int lg = 0;
    while ((1ull << lg) < n) ++lg;
    struct Setting {
      int kind, bits, sched;
    };
    const Setting first{b.index_kind, b.index_kind == 0 ? b.bucket_bits : -1,
                        o.schedule_built ? 1 : 0};
    auto apply = [&](const Setting &st) {
      {
        std::lock_guard<std::mutex> lk(b.mu);
        b.index_kind_request = st.kind;
        b.bucket_bits_request = st.bits;
        build_index(b);
      }
      {
        std::lock_guard<std::mutex> lk(o.mu);
        o.schedule_request = st.sched;
        o.schedule_built = false;
        o.sched.release();
        drop_gather_cache(o);
      }
      ensure_schedule(o);
    };
    // the tuner times the orbit-and-index kernels: the gather cache is off
    // while it runs, and its request is restored afterwards
    struct CacheOff {
      GatherCache &gc;
      int saved;
      explicit CacheOff(GatherCache &g) : gc(g) {
        std::lock_guard<std::mutex> lk(gc.mu);
        saved = gc.request;
        gc.request = 0;
        gc.drop();
      }
      ~CacheOff() {
        std::lock_guard<std::mutex> lk(gc.mu);
        gc.request = saved;
      }
    } cache_off(*o.cache);
    report->ms_per_matvec_first = time_matvec(o, 1, repeats, x.ptr, y.ptr);
    // bucket searches that need no confirming load beat the hash table
    const bool exact = b.complete && (b.w < 0 || o.tables.keeps_weight);
    std::vector<Setting> cands;
    for (int bits = lg + 1; bits <= std::min(lg + 4, kMaxBucketBits); ++bits)
      cands.push_back({0, bits, 1});
    cands.push_back({0, lg + 3, 0});
    if (!exact) cands.push_back({1, -1, 1});
    if (rank_index_mode(b)) cands.push_back({2, -1, 1});
    // an index may take a quarter of the memory free outside the tuner's
    // own vectors and the current index: the rest is the eigensolver's
    const uint64_t budget =
        (free_device_bytes() + 16 * n * kmax + (b.buckets.n + b.slots.n * 2) * 4) / 4;
    Setting best = first;
    double best_ms = report->ms_per_matvec_first;
    for (const Setting &c : cands) {
      uint64_t need = (2ull << lg) * 8;                                // hash slots
      if (c.kind == 0) {
        const uint64_t nb = level2_size(b, c.bits, nullptr);
        need = 4 * (nb ? nb + 1 : (2ull << c.bits));
      }
      if (need > budget) continue;
      apply(c);
      const double ms = time_matvec(o, 1, repeats, x.ptr, y.ptr);
      ++report->candidates;
      if (ms < best_ms) {
        best_ms = ms;
        best = Setting{c.kind, c.kind == 0 ? b.bucket_bits : -1, c.sched};
      }
    }
    apply(best);
    report->index_kind = b.index_kind;
    report->bucket_bits = b.bucket_bits;
    report->schedule = o.schedule_built ? 1 : 0;
    // rows per wave of the exchange-bond kernel (lanes per row 1, 2, 4) on
    // the chosen index; the sums are the same for each
    o.lanes_per_row = 1;
    report->ms_per_matvec = time_matvec(o, 1, repeats, x.ptr, y.ptr);
    report->rows_per_wave = kRowsPerWaveMax;
    report->ms_per_rows_per_wave[0] = report->ms_per_matvec;
    if (o.tables.bonds && o.tables.h_bond_sites.size() <= 128) {
      int best_l = 1;
      for (int q = 1; q < 3; ++q) {
        const int l = 1 << q;
        o.lanes_per_row = l;
        const double ms = time_matvec(o, 1, repeats, x.ptr, y.ptr);
        report->ms_per_rows_per_wave[q] = ms;
        if (ms < report->ms_per_matvec) {
          report->ms_per_matvec = ms;
          best_l = l;
        }
      }
      o.lanes_per_row = best_l;
      report->rows_per_wave = kRowsPerWaveMax / best_l;
    }
    double best_pv = 0.0;
    const int Ks[4] = {1, 2, 4, 8};
    for (int q = 0; q < 4; ++q) {
      if (Ks[q] > kmax) break;
      const double pv =
          Ks[q] == 1 ? report->ms_per_matvec : time_matvec(o, Ks[q], repeats, x.ptr, y.ptr) / Ks[q];
      report->ms_per_vector_block[q] = pv;
      if (best_pv == 0.0 || pv < best_pv) {
        best_pv = pv;
        report->best_block = Ks[q];
      }
    }
  });
}

int spined_operator_count_offdiag(const ls_operator *op, uint64_t *out) {
  return guarded([&] {
    Operator &o = prepare_operator(op);
    if (!out) throw Error(LS_INVALID_ARGUMENT, "NULL argument");
    *out = count_offdiag_device(o);
  });
}

int spined_operator_orbit_method(const ls_operator *op) {
  return op ? op->p->orbit_method() : -1;
}

int spined_operator_set_orbit_method(ls_operator *op, int method) {
  return guarded([&] {
    if (!op || method < -1 || method > 1) throw Error(LS_INVALID_ARGUMENT, "bad argument");
    op->p->orbit_method_override = method;
  });
}

int spined_operator_set_schedule(ls_operator *op, int mode) {
  return guarded([&] {
    if (!op || mode < -1 || mode > 1) throw Error(LS_INVALID_ARGUMENT, "bad argument");
    Operator &o = *op->p;
    {
      std::lock_guard<std::mutex> lk(o.mu);
      o.schedule_request = mode;
      o.schedule_built = false;
      o.sched.release();
      drop_gather_cache(o);
    }
  });
}

int spined_operator_get_schedule(const ls_operator *op) {
  return op ? (op->p->schedule_built ? 1 : 0) : -1;
}

int spined_operator_set_rows_per_wave(ls_operator *op, int rows) {
  return guarded([&] {
    if (!op || !(rows == -1 || rows == 64 || rows == 32 || rows == 16))
      throw Error(LS_INVALID_ARGUMENT, "rows per wave must be 64, 32, 16 or -1");
    Operator &o = *op->p;
    std::lock_guard<std::mutex> lk(o.mu);
    o.lanes_per_row = rows == -1 ? 1 : kRowsPerWaveMax / rows;
  });
}

int spined_operator_get_rows_per_wave(const ls_operator *op) {
  return op ? kRowsPerWaveMax / op->p->lanes_per_row : -1;
}

int spined_operator_set_gather_cache(ls_operator *op, int mode) {
  return guarded([&] {
    if (!op || mode < -1 || mode > 1) throw Error(LS_INVALID_ARGUMENT, "bad argument");
    GatherCache &gc = *op->p->cache;
    std::lock_guard<std::mutex> lk(gc.mu);
    gc.request = mode;
    if (mode == 0) gc.drop();
  });
}

int spined_operator_get_gather_cache(const ls_operator *op) {
  if (!op) return -1;
  GatherCache &gc = *op->p->cache;
  std::lock_guard<std::mutex> lk(gc.mu);
  return gc.built ? 1 : 0;
}

int spined_basis_set_bucket_bits(ls_spin_basis *basis, int bits) {
  return guarded([&] {
    if (!basis || bits < -1 || bits > kMaxBucketBits)
      throw Error(LS_INVALID_ARGUMENT, "bad argument");
    Basis &b = *basis->p;
    std::lock_guard<std::mutex> lk(b.mu);
    b.bucket_bits_request = bits;
    if (b.built) build_index(b);
  });
}

int spined_basis_get_bucket_bits(const ls_spin_basis *basis) {
  return basis ? basis->p->bucket_bits : -1;
}

int spined_basis_set_index_kind(ls_spin_basis *basis, int kind) {
  return guarded([&] {
    if (!basis || kind < -1 || kind > 2) throw Error(LS_INVALID_ARGUMENT, "bad argument");
    Basis &b = *basis->p;
    std::lock_guard<std::mutex> lk(b.mu);
    if (kind == 2 && b.built && !rank_index_mode(b))
      throw Error(LS_INVALID_ARGUMENT, "the rank index needs a complete basis whose symmetry "
                                       "group has no permutation but the identity");
    b.index_kind_request = kind;
    if (b.built) build_index(b);
  });
}

int spined_basis_get_index_kind(const ls_spin_basis *basis) {
  if (!basis) return -1;
  const Basis &b = *basis->p;
  return b.built ? b.index_kind : (b.index_kind_request >= 0 ? b.index_kind_request : 0);
}

int spined_basis_is_complete(const ls_spin_basis *basis) {
  if (!basis) return -1;
  const Basis &b = *basis->p;
  return b.built ? (b.complete ? 1 : 0) : -1;
}

int spined_basis_sector_dimension(const ls_spin_basis *basis, double *out) {
  return guarded([&] {
    if (!basis || !out) throw Error(LS_INVALID_ARGUMENT, "NULL argument");
    double im = 0.0;
    *out = sector_dimension(*basis->p, &im);
  });
}

// Host execution of the device orbit code (same inline functions) so that
// the generic and ring algorithms can be checked against the oracle without
// a GPU.  Not used by any compute path.
int spined_debug_orbit(const ls_spin_basis *basis, int method, uint64_t count,
                       const uint64_t *states, uint64_t *reps, int32_t *phases, int32_t *counts,
                       uint8_t *valid) {
  return guarded([&] {
    if (!basis || (count && (!states || !reps))) throw Error(LS_INVALID_ARGUMENT, "NULL");
    const Basis &b = *basis->p;
    if (method == 1 && !b.tables.ring_available)
      throw Error(LS_INVALID_ARGUMENT, "ring method unavailable for this group");
    auto ga = b.tables.generic_args(false);
    if (method == 2) ga.torus_l1 = 0; // one program per element, no torus factorisation
    const auto ra = b.tables.ring_args(false);
    for (uint64_t i = 0; i < count; ++i) {
      if (method == 3) { // the minimum alone (orbit_min_generic); no phase, count or validity
        reps[i] = orbit_min_generic(ga, states[i]);
        continue;
      }
      const OrbitResult r = method == 1 ? orbit_ring(ra, states[i]) : orbit_generic(ga, states[i]);
      reps[i] = r.rep;
      if (phases) phases[i] = r.phase;
      if (counts) counts[i] = r.count;
      if (valid) valid[i] = r.valid ? 1 : 0;
    }
  });
}

int spined_debug_orbit_hop(const ls_spin_basis *basis, uint64_t count, const uint64_t *states,
                           const int32_t *sites, uint64_t *reps, int32_t *phases, int32_t *counts,
                           uint8_t *valid) {
  return guarded([&] {
    if (!basis || (count && (!states || !sites || !reps))) throw Error(LS_INVALID_ARGUMENT, "NULL");
    const Basis &b = *basis->p;
    if (!b.tables.ring_available) throw Error(LS_INVALID_ARGUMENT, "ring method unavailable");
    const auto ra = b.tables.ring_args(false);
    const int n = ra.n;
    for (uint64_t q = 0; q < count; ++q) {
      const uint64_t r = states[q];
      const int i = sites[q];
      const int j = i + 1 == n ? 0 : i + 1;
      if (i < 0 || i >= n || (((r >> i) ^ (r >> j)) & 1) == 0)
        throw Error(LS_INVALID_ARGUMENT, "sites i and i+1 must hold different spins");
      const RunInfo ri0 = run_info(~r & ra.mask, n, ra.mask);
      const RunInfo ri1 = run_info(r, n, ra.mask);
      const OrbitResult o = orbit_ring_hop(ra, ri0, ri1, r, i);
      reps[q] = o.rep;
      if (phases) phases[q] = o.phase;
      if (counts) counts[q] = o.count;
      if (valid) valid[q] = o.valid ? 1 : 0;
    }
  });
}

int spined_debug_copy_device(void *dst, const void *src, uint64_t bytes) {
  return guarded([&] {
    if (!dst || !src || bytes % 16) throw Error(LS_INVALID_ARGUMENT, "bytes must be a multiple of 16");
    launch_copy_calibration(src, dst, bytes, stream());
    SPINED_HIP(hipStreamSynchronize(stream()));
  });
}

int spined_debug_gather_device(void *dst, uint64_t dst_words, const void *src, uint64_t n_lines,
                               int32_t words) {
  return guarded([&] {
    if (!dst || !src || n_lines == 0 || (n_lines & (n_lines - 1)) || words < 1 || words > 32)
      throw Error(LS_INVALID_ARGUMENT,
                  "n_lines must be a power of two and words per line in [1, 32]");
    if (dst_words < gather_calibration_out_words())
      throw Error(LS_INVALID_ARGUMENT, "dst must hold spined_debug_gather_out_words() words");
    launch_gather_calibration(src, dst, n_lines, words, stream());
    SPINED_HIP(hipStreamSynchronize(stream()));
  });
}

uint64_t spined_debug_gather_out_words(void) { return gather_calibration_out_words(); }

int spined_debug_ring_available(const ls_spin_basis *basis) {
  return basis && basis->p->tables.ring_available ? 1 : 0;
}

int spined_debug_torus_shape(const ls_spin_basis *basis, int32_t *l1, int32_t *l2) {
  return guarded([&] {
    if (!basis || !l1 || !l2) throw Error(LS_INVALID_ARGUMENT, "NULL");
    *l1 = basis->p->tables.torus_l1;
    *l2 = basis->p->tables.torus_l2;
  });
}

} // extern "C"
