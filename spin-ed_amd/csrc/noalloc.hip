// noalloc.hip -- spined_operator_expectation_device_noalloc: the device
// expectation <x_c|H|x_c> (spined_operator_expectation_device) for callers
// that cannot allocate device memory (code under HIP graph capture, tight
// serving loops, harnesses that forbid allocation).
//
// The dot's partial sums and the per-column results live in `scratch`, a
// file-scope device array of fixed size (kScratchDoubles doubles, 33 KB, held
// by every GPU the library runs on until the process ends).  Host code never
// passes the array's name to a kernel: the name is a host-side stand-in, not
// a device address.  The first call on each device, with that device
// current, looks up the array's address there (hipGetSymbolAddress) and keeps
// it; kernels and copies get that pointer.  Every call on one GPU shares that
// GPU's array: calls that overlap on one GPU race.  Each partial slot is
// written before it is read, so nothing has to start at zero.
//
// The call allocates nothing: it refuses an operator whose basis or tables
// are not on the device yet (one ls_operator_matmat or spined_eigsh sets
// them up), uses the gather cache only when it is already built, and runs
// the row schedule only when it already exists.
#include <mutex>
#include <vector>

#include "handles.hpp"

namespace spined {
namespace {

constexpr uint64_t kMaxColumns = 64;
constexpr uint64_t kScratchDoubles = 2 * static_cast<uint64_t>(kVdotMaxBlocks) + 2 * kMaxColumns;

__device__ __align__(256) double scratch[kScratchDoubles];

std::mutex g_scratch_mu;
std::vector<double *> g_scratch_ptr; // per device: the address of `scratch` there

double *scratch_on_current_device() {
  const int d = current_device();
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (static_cast<int>(g_scratch_ptr.size()) <= d) g_scratch_ptr.resize(d + 1, nullptr);
  if (!g_scratch_ptr[d]) {
    void *p = nullptr;
    SPINED_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw Error(LS_DEVICE_ERROR, "hipGetSymbolAddress returned no address for scratch");
    g_scratch_ptr[d] = static_cast<double *>(p);
  }
  return g_scratch_ptr[d];
}

VecType vec_type_of(int dtype) {
  switch (dtype) {
  case LS_FLOAT32: return VecType::F32;
  case LS_FLOAT64: return VecType::F64;
  case LS_COMPLEX64: return VecType::C64;
  case LS_COMPLEX128: return VecType::C128;
  default: throw Error(LS_INVALID_DATATYPE, "datatype must be 0..3");
  }
}

size_t elem_bytes(int dtype) { return dtype == LS_FLOAT32 ? 4 : dtype == LS_COMPLEX128 ? 16 : 8; }

} // namespace
} // namespace spined

using namespace spined;

extern "C" int spined_operator_expectation_device_noalloc(const ls_operator *op, int dtype,
                                                          uint64_t size, uint64_t block_size,
                                                          const void *x, uint64_t x_stride,
                                                          void *y_scratch, double *out) {
  return guarded([&] {
    if (block_size > kMaxColumns)
      throw Error(LS_INVALID_ARGUMENT, "at most " + std::to_string(kMaxColumns) +
                                           " columns per call that allocates nothing");
    if (!op || !op->p) throw Error(LS_INVALID_ARGUMENT, "operator is NULL");
    const VecType vt = vec_type_of(dtype);
    Operator &o = *op->p;
    const Basis &b = *o.basis;
    if (!b.built) throw Error(LS_CACHE_NOT_BUILT, "basis has not been built");
    if (!b.tables.on_device || !o.tables.on_device)
      throw Error(LS_INVALID_ARGUMENT, "operator tables not on the device yet: apply the operator "
                                       "once (ls_operator_matmat) before a call that allocates "
                                       "nothing");
    if (size != b.D) throw Error(LS_DIMENSION_MISMATCH, "size differs from dimension");
    if (block_size == 0) return;
    if (!x || !out || !y_scratch) throw Error(LS_INVALID_ARGUMENT, "NULL argument");
    if (x_stride < size) throw Error(LS_INVALID_ARGUMENT, "stride smaller than size");
    double *sp = scratch_on_current_device();
    double *partial = sp;                                   // 2 * kVdotMaxBlocks
    double *res = sp + 2 * static_cast<uint64_t>(kVdotMaxBlocks); // 2 * kMaxColumns
    const unsigned nb = vdot_blocks(size);
    if (nb > kVdotMaxBlocks) throw Error(LS_INTERNAL_ERROR, "dot grid exceeds the scratch array");
    const size_t es = elem_bytes(dtype);
    hipStream_t s = stream();
    for (uint64_t c = 0; c < block_size; ++c) {
      const char *xc = static_cast<const char *>(x) + c * x_stride * es;
      MatvecCall m{&o, vt, vt, xc, y_scratch, size, s};
      m.no_alloc = true;
      launch_matvec(m);
      launch_vdot(vt, xc, y_scratch, size, partial, nb, res + 2 * c, s);
    }
    SPINED_HIP(hipMemcpyAsync(out, res, 16 * block_size, hipMemcpyDeviceToHost, s));
    SPINED_HIP(hipStreamSynchronize(s));
  });
}
