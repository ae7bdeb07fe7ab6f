// group.cpp -- symmetry groups, orbit programs and operator compilation
// (host side).  Restates what ls_create_symmetry / ls_create_group /
// ls_create_spin_basis / ls_create_interactionN / ls_create_operator do
// behind src/SpinED/Internal.hs:69-402.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>
#include <unordered_map>

#include "engine.hpp"

namespace spined {

unsigned periodicity_of(const std::vector<unsigned> &perm) {
  const size_t n = perm.size();
  if (n == 0 || n > 64) return 0;
  std::vector<char> seen(n, 0);
  for (unsigned p : perm) {
    if (p >= n || seen[p]) return 0;
    seen[p] = 1;
  }
  std::vector<char> vis(n, 0);
  uint64_t l = 1;
  for (size_t i = 0; i < n; ++i) {
    if (vis[i]) continue;
    unsigned len = 0;
    for (size_t j = i; !vis[j]; j = perm[j]) {
      vis[j] = 1;
      ++len;
    }
    l = l / std::gcd(l, static_cast<uint64_t>(len)) * len;
    if (l > 0xffffffffull) return 0;
  }
  return static_cast<unsigned>(l);
}

namespace {
struct PermHash {
  size_t operator()(const std::vector<uint8_t> &p) const {
    uint64_t h = 0x9ae16a3b2f90404full;
    for (uint8_t v : p) h = mix64(h ^ v);
    return static_cast<size_t>(h);
  }
};
} // namespace

// Closure by breadth-first products gen o e, (gen o e)[i] = e[gen[i]]; a
// product reached twice with different phases means the sectors do not
// form a one-dimensional representation (test/Spec.hs:75-79).
std::shared_ptr<Group> close_group(const std::vector<Symmetry> &gens) {
  auto g = std::make_shared<Group>();
  g->generators = gens;
  if (gens.empty()) {
    g->n = 0;
    g->perms.push_back({});
    g->phases.push_back(0);
    return g;
  }
  const size_t n = gens[0].perm.size();
  int64_t L = 2;
  for (const auto &s : gens) {
    if (s.perm.size() != n)
      throw Error(LS_INCOMPATIBLE_SYMMETRIES, "generators act on different numbers of sites");
    L = L / std::gcd(L, static_cast<int64_t>(s.periodicity)) * s.periodicity;
    if (L > (int64_t(1) << 30))
      throw Error(LS_INCOMPATIBLE_SYMMETRIES, "phase denominator too large");
  }
  g->n = static_cast<int>(n);
  g->L = L;
  std::vector<uint8_t> id(n);
  for (size_t i = 0; i < n; ++i) id[i] = static_cast<uint8_t>(i);
  std::unordered_map<std::vector<uint8_t>, size_t, PermHash> where;
  g->perms.push_back(id);
  g->phases.push_back(0);
  where.emplace(id, 0);
  std::vector<uint8_t> tmp(n);
  for (size_t e = 0; e < g->perms.size(); ++e) {
    for (const auto &s : gens) {
      const auto &cur = g->perms[e];
      for (size_t i = 0; i < n; ++i) tmp[i] = cur[s.perm[i]];
      const int64_t p =
          (g->phases[e] + static_cast<int64_t>(s.sector) * (L / s.periodicity)) % L;
      auto it = where.find(tmp);
      if (it != where.end()) {
        if (g->phases[it->second] != p)
          throw Error(LS_INCOMPATIBLE_SYMMETRIES,
                      "symmetry sectors are incompatible with the group structure");
        continue;
      }
      if (g->perms.size() >= (size_t(1) << 20))
        throw Error(LS_INVALID_ARGUMENT, "symmetry group too large");
      where.emplace(tmp, g->perms.size());
      g->perms.push_back(tmp);
      g->phases.push_back(p);
    }
  }
  return g;
}

GenericOrbitArgs OrbitTables::generic_args(bool device) const {
  GenericOrbitArgs a{};
  if (device) {
    a.masks = masks.ptr;
    a.shifts = shifts.ptr;
    a.offsets = offsets.ptr;
    a.flags = flags.ptr;
    a.phases = phases.ptr;
  } else {
    a.masks = h_masks.data();
    a.shifts = h_shifts.data();
    a.offsets = h_offsets.data();
    a.flags = h_flags.data();
    a.phases = h_phases.data();
  }
  a.n_perm = nperm;
  a.n_sites = n;
  a.site_mask = site_mask_of(n);
  a.has_flip = has_flip;
  a.flip_phase = flip_phase;
  a.L = L;
  a.trivial = trivial_chars ? 1 : 0;
  if (torus_l1) {
    a.torus_l1 = torus_l1;
    a.torus_l2 = torus_l2;
    a.n_point = static_cast<int32_t>(h_point_elem.size());
    a.point_elem = device ? point_elem.ptr : h_point_elem.data();
    a.tp_phase = device ? tp_phase.ptr : h_tp_phase.data();
    a.row_lo = device ? row_lo.ptr : h_row_lo.data();
    a.row_hi = device ? row_hi.ptr : h_row_hi.data();
  }
  return a;
}

RingOrbitArgs OrbitTables::ring_args(bool device) const {
  RingOrbitArgs a{};
  a.phase_tab = device ? ring_phases.ptr : h_ring_phases.data();
  a.n = n;
  a.mask = site_mask_of(n);
  a.has_rev = ring_has_rev;
  a.has_flip = has_flip;
  a.flip_phase = flip_phase;
  a.L = L;
  a.trivial = trivial_chars ? 1 : 0;
  return a;
}

namespace {
// (mask, shift) decomposition of (g.x)_k = src_{q(k)}: the bit at source
// position q(k) moves by k - q(k).
std::map<int, uint64_t> shift_terms(const std::vector<int> &src_of_dest) {
  std::map<int, uint64_t> terms;
  for (size_t k = 0; k < src_of_dest.size(); ++k) {
    const int q = src_of_dest[k];
    terms[static_cast<int>(k) - q] |= 1ull << q;
  }
  return terms;
}

// Torus factorisation of the group (GenericOrbitArgs): the first l1 x l2
// = n (l1, l2 >= 2) whose translations t_{dx,dy}, t[i] = site((x_i + dx)
// mod l1, (y_i + dy) mod l2) with site = l1 y + x, all belong to the group,
// and a transversal of their cosets: every element is p[t[i]] for exactly
// one (p, t).  Taken when it needs fewer instructions than one program per
// element (about 4 per translation against the programs' terms).
void detect_torus(const Group &g, int n, OrbitTables &t) {
  t.torus_l1 = t.torus_l2 = 0;
  t.h_point_elem.clear();
  t.h_tp_phase.clear();
  t.h_row_lo.clear();
  t.h_row_hi.clear();
  if (g.n != n || n < 4 || t.ring_available) return;
  std::map<std::vector<uint8_t>, int> index;
  for (int e = 0; e < t.nperm; ++e) index.emplace(g.perms[e], e);
  for (int l1 = 2; l1 <= n / 2; ++l1) {
    if (n % l1) continue;
    const int l2 = n / l1, nt = n;
    std::vector<std::vector<uint8_t>> tr(nt, std::vector<uint8_t>(n));
    bool all = true;
    for (int dy = 0; dy < l2 && all; ++dy)
      for (int dx = 0; dx < l1 && all; ++dx) {
        auto &p = tr[dy * l1 + dx];
        for (int i = 0; i < n; ++i)
          p[i] = static_cast<uint8_t>(((i / l1 + dy) % l2) * l1 + (i % l1 + dx) % l1);
        all = index.count(p) != 0;
      }
    if (!all || t.nperm % nt) continue;
    std::vector<char> covered(t.nperm, 0);
    std::vector<uint32_t> points;
    std::vector<int32_t> phases;
    bool ok = true;
    for (int e = 0; e < t.nperm && ok; ++e) {
      if (covered[e]) continue;
      points.push_back(static_cast<uint32_t>(e));
      std::vector<uint8_t> c(n);
      for (int k = 0; k < nt && ok; ++k) {
        for (int i = 0; i < n; ++i) c[i] = g.perms[e][tr[k][i]];
        const auto it = index.find(c);
        if (it == index.end() || covered[it->second]) {
          ok = false;
          break;
        }
        covered[it->second] = 1;
        phases.push_back(static_cast<int32_t>(g.phases[it->second]));
      }
    }
    if (!ok || static_cast<int>(points.size()) * nt != t.nperm) continue;
    uint64_t point_terms = 0;
    for (uint32_t e : points) point_terms += t.h_offsets[e + 1] - t.h_offsets[e];
    const uint64_t direct = 2ull * t.total_terms + 2ull * t.nperm;
    const uint64_t torus = 2ull * point_terms + points.size() * (2ull * l1 + 4ull * nt);
    if (torus >= direct) return;
    t.torus_l1 = l1;
    t.torus_l2 = l2;
    t.h_point_elem = points;
    t.h_tp_phase = phases;
    const uint64_t rows = [&] {
      uint64_t r = 0;
      for (int y = 0; y < l2; ++y) r |= 1ull << (y * l1);
      return r;
    }();
    for (int dx = 0; dx < l1; ++dx) {
      const uint64_t lo = (1ull << (l1 - dx)) - 1;  // row bits 0 .. l1-dx-1
      const uint64_t hi = ((1ull << l1) - 1) & ~lo; // row bits l1-dx .. l1-1
      t.h_row_lo.push_back(lo * rows);
      t.h_row_hi.push_back(hi * rows);
    }
    return;
  }
}

template <class T> void upload(DeviceBuffer<T> &d, const std::vector<T> &h) {
  d.allocate(std::max<size_t>(h.size(), 1));
  if (!h.empty())
    SPINED_HIP(hipMemcpy(d.ptr, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
}
} // namespace

void build_orbit_tables(const Group &g, int n, int spin_inversion, OrbitTables &t) {
  t.n = n;
  t.nperm = static_cast<int>(g.perms.size());
  t.has_flip = spin_inversion != 0;
  t.L = static_cast<int32_t>(g.L);
  t.flip_phase = spin_inversion == -1 ? t.L / 2 : 0;
  t.h_masks.clear();
  t.h_shifts.clear();
  t.h_offsets.assign(1, 0);
  t.h_flags.clear();
  t.h_phases.clear();
  t.trivial_chars = t.flip_phase == 0;
  t.real_chars = true;
  for (int e = 0; e < t.nperm; ++e) {
    std::vector<int> direct(n), reversed(n);
    for (int k = 0; k < n; ++k) {
      const int q = g.n == 0 ? k : g.perms[e][k];
      direct[k] = q;
      reversed[k] = n - 1 - q; // position of x_q inside reverse_n(x)
    }
    auto td = shift_terms(direct), tr = shift_terms(reversed);
    const bool use_rev = tr.size() < td.size();
    const auto &terms = use_rev ? tr : td;
    for (const auto &kv : terms) {
      t.h_masks.push_back(kv.second);
      t.h_shifts.push_back(kv.first);
    }
    t.h_offsets.push_back(static_cast<uint32_t>(t.h_masks.size()));
    t.h_flags.push_back(use_rev ? 1u : 0u);
    const int32_t ph = static_cast<int32_t>(g.phases[e]);
    t.h_phases.push_back(ph);
    if (ph != 0) t.trivial_chars = false;
    if (ph != 0 && 2 * ph != t.L) t.real_chars = false;
  }
  t.total_terms = static_cast<uint32_t>(t.h_masks.size());

  // Ring detection: every permutation is a rotation p[i] = (i + k) mod n or
  // a reflection p[i] = (c - i) mod n, and all n rotations are present.
  t.ring_available = false;
  t.ring_has_rev = 0;
  t.h_ring_phases.assign(2 * std::max(n, 1), 0);
  if (n >= 3 && g.n == n) {
    std::vector<int> seen_rot(n, 0), seen_ref(n, 0);
    bool ok = true;
    for (int e = 0; e < t.nperm && ok; ++e) {
      const auto &p = g.perms[e];
      const int k = p[0];
      bool is_rot = true, is_ref = true;
      for (int i = 0; i < n; ++i) {
        if (p[i] != (i + k) % n) is_rot = false;
        if (p[i] != ((k - i) % n + n) % n) is_ref = false;
      }
      if (is_rot) {
        seen_rot[k] = 1;
        t.h_ring_phases[k] = static_cast<int32_t>(g.phases[e]);
      } else if (is_ref) {
        const int c = k; // p[i] = (c - i) mod n  <=>  rotation by (n-1-c) of reverse_n
        const int kk = ((n - 1 - c) % n + n) % n;
        seen_ref[kk] = 1;
        t.h_ring_phases[n + kk] = static_cast<int32_t>(g.phases[e]);
      } else {
        ok = false;
      }
    }
    const int nrot = static_cast<int>(std::count(seen_rot.begin(), seen_rot.end(), 1));
    const int nref = static_cast<int>(std::count(seen_ref.begin(), seen_ref.end(), 1));
    if (ok && nrot == n && (nref == 0 || nref == n) && nrot + nref == t.nperm) {
      t.ring_available = true;
      t.ring_has_rev = nref == n ? 1 : 0;
    }
  }

  detect_torus(g, n, t);

  // stabiliser sizes divide |G|: tabulate sqrt(c / s) over its divisors,
  // computed exactly as the row gather would (IEEE division and sqrt)
  const int64_t G = static_cast<int64_t>(t.nperm) * (t.has_flip ? 2 : 1);
  t.nd = 0;
  t.h_div_idx.clear();
  t.h_ratio.clear();
  if (G <= 8192) {
    std::vector<int64_t> divs;
    for (int64_t d = 1; d <= G; ++d)
      if (G % d == 0) divs.push_back(d);
    t.nd = static_cast<int>(divs.size());
    t.h_div_idx.assign(G + 1, 0);
    for (int i = 0; i < t.nd; ++i) t.h_div_idx[divs[i]] = static_cast<uint16_t>(i);
    t.h_ratio.resize(static_cast<size_t>(t.nd) * t.nd);
    for (int i = 0; i < t.nd; ++i)
      for (int j = 0; j < t.nd; ++j)
        t.h_ratio[static_cast<size_t>(i) * t.nd + j] =
            std::sqrt(static_cast<double>(divs[i]) / static_cast<double>(divs[j]));
  }

}

void ensure_device(OrbitTables &t) {
  if (t.on_device) return;
  upload(t.masks, t.h_masks);
  upload(t.shifts, t.h_shifts);
  upload(t.offsets, t.h_offsets);
  upload(t.flags, t.h_flags);
  upload(t.phases, t.h_phases);
  upload(t.ring_phases, t.h_ring_phases);
  upload(t.point_elem, t.h_point_elem);
  upload(t.tp_phase, t.h_tp_phase);
  upload(t.row_lo, t.h_row_lo);
  upload(t.row_hi, t.h_row_hi);
  upload(t.div_idx, t.h_div_idx);
  upload(t.ratio, t.h_ratio);
  t.on_device = true;
}

// (1/|G|) sum_g chi(g)^* Fix(g): the trace of the sector projector.  Fix of a
// permutation with cycles c_j is the t^w coefficient of prod (1 + t^|c_j|)
// (2^#cycles without a weight); with spin inversion it is 2^#cycles when all
// cycles are even and w = n/2 (or no weight), else 0.
double sector_dimension(const Basis &b, double *imag) {
  const int n = b.n;
  const auto &g = *b.group;
  const int64_t L = b.tables.L;
  long double re = 0.0L, im = 0.0L;
  __int128 exact = 0;
  const bool real = b.tables.real_chars;
  const int nflip = b.z != 0 ? 2 : 1;
  for (size_t e = 0; e < g.perms.size(); ++e) {
    std::vector<int> lens;
    std::vector<char> vis(n, 0);
    for (int i = 0; i < n; ++i) {
      if (vis[i]) continue;
      int len = 0;
      for (int j = i; !vis[j]; j = g.n == 0 ? j : g.perms[e][j]) {
        vis[j] = 1;
        ++len;
        if (g.n == 0) break;
      }
      lens.push_back(len);
    }
    for (int f = 0; f < nflip; ++f) {
      __int128 fix = 0;
      if (f == 0) {
        if (b.w < 0) {
          fix = static_cast<__int128>(1) << lens.size();
        } else {
          std::vector<__int128> poly(n + 1, 0);
          poly[0] = 1;
          for (int len : lens)
            for (int d = n; d >= len; --d) poly[d] += poly[d - len];
          fix = poly[b.w];
        }
      } else {
        bool even = true;
        for (int len : lens) even = even && (len % 2 == 0);
        if (even && (b.w < 0 || 2 * b.w == n)) fix = static_cast<__int128>(1) << lens.size();
      }
      int64_t ph = g.phases[e] + (f ? b.tables.flip_phase : 0);
      ph %= L;
      if (real) {
        exact += ph == 0 ? fix : -fix;
      } else {
        const long double ang = -2.0L * 3.14159265358979323846264338327950288L *
                                static_cast<long double>(ph) / static_cast<long double>(L);
        re += static_cast<long double>(fix) * cosl(ang);
        im += static_cast<long double>(fix) * sinl(ang);
      }
    }
  }
  const long double G = static_cast<long double>(g.perms.size() * nflip);
  if (real) {
    if (imag) *imag = 0.0;
    return static_cast<double>(static_cast<long double>(exact) / G);
  }
  if (imag) *imag = static_cast<double>(im / G);
  return static_cast<double>(re / G);
}

// Operator compilation: tuples in term order; for each tuple and local
// state a the diagonal entry M[a][a] and the templates b != a with
// M[b][a] != 0, carrying the XOR mask that turns a into b and the
// coefficient conj(M[b][a]) of the row gather.
void compile_operator(Operator &op) {
  auto &t = op.tables;
  t = OperatorTables{};
  t.is_real = op.basis->tables.real_chars;
  t.h_range.clear();
  for (const auto &term : op.terms) {
    if (!term->is_real) t.is_real = false;
    const int n = term->n;
    const int dim = 1 << n;
    t.max_local = std::max(t.max_local, dim);
    for (const auto &s : term->sites) {
      uint32_t packed = 0;
      for (int k = 0; k < n; ++k) {
        if (s[k] >= op.basis->n)
          throw Error(LS_INVALID_ARGUMENT, "interaction site index out of range");
        packed |= static_cast<uint32_t>(s[k]) << (8 * k);
      }
      t.h_tuple_sites.push_back(packed);
      t.h_tuple_n.push_back(static_cast<uint32_t>(n));
      t.h_tuple_off.push_back(static_cast<uint32_t>(t.h_diag_re.size()));
      for (int a = 0; a < dim; ++a) {
        const cplx d = term->matrix[a * dim + a];
        t.h_diag_re.push_back(d.real());
        t.h_diag_im.push_back(-d.imag());
      }
      for (int a = 0; a < dim; ++a) {
        t.h_range.push_back(static_cast<uint32_t>(t.h_flip.size()));
        for (int b = 0; b < dim; ++b) {
          if (b == a) continue;
          const cplx m = term->matrix[b * dim + a];
          if (m == cplx(0.0, 0.0)) continue;
          if (__builtin_popcount(static_cast<unsigned>(a)) !=
              __builtin_popcount(static_cast<unsigned>(b)))
            t.keeps_weight = false;
          uint64_t fl = 0;
          for (int k = 0; k < n; ++k) {
            const int bit = n - 1 - k; // first site = most significant bit
            if (((a ^ b) >> bit) & 1) fl |= 1ull << s[k];
          }
          t.h_flip.push_back(fl);
          t.h_coef_re.push_back(m.real());
          t.h_coef_im.push_back(-m.imag());
        }
      }
      t.h_range.push_back(static_cast<uint32_t>(t.h_flip.size()));
    }
  }
  t.ntuples = static_cast<int>(t.h_tuple_sites.size());
  t.ntemplates = static_cast<int>(t.h_flip.size());
  // the range table holds 2^n + 1 entries per tuple: the entries of tuple i
  // start at tuple_off[i] + i.

  // Exchange-bond form: every tuple is a 2-site real matrix with
  // M00 = M33 (parallel), M11 = M22 (antiparallel), M12 = M21 (swap) and no
  // other off-diagonal entry -- Heisenberg and XXZ bonds of any strength.
  // Then the diagonal is one select per bond and the (single) active
  // template of an antiparallel bond is the bond itself, in tuple order.
  t.bonds = t.is_real && t.ntuples > 0;
  t.h_bond_sites.clear();
  t.h_bond_dp.clear();
  t.h_bond_da.clear();
  t.h_bond_o.clear();
  t.h_bond_flip.clear();
  for (const auto &term : op.terms) {
    if (!t.bonds) break;
    if (term->n != 2 || !term->is_real) {
      t.bonds = false;
      break;
    }
    const auto &M = term->matrix;
    auto at = [&](int r, int c) { return M[r * 4 + c].real(); };
    bool ok = at(0, 0) == at(3, 3) && at(1, 1) == at(2, 2) && at(1, 2) == at(2, 1);
    for (int r = 0; r < 4 && ok; ++r)
      for (int c = 0; c < 4 && ok; ++c) {
        if (r == c) continue;
        if ((r == 1 && c == 2) || (r == 2 && c == 1)) continue;
        if (at(r, c) != 0.0) ok = false;
      }
    if (!ok) {
      t.bonds = false;
      break;
    }
    for (const auto &s : term->sites) {
      t.h_bond_sites.push_back(static_cast<uint32_t>(s[0]) | (static_cast<uint32_t>(s[1]) << 8));
      t.h_bond_dp.push_back(at(0, 0));
      t.h_bond_da.push_back(at(1, 1));
      t.h_bond_o.push_back(at(1, 2));
      t.h_bond_flip.push_back((1ull << s[0]) | (1ull << s[1]));
    }
  }
  if (!t.bonds) {
    t.h_bond_sites.clear();
    t.h_bond_dp.clear();
    t.h_bond_da.clear();
    t.h_bond_o.clear();
    t.h_bond_flip.clear();
  }
  // nearest-neighbour ring bonds: the row gather may take the hop orbit path
  t.nn_ring = t.bonds && !t.h_bond_sites.empty();
  t.h_bond_lo.clear();
  const int n = op.basis ? op.basis->n : 0;
  for (uint32_t bs : t.h_bond_sites) {
    if (!t.nn_ring) break;
    const int s0 = static_cast<int>(bs & 0xffu), s1 = static_cast<int>(bs >> 8);
    if (n >= 3 && s1 == (s0 + 1) % n)
      t.h_bond_lo.push_back(static_cast<uint8_t>(s0));
    else if (n >= 3 && s0 == (s1 + 1) % n)
      t.h_bond_lo.push_back(static_cast<uint8_t>(s1));
    else
      t.nn_ring = false;
  }
  if (!t.nn_ring) t.h_bond_lo.clear();
  t.nn_identity = t.nn_ring;
  for (size_t k = 0; k < t.h_bond_lo.size(); ++k)
    t.nn_identity = t.nn_identity && t.h_bond_lo[k] == k;
}

void ensure_device(OperatorTables &t) {
  if (t.on_device) return;
  upload(t.tuple_sites, t.h_tuple_sites);
  upload(t.tuple_n, t.h_tuple_n);
  upload(t.tuple_off, t.h_tuple_off);
  upload(t.range, t.h_range);
  upload(t.diag_re, t.h_diag_re);
  upload(t.diag_im, t.h_diag_im);
  upload(t.flip, t.h_flip);
  upload(t.coef_re, t.h_coef_re);
  upload(t.coef_im, t.h_coef_im);
  upload(t.bond_sites, t.h_bond_sites);
  upload(t.bond_dp, t.h_bond_dp);
  upload(t.bond_da, t.h_bond_da);
  upload(t.bond_o, t.h_bond_o);
  upload(t.bond_flip, t.h_bond_flip);
  upload(t.bond_lo, t.h_bond_lo);
  t.on_device = true;
}

int Operator::orbit_method() const {
  const auto &bt = basis->tables;
  if (orbit_method_override == 0) return 0;
  if (orbit_method_override == 1) return bt.ring_available ? 1 : 0;
  return bt.ring_available ? 1 : 0;
}

} // namespace spined
