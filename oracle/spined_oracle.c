/*
 * spined_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY; see the header).
 *
 * Plain C11 + optional OpenMP.  Every function states which reference
 * interface it restates.  The arithmetic itself lives in lattice-symmetries
 * and PRIMME, which the reference fetches un-pinned from git
 * (configure:21-33, cabal.project:13-15) and which are absent here; the
 * algorithm below is restated from their documented behaviour and pinned by
 * the checks in tests/ (closed forms, brute force, golden fixtures).
 *
 * Pins per function (what makes a plausible mistake fail a test):
 *   orc_periodicity / orc_create  -- test/Spec.hs:38-43, :72-79 semantics
 *                                    (periodicity 2, invalid perm, sector
 *                                    range, group sizes 8/16, inconsistent
 *                                    characters rejected); brute-force closure.
 *   orc_sector_dimension          -- equals the brute-force basis size for many
 *                                    sectors; known dimensions 28,968 (N=24),
 *                                    15,804,956 (6x6), 861,725,794 (N=40).
 *   orc_build                     -- brute force over all states in Python for
 *                                    N<=12; sum_r 1/N_r^2 = C(N,w) identity.
 *   orc_matvec_*                  -- dense projected Hamiltonian built from
 *                                    explicit permutation matrices (numpy,
 *                                    N<=10, real and complex sectors); spectra
 *                                    of all sectors union to the full spectrum.
 *   orc_tridiag_eigh              -- scipy.linalg.eigh_tridiagonal (LAPACK).
 *   orc_lanczos                   -- numpy.linalg.eigh on the dense sector
 *                                    matrix; E0 = -8 (README.md:83-95);
 *                                    Bethe-ansatz E0 for N=10, 24.
 *   orc_operator_expectation      -- parity unpinned (normalisation of
 *                                    ls_operator_expectation is not fixed by the
 *                                    reference; <x|H|x> without division).
 */
#include "spined_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_MAX_GROUP (1 << 20)

typedef struct {
  int n;                   /* sites in the tuple: 1..4 */
  unsigned sites[4];
  int interaction;         /* index into matrices */
} orc_tuple;

typedef struct {
  int n;
  double *m;               /* 2^n x 2^n complex interleaved row-major */
} orc_matrix;

struct orc_system {
  int n, w, z;
  int gperm;               /* group size without spin inversion */
  int G;                   /* full group size */
  int64_t L;               /* phase denominator */
  unsigned *perm;          /* G x n */
  int *flip;
  int64_t *phase;
  uint64_t *lut;           /* G x nbytes x 256 */
  int nbytes;
  uint64_t mask;
  /* basis */
  int built;
  int64_t D;
  uint64_t *reps;
  uint32_t *stab;
  /* operator */
  int ntuples;
  orc_tuple *tuples;
  int nmat;
  orc_matrix *mats;
};

/* ------------------------------------------------------------------------ */
/* small helpers                                                             */

static int64_t gcd64(int64_t a, int64_t b) {
  while (b) {
    int64_t t = a % b;
    a = b;
    b = t;
  }
  return a < 0 ? -a : a;
}

static uint64_t orc_mix64(uint64_t z) {
  /* SplitMix64 finaliser (Steele, Lea & Flood, OOPSLA 2014). */
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

void orc_random_vector(uint64_t seed, int64_t n, double *v) {
  for (int64_t r = 0; r < n; ++r) {
    uint64_t h = orc_mix64(seed ^ orc_mix64((uint64_t)r));
    double u = (double)(h >> 11) * 0x1.0p-53;
    v[r] = 2.0 * u - 1.0;
  }
}

/* Restates ls_create_symmetry's validation (Internal.hs:69-116; semantics of
 * test/Spec.hs:38-43): p must be a permutation of 0..n-1; the periodicity is
 * the least m >= 1 with p^m = id. */
unsigned orc_periodicity(unsigned n, const unsigned *perm) {
  if (n == 0 || n > 64) return 0;
  unsigned char seen[64] = {0};
  for (unsigned i = 0; i < n; ++i) {
    if (perm[i] >= n || seen[perm[i]]) return 0;
    seen[perm[i]] = 1;
  }
  /* lcm of cycle lengths */
  unsigned char visited[64] = {0};
  uint64_t l = 1;
  for (unsigned i = 0; i < n; ++i) {
    if (visited[i]) continue;
    unsigned len = 0, j = i;
    while (!visited[j]) {
      visited[j] = 1;
      j = perm[j];
      ++len;
    }
    l = l / (uint64_t)gcd64((int64_t)l, len) * len;
    if (l > 0xffffffffull) return 0;
  }
  return (unsigned)l;
}

/* ------------------------------------------------------------------------ */
/* group closure                                                             */

static uint64_t perm_hash(const unsigned *p, int n) {
  uint64_t h = 0x12345678abcdefull;
  for (int i = 0; i < n; ++i) h = orc_mix64(h ^ ((uint64_t)p[i] + ((uint64_t)i << 8)));
  return h;
}

static void build_lut(orc_system *s) {
  int n = s->n;
  s->nbytes = (n + 7) / 8;
  s->lut = (uint64_t *)calloc((size_t)s->G * s->nbytes * 256, sizeof(uint64_t));
  for (int g = 0; g < s->G; ++g) {
    const unsigned *p = s->perm + (size_t)g * n;
    /* (g.x)_k = x_{p[k]}: bit i of x lands on site k with p[k] = i */
    int inv[64];
    for (int k = 0; k < n; ++k) inv[p[k]] = k;
    for (int j = 0; j < s->nbytes; ++j) {
      uint64_t *t = s->lut + ((size_t)g * s->nbytes + j) * 256;
      for (int v = 0; v < 256; ++v) {
        uint64_t img = 0;
        for (int b = 0; b < 8; ++b) {
          int i = 8 * j + b;
          if (i < n && ((v >> b) & 1)) img |= 1ull << inv[i];
        }
        t[v] = img;
      }
    }
  }
}

uint64_t orc_apply(const orc_system *s, int g, uint64_t x) {
  const uint64_t *t = s->lut + (size_t)g * s->nbytes * 256;
  uint64_t y = 0;
  for (int j = 0; j < s->nbytes; ++j) y |= t[j * 256 + ((x >> (8 * j)) & 0xff)];
  if (s->flip[g]) y ^= s->mask;
  return y;
}

/* Restates ls_create_group (Internal.hs:120-150) and the spin-inversion part
 * of ls_create_spin_basis (Internal.hs:172-228): closure of the generators
 * with exact phases; inconsistent characters are an error
 * (test/Spec.hs:75-79). */
orc_system *orc_create(int n_sites, int hamming_weight, int spin_inversion, int n_gens,
                       const unsigned *perms, const unsigned *sectors, int *status) {
  *status = ORC_OK;
  if (n_sites <= 0 || n_sites > 64) {
    *status = ORC_INVALID_NUMBER_SPINS;
    return NULL;
  }
  if (hamming_weight < -1 || hamming_weight > n_sites) {
    *status = ORC_INVALID_HAMMING_WEIGHT;
    return NULL;
  }
  if (spin_inversion < -1 || spin_inversion > 1) {
    *status = ORC_INVALID_SPIN_INVERSION;
    return NULL;
  }
  if (spin_inversion != 0 && hamming_weight != -1 && 2 * hamming_weight != n_sites) {
    *status = ORC_INVALID_SPIN_INVERSION;
    return NULL;
  }
  int n = n_sites;
  int64_t L = 2;
  unsigned *per = (unsigned *)malloc(sizeof(unsigned) * (n_gens > 0 ? n_gens : 1));
  for (int j = 0; j < n_gens; ++j) {
    per[j] = orc_periodicity((unsigned)n, perms + (size_t)j * n);
    if (per[j] == 0) {
      free(per);
      *status = ORC_INVALID_PERMUTATION;
      return NULL;
    }
    if (sectors[j] >= per[j]) {
      free(per);
      *status = ORC_INVALID_SECTOR;
      return NULL;
    }
    L = L / gcd64(L, per[j]) * per[j];
    if (L > (1ll << 40)) {
      free(per);
      *status = ORC_INCOMPATIBLE_SYMMETRIES;
      return NULL;
    }
  }
  /* BFS closure: elements e, products gen o e with (gen o e)[i] = e[gen[i]] */
  size_t cap = 64;
  unsigned *elems = (unsigned *)malloc(sizeof(unsigned) * n * cap);
  int64_t *ph = (int64_t *)malloc(sizeof(int64_t) * cap);
  size_t count = 1;
  for (int i = 0; i < n; ++i) elems[i] = (unsigned)i;
  ph[0] = 0;
  size_t hcap = 1 << 10;
  int *htab = (int *)malloc(sizeof(int) * hcap);
  for (size_t i = 0; i < hcap; ++i) htab[i] = -1;
  htab[perm_hash(elems, n) & (hcap - 1)] = 0;
  unsigned tmp[64];
  for (size_t e = 0; e < count; ++e) {
    for (int j = 0; j < n_gens; ++j) {
      const unsigned *gen = perms + (size_t)j * n;
      for (int i = 0; i < n; ++i) tmp[i] = elems[e * n + gen[i]];
      int64_t p = (ph[e] + (int64_t)sectors[j] * (L / per[j])) % L;
      uint64_t h = perm_hash(tmp, n);
      size_t pos = h & (hcap - 1);
      int found = -1;
      while (htab[pos] != -1) {
        if (memcmp(elems + (size_t)htab[pos] * n, tmp, sizeof(unsigned) * n) == 0) {
          found = htab[pos];
          break;
        }
        pos = (pos + 1) & (hcap - 1);
      }
      if (found >= 0) {
        if (ph[found] != p) {
          free(per); free(elems); free(ph); free(htab);
          *status = ORC_INCOMPATIBLE_SYMMETRIES;
          return NULL;
        }
        continue;
      }
      if (count >= ORC_MAX_GROUP) {
        free(per); free(elems); free(ph); free(htab);
        *status = ORC_OUT_OF_MEMORY;
        return NULL;
      }
      if (count == cap) {
        cap *= 2;
        elems = (unsigned *)realloc(elems, sizeof(unsigned) * n * cap);
        ph = (int64_t *)realloc(ph, sizeof(int64_t) * cap);
      }
      memcpy(elems + count * n, tmp, sizeof(unsigned) * n);
      ph[count] = p;
      htab[pos] = (int)count;
      ++count;
      if (2 * count > hcap) { /* grow and rehash */
        size_t ncap = hcap * 4;
        int *nt = (int *)malloc(sizeof(int) * ncap);
        for (size_t i = 0; i < ncap; ++i) nt[i] = -1;
        for (size_t q = 0; q < count; ++q) {
          size_t pp = perm_hash(elems + q * n, n) & (ncap - 1);
          while (nt[pp] != -1) pp = (pp + 1) & (ncap - 1);
          nt[pp] = (int)q;
        }
        free(htab);
        htab = nt;
        hcap = ncap;
      }
    }
  }
  free(htab);
  free(per);

  orc_system *s = (orc_system *)calloc(1, sizeof(orc_system));
  s->n = n;
  s->w = hamming_weight;
  s->z = spin_inversion;
  s->gperm = (int)count;
  s->G = (int)count * (spin_inversion != 0 ? 2 : 1);
  s->L = L;
  s->mask = n == 64 ? ~0ull : ((1ull << n) - 1);
  s->perm = (unsigned *)malloc(sizeof(unsigned) * n * s->G);
  s->flip = (int *)malloc(sizeof(int) * s->G);
  s->phase = (int64_t *)malloc(sizeof(int64_t) * s->G);
  for (int g = 0; g < s->G; ++g) {
    size_t src = (size_t)(g % count);
    memcpy(s->perm + (size_t)g * n, elems + src * n, sizeof(unsigned) * n);
    s->flip[g] = g >= (int)count;
    s->phase[g] = (ph[src] + ((g >= (int)count && spin_inversion == -1) ? L / 2 : 0)) % L;
  }
  free(elems);
  free(ph);
  build_lut(s);
  return s;
}

void orc_destroy(orc_system *s) {
  if (!s) return;
  free(s->perm);
  free(s->flip);
  free(s->phase);
  free(s->lut);
  free(s->reps);
  free(s->stab);
  for (int i = 0; i < s->nmat; ++i) free(s->mats[i].m);
  free(s->mats);
  free(s->tuples);
  free(s);
}

int orc_group_size(const orc_system *s) { return s->gperm; }
int orc_full_group_size(const orc_system *s) { return s->G; }
int orc_phase_denominator(const orc_system *s) { return (int)s->L; }

void orc_get_elements(const orc_system *s, unsigned *perms, int *flips, int *phases) {
  if (perms) memcpy(perms, s->perm, sizeof(unsigned) * s->n * s->G);
  for (int g = 0; g < s->G; ++g) {
    if (flips) flips[g] = s->flip[g];
    if (phases) phases[g] = (int)s->phase[g];
  }
}

/* Orbit minimum with the coset-phase test.  For rep = min_g g.x, the
 * elements reaching rep form the coset Stab(rep) h; the symmetric basis
 * state of rep exists iff all of them carry the same phase (the character
 * is trivial on the stabiliser), and that phase is chi(h). */
void orc_orbit(const orc_system *s, uint64_t x, uint64_t *rep, int *phase, int *count,
               int *valid) {
  uint64_t best = ~0ull;
  int64_t bp = 0;
  int c = 0, ok = 1;
  for (int g = 0; g < s->G; ++g) {
    uint64_t y = orc_apply(s, g, x);
    if (y < best) {
      best = y;
      bp = s->phase[g];
      c = 1;
      ok = 1;
    } else if (y == best) {
      ++c;
      if (s->phase[g] != bp) ok = 0;
    }
  }
  *rep = best;
  *phase = (int)bp;
  *count = c;
  *valid = ok;
}

/* ------------------------------------------------------------------------ */
/* sector dimension by character counting                                    */

/* exact binomial coefficients C(n, k) for n <= 64 (Pascal's rule) */
static uint64_t binom_u(int n, int k) {
  static uint64_t table[65][65];
  static int ready = 0;
  if (!ready) {
    for (int a = 0; a <= 64; ++a) {
      table[a][0] = 1;
      for (int b = 1; b <= a; ++b) table[a][b] = table[a - 1][b - 1] + (b <= a - 1 ? table[a - 1][b] : 0);
    }
    ready = 1;
  }
  if (k < 0 || k > n || n > 64) return 0;
  return table[n][k];
}

double orc_sector_dimension(const orc_system *s, double *imag) {
  int n = s->n;
  double re = 0.0, im = 0.0;
  /* exact path for real characters */
  int real_chars = 1;
  for (int g = 0; g < s->G; ++g)
    if (s->phase[g] != 0 && 2 * s->phase[g] != s->L) real_chars = 0;
  __int128 acc = 0;
  for (int g = 0; g < s->G; ++g) {
    const unsigned *p = s->perm + (size_t)g * n;
    unsigned char vis[64] = {0};
    int lens[64], nc = 0;
    for (int i = 0; i < n; ++i) {
      if (vis[i]) continue;
      int len = 0, j = i;
      while (!vis[j]) {
        vis[j] = 1;
        j = (int)p[j];
        ++len;
      }
      lens[nc++] = len;
    }
    __int128 fix = 0;
    if (!s->flip[g]) {
      if (s->w < 0) {
        fix = (__int128)1 << nc;
      } else {
        /* coefficient of t^w in prod (1 + t^len) */
        __int128 poly[65];
        memset(poly, 0, sizeof(poly));
        poly[0] = 1;
        for (int c = 0; c < nc; ++c)
          for (int d = n; d >= lens[c]; --d) poly[d] += poly[d - lens[c]];
        fix = poly[s->w];
      }
    } else {
      int all_even = 1;
      for (int c = 0; c < nc; ++c)
        if (lens[c] % 2) all_even = 0;
      if (all_even && (s->w < 0 || 2 * s->w == n)) fix = (__int128)1 << nc;
    }
    if (real_chars) {
      acc += (s->phase[g] == 0) ? fix : -fix;
    } else {
      double ang = -2.0 * M_PI * (double)s->phase[g] / (double)s->L;
      re += (double)fix * cos(ang);
      im += (double)fix * sin(ang);
    }
  }
  if (real_chars) {
    if (imag) *imag = 0.0;
    return (double)(acc / s->G) + (double)(acc % s->G) / s->G;
  }
  if (imag) *imag = im / s->G;
  return re / s->G;
}

/* ------------------------------------------------------------------------ */
/* basis build (ls_build, Internal.hs:178-182 / SpinED.hs:315-337)           */

static uint64_t gosper_next(uint64_t v) {
  uint64_t t = v | (v - 1);
  return (t + 1) | (((~t & -~t) - 1) >> (__builtin_ctzll(v) + 1));
}

/* Test whether x is a representative; returns stabiliser size or 0. */
static uint32_t rep_test(const orc_system *s, uint64_t x) {
  uint32_t c = 0;
  for (int g = 0; g < s->G; ++g) {
    uint64_t y = orc_apply(s, g, x);
    if (y < x) return 0;
    if (y == x) {
      if (s->phase[g] != 0) return 0;
      ++c;
    }
  }
  return c;
}

typedef struct {
  uint64_t *r;
  uint32_t *c;
  int64_t n, cap;
} orc_buf;

static void buf_push(orc_buf *b, uint64_t r, uint32_t c) {
  if (b->n == b->cap) {
    b->cap = b->cap ? 2 * b->cap : 1024;
    b->r = (uint64_t *)realloc(b->r, sizeof(uint64_t) * b->cap);
    b->c = (uint32_t *)realloc(b->c, sizeof(uint32_t) * b->cap);
  }
  b->r[b->n] = r;
  b->c[b->n] = c;
  ++b->n;
}

/* Colex unranking of a weight-w combination (rank counted in ascending
 * integer order, which Gosper's successor follows). */
static uint64_t unrank_combination(int n, int w, uint64_t rank) {
  uint64_t x = 0;
  for (int k = w; k >= 1; --k) {
    int c = k - 1;
    while (c + 1 < n && binom_u(c + 1, k) <= rank) ++c;
    /* c is largest with C(c, k) <= rank */
    rank -= binom_u(c, k);
    x |= 1ull << c;
  }
  return x;
}

int64_t orc_build_range(orc_system *s, uint64_t first, uint64_t last) {
  free(s->reps);
  free(s->stab);
  s->reps = NULL;
  s->stab = NULL;
  int n = s->n;
  uint64_t total;
  if (s->w < 0)
    total = n == 64 ? ~0ull : (1ull << n);
  else
    total = binom_u(n, s->w);
  int nthreads = 1;
#ifdef _OPENMP
  nthreads = omp_get_max_threads();
#endif
  int nchunks = nthreads * 8;
  orc_buf *bufs = (orc_buf *)calloc(nchunks, sizeof(orc_buf));
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (int ch = 0; ch < nchunks; ++ch) {
    uint64_t base = total / (uint64_t)nchunks, rem = total % (uint64_t)nchunks;
    uint64_t uch = (uint64_t)ch;
    uint64_t lo = base * uch + (uch < rem ? uch : rem);
    uint64_t cnt = base + (uch < rem ? 1 : 0);
    if (cnt == 0) continue;
    uint64_t x = s->w < 0 ? lo : unrank_combination(n, s->w, lo);
    for (uint64_t i = 0; i < cnt; ++i) {
      if (x >= first && x <= last) {
        uint32_t c = rep_test(s, x);
        if (c) buf_push(&bufs[ch], x, c);
      }
      if (i + 1 < cnt) x = s->w < 0 ? x + 1 : gosper_next(x);
    }
  }
  int64_t D = 0;
  for (int ch = 0; ch < nchunks; ++ch) D += bufs[ch].n;
  s->reps = (uint64_t *)malloc(sizeof(uint64_t) * (D ? D : 1));
  s->stab = (uint32_t *)malloc(sizeof(uint32_t) * (D ? D : 1));
  int64_t off = 0;
  for (int ch = 0; ch < nchunks; ++ch) {
    if (bufs[ch].n) {
      memcpy(s->reps + off, bufs[ch].r, sizeof(uint64_t) * bufs[ch].n);
      memcpy(s->stab + off, bufs[ch].c, sizeof(uint32_t) * bufs[ch].n);
    }
    off += bufs[ch].n;
    free(bufs[ch].r);
    free(bufs[ch].c);
  }
  free(bufs);
  s->D = D;
  s->built = 1;
  return D;
}

int64_t orc_build(orc_system *s) { return orc_build_range(s, 0, ~0ull); }

/* Restates ls_build_unsafe (Internal.hs:181-182, :230-236): adopt the given
 * representatives without validation; stabiliser sizes are recomputed. */
int orc_set_basis(orc_system *s, int64_t n, const uint64_t *reps) {
  free(s->reps);
  free(s->stab);
  s->reps = (uint64_t *)malloc(sizeof(uint64_t) * (n ? n : 1));
  s->stab = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
  memcpy(s->reps, reps, sizeof(uint64_t) * n);
  for (int64_t i = 0; i < n; ++i) {
    uint64_t r;
    int p, c, v;
    orc_orbit(s, reps[i], &r, &p, &c, &v);
    s->stab[i] = (uint32_t)c;
  }
  s->D = n;
  s->built = 1;
  return ORC_OK;
}

int orc_set_basis_with_stab(orc_system *s, int64_t n, const uint64_t *reps,
                            const uint32_t *stab) {
  free(s->reps);
  free(s->stab);
  s->reps = (uint64_t *)malloc(sizeof(uint64_t) * (n ? n : 1));
  s->stab = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
  if (!s->reps || !s->stab) return ORC_OUT_OF_MEMORY;
  memcpy(s->reps, reps, sizeof(uint64_t) * n);
  memcpy(s->stab, stab, sizeof(uint32_t) * n);
  s->D = n;
  s->built = 1;
  return ORC_OK;
}

int64_t orc_dimension(const orc_system *s) { return s->built ? s->D : -1; }

void orc_get_basis(const orc_system *s, uint64_t *reps, uint32_t *stab) {
  if (reps) memcpy(reps, s->reps, sizeof(uint64_t) * s->D);
  if (stab) memcpy(stab, s->stab, sizeof(uint32_t) * s->D);
}

int64_t orc_index(const orc_system *s, uint64_t rep) {
  int64_t lo = 0, hi = s->D;
  while (lo < hi) {
    int64_t mid = lo + (hi - lo) / 2;
    if (s->reps[mid] < rep)
      lo = mid + 1;
    else
      hi = mid;
  }
  return (lo < s->D && s->reps[lo] == rep) ? lo : -1;
}

/* ------------------------------------------------------------------------ */
/* operator (ls_create_interactionN / ls_create_operator, Internal.hs:256-402) */

int orc_add_interaction(orc_system *s, int n, const double *matrix, int n_tuples,
                        const unsigned *sites) {
  if (n < 1 || n > 4 || n_tuples < 0) return ORC_INVALID_ARGUMENT;
  for (int t = 0; t < n_tuples * n; ++t)
    if ((int)sites[t] >= s->n) return ORC_INVALID_ARGUMENT;
  int dim = 1 << n;
  s->mats = (orc_matrix *)realloc(s->mats, sizeof(orc_matrix) * (s->nmat + 1));
  s->mats[s->nmat].n = n;
  s->mats[s->nmat].m = (double *)malloc(sizeof(double) * 2 * dim * dim);
  memcpy(s->mats[s->nmat].m, matrix, sizeof(double) * 2 * dim * dim);
  s->tuples = (orc_tuple *)realloc(s->tuples, sizeof(orc_tuple) * (s->ntuples + n_tuples));
  for (int t = 0; t < n_tuples; ++t) {
    orc_tuple *tu = &s->tuples[s->ntuples + t];
    tu->n = n;
    for (int k = 0; k < n; ++k) tu->sites[k] = sites[t * n + k];
    tu->interaction = s->nmat;
  }
  s->ntuples += n_tuples;
  s->nmat += 1;
  return ORC_OK;
}

int orc_operator_is_real(const orc_system *s) {
  for (int i = 0; i < s->nmat; ++i) {
    int dim = 1 << s->mats[i].n;
    for (int j = 0; j < dim * dim; ++j)
      if (s->mats[i].m[2 * j + 1] != 0.0) return 0;
  }
  for (int g = 0; g < s->G; ++g)
    if (s->phase[g] != 0 && 2 * s->phase[g] != s->L) return 0;
  return 1;
}

int orc_number_offdiag_templates(const orc_system *s) {
  int c = 0;
  for (int t = 0; t < s->ntuples; ++t) {
    const orc_matrix *m = &s->mats[s->tuples[t].interaction];
    int dim = 1 << m->n;
    for (int a = 0; a < dim; ++a)
      for (int b = 0; b < dim; ++b)
        if (a != b && (m->m[2 * (b * dim + a)] != 0.0 || m->m[2 * (b * dim + a) + 1] != 0.0)) ++c;
  }
  return c;
}

static inline int local_index(const orc_tuple *t, uint64_t x) {
  int a = 0;
  for (int k = 0; k < t->n; ++k) a = (a << 1) | (int)((x >> t->sites[k]) & 1);
  return a;
}

static inline uint64_t set_local(const orc_tuple *t, uint64_t x, int b) {
  for (int k = 0; k < t->n; ++k) {
    int bit = (b >> (t->n - 1 - k)) & 1;
    x = (x & ~(1ull << t->sites[k])) | ((uint64_t)bit << t->sites[k]);
  }
  return x;
}

/* y_r = sum_sigma conj(c_sigma) chi(h) sqrt(|Stab r'| / |Stab r|) x_{r'} with
 * H|r> = sum c_sigma |sigma>, h.sigma = r' = min orbit, derived from the
 * projector P = (1/|G|) sum_g chi(g)* g.  Summation order: diagonal first,
 * then tuples in order, and for each tuple target local states b ascending;
 * every off-diagonal term is accumulated with one fma. */
int orc_matvec_real(const orc_system *s, const double *x, double *y, int64_t row_begin,
                    int64_t row_end) {
  if (!s->built) return ORC_NOT_BUILT;
  if (!orc_operator_is_real(s)) return ORC_INVALID_ARGUMENT;
  if (row_end < 0 || row_end > s->D) row_end = s->D;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 256)
#endif
  for (int64_t i = row_begin; i < row_end; ++i) {
    uint64_t r = s->reps[i];
    double diag = 0.0;
    for (int t = 0; t < s->ntuples; ++t) {
      const orc_tuple *tu = &s->tuples[t];
      const orc_matrix *m = &s->mats[tu->interaction];
      int dim = 1 << m->n, a = local_index(tu, r);
      diag += m->m[2 * (a * dim + a)];
    }
    double acc = diag * x[i];
    double sr = (double)s->stab[i];
    for (int t = 0; t < s->ntuples; ++t) {
      const orc_tuple *tu = &s->tuples[t];
      const orc_matrix *m = &s->mats[tu->interaction];
      int dim = 1 << m->n, a = local_index(tu, r);
      for (int b = 0; b < dim; ++b) {
        if (b == a) continue;
        double mba = m->m[2 * (b * dim + a)];
        if (mba == 0.0) continue;
        uint64_t sigma = set_local(tu, r, b);
        uint64_t rp;
        int ph, cnt, ok;
        orc_orbit(s, sigma, &rp, &ph, &cnt, &ok);
        if (!ok) continue;
        int64_t j = orc_index(s, rp);
        if (j < 0) continue;
        double coef = mba * sqrt((double)cnt / sr);
        if (ph != 0) coef = -coef;
        acc = fma(coef, x[j], acc);
      }
    }
    y[i - row_begin] = acc;
  }
  return ORC_OK;
}

static void phase_to_complex(int64_t ph, int64_t L, double *c, double *sn) {
  int64_t p = ((ph % L) + L) % L;
  if (p == 0) { *c = 1.0; *sn = 0.0; }
  else if (2 * p == L) { *c = -1.0; *sn = 0.0; }
  else if (4 * p == L) { *c = 0.0; *sn = 1.0; }
  else if (4 * p == 3 * L) { *c = 0.0; *sn = -1.0; }
  else {
    double ang = 2.0 * M_PI * (double)p / (double)L;
    *c = cos(ang);
    *sn = sin(ang);
  }
}

int orc_matvec_complex(const orc_system *s, const double *x, double *y, int64_t row_begin,
                       int64_t row_end) {
  if (!s->built) return ORC_NOT_BUILT;
  if (row_end < 0 || row_end > s->D) row_end = s->D;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 256)
#endif
  for (int64_t i = row_begin; i < row_end; ++i) {
    uint64_t r = s->reps[i];
    double dre = 0.0, dim_ = 0.0;
    for (int t = 0; t < s->ntuples; ++t) {
      const orc_tuple *tu = &s->tuples[t];
      const orc_matrix *m = &s->mats[tu->interaction];
      int dim = 1 << m->n, a = local_index(tu, r);
      dre += m->m[2 * (a * dim + a)];
      dim_ -= m->m[2 * (a * dim + a) + 1]; /* conj */
    }
    /* explicit fma everywhere: the summation order and roundings are part
     * of the contract the engine reproduces */
    double are = fma(dre, x[2 * i], -(dim_ * x[2 * i + 1]));
    double aim = fma(dre, x[2 * i + 1], dim_ * x[2 * i]);
    double sr = (double)s->stab[i];
    for (int t = 0; t < s->ntuples; ++t) {
      const orc_tuple *tu = &s->tuples[t];
      const orc_matrix *m = &s->mats[tu->interaction];
      int dim = 1 << m->n, a = local_index(tu, r);
      for (int b = 0; b < dim; ++b) {
        if (b == a) continue;
        double mre = m->m[2 * (b * dim + a)], mim = -m->m[2 * (b * dim + a) + 1];
        if (mre == 0.0 && mim == 0.0) continue;
        uint64_t sigma = set_local(tu, r, b);
        uint64_t rp;
        int ph, cnt, ok;
        orc_orbit(s, sigma, &rp, &ph, &cnt, &ok);
        if (!ok) continue;
        int64_t j = orc_index(s, rp);
        if (j < 0) continue;
        double f = sqrt((double)cnt / sr), cre, cim;
        phase_to_complex(ph, s->L, &cre, &cim);
        double kre = fma(mre, cre, -(mim * cim)) * f;
        double kim = fma(mre, cim, mim * cre) * f;
        are = fma(kre, x[2 * j], are);
        are = fma(-kim, x[2 * j + 1], are);
        aim = fma(kre, x[2 * j + 1], aim);
        aim = fma(kim, x[2 * j], aim);
      }
    }
    y[2 * (i - row_begin)] = are;
    y[2 * (i - row_begin) + 1] = aim;
  }
  return ORC_OK;
}

/* Restates ls_operator_expectation (Internal.hs:380-381, :437-452) for one
 * real column: <x|H|x>, not divided by <x|x> (normalisation parity
 * unpinned). */
int orc_expectation_real(const orc_system *s, const double *x, double *out) {
  if (!s->built) return ORC_NOT_BUILT;
  double *y = (double *)malloc(sizeof(double) * (s->D ? s->D : 1));
  int st = orc_matvec_real(s, x, y, 0, -1);
  if (st == ORC_OK) {
    double acc = 0.0;
    for (int64_t i = 0; i < s->D; ++i) acc += x[i] * y[i];
    *out = acc;
  }
  free(y);
  return st;
}

int64_t orc_count_offdiag(const orc_system *s, int64_t row_begin, int64_t row_end) {
  if (row_end < 0 || row_end > s->D) row_end = s->D;
  int64_t c = 0;
  for (int64_t i = row_begin; i < row_end; ++i) {
    uint64_t r = s->reps[i];
    for (int t = 0; t < s->ntuples; ++t) {
      const orc_tuple *tu = &s->tuples[t];
      const orc_matrix *m = &s->mats[tu->interaction];
      int dim = 1 << m->n, a = local_index(tu, r);
      for (int b = 0; b < dim; ++b)
        if (b != a && (m->m[2 * (b * dim + a)] != 0.0 || m->m[2 * (b * dim + a) + 1] != 0.0)) ++c;
    }
  }
  return c;
}

/* ------------------------------------------------------------------------ */
/* tridiagonal eigensolver: implicit QL with Wilkinson-type shift            */
/* (Bowdler, Martin, Reinsch & Wilkinson, Numer. Math. 11 (1968) 293)        */

int orc_tridiag_eigh(int n, const double *alpha, const double *beta, double *evals,
                     double *evecs) {
  if (n <= 0) return ORC_INVALID_ARGUMENT;
  double *d = (double *)malloc(sizeof(double) * n);
  double *e = (double *)malloc(sizeof(double) * n);
  memcpy(d, alpha, sizeof(double) * n);
  for (int i = 0; i < n - 1; ++i) e[i] = beta[i];
  e[n - 1] = 0.0;
  double *z = evecs;
  if (z) {
    memset(z, 0, sizeof(double) * n * n);
    for (int i = 0; i < n; ++i) z[i * n + i] = 1.0;
  }
  int status = ORC_OK;
  for (int l = 0; l < n && status == ORC_OK; ++l) {
    int iter = 0;
    for (;;) {
      int m = l;
      for (; m < n - 1; ++m) {
        double dd = fabs(d[m]) + fabs(d[m + 1]);
        if (fabs(e[m]) <= DBL_EPSILON * dd) break;
      }
      if (m == l) break;
      if (++iter > 100) {
        status = ORC_NOT_CONVERGED;
        break;
      }
      /* shift from the leading 2x2 block */
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = hypot(g, 1.0);
      g = d[m] - d[l] + e[l] / (g + copysign(r, g));
      double sn = 1.0, c = 1.0, p = 0.0;
      int i = m - 1;
      int deflated = 0;
      for (; i >= l; --i) {
        double f = sn * e[i], b = c * e[i];
        r = hypot(f, g);
        e[i + 1] = r;
        if (r == 0.0) {
          d[i + 1] -= p;
          e[m] = 0.0;
          deflated = 1;
          break;
        }
        sn = f / r;
        c = g / r;
        g = d[i + 1] - p;
        r = (d[i] - g) * sn + 2.0 * c * b;
        p = sn * r;
        d[i + 1] = g + p;
        g = c * r - b;
        if (z) {
          double *zi = z + (size_t)i * n, *zi1 = z + (size_t)(i + 1) * n;
          for (int k = 0; k < n; ++k) {
            double t = zi1[k];
            zi1[k] = sn * zi[k] + c * t;
            zi[k] = c * zi[k] - sn * t;
          }
        }
      }
      if (deflated) continue;
      d[l] -= p;
      e[l] = g;
      e[m] = 0.0;
    }
  }
  /* selection sort ascending (n is small) */
  for (int i = 0; i < n; ++i) {
    int k = i;
    for (int j = i + 1; j < n; ++j)
      if (d[j] < d[k]) k = j;
    if (k != i) {
      double t = d[i];
      d[i] = d[k];
      d[k] = t;
      if (z)
        for (int q = 0; q < n; ++q) {
          t = z[(size_t)i * n + q];
          z[(size_t)i * n + q] = z[(size_t)k * n + q];
          z[(size_t)k * n + q] = t;
        }
    }
  }
  memcpy(evals, d, sizeof(double) * n);
  free(d);
  free(e);
  return status;
}

/* ------------------------------------------------------------------------ */
/* Lanczos with full reorthogonalisation (role of PRIMME eigh, SpinED.hs:404) */

static double dot(int64_t n, const double *a, const double *b) {
  double s = 0.0;
#ifdef _OPENMP
#pragma omp parallel for reduction(+ : s) schedule(static)
#endif
  for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

int orc_lanczos(const orc_system *s, int k, int max_iter, double tol, uint64_t seed,
                const double *v0, double *evals, double *evecs, double *rnorms, int *iterations,
                double *alpha_out, double *beta_out) {
  if (!s->built) return ORC_NOT_BUILT;
  int64_t D = s->D;
  if (k < 1 || k > D) return ORC_INVALID_ARGUMENT;
  if (max_iter > D) max_iter = (int)D;
  if (max_iter < k) return ORC_INVALID_ARGUMENT;
  double *V = (double *)malloc(sizeof(double) * (size_t)D * (max_iter + 1));
  double *w = (double *)malloc(sizeof(double) * D);
  double *alpha = (double *)calloc(max_iter + 1, sizeof(double));
  double *beta = (double *)calloc(max_iter + 1, sizeof(double));
  double *theta = (double *)malloc(sizeof(double) * (max_iter + 1));
  double *S = (double *)malloc(sizeof(double) * (size_t)(max_iter + 1) * (max_iter + 1));
  double *h = (double *)malloc(sizeof(double) * (max_iter + 1));
  if (!V || !w || !S) return ORC_OUT_OF_MEMORY;
  if (v0)
    memcpy(V, v0, sizeof(double) * D);
  else
    orc_random_vector(seed, D, V);
  double nrm = sqrt(dot(D, V, V));
  for (int64_t i = 0; i < D; ++i) V[i] /= nrm;
  int status = ORC_NOT_CONVERGED, m = 0;
  for (int j = 0; j < max_iter; ++j) {
    double *vj = V + (size_t)j * D;
    orc_matvec_real(s, vj, w, 0, -1);
    double a = 0.0;
    for (int pass = 0; pass < 2; ++pass) {
      for (int i = 0; i <= j; ++i) h[i] = dot(D, V + (size_t)i * D, w);
      for (int i = 0; i <= j; ++i) {
        const double *vi = V + (size_t)i * D;
        double hi = h[i];
        for (int64_t q = 0; q < D; ++q) w[q] -= hi * vi[q];
      }
      a += h[j];
    }
    alpha[j] = a;
    double b = sqrt(dot(D, w, w));
    beta[j] = b;
    m = j + 1;
    int done = (b <= 1e-14 * fabs(a) + 1e-300) || m == max_iter || m == D;
    if (m >= k) {
      orc_tridiag_eigh(m, alpha, beta, theta, S);
      int conv = 1;
      for (int i = 0; i < k; ++i)
        if (fabs(b * S[(size_t)i * m + (m - 1)]) > tol) conv = 0;
      if (conv || done) {
        /* Ritz vectors and explicit residuals */
        double *y = (double *)malloc(sizeof(double) * D);
        double *hy = (double *)malloc(sizeof(double) * D);
        int ok = 1;
        for (int i = 0; i < k; ++i) {
          memset(y, 0, sizeof(double) * D);
          for (int q = 0; q < m; ++q) {
            double c = S[(size_t)i * m + q];
            const double *vq = V + (size_t)q * D;
            for (int64_t p = 0; p < D; ++p) y[p] += c * vq[p];
          }
          double yn = sqrt(dot(D, y, y));
          for (int64_t p = 0; p < D; ++p) y[p] /= yn;
          orc_matvec_real(s, y, hy, 0, -1);
          double rq = dot(D, y, hy);
          for (int64_t p = 0; p < D; ++p) hy[p] -= rq * y[p];
          double rn = sqrt(dot(D, hy, hy));
          if (rn > tol) ok = 0;
          evals[i] = rq;
          if (rnorms) rnorms[i] = rn;
          if (evecs) memcpy(evecs + (size_t)i * D, y, sizeof(double) * D);
        }
        free(y);
        free(hy);
        if (ok || done) {
          status = ok ? ORC_OK : ORC_NOT_CONVERGED;
          break;
        }
      }
    }
    double *vn = V + (size_t)(j + 1) * D;
    for (int64_t q = 0; q < D; ++q) vn[q] = w[q] / b;
  }
  if (iterations) *iterations = m;
  if (alpha_out) memcpy(alpha_out, alpha, sizeof(double) * m);
  if (beta_out) memcpy(beta_out, beta, sizeof(double) * m);
  free(V);
  free(w);
  free(alpha);
  free(beta);
  free(theta);
  free(S);
  free(h);
  return status;
}

/*
 * parity unpinned (no reference output exists and lattice-symmetries/PRIMME
 * are absent): the character sign convention for complex sectors, the
 * site<->bit and local-index bit order for non-swap-symmetric matrices,
 * the normalisation of ls_operator_expectation and PRIMME's iteration
 * trajectory.  Each is a documented choice, shared with the engine.
 */
