"""Pins of the CPU oracle (runs without a GPU).

The reference ships no golden vectors and its arithmetic (lattice-symmetries,
PRIMME) is not in the image, so the oracle is pinned by: the semantic checks
of test/Spec.hs, README.md's E0 = -8, brute force written independently in
Python (explicit permutations, dense Kronecker Hamiltonians, explicit
projectors), closed forms (character counting, orbit-stabiliser, Bethe
ansatz, XX free fermions, Marshall sign) and LAPACK (numpy / scipy).
"""
import itertools
import math

import numpy as np
import pytest
import scipy.linalg

import pyoracle as O
from spined import lattice

H2 = O.heisenberg_matrix()


# ---------------------------------------------------------------- Spec.hs ----
def test_periodicity_and_invalid_symmetries():
    # test/Spec.hs:38-43
    assert O.periodicity([3, 2, 1, 0]) == 2
    assert O.periodicity([4, 3, 4, 1]) == 0            # not a permutation
    with pytest.raises(O.OracleError) as e:
        O.System(5, None, None, [([4, 3, 2, 1, 0], 3)])  # sector >= periodicity 2
    assert e.value.code == 2


def test_group_sizes_and_inconsistent_sectors():
    # test/Spec.hs:72-79, restated under the current API (flip is spin_inversion)
    refl, trans = [3, 2, 1, 0], [1, 2, 3, 0]
    s = O.System(4, 2, 1, [(refl, 0), (trans, 0)])
    assert s.group_size == 8
    assert s.full_group_size == 16
    with pytest.raises(O.OracleError) as e:
        O.System(4, 2, 1, [(refl, 0), (trans, 1)])
    assert e.value.code == 3


# ------------------------------------------------------------ brute force ----
def apply_perm(x, perm):
    """(g.x)_i = x_{perm[i]}"""
    return sum(((x >> p) & 1) << i for i, p in enumerate(perm))


def brute_group(n, gens, z):
    """Closure with complex characters (independent of the oracle's integers)."""
    ident = tuple(range(n))
    elems = {ident: 1.0 + 0j}
    frontier = [ident]
    chars = [np.exp(2j * np.pi * s / O.periodicity(p)) for p, s in gens]
    while frontier:
        nxt = []
        for e in frontier:
            for (p, _), c in zip(gens, chars):
                q = tuple(e[p[i]] for i in range(n))
                ch = elems[e] * c
                if q in elems:
                    assert abs(elems[q] - ch) < 1e-9
                else:
                    elems[q] = ch
                    nxt.append(q)
        frontier = nxt
    out = [(list(q), False, ch) for q, ch in elems.items()]
    if z:
        out += [(p, True, ch * z) for p, _, ch in out]
    return out


def brute_basis(n, w, z, gens):
    G = brute_group(n, gens, z)
    mask = (1 << n) - 1
    states = range(1 << n) if w is None else (
        sum(1 << i for i in c) for c in itertools.combinations(range(n), w))
    reps = []
    for x in sorted(states):
        imgs = [(apply_perm(x, p) ^ (mask if f else 0), ch) for p, f, ch in G]
        if min(y for y, _ in imgs) != x:
            continue
        s = sum(ch for y, ch in imgs if y == x)
        if abs(s) > 1e-9:
            reps.append((x, sum(1 for y, _ in imgs if y == x)))
    return reps


SMALL = [
    (6, 3, 1, lattice.ring_symmetries(6, 0, 0)),
    (6, None, None, lattice.ring_symmetries(6, 2, None)),
    (8, 4, -1, lattice.ring_symmetries(8, 4, 1)),
    (8, 4, None, lattice.ring_symmetries(8, 1, None)),
    (9, 4, None, lattice.ring_symmetries(9, 3, 0)[:1]),
    (10, 5, -1, lattice.ring_symmetries(10, 5, 1)),      # example/heisenberg_chain_10.yaml
    (9, None, None, [([1, 2, 0, 4, 5, 3, 7, 8, 6], 1), ([3, 4, 5, 6, 7, 8, 0, 1, 2], 0)]),
    (8, 4, 1, []),
]


@pytest.mark.parametrize("n,w,z,gens", SMALL)
def test_basis_brute_force(n, w, z, gens):
    s = O.System(n, w, z, gens)
    d = s.build()
    reps, stab = s.basis()
    ref = brute_basis(n, w, z, gens)
    assert [r for r, _ in ref] == reps.tolist()
    assert [c for _, c in ref] == stab.tolist()
    dim, im = s.sector_dimension()
    assert abs(im) < 1e-9 and round(dim) == d


def test_chain10_example_has_13_states():
    s = O.System(10, 5, -1, lattice.ring_symmetries(10, 5, 1))
    assert s.build() == 13


@pytest.mark.parametrize("n,w,gens,expected", [
    (24, 12, lattice.ring_symmetries(24, 0, 0), 28968),
    (36, 18, lattice.square_symmetries(6), 15804956),
    (40, 20, lattice.ring_symmetries(40, 0, 0), 861725794),
    (16, 8, lattice.square_symmetries(4), 107),
])
def test_known_sector_dimensions(n, w, gens, expected):
    s = O.System(n, w, 1, gens)
    assert s.sector_dimension()[0] == expected


def test_orbit_stabiliser_identity():
    # sum_r 1/N_r^2 = sum_r |G| / |Stab r| = C(N, w) for all-trivial sectors
    for n, gens in [(24, lattice.ring_symmetries(24, 0, 0)), (16, lattice.square_symmetries(4))]:
        s = O.System(n, n // 2, 1, gens)
        s.build()
        _, stab = s.basis()
        G = s.full_group_size
        assert int(np.sum(G // stab.astype(np.int64))) == math.comb(n, n // 2)


# ------------------------------------------------- dense projected matrix ----
def dense_hamiltonian(n, terms):
    """Full 2^n matrix from Kronecker products; basis index = integer state,
    local index of a tuple = bits of its sites, first site most significant."""
    dim = 1 << n
    H = np.zeros((dim, dim), dtype=np.complex128)
    for m, sites in terms:
        m = np.asarray(m, dtype=np.complex128)
        k = len(sites[0])
        for t in sites:
            for x in range(dim):
                a = 0
                for s in t:
                    a = (a << 1) | ((x >> s) & 1)
                for b in range(1 << k):
                    if m[b, a] == 0:
                        continue
                    y = x
                    for j, s in enumerate(t):
                        bit = (b >> (k - 1 - j)) & 1
                        y = (y & ~(1 << s)) | (bit << s)
                    H[y, x] += m[b, a]
    return H


def projected_basis_vectors(n, w, z, gens, reps):
    """|r~> = P|r> / |P|r>| with P = (1/|G|) sum chi(g)* g (explicit vectors)."""
    G = brute_group(n, gens, z)
    mask = (1 << n) - 1
    B = np.zeros((1 << n, len(reps)), dtype=np.complex128)
    for j, r in enumerate(reps):
        for p, f, ch in G:
            B[apply_perm(int(r), p) ^ (mask if f else 0), j] += np.conj(ch) / len(G)
        B[:, j] /= np.linalg.norm(B[:, j])
    return B


def random_hermitian(k, seed, real=False, conserve_weight=False):
    """Random Hermitian 2^k x 2^k matrix with no site-exchange symmetry;
    with conserve_weight, zero between local states of different popcount."""
    rng = np.random.default_rng(seed)
    d = 1 << k
    a = rng.normal(size=(d, d)) + (0 if real else 1j * rng.normal(size=(d, d)))
    h = (a + a.conj().T) / 2
    if conserve_weight:
        pc = np.array([bin(i).count("1") for i in range(d)])
        h = np.where(pc[:, None] == pc[None, :], h, 0)
    return h


def ring_tuples(n, k):
    return [[(i + j) % n for j in range(k)] for i in range(n)]


# 3- and 4-site terms (ls_create_interaction3/4): random Hermitian matrices
# that no site permutation leaves alone, so the local-index bit order (the
# first listed site is the most significant bit) is exercised; translation
# invariant sets of tuples, so they commute with the ring's translations
MULTI = [
    (8, None, None, lattice.ring_symmetries(8, 0, None),
     [(random_hermitian(3, 31, real=True), ring_tuples(8, 3))]),
    (8, None, None, lattice.ring_symmetries(8, 3, None),                            # complex
     [(random_hermitian(3, 32), ring_tuples(8, 3)), (random_hermitian(4, 41), ring_tuples(8, 4))]),
    (9, 4, None, lattice.ring_symmetries(9, 0, None),
     [(random_hermitian(4, 42, real=True, conserve_weight=True), ring_tuples(9, 4)),
      (H2, lattice.ring_bonds(9))]),
    (8, 4, None, [], [(random_hermitian(3, 33, conserve_weight=True), ring_tuples(8, 3))]),
]

DENSE = [
    (8, 4, 1, lattice.ring_symmetries(8, 0, 0), [(H2, lattice.ring_bonds(8))]),
    (8, 4, -1, lattice.ring_symmetries(8, 4, 1), [(H2, lattice.ring_bonds(8))]),
    (8, None, None, lattice.ring_symmetries(8, 1, None), [(H2, lattice.ring_bonds(8))]),   # complex
    (9, None, None, lattice.ring_symmetries(9, 2, None), [(H2, lattice.ring_bonds(9))]),   # complex
    (8, 4, 1, [], [(lattice.xxz_matrix(0.7, 1.1), lattice.ring_bonds(8))]),
    (9, None, None, [([1, 2, 0, 4, 5, 3, 7, 8, 6], 1)],     # H must commute with the group
     [(H2, [[0, 3], [1, 4], [2, 5], [3, 6], [4, 7], [5, 8], [0, 1], [1, 2], [2, 0]]),
      (np.array([[0.3, 1 - 0.5j], [1 + 0.5j, -0.2]]), [[0], [1], [2], [3], [4], [5], [6], [7], [8]])]),
] + MULTI


@pytest.mark.parametrize("n,w,z,gens,terms", DENSE)
def test_matvec_equals_dense_projection(n, w, z, gens, terms):
    s = O.System(n, w, z, gens)
    for m, sites in terms:
        s.add_interaction(np.asarray(m, dtype=np.complex128), sites)
    s.build()
    reps, _ = s.basis()
    H = dense_hamiltonian(n, terms)
    # restrict to the Hamming sector only through the basis vectors
    B = projected_basis_vectors(n, w, z, gens, reps)
    Ht = B.conj().T @ H @ B
    D = len(reps)
    M = np.zeros((D, D), dtype=np.complex128)
    for j in range(D):
        e = np.zeros(D, dtype=np.complex128)
        e[j] = 1
        M[:, j] = s.matvec(e if not s.is_real else e.real)
    np.testing.assert_allclose(M, Ht, atol=1e-12)


def test_sector_spectra_union_to_full_spectrum():
    n = 8
    terms = [(H2, lattice.ring_bonds(n))]
    full = np.linalg.eigvalsh(dense_hamiltonian(n, terms).real)
    collected = []
    for k in range(n):
        s = O.System(n, None, None, lattice.ring_symmetries(n, k, None))
        s.add_interaction(H2, lattice.ring_bonds(n))
        D = s.build()
        M = np.array([s.matvec(np.eye(D, dtype=np.complex128)[j]) for j in range(D)]).T
        collected.extend(np.linalg.eigvalsh(M))
    np.testing.assert_allclose(np.sort(collected), full, atol=1e-10)


# ------------------------------------------------------------ eigensolver ----
def test_tridiagonal_against_lapack():
    rng = np.random.default_rng(0)
    for n in (1, 2, 5, 40, 200):
        a, b = rng.normal(size=n), rng.normal(size=max(n - 1, 0))
        ev, Z = O.tridiag_eigh(a, b)
        ref = scipy.linalg.eigh_tridiagonal(a, b, eigvals_only=True)
        np.testing.assert_allclose(ev, ref, atol=1e-12 * max(1, np.max(np.abs(ref))))
        T = np.diag(a) + np.diag(b, 1) + np.diag(b, -1)
        np.testing.assert_allclose(T @ Z, Z * ev, atol=1e-10)


def test_readme_four_site_ring():
    s = O.System(4)
    s.add_interaction(H2, [[0, 1], [1, 2], [2, 3], [3, 0]])
    assert s.build() == 16
    r = s.lanczos(k=1)
    assert abs(r["evals"][0] - (-8.0)) < 1e-12      # README.md:83-95


def bethe_ground_energy(n):
    """XXX ring, M = n/2 magnons, sigma.sigma convention: E = n - 8 sum 1/(4 l^2 + 1)
    with n atan(2 l_j) = pi I_j + sum_k atan(l_j - l_k) (log form; Newton)."""
    M = n // 2
    I = np.arange(M) - (M - 1) / 2.0
    lam = np.tan(np.pi * I / n) / 2.0
    for _ in range(200):
        d = lam[:, None] - lam[None, :]
        F = n * np.arctan(2 * lam) - np.pi * I - np.sum(np.arctan(d), axis=1)
        K = 1.0 / (1.0 + d ** 2)              # dF_j/dl_k for k != j
        J = K.copy()
        np.fill_diagonal(J, 2 * n / (1 + 4 * lam ** 2) - (K.sum(axis=1) - 1.0))
        step = np.linalg.solve(J, F)
        lam -= step
        if np.max(np.abs(step)) < 1e-15:
            break
    return n - 8.0 * np.sum(1.0 / (4 * lam ** 2 + 1.0))


def test_bethe_ansatz_solver_reproduces_dense_ed():
    for n in (4, 6, 8, 10):
        H = dense_hamiltonian(n, [(H2, lattice.ring_bonds(n))]).real
        assert abs(bethe_ground_energy(n) - np.linalg.eigvalsh(H)[0]) < 1e-11


def test_lanczos_config1_and_marshall_sign():
    n = 10
    s = O.System(n)
    s.add_interaction(H2, lattice.ring_bonds(n))
    s.build()
    r = s.lanczos(k=1, tol=1e-11)
    assert abs(r["evals"][0] - bethe_ground_energy(n)) < 1e-10 * 18.06
    assert abs(r["evals"][0] - (-18.061785417968172)) < 1e-10 * 18.06
    assert r["rnorms"][0] <= 1e-10
    v = r["evecs"][:, 0]
    reps, _ = s.basis()
    # Marshall: sign = (-1)^(#up spins on the even sublattice) up to a global sign
    a_mask = sum(1 << i for i in range(0, n, 2))
    marshall = np.array([(-1) ** bin(int(x) & a_mask).count("1") for x in reps])
    big = np.abs(v) > 1e-8
    prod = np.sign(v[big]) * marshall[big]
    assert np.all(prod == prod[0])


def test_lanczos_ring24_sector_against_bethe_and_dense():
    s = O.System(24, 12, 1, lattice.ring_symmetries(24, 0, 0))
    s.add_interaction(H2, lattice.ring_bonds(24))
    s.build()
    r = s.lanczos(k=1, tol=1e-10, max_iter=300)
    e_bethe = bethe_ground_energy(24)
    assert abs(e_bethe - (-42.680058066148944)) < 1e-11
    assert abs(r["evals"][0] - e_bethe) < 1e-10 * abs(e_bethe)


def test_lanczos_against_dense_sector_matrix():
    s = O.System(12, 6, 1, lattice.ring_symmetries(12, 0, 0))
    s.add_interaction(H2, lattice.ring_bonds(12))
    D = s.build()
    M = np.array([s.matvec(np.eye(D)[j]) for j in range(D)]).T
    ref = np.linalg.eigvalsh(M)
    r = s.lanczos(k=3, tol=1e-11)
    np.testing.assert_allclose(r["evals"], ref[:3], atol=1e-10)


def xx_free_fermion_energy(J):
    """XX ring (Delta = 0) with bond couplings J_b: Jordan-Wigner hopping 2 J_b,
    boundary sign (-1)^(M-1) for M = n/2 fermions."""
    n = len(J)
    M = n // 2
    T = np.zeros((n, n))
    for b in range(n):
        i, j = b, (b + 1) % n
        sgn = (-1) ** (M - 1) if j == 0 else 1
        T[i, j] = T[j, i] = 2 * J[b] * sgn
    return np.sum(np.linalg.eigvalsh(T)[:M])


def test_xx_point_free_fermions():
    rng = np.random.default_rng(4)
    for n in (8, 12):
        J = rng.uniform(0.8, 1.2, n)
        s = O.System(n, n // 2, 1, [])
        for b in range(n):
            s.add_interaction(np.asarray(lattice.xxz_matrix(0.0, J[b]), dtype=np.complex128),
                              [[b, (b + 1) % n]])
        s.build()
        r = s.lanczos(k=1, tol=1e-11)
        assert abs(r["evals"][0] - xx_free_fermion_energy(J)) < 1e-10


def test_random_vector_is_counter_hash():
    a = O.random_vector(11, 1000)
    b = O.random_vector(11, 10)
    np.testing.assert_array_equal(a[:10], b)
    assert np.all(np.abs(a) <= 1) and abs(np.mean(a)) < 0.1
