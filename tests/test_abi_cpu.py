"""C-ABI and host-logic checks that need no GPU: the library loads and
exports every declared symbol, object creation and validation follow the
reference's error behaviour (src/SpinED/Internal.hs, test/Spec.hs), and the
device orbit code (run on the host through a test hook) matches the oracle."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import pyoracle as O
import spined
from spined import _lib, lattice
from spined.internal import LatticeSymmetriesException, SpinEDException

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    names = []
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"\b([a-z_][a-z0-9_]*)\s*\(", text):
            name = m.group(1)
            if name.startswith(("ls_", "spined_")) and not name.endswith("_fn"):
                names.append(name)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    names = declared_functions()
    assert len([n for n in names if n.startswith("ls_")]) == 32
    for n in names:
        assert hasattr(L, n), n
    assert set(_lib.LS_SYMBOLS) <= set(names)
    assert set(_lib.SPINED_SYMBOLS) <= set(names)


def test_error_strings_are_owned_and_freed():
    for code in range(0, 16):
        s = _lib.error_message(code)
        assert isinstance(s, str) and s
    assert _lib.error_message(0) == "no error"


def test_symmetry_semantics_spec_hs():
    s1 = spined.mk_symmetry([3, 2, 1, 0], 0)
    assert spined.get_sector(s1) == 0 and spined.get_periodicity(s1) == 2
    assert spined.get_phase(spined.mk_symmetry([1, 2, 3, 0], 1)) == 0.25
    with pytest.raises(LatticeSymmetriesException) as e:
        spined.mk_symmetry([4, 3, 4, 1], 0)
    assert e.value.code == 1
    with pytest.raises(LatticeSymmetriesException) as e:
        spined.mk_symmetry([4, 3, 2, 1, 0], 3)
    assert e.value.code == 2
    with pytest.raises(SpinEDException):
        spined.mk_symmetry([0, -1], 0)
    with pytest.raises(SpinEDException):
        spined.mk_symmetry([0, 1], -5)


def test_group_semantics_spec_hs():
    r = spined.mk_symmetry([3, 2, 1, 0], 0)
    t = spined.mk_symmetry([1, 2, 3, 0], 0)
    t1 = spined.mk_symmetry([1, 2, 3, 0], 1)
    assert spined.get_group_size(spined.mk_group([r, t])) == 8
    assert spined.get_group_size(spined.mk_group([])) == 1
    with pytest.raises(LatticeSymmetriesException) as e:
        spined.mk_group([r, t1])
    assert e.value.code == 3


def test_basis_validation():
    g = spined.mk_group([spined.mk_symmetry([1, 2, 3, 0], 0), spined.mk_symmetry([3, 2, 1, 0], 0)])
    b = spined.mk_basis(g, 4, 2)                         # test/Spec.hs:80-82
    with pytest.raises(SpinEDException):
        spined.mk_basis(g, -2)                           # test/Spec.hs:83-84
    with pytest.raises(SpinEDException):
        spined.mk_basis(g, 4, -1)
    with pytest.raises(SpinEDException):
        spined.mk_basis(g, 4, 2, 0)
    for args, code in [((5, 2), 4), ((4, 5), 5), ((4, 1, 1), 6)]:
        with pytest.raises(LatticeSymmetriesException) as e:
            spined.mk_basis(g, *args)
        assert e.value.code == code
    with pytest.raises(LatticeSymmetriesException) as e:
        spined.get_number_states(b)                      # not built
    assert e.value.code == 8


def test_interaction_validation_and_reality():
    x = spined.mk_interaction([[1.0, 0.0], [0.0, -1.0]], [[0], [2]])   # test/Spec.hs:96-104
    assert spined.is_real_interaction(x)
    y = spined.mk_interaction([[0, complex(0, -1)], [complex(0, 1), 0]], [[0]])
    assert not spined.is_real_interaction(y)
    with pytest.raises(SpinEDException):
        spined.mk_interaction(lattice.HEISENBERG, [])
    with pytest.raises(SpinEDException):
        spined.mk_interaction(lattice.HEISENBERG, [[0, 1, 2, 3, 4]])
    with pytest.raises(SpinEDException):
        spined.mk_interaction(lattice.HEISENBERG, [[0, 1], [2]])
    with pytest.raises(SpinEDException):
        spined.mk_interaction([[1, 2, 1], [4, 2, 1], [1, -2, 3]], [[0, 0, 0]])
    with pytest.raises(SpinEDException):
        spined.mk_interaction(lattice.HEISENBERG, [[0, -1]])
    with pytest.raises(LatticeSymmetriesException):
        spined.mk_interaction(lattice.HEISENBERG, [[1, 1]])


def test_operator_reality_and_lazy_binding():
    g = spined.mk_group(
        [spined.mk_symmetry(p, s) for p, s in lattice.ring_symmetries(8, 0, 0)])
    b = spined.mk_basis(g, 8, 4, 1)
    op = spined.mk_operator(b, [spined.mk_interaction(lattice.HEISENBERG, lattice.ring_bonds(8))])
    assert spined.is_operator_real(op)
    gc = spined.mk_group([spined.mk_symmetry(lattice.ring_translation(8), 1)])
    bc = spined.mk_basis(gc, 8)
    opc = spined.mk_operator(bc, [spined.mk_interaction(lattice.HEISENBERG, lattice.ring_bonds(8))])
    assert not spined.is_operator_real(opc)
    # before ls_build a matvec is an error, not a crash
    with pytest.raises(LatticeSymmetriesException) as e:
        spined.apply(op, np.zeros(7))
    assert e.value.code == 8


def test_sector_dimension_api_matches_oracle():
    cases = [(24, 12, 1, lattice.ring_symmetries(24, 0, 0)),
             (36, 18, 1, lattice.square_symmetries(6)),
             (40, 20, 1, lattice.ring_symmetries(40, 0, 0)),
             (12, None, None, lattice.ring_symmetries(12, 5, None))]
    for n, w, z, gens in cases:
        b = spined.mk_basis(spined.mk_group([spined.mk_symmetry(p, s) for p, s in gens]), n, w, z)
        assert spined.sector_dimension(b) == pytest.approx(O.System(n, w, z, gens).sector_dimension()[0], abs=1e-6)


@pytest.mark.parametrize("n,w,z,gens", [
    (40, 20, 1, lattice.ring_symmetries(40, 0, 0)),
    (24, 12, -1, lattice.ring_symmetries(24, 12, 1)),
    (17, None, None, lattice.ring_symmetries(17, 3, None)),
    (36, 18, 1, lattice.square_symmetries(6)),
    (26, 13, 1, []),
])
def test_device_orbit_code_matches_oracle(n, w, z, gens):
    b = spined.mk_basis(spined.mk_group([spined.mk_symmetry(p, s) for p, s in gens]), n, w, z)
    o = O.System(n, w, z, gens)
    rng = np.random.default_rng(n)
    if w is None:
        st = rng.integers(0, 1 << n, 3000, dtype=np.uint64)
    else:
        st = np.array([sum(1 << int(i) for i in rng.choice(n, w, replace=False)) for _ in range(3000)],
                      dtype=np.uint64)
    ref = [o.orbit(int(x)) for x in st]
    methods = [0] + ([1] if spined.ring_available(b) else [])
    assert spined.ring_available(b) == (len(gens) > 0 and n in (40, 24, 17))
    for m in methods:
        reps, ph, cnt, val = spined.debug_orbit(b, st, m)
        assert reps.tolist() == [r[0] for r in ref]
        assert cnt.tolist() == [r[2] for r in ref]
        assert val.tolist() == [r[3] for r in ref]
        assert all(p == r[1] for p, r, v in zip(ph, ref, val) if v)


def _torus_gens(l1, l2, kx, ky, reflect=False):
    """Translations of a row-major l1 x l2 torus (site = l1 y + x) with
    momenta (kx, ky), optionally the in-row reflection (sector 0)."""
    idx = lambda x, y: (y % l2) * l1 + (x % l1)
    gens = [([idx(x + 1, y) for y in range(l2) for x in range(l1)], kx),
            ([idx(x, y + 1) for y in range(l2) for x in range(l1)], ky)]
    if reflect:
        gens.append(([idx(l1 - 1 - x, y) for y in range(l2) for x in range(l1)], 0))
    return gens


@pytest.mark.parametrize("n,w,z,gens,shape", [
    (36, 18, 1, lattice.square_symmetries(6), (6, 6)),
    (16, 8, -1, lattice.square_symmetries(4), (4, 4)),
    (25, 12, None, lattice.square_symmetries(5), (5, 5)),
    (9, None, 1, lattice.square_symmetries(3), (3, 3)),
    (24, 12, 1, _torus_gens(4, 6, 0, 0, True), (4, 6)),
    (24, None, None, _torus_gens(4, 6, 1, 2), (4, 6)),
    (24, 12, -1, _torus_gens(6, 4, 3, 2), (6, 4)),
    (20, 10, 1, _torus_gens(5, 4, 0, 2, True), (5, 4)),
    (12, 6, 1, lattice.ring_symmetries(12, 0, 0), (0, 0)),
    (12, 6, None, [(lattice.ring_translation(12), 0)] + [([(i + 2) % 12 for i in range(12)], 0)],
     (0, 0)),
    # translations along the rows only: no torus, the per-element scan
    (16, 8, 1, _torus_gens(4, 4, 0, 0)[:1], (0, 0)),
    # a 2-leg ladder (2 x 8 torus): rung swap and leg translations
    (16, 8, -1, _torus_gens(2, 8, 0, 0), (2, 8)),
])
def test_torus_factorised_orbit(n, w, z, gens, shape):
    """The generic orbit over the torus factorisation (a point element's
    program, then in-row and row rotations) gives the scan over every
    element's program exactly: representative, stabiliser count and
    validity, and the phase wherever the orbit is valid; both equal the
    oracle, and the minimum-only scan finds the same representatives.  Cases: square tori with their point groups, rectangles with
    non-zero momenta (complex phases) and spin inversion; ring groups keep
    the scan (no torus)."""
    b = spined.mk_basis(spined.mk_group([spined.mk_symmetry(p, s) for p, s in gens]), n, w, z)
    assert spined.torus_shape(b) == shape
    rng = np.random.default_rng(n + 7)
    if w is None:
        st = rng.integers(0, 1 << n, 1500, dtype=np.uint64)
    else:
        st = np.array([sum(1 << int(i) for i in rng.choice(n, w, replace=False))
                       for _ in range(1500)], dtype=np.uint64)
    fac = spined.debug_orbit(b, st, 0)
    scan = spined.debug_orbit(b, st, 2)
    for q in (0, 2, 3):
        np.testing.assert_array_equal(fac[q], scan[q])
    # the minimum-only scan (gather-cache fill of trivial-character sectors)
    np.testing.assert_array_equal(spined.debug_orbit(b, st, 3)[0], scan[0])
    v = fac[3].astype(bool)
    np.testing.assert_array_equal(fac[1][v], scan[1][v])
    o = O.System(n, w, z, gens)
    ref = [o.orbit(int(x)) for x in st]
    assert fac[0].tolist() == [r[0] for r in ref]
    assert fac[2].tolist() == [r[2] for r in ref]
    assert fac[3].tolist() == [r[3] for r in ref]
    assert all(p == r[1] for p, r, ok in zip(fac[1], ref, v) if ok)


def test_build_unsafe_rejects_unsorted_input():
    g = spined.mk_group([])
    b = spined.mk_basis(g, 6, 3)
    with pytest.raises(LatticeSymmetriesException) as e:
        spined.build_basis(b, np.array([7, 11, 13, 12], dtype=np.uint64))
    assert e.value.code == 9


@pytest.mark.parametrize("n,w,z,gens", [
    (4, None, 1, lattice.ring_symmetries(4, 0, 0)),
    (5, None, None, lattice.ring_symmetries(5, 2, None)),
    (6, None, -1, lattice.ring_symmetries(6, 3, 1)),
    (8, None, 1, lattice.ring_symmetries(8, 0, None)),
    (13, None, None, lattice.ring_symmetries(13, 0, 0)),
    (13, None, 1, lattice.ring_symmetries(13, 4, None)),
    (16, None, 1, lattice.ring_symmetries(16, 8, 1)),
    (24, None, -1, lattice.ring_symmetries(24, 12, 1)),
    (31, None, None, lattice.ring_symmetries(31, 7, None)),
    (40, None, 1, lattice.ring_symmetries(40, 0, 0)),
    (64, None, 1, lattice.ring_symmetries(64, 0, 0)),
])
def test_hop_orbit_matches_full_scan(n, w, z, gens):
    """The nearest-neighbour hop path (run data of the unmoved state, updated
    for the two runs at the bond) gives exactly the full ring scan's
    (representative, phase, count, valid), for every bond of states of every
    weight, including the one-run cases that fall back to the full scan."""
    b = spined.mk_basis(spined.mk_group([spined.mk_symmetry(p, s) for p, s in gens]), n, w, z)
    assert spined.ring_available(b)
    rng = np.random.default_rng(n)
    mask = (1 << n) - 1
    states, sites = [], []
    if n <= 13:
        pool = range(1 << n)
    else:
        pool = []
        for wt in list(range(1, 4)) + list(range(n // 2 - 2, n // 2 + 3)) + list(range(n - 3, n)):
            for _ in range(60):
                idx = rng.choice(n, wt, replace=False)
                pool.append(sum(1 << int(q) for q in idx))
    for r in pool:
        for i in range(n):
            j = (i + 1) % n
            if ((r >> i) ^ (r >> j)) & 1:
                states.append(r)
                sites.append(i)
    states = np.array(states, dtype=np.uint64)
    sites = np.array(sites, dtype=np.int32)
    moved = states ^ ((np.uint64(1) << sites.astype(np.uint64)) |
                      (np.uint64(1) << ((sites.astype(np.uint64) + np.uint64(1)) % np.uint64(n))))
    h = spined.debug_orbit_hop(b, states, sites)
    f = spined.debug_orbit(b, moved, 1)
    np.testing.assert_array_equal(h[3], f[3])
    np.testing.assert_array_equal(h[0], f[0])
    np.testing.assert_array_equal(h[2], f[2])
    v = f[3]
    np.testing.assert_array_equal(h[1][v], f[1][v])
    assert len(states) > 0


def _c_layout(ctype_name, struct):
    """sizeof and the field offsets of a struct of include/spined.h, from
    the system C compiler; None without one."""
    import shutil
    import subprocess
    import tempfile
    from pathlib import Path
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        return None
    inc = Path(__file__).resolve().parent.parent / "include"
    fields = [f for f, _ in struct._fields_]
    src = "#include <stddef.h>\n#include <stdio.h>\n#include \"spined.h\"\nint main(void) {\n"
    src += f'  printf("%zu\\n", sizeof({ctype_name}));\n'
    for f in fields:
        src += f'  printf("%zu\\n", offsetof({ctype_name}, {f}));\n'
    src += "  return 0;\n}\n"
    with tempfile.TemporaryDirectory() as d:
        c = Path(d) / "layout.c"
        c.write_text(src)
        exe = Path(d) / "layout"
        subprocess.run([cc, f"-I{inc}", str(c), "-o", str(exe)], check=True)
        out = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True,
                                               text=True).stdout.split()]
    return out[0], dict(zip(fields, out[1:]))


@pytest.mark.parametrize("ctype_name,attr", [("spined_eigsh_options", "SpinedEigshOptions"),
                                             ("spined_autotune_report", "SpinedAutotuneReport")])
def test_struct_layouts_match_header(ctype_name, attr):
    """The ctypes mirrors of the option and report structs have the C layout
    (size and every field offset a host reads or sets through peekByteOff /
    pokeByteOff), compiled from include/spined.h."""
    import ctypes as C
    from spined import _lib
    struct = getattr(_lib, attr)
    got = _c_layout(ctype_name, struct)
    if got is None:
        pytest.skip("no C compiler")
    size, offs = got
    assert size == C.sizeof(struct)
    for f, off in offs.items():
        assert getattr(struct, f).offset == off, f
