"""Host setup (mesh, element COO, HRZ lumped mass, canonical CSR) against the
reference.

Pins, strongest first:
  * the reference's own sources compiled unchanged (oracle/_ref), bit for bit;
  * the CSR sizes the reference produced for full-size meshes (counted with a
    build of the reference and written into this file), which pin its
    exact-zero elimination;
  * the coo2csr known-answer test of sparse_matrix/test_coo2csr_lib.c:10-12,
    with the expected CSR worked out by hand;
  * SHA-256 digests of small-mesh CSR arrays written from oracle/_ref
    (tests/golden/make_setup_golden.py), for machines without the reference.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle
from seisfem_amd import host

GOLD = os.path.join(os.path.dirname(__file__), "golden")
needs_ref = pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built")

SMALL = [(host.Q4, 20, 20), (host.Q9, 10, 10), (host.T6, 10, 10), (host.T3, 10, 10),
         (host.Q4, 7, 5), (host.Q9, 6, 4), (host.T6, 5, 8)]


def _mesh(tc, nx, ny, edge=10.0):
    return host.mesh_structured(tc, nx, ny, 0.0, edge * nx, 0.0, edge * ny)


def _rho(n, seed=1):
    return np.random.default_rng(seed).uniform(1500.0, 2500.0, n)


def _bits_equal(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


@needs_ref
@pytest.mark.parametrize("tc,nx,ny", SMALL)
def test_mesh_matches_reference(tc, nx, ny):
    o = host.element_order(tc)
    en, x, y = _mesh(tc, nx, ny)
    ren, rx, ry = oracle.ref_mesh(tc, nx, ny, 0.0, 10.0 * nx, 0.0, 10.0 * ny, x.size, en.size // o, o)
    assert np.array_equal(en, ren)
    assert _bits_equal(x, rx) and _bits_equal(y, ry)


@needs_ref
@pytest.mark.parametrize("tc,nx,ny", SMALL)
def test_coo_lump_csr_match_reference(tc, nx, ny):
    en, x, y = _mesh(tc, nx, ny)
    rho = _rho(x.size)
    ci, cj, vals = host.assemble_coo(tc, en, x, y, rho)
    ref = oracle.ref_assemble(tc, en, x, y, rho)
    assert _bits_equal(host.lumped_mass(tc, en, x, y, rho), ref["lump"])
    for m in range(7):
        ri, rj, rv = ref[m]
        assert np.array_equal(ci, ri) and np.array_equal(cj, rj), host.MATRIX_NAMES[m]
        assert _bits_equal(vals[m], rv), host.MATRIX_NAMES[m]
        mine = host.coo2csr(x.size, ci, cj, vals[m])
        theirs = oracle.ref_coo2csr(x.size, ri, rj, rv)
        for a, b in zip(mine, theirs):
            assert _bits_equal(a, b), host.MATRIX_NAMES[m]


@needs_ref
def test_coo2csr_sort_paths_match_reference():
    """Rows long enough for the partition sort (>= 21 entries), rows with many
    duplicates and explicit zeros, and a row whose duplicate sum is -0.0."""
    rng = np.random.default_rng(7)
    n = 40
    rows, cols, vals = [], [], []
    for i in range(n):
        k = int(rng.integers(0, 60))
        rows += [i] * k
        cols += list(rng.integers(0, n, k))
        vals += list(rng.choice([0.0, 1.0, -1.0, 0.1, 1e-17, -0.3], k))
    rows += [3, 3]
    cols += [5, 5]
    vals += [-0.0, -0.0]
    perm = rng.permutation(len(rows))
    ai = np.array(rows, np.int32)[perm]
    aj = np.array(cols, np.int32)[perm]
    ax = np.array(vals, np.float64)[perm]
    mine = host.coo2csr(n, ai, aj, ax)
    theirs = oracle.ref_coo2csr(n, ai, aj, ax)
    for a, b in zip(mine, theirs):
        assert _bits_equal(a, b)


def test_coo2csr_known_answer():
    """sparse_matrix/test_coo2csr_lib.c:10-12: duplicates summed in sorted
    order, exact zeros dropped; expected CSR derived by hand."""
    ai = np.array([0, 0, 0, 1, 1, 1, 2, 2, 2, 3, 3, 3, 2, 2, 2, 1, 1, 1], np.int32)
    aj = np.array([0, 1, 2, 0, 1, 2, 0, 1, 2, 3, 2, 1, 3, 2, 1, 3, 2, 1], np.int32)
    ax = np.array([1.0, -0.5, -0.5, -0.5, 0.5, 0.0, -0.5, 0.0, 0.5, 1.0, -0.5, -0.5, -0.5, 0.5, 0.0,
                   -0.5, 0.0, 0.5])
    bp, bj, bx = host.coo2csr(4, ai, aj, ax)
    assert bp.tolist() == [0, 3, 6, 9, 12]
    assert bj.tolist() == [0, 1, 2, 0, 1, 3, 0, 2, 3, 1, 2, 3]
    assert bx.tolist() == [1.0, -0.5, -0.5, -0.5, 1.0, -0.5, -0.5, 1.0, -0.5, -0.5, -0.5, 1.0]
    if oracle.ref_available():
        for a, b in zip((bp, bj, bx), oracle.ref_coo2csr(4, ai, aj, ax)):
            assert _bits_equal(a, b)


# CSR sizes (M, S1..S6) the reference produced (its coo2csr_canonical output).
REF_NNZ = {
    (host.Q4, 20, 20, 10.0): [3721, 3721, 3721, 3677, 3677, 3685, 3706],
    (host.Q4, 200, 200, 10.0): [361201, 361201, 361201, 358189, 358189, 359985, 360397],
    (host.Q4, 200, 200, 5.0): [361201, 361201, 361201, 358189, 358189, 359985, 360397],
    (host.T3, 200, 200, 5.0): [281201, 120801, 120801, 280399, 280399, 241200, 241200],
    (host.T6, 200, 200, 5.0): [1843201, 1522001, 1522001, 1681599, 1681599, 1603200, 1603200],
    (host.Q9, 200, 200, 5.0): [2563201, 2563201, 2563201, 2558129, 2558129, 2559997, 2560887],
}


@pytest.mark.parametrize("key", list(REF_NNZ))
def test_csr_sizes_match_reference_runs(key):
    tc, nx, ny, edge = key
    en, x, y = _mesh(tc, nx, ny, edge)
    rho = np.full(x.size, 2200.0)
    ci, cj, vals = host.assemble_coo(tc, en, x, y, rho)
    got = [int(host.coo2csr(x.size, ci, cj, vals[m])[0][-1]) for m in range(7)]
    assert got == REF_NNZ[key]


def _digest(arrs):
    h = hashlib.sha256()
    for a in arrs:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def test_setup_digests_match_golden():
    with open(os.path.join(GOLD, "setup_digests.json")) as f:
        gold = json.load(f)
    for key, want in gold.items():
        tc, nx, ny = (int(v) for v in key.split("_"))
        en, x, y = _mesh(tc, nx, ny)
        rho = _rho(x.size)
        ci, cj, vals = host.assemble_coo(tc, en, x, y, rho)
        got = {"lump": _digest([host.lumped_mass(tc, en, x, y, rho)])}
        for m in range(7):
            got[host.MATRIX_NAMES[m]] = _digest(host.coo2csr(x.size, ci, cj, vals[m]))
        assert got == want, key


def test_total_mass_and_partition_of_unity():
    """Closed forms: sum of M = sum of lumped mass = total mass (homogeneous rho,
    rectangle); S1..S6 rows sum to ~0 (derivatives of a partition of unity)."""
    for tc in (host.Q4, host.Q9, host.T6, host.T3):
        en, x, y = _mesh(tc, 8, 6)
        rho = np.full(x.size, 2000.0)
        ci, cj, vals = host.assemble_coo(tc, en, x, y, rho)
        area = 80.0 * 60.0
        assert abs(vals[0].sum() - 2000.0 * area) < 1e-6 * 2000.0 * area
        assert abs(host.lumped_mass(tc, en, x, y, rho).sum() - 2000.0 * area) < 1e-6 * 2000.0 * area
        for m in range(1, 7):
            rows = np.bincount(ci, weights=vals[m], minlength=x.size)
            assert np.max(np.abs(rows)) < 1e-9 * max(1.0, np.max(np.abs(vals[m])))
        # S1, S2 symmetric; S3 = S4^T
        def dense(m):
            d = np.zeros((x.size, x.size))
            np.add.at(d, (ci, cj), vals[m])
            return d
        assert np.allclose(dense(1), dense(1).T, atol=1e-12)
        assert np.allclose(dense(3), dense(4).T, atol=1e-12)


def test_locality_order_is_a_hilbert_walk():
    """On a 64 x 64 lattice every cell gets its own level-6 Hilbert cell, so
    consecutive nodes of the order are lattice neighbours (a Z/Morton order or
    a row-major order would jump); on an irregular cloud it is a permutation."""
    x, y = np.meshgrid(np.arange(64.0), np.arange(64.0), indexing="ij")
    x, y = x.ravel(), y.ravel()
    o = host.locality_order(x, y)
    assert np.array_equal(np.sort(o), np.arange(x.size))
    assert np.all(np.abs(np.diff(x[o])) + np.abs(np.diff(y[o])) == 1.0)
    rng = np.random.default_rng(3)
    xr, yr = rng.uniform(-5.0, 7.0, 1000), rng.uniform(100.0, 101.0, 1000)
    o = host.locality_order(xr, yr)
    assert np.array_equal(np.sort(o), np.arange(1000))


def _fortran_abi(fn, n, ai, aj, ax):
    """Calls a coo2csr_canonical with the Fortran ABI (every argument by
    pointer); returns every output in full (Bp, Bj, Bx over all nnz entries,
    csr_size) and the Ai / Aj it leaves behind."""
    import ctypes as C
    ai = np.array(ai, np.int32)
    aj = np.array(aj, np.int32)
    ax = np.array(ax, np.float64)
    nnz = C.c_int(ai.size)
    psz = C.c_int(n + 1)
    csz = C.c_int(-7)
    bp = np.full(n + 1, 12345, np.int32)
    bj = np.full(max(ai.size, 1), 777, np.int32)
    bx = np.full(max(ai.size, 1), np.nan)
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))

    def go():
        fn(C.byref(nnz), C.byref(psz), C.byref(csz), ip(ai), ip(aj), dp(ax), ip(bp), ip(bj), dp(bx))
        return csz.value

    k = oracle._big_stack(go)
    return dict(bp=bp, bj=bj[:ai.size], bx=bx[:ai.size], csr_size=k, ai=ai, aj=aj, ax=ax)


def _kat():
    ai = np.array([0, 0, 0, 1, 1, 1, 2, 2, 2, 3, 3, 3, 2, 2, 2, 1, 1, 1], np.int32)
    aj = np.array([0, 1, 2, 0, 1, 2, 0, 1, 2, 3, 2, 1, 3, 2, 1, 3, 2, 1], np.int32)
    ax = np.array([1.0, -0.5, -0.5, -0.5, 0.5, 0.0, -0.5, 0.0, 0.5, 1.0, -0.5, -0.5, -0.5, 0.5, 0.0,
                   -0.5, 0.0, 0.5])
    return ai, aj, ax


def test_fortran_abi_coo2csr_known_answer():
    """The library's __coo2csr_lib_MOD_coo2csr_canonical on the known-answer
    input of sparse_matrix/test_coo2csr_lib.c:10-12: the canonical CSR, the
    tail of Bj at -1 and of Bx at 0 (coo2csr_lib.f90:23-25, 36-37), and Ai / Aj
    incremented in place (:26-27)."""
    from seisfem_amd import lib
    ai, aj, ax = _kat()
    out = _fortran_abi(lib().__coo2csr_lib_MOD_coo2csr_canonical, 4, ai, aj, ax)
    assert out["csr_size"] == 12
    assert out["bp"].tolist() == [0, 3, 6, 9, 12]
    assert out["bj"].tolist() == [0, 1, 2, 0, 1, 3, 0, 2, 3, 1, 2, 3] + [-1] * 6
    assert out["bx"].tolist() == [1.0, -0.5, -0.5, -0.5, 1.0, -0.5, -0.5, 1.0, -0.5, -0.5, -0.5, 1.0] + [0.0] * 6
    assert np.array_equal(out["ai"], ai + 1) and np.array_equal(out["aj"], aj + 1)
    assert _bits_equal(out["ax"], ax)


def test_fortran_abi_coo2csr_rejects_rows_out_of_range():
    from seisfem_amd import lib
    ai, aj, ax = _kat()
    ai[4] = 4                                          # row 4 of a 4-row matrix
    out = _fortran_abi(lib().__coo2csr_lib_MOD_coo2csr_canonical, 4, ai, aj, ax)
    assert out["csr_size"] == -1
    assert np.array_equal(out["ai"], ai) and np.array_equal(out["aj"], aj)


@needs_ref
@pytest.mark.parametrize("case", ["kat", "sorts", "q9", "t6"])
def test_fortran_abi_coo2csr_matches_reference_fortran(case):
    """Every output of the library's Fortran-ABI entry point, including the
    tails of Bj / Bx and the mutated Ai / Aj, equals the reference's own
    coo2csr_lib.f90 compiled unchanged (oracle/_ref), bit for bit."""
    from seisfem_amd import lib
    if case == "kat":
        n, (ai, aj, ax) = 4, _kat()
    elif case == "sorts":
        rng = np.random.default_rng(11)
        n = 30
        k = 900
        ai = rng.integers(0, n, k).astype(np.int32)
        aj = rng.integers(0, n, k).astype(np.int32)
        ax = rng.choice([0.0, -0.0, 1.0, -1.0, 0.25, 1e-300], k)
    else:
        tc = host.Q9 if case == "q9" else host.T6
        en, x, y = _mesh(tc, 6, 5)
        ci, cj, vals = host.assemble_coo(tc, en, x, y, _rho(x.size))
        n, ai, aj, ax = x.size, ci, cj, vals[3]
    mine = _fortran_abi(lib().__coo2csr_lib_MOD_coo2csr_canonical, n, ai, aj, ax)
    theirs = _fortran_abi(oracle.rlib()._QMcoo2csr_libPcoo2csr_canonical, n, ai, aj, ax)
    assert mine["csr_size"] == theirs["csr_size"] > 0
    for key in ("bp", "bj", "bx", "ai", "aj", "ax"):
        assert _bits_equal(mine[key], theirs[key]), key
