"""GPU parity of the HIP path (through the C-ABI) against the oracle and the reference goldens.

Bars: integer/index work (basis, CSR structure) and the builder's values bit-exact; SpMV
within 1e-14 relative (summation order differs); time evolution within the rigorous error
bound: ||psi_gpu - psi_oracle|| <= err_gpu + err_oracle + 1e-10, observables within the
requested tol, and the reference ctests' own 1e-8 on the goldens.
"""
import os

import numpy as np
import pytest

from models import blackhole, number_operator_terms, two_oscillators
from timeevolver_amd import capi
from timeevolver_amd.hamiltonian import as_tuples, createNumberOperator
from timeevolver_amd.workloads import (fock_state, random_state, ring_bose_hubbard_terms, ring_couplings,
                                       ring_dimension, ring_nnz)

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _coo_equal(gpu_m, ref):
    v, c, r = gpu_m.to_coo()
    rv, rc, rr = ref
    assert len(v) == len(rv)
    assert np.array_equal(c, rc) and np.array_equal(r, rr)
    assert np.array_equal(v.view(np.uint64), np.ascontiguousarray(rv).view(np.uint64))   # bitwise


# ---------------------------------------------------------------------------------------- basis
@pytest.mark.parametrize("args", [(2, 3, 0, 0, 1), (8, 8, 0, 0, 1), (6, 4, 2, 2, 1), (5, 4, 4, 2, 1), (7, 3, 1, 3, 2),
                                  (3, 1, 0, 0, 1), (0, 4, 0, 0, 1), (2, 8, 8, 1, 1)])
def test_basis_bitexact(gpu, oracle_lib, args):
    g = capi.Basis.basis(gpu, *args)
    o = oracle_lib.Basis.basis(*args)
    assert g.D == o.D
    so = o.states()
    assert np.array_equal(g.states(), so)
    idx = g.find(so)
    assert np.array_equal(idx, np.arange(o.D))
    # absent states: one particle too many, a negative entry
    bad = so.copy()
    bad[:, 0] += 1
    assert (g.find(bad) == -1).all()


def test_tensor_basis_bitexact(gpu, oracle_lib):
    for args in [(20, 2, 2, 8, 1), (1, 2, 1, 2, 1), (3, 2, 2, 4, 1)]:
        g = capi.Basis.tensor(gpu, *args)
        o = oracle_lib.Basis.tensor(*args)
        assert g.D == o.D
        s = o.states()
        assert np.array_equal(g.states(), s)
        assert np.array_equal(g.find(s), np.arange(o.D))


# ---------------------------------------------------------------------------------------- operators
def _ring_terms(L, seed=11, U=0.5):
    eps, J = ring_couplings(L, seed)
    return ring_bose_hubbard_terms(L, eps, J, U)


@pytest.mark.parametrize("L,N", [(3, 3), (6, 6), (8, 8), (10, 10), (12, 12)])
def test_ring_operator_bitexact(gpu, oracle_lib, L, N):
    terms = as_tuples(_ring_terms(L))
    gb = capi.Basis.basis(gpu, N, L)
    ob = oracle_lib.Basis.basis(N, L)
    gm = capi.operator_build(gpu, gb, terms)
    ref = oracle_lib.build_operator(ob, terms)
    assert gm.nnz == ring_nnz(L, N) and gb.D == ring_dimension(L, N)
    _coo_equal(gm, ref)
    assert gm.norm1() == oracle_lib.norm1(ob.D, ref[0], ref[1])      # exact COO-order column sums


def test_blackhole_operators_bitexact(gpu, oracle_lib):
    args, ham, init = blackhole()
    gb = capi.Basis.tensor(gpu, *args)
    ob = oracle_lib.Basis.tensor(*args)
    _coo_equal(capi.operator_build(gpu, gb, ham), oracle_lib.build_operator(ob, as_tuples(ham)))
    for t in number_operator_terms(ob.K):
        _coo_equal(capi.operator_build(gpu, gb, t), oracle_lib.build_operator(ob, as_tuples(t)))


def test_general_terms_bitexact(gpu, oracle_lib):
    from timeevolver_amd.hamiltonian import createFourPointA, createFourPointC, createSixPoint, linInteraction, OpTerm
    terms = [createNumberOperator(0, 0.3), linInteraction(0, 3, 0, 2, False, -0.7),
             linInteraction(0, 3, 0, 2, True, -0.7), createFourPointA(0, 1, 2, 3, 0, 0, 2, 2, 0.11),
             createFourPointC(1, 0, 1, 0, 0, 0, 0, 0, 0.05), createSixPoint(0, 1, 2, 3, 0, 1, 0, 0, 2, 2, 0, 0, 0.2),
             OpTerm(complex(0.25, -0.5), [-1, 1], [1, 2], [0, 2])]
    for args in [(4, 4, 2, 2, 1), (6, 4, 0, 0, 2)]:
        gb = capi.Basis.basis(gpu, *args)
        ob = oracle_lib.Basis.basis(*args)
        _coo_equal(capi.operator_build(gpu, gb, terms), oracle_lib.build_operator(ob, as_tuples(terms)))


# ---------------------------------------------------------------------------------------- spmv
def test_spmv_vs_oracle(gpu, oracle_lib):
    L = N = 8
    terms = as_tuples(_ring_terms(L))
    ob = oracle_lib.Basis.basis(N, L)
    v, c, r = oracle_lib.build_operator(ob, terms)
    A = capi.Matrix.from_coo(gpu, v, c, r, ob.D)
    x = random_state(ob.D, 5)
    for alpha in (1.0, -0.3j, 0.2 - 0.7j):
        y = A.spmv(x, alpha)
        yr = oracle_lib.spmv(ob.D, v, c, r, x, alpha)
        assert np.linalg.norm(y - yr) <= 1e-14 * np.linalg.norm(yr)


def _spmv_check(oracle_lib, M, D, v, c, r, seed=5):
    x = random_state(D, seed)
    for alpha in (1.0, -0.3j, 0.2 - 0.7j):
        y = M.spmv(x, alpha)
        yr = oracle_lib.spmv(D, v, c, r, x, alpha)
        assert np.linalg.norm(y - yr) <= 1e-14 * np.linalg.norm(yr)


def _aligned_slots(rows):
    """Slots of the sliced-ELL layout: per slice of 64 rows (each row's off-diagonal columns
    ascending), the next slot takes each row's next column if it lies within 65536 of the
    smallest next column over the slice. Returns 64 * (number of slots)."""
    total = 0
    for s0 in range(0, len(rows), 64):
        sl = rows[s0:s0 + 64]
        head = [0] * len(sl)
        while True:
            nxt = [row[h] for row, h in zip(sl, head) if h < len(row)]
            if not nxt:
                break
            lo = min(nxt)
            head = [h + (h < len(row) and row[h] - lo < 65536) for row, h in zip(sl, head)]
            total += 64
    return total


def _sell_rows(c, r, order, v=None):
    """Off-diagonal columns of each row in the row order `order` (new -> old), columns renumbered
    in the same order and ascending; with v, (column, value bits) pairs."""
    inv = np.empty_like(order)
    inv[order] = np.arange(len(order))
    off = c != r
    rr, cc = inv[r[off].astype(np.int64)], inv[c[off].astype(np.int64)]
    idx = np.lexsort((cc, rr))
    rr, cc = rr[idx], cc[idx]
    bounds = np.searchsorted(rr, np.arange(len(order) + 1))
    if v is None:
        return [cc[bounds[i]:bounds[i + 1]].tolist() for i in range(len(order))]
    bits = np.ascontiguousarray(v[off]).view(np.uint64).reshape(-1, 2)[idx]
    key = (bits[:, 0] ^ (bits[:, 1] * np.uint64(0x9E3779B97F4A7C15))).tolist()
    return [list(zip(cc[bounds[i]:bounds[i + 1]].tolist(), key[bounds[i]:bounds[i + 1]])) for i in range(len(order))]


def _sell_layout(rows):
    """The sliced-ELL layout of rows of (column, value) pairs: in every full 64-row slice the
    leading and trailing entry positions at which all rows have the same column - row and the same
    value are shared records; the rows' middle entries take aligned slots. Returns (64 * slots,
    shared records, records past the 12 kept at fixed places per slice)."""
    slots = shared = over = 0
    for s0 in range(0, len(rows), 64):
        sl = rows[s0:s0 + 64]
        lo = hi = 0
        if len(sl) == 64:
            mn = min(len(row) for row in sl)
            same = lambda pos: len({(row[p][0] - (s0 + i), row[p][1]) for i, (row, p) in enumerate(zip(sl, pos))}) == 1
            while lo < mn and same([lo] * 64):
                lo += 1
            while lo + hi < mn and same([len(row) - 1 - hi for row in sl]):
                hi += 1
        shared += lo + hi
        over += max(0, lo + hi - 12)
        slots += _aligned_slots([[e[0] for e in row[lo:len(row) - hi]] for row in sl])
    return slots, shared, over


def _sell_bytes(rows, D, diag_bytes):
    """Matrix bytes one sliced-ELL product streams (te_matrix_storage)."""
    slots, shared, over = _sell_layout(rows)
    ns = (D + 63) // 64
    return 4 * slots + 4 * (slots // 64) + 16 * (ns + 1) + diag_bytes * D + (8 * 12 * ns + 8 * over if shared else 0)


def _mode_significance(n_modes, terms):
    """Significance order of the modes for the locality order: a Cuthill-McKee numbering of the
    graph joining the modes of one term, from the start mode with the smallest bandwidth."""
    adj = [set() for _ in range(n_modes)]
    for t in terms:
        ms = list(t[2])
        for a in ms:
            for b in ms:
                if a != b:
                    adj[a].add(b)
    if not any(adj):
        return list(range(n_modes))
    best, best_bw = None, n_modes + 1
    for start in range(n_modes):
        order, pos = [], {}
        seed = start
        while len(order) < n_modes:
            if seed not in pos:
                pos[seed] = len(order)
                order.append(seed)
                h = len(order) - 1
                while h < len(order):
                    nb = sorted((u for u in adj[order[h]] if u not in pos), key=lambda u: (len(adj[u]), u))
                    for u in nb:
                        pos[u] = len(order)
                        order.append(u)
                    h += 1
            seed = next((m for m in range(n_modes) if m not in pos), None)
            if seed is None:
                break
        bw = max((abs(pos[a] - pos[b]) for a in range(n_modes) for b in adj[a]), default=0)
        if bw < best_bw:
            best, best_bw = order, bw
    return best


def _distinct_offdiag(v, c, r):
    off = c != r
    bits = np.ascontiguousarray(v[off]).view(np.uint64).reshape(-1, 2)
    return len(np.unique(bits, axis=0)), int(off.sum())


@pytest.mark.parametrize("L", [6, 8, 10])
def test_value_dictionary_storage(gpu, oracle_lib, monkeypatch, L):
    """Ring operators are stored as a value dictionary whose table holds exactly the distinct
    off-diagonal values (counted bit for bit on the host): in the basis order for a COO matrix,
    in the lexicographic locality order for a built operator. SpMV matches the oracle and the
    plain CSR row-block kernel, and an evolution takes the same steps with either."""
    terms = as_tuples(_ring_terms(L))
    ob = oracle_lib.Basis.basis(L, L)
    v, c, r = oracle_lib.build_operator(ob, terms)
    n_distinct, n_off = _distinct_offdiag(v, c, r)
    monkeypatch.setenv("TE_SPMV_SELL", "0")
    gb = capi.Basis.basis(gpu, L, L)
    Hg = capi.operator_build(gpu, gb, terms)
    A = capi.Matrix.from_coo(gpu, v, c, r, ob.D)
    monkeypatch.delenv("TE_SPMV_SELL")
    for M, want in ((A, 2), (Hg, 3)):
        fmt, nv, nbytes = M.storage()
        assert fmt == want and nv == n_distinct
        assert nbytes == 6 * n_off + 4 * (ob.D + 1) + 16 * ob.D
        _spmv_check(oracle_lib, M, ob.D, v, c, r)
    # sliced ELL (default): 64-row slices, rows in the basis order (COO matrix) or in the
    # lexicographic order of the states with the modes ranked 0, 1, L-1, 2, L-2, ... (built
    # operator: Cuthill-McKee on the ring); shared records, 4-byte words, a 4-byte base per slot
    # of a slice; the real diagonal as 8-byte values
    sig = _mode_significance(L, terms)
    assert sig[:5] == [0, 1, L - 1, 2, L - 2]
    states = ob.states().astype(np.int64)[:, sig]
    lex = np.lexsort(states.T[::-1])                  # sig[0] most significant
    for M, order in ((capi.Matrix.from_coo(gpu, v, c, r, ob.D), np.arange(ob.D)),
                     (capi.operator_build(gpu, gb, terms), lex)):
        fmt, nv, nbytes = M.storage()
        assert fmt == 4 and nv == n_distinct
        rows = _sell_rows(c, r, order, v)
        assert nbytes == _sell_bytes(rows, ob.D, 8)
        if order is lex and L >= 8:                   # a share of the ring's transitions: 9 % at L = 8, 28 % at 10
            assert _sell_layout(rows)[1] * 64 > (0.05 if L == 8 else 0.2) * n_off
        _spmv_check(oracle_lib, M, ob.D, v, c, r)
    monkeypatch.setenv("TE_SPMV_VD", "0")
    P = capi.Matrix.from_coo(gpu, v, c, r, ob.D)
    monkeypatch.delenv("TE_SPMV_VD")
    assert P.storage()[0] == 1 and P.storage()[2] == 20 * len(v) + 4 * (ob.D + 1)
    x = random_state(ob.D, 3)
    yv, yp = A.spmv(x, -1j), P.spmv(x, -1j)
    assert np.linalg.norm(yv - yp) <= 1e-14 * np.linalg.norm(yp)
    if L == 8:
        rv = capi.Evolver(gpu, A, fock_state(ob.D), 3.0, 0.1, [], tol=1e-8).run()
        rp = capi.Evolver(gpu, P, fock_state(ob.D), 3.0, 0.1, [], tol=1e-8).run()
        assert rv["n_steps"] == rp["n_steps"] and rv["statusCode"] == rp["statusCode"]
        assert np.linalg.norm(rv["state"] - rp["state"]) <= 1e-12
        assert abs(rv["err"] - rp["err"]) <= 1e-12


def test_sliced_ell_slots_beyond_16_bit_span(gpu, oracle_lib):
    """Rows whose far entries (70,000 and 140,000 columns away) are present in some rows of a
    slice and not in others: the slots that hold them get padding in the other rows, since one
    slot spans fewer than 65536 columns. Slot count as the alignment rule gives, larger than the
    longest rows' count; the product matches the oracle's."""
    rng = np.random.default_rng(7)
    D = 150_000
    table = rng.normal(size=40) + 1j * rng.normal(size=40)
    rows, cols = [], []
    i = np.arange(D)
    for d, p in ((1, 1.0), (-1, 1.0), (3, 0.7), (70_000, 0.5), (-70_000, 0.5), (140_000, 0.3)):
        keep = rng.random(D) < p
        rows.append(i[keep])
        cols.append((i[keep] + d) % D)
    rows.append(i)
    cols.append(i)
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    key = np.unique(r * D + c, return_index=True)[1]   # one entry per (row, column)
    r, c = r[key].astype(np.uint64), c[key].astype(np.uint64)
    v = table[rng.integers(0, len(table), len(r))]
    A = capi.Matrix.from_coo(gpu, v, c, r, D)
    fmt, nv, nbytes = A.storage()
    assert fmt == 4 and nv <= len(table)
    ns = (D + 63) // 64
    slots = _aligned_slots(_sell_rows(c, r, np.arange(D)))
    longest = 64 * int(np.bincount(r[c != r].astype(np.int64), minlength=ns * 64).reshape(ns, 64).max(1).sum())
    assert slots > longest
    rows = _sell_rows(c, r, np.arange(D), v)
    assert _sell_layout(rows)[0] == slots and nbytes == _sell_bytes(rows, D, 16)
    _spmv_check(oracle_lib, A, D, v, c, r)


def test_sliced_ell_diagonal_kinds(gpu, oracle_lib, monkeypatch):
    """The sliced-ELL form streams the diagonal as 8-byte reals when every imaginary part is +0.0
    bit for bit, and as 16-byte complex values otherwise: a complex diagonal, or a single -0.0
    imaginary part (multiplying by (re, +0.0) instead of (re, -0.0) could change the sign of a
    zero). Products match the oracle in every case, and the row sums match the value-dictionary
    row-block kernel's bit for bit (same rows, diagonal first, then the entries in column order):
    compared at alpha = 1 and alpha = -0.3i, where y = alpha * sum rounds the same whichever of
    its two products the compiler fuses into a multiply-add (with a general complex alpha that
    choice may differ between the two kernels by one rounding). Rows in the basis order, and in
    the locality order of a built ring operator, where most entries are shared records (the
    same column offset and value on every row of a slice: leading, per-row and trailing parts
    of each row's sum)."""
    L = N = 8
    ob = oracle_lib.Basis.basis(N, L)
    terms = as_tuples(_ring_terms(L))
    v0, c0, r0 = oracle_lib.build_operator(ob, terms)
    D = ob.D
    states = ob.states().astype(np.int64)[:, _mode_significance(L, terms)]
    order = np.lexsort(states.T[::-1])                # new -> old
    inv = np.empty_like(order)
    inv[order] = np.arange(D)
    c1, r1 = inv[c0.astype(np.int64)], inv[r0.astype(np.int64)]
    srt = np.lexsort((c1, r1))                        # rows ascending, columns ascending
    x = random_state(D, 4)
    for v, c, r in ((v0, c0, r0), (v0[srt], c1[srt].astype(np.uint64), r1[srt].astype(np.uint64))):
        _sell_diagonal_kinds_case(gpu, oracle_lib, monkeypatch, v, c, r, D, x)


def _sell_diagonal_kinds_case(gpu, oracle_lib, monkeypatch, v, c, r, D, x):
    on = c == r
    assert np.all(np.ascontiguousarray(v[on].imag).view(np.uint64) == 0)
    rows = _sell_rows(c, r, np.arange(D), v)      # the diagonal does not enter the layout
    v_cplx = v.copy()
    v_cplx[on] = v[on] * (1.0 + 0.25j)
    v_negz = v.copy()
    k = int(np.flatnonzero(on)[D // 2])
    v_negz[k] = complex(v[k].real, -0.0)
    for vals, diag_bytes in ((v, 8), (v_cplx, 16), (v_negz, 16)):
        A = capi.Matrix.from_coo(gpu, vals, c, r, D)
        fmt, _, nbytes = A.storage()
        assert fmt == 4 and nbytes == _sell_bytes(rows, D, diag_bytes)
        _spmv_check(oracle_lib, A, D, vals, c, r)
        monkeypatch.setenv("TE_SPMV_SELL", "0")
        B = capi.Matrix.from_coo(gpu, vals, c, r, D)
        monkeypatch.delenv("TE_SPMV_SELL")
        assert B.storage()[0] == 2
        for alpha in (1.0, -0.3j):
            ya, yb = A.spmv(x, alpha), B.spmv(x, alpha)
            ndiff = int(np.sum(ya.view(np.uint64) != yb.view(np.uint64)))
            assert ndiff == 0, (diag_bytes, alpha, ndiff, float(np.abs(ya - yb).max()))


def test_scattered_columns_keep_row_blocks(gpu, oracle_lib):
    """A Hermitian matrix whose rows reach columns anywhere in 0..D (D = 1,000,000, no
    locality): one 16-bit span per slot would pad the aligned slices about sixfold (host count
    of the alignment rule on the first 300 slices), so the matrix keeps the value-dictionary
    row blocks (format 2); products and a short evolution match the oracle."""
    rng = np.random.default_rng(11)
    D = 1_000_000
    rows = np.repeat(np.arange(D), 3)
    cols = rng.integers(0, D, len(rows))
    keep = rows != cols
    rows, cols = rows[keep], cols[keep]
    key = np.unique(rows * D + cols)
    rows, cols = key // D, key % D
    vals = np.array([0.3, -0.7, 0.2 + 0.4j, 1.1j])[rng.integers(0, 4, len(rows))]
    r = np.concatenate([rows, cols, np.arange(D)]).astype(np.uint64)
    c = np.concatenate([cols, rows, np.arange(D)]).astype(np.uint64)
    v = np.concatenate([vals, np.conj(vals), rng.normal(size=D)]).astype(complex)
    A = capi.Matrix.from_coo(gpu, v, c, r, D)
    assert A.storage()[0] == 2
    _spmv_check(oracle_lib, A, D, v, c, r)
    psi = random_state(D, 2)
    ro = oracle_lib.evolve(D, (v, c, r), psi, 0.3, 0.1, [], tol=1e-8, m=30)
    rg = capi.Evolver(gpu, A, psi, 0.3, 0.1, [], tol=1e-8, m=30).run()
    assert rg["statusCode"] == ro["statusCode"]
    assert np.linalg.norm(rg["state"] - ro["state"]) <= ro["err"] + rg["err"] + 1e-10


def test_locality_order_evolution(gpu, oracle_lib, monkeypatch):
    """An operator built on a basis evolves in the lexicographic order of the states; every
    observable kind (diagonal, non-diagonal sparse, vector, host callback), the sampled state
    and the returned state come back in the basis order: same results as the basis-order
    evolver (to rounding) and as the oracle (within the error bound)."""
    L = N = 6
    terms = as_tuples(_ring_terms(L))
    ob = oracle_lib.Basis.basis(N, L)
    Hc = oracle_lib.build_operator(ob, terms)
    psi = random_state(ob.D, 21)
    proj = random_state(ob.D, 22)
    n2 = oracle_lib.build_operator(ob, as_tuples([createNumberOperator(2, 1)]))
    ro = oracle_lib.evolve(ob.D, Hc, psi, 2.0, 0.1, [("sparse", Hc), ("vector", proj), ("sparse", n2)], tol=1e-8)
    diag2 = np.zeros(ob.D)
    diag2[n2[2]] = n2[0].real
    res = {}
    configs = {"sell+lex": ({}, 4), "rows+lex": ({"TE_SPMV_SELL": "0"}, 3),
               "rows+basis": ({"TE_SPMV_SELL": "0", "TE_SPMV_ORDER": "0"}, 2)}
    for name, (env, fmt) in configs.items():
        for k, val in env.items():
            monkeypatch.setenv(k, val)
        gb = capi.Basis.basis(gpu, N, L)
        H = capi.operator_build(gpu, gb, terms)
        N2 = capi.operator_build(gpu, gb, [createNumberOperator(2, 1)])
        for k in env:
            monkeypatch.delenv(k)
        assert H.storage()[0] == fmt
        host = ("host", lambda st: float(np.sum(diag2 * np.abs(st) ** 2)))
        res[name] = capi.Evolver(gpu, H, psi, 2.0, 0.1, [H, ("vector", proj), host, N2], tol=1e-8).run()
    rg, rb = res["sell+lex"], res["rows+basis"]
    for other in (res["rows+lex"], rb):
        assert rg["n_steps"] == other["n_steps"] and rg["statusCode"] == other["statusCode"]
        assert np.linalg.norm(rg["state"] - other["state"]) <= 1e-12
        assert np.abs(rg["observables"] - other["observables"]).max() <= 1e-12
    bound = ro["err"] + rg["err"]
    assert np.linalg.norm(rg["state"] - ro["state"]) <= bound + 1e-10
    for k, ko in ((0, 0), (1, 1), (2, 2), (3, 2)):
        assert np.abs(rg["observables"][k] - ro["observables"][ko]).max() <= 1e-9


def test_value_dictionary_fallbacks(gpu, oracle_lib):
    """Matrices the dictionary cannot hold keep the plain CSR kernels, with correct products:
    too many distinct values, repeated diagonal entries (COO duplicates are summed), rows longer
    than a row block, a diagonal-only matrix (the evolver reads those as dense diagonals);
    rows without a diagonal entry and a table of one value do qualify."""
    rng = np.random.default_rng(0)
    n = 3000
    # random values: ~20000 distinct
    r = rng.integers(0, n, 20000).astype(np.uint64)
    c = rng.integers(0, n, 20000).astype(np.uint64)
    v = rng.normal(size=20000) + 1j * rng.normal(size=20000)
    M = capi.Matrix.from_coo(gpu, v, c, r, n)
    assert M.storage()[0] == 1
    _spmv_check(oracle_lib, M, n, v, c, r)
    # few distinct values, but every diagonal entry given twice
    idx = np.arange(n, dtype=np.uint64)
    nb = (idx + 1) % n
    r2 = np.concatenate([idx, idx, idx, nb])
    c2 = np.concatenate([idx, idx, nb, idx])
    v2 = np.concatenate([np.full(n, 0.5), np.full(n, 0.25), np.full(n, -1.0 + 0.5j), np.full(n, -1.0 - 0.5j)])
    M = capi.Matrix.from_coo(gpu, v2, c2, r2, n)
    assert M.storage()[0] == 1
    _spmv_check(oracle_lib, M, n, v2, c2, r2)
    # one row with 400 off-diagonal entries
    r3 = np.concatenate([idx, np.zeros(400, np.uint64)])
    c3 = np.concatenate([idx, np.arange(1, 401, dtype=np.uint64)])
    v3 = np.concatenate([np.linspace(-1, 1, n), np.full(400, 0.125j)])
    M = capi.Matrix.from_coo(gpu, v3, c3, r3, n)
    assert M.storage()[0] == 0
    _spmv_check(oracle_lib, M, n, v3, c3, r3)
    # diagonal only
    vv = np.linspace(-2, 3, n) + 0.5j
    M = capi.Matrix.from_coo(gpu, vv, idx, idx, n)
    assert M.storage()[0] == 1
    _spmv_check(oracle_lib, M, n, vv, idx, idx)
    # off-diagonal band with no diagonal in half of the rows
    rr, cc = np.concatenate([idx[::2], idx, nb]), np.concatenate([idx[::2], nb, idx])
    vv = np.concatenate([np.arange(n // 2) * 0.01, np.full(n, 0.3), np.full(n, 0.3)])
    M = capi.Matrix.from_coo(gpu, vv, cc, rr, n)
    fmt, nv, _ = M.storage()
    assert fmt == 4 and nv == _distinct_offdiag(vv, cc, rr)[0] == 1
    _spmv_check(oracle_lib, M, n, vv, cc, rr)


@pytest.mark.parametrize("storage", ["sell", "rows", "csr"])
def test_irregular_rows_evolution(gpu, oracle_lib, monkeypatch, storage):
    """A Hermitian matrix with rows of 0 to 40 off-diagonal entries (few distinct values),
    evolved with each storage of the Lanczos product: sliced ELL with length-sorted windows and
    padded slots, value-dictionary row blocks, plain CSR row blocks. All within the error bound
    of the oracle, and the same steps and states to rounding."""
    rng = np.random.default_rng(7)
    n = 4099                                             # not a multiple of the 64-row slice
    deg = rng.integers(0, 21, n)
    rows = np.repeat(np.arange(n), deg)
    cols = rng.integers(0, n, len(rows))
    keep = rows != cols
    rows, cols = rows[keep], cols[keep]
    key = np.unique(rows * n + cols)
    rows, cols = key // n, key % n
    vals = np.array([0.3, -0.7, 0.2 + 0.4j, 1.1j])[rng.integers(0, 4, len(rows))]
    # Hermitian: A + A^H, then a diagonal
    r = np.concatenate([rows, cols, np.arange(n)])
    c = np.concatenate([cols, rows, np.arange(n)])
    v = np.concatenate([vals, np.conj(vals), rng.normal(size=n)])
    r, c, v = r.astype(np.uint64), c.astype(np.uint64), v.astype(complex)
    psi = random_state(n, 3)
    env = {"sell": {}, "rows": {"TE_SPMV_SELL": "0"}, "csr": {"TE_SPMV_VD": "0"}}[storage]
    for k, val in env.items():
        monkeypatch.setenv(k, val)
    H = capi.Matrix.from_coo(gpu, v, c, r, n)
    for k in env:
        monkeypatch.delenv(k)
    fmt = H.storage()[0]
    assert fmt == {"sell": 4, "rows": 2, "csr": 1}[storage]
    _spmv_check(oracle_lib, H, n, v, c, r)
    rg = capi.Evolver(gpu, H, psi, 1.0, 0.1, [("vector", random_state(n, 4))], tol=1e-8, m=30).run()
    ro = oracle_lib.evolve(n, (v, c, r), psi, 1.0, 0.1, [("vector", random_state(n, 4))], tol=1e-8, m=30)
    assert rg["n_steps"] == ro["n_steps"] and rg["statusCode"] == ro["statusCode"]
    assert np.linalg.norm(rg["state"] - ro["state"]) <= rg["err"] + ro["err"] + 1e-10
    assert np.abs(rg["observables"] - ro["observables"]).max() <= 1e-9


# ---------------------------------------------------------------------------------------- evolution
def _gpu_model(gpu, basis_kind, args, terms, obs_terms):
    b = capi.Basis.tensor(gpu, *args) if basis_kind == "tensor" else capi.Basis.basis(gpu, *args)
    H = capi.operator_build(gpu, b, terms)
    obs = [capi.operator_build(gpu, b, t) for t in obs_terms]
    return b, H, obs


def _bh_gpu(gpu, N0, Nm, K, maxT, step):
    args, ham, init = blackhole(N0=N0, Nm=Nm, K=K)
    b, H, obs = _gpu_model(gpu, "tensor", args, ham, number_operator_terms(args[1] + args[3]))
    psi = np.zeros(b.D, complex)
    psi[b.find(init)[0]] = 1
    ev = capi.Evolver(gpu, H, psi, maxT, step, obs, tol=1e-8, m=40)
    return ev.run()


def test_blackhole_golden_gpu(gpu):
    r = _bh_gpu(gpu, 20, 2, 4, 10.0, 0.01)
    g = np.loadtxt(os.path.join(GOLD, "blackhole_N20_Nm2_K4.txt"))
    print("black hole: n_steps", r["n_steps"], "err", r["err"], "max |diff|", np.abs(r["observables"] - g).max())
    assert np.abs(r["observables"] - g).max() <= 1e-8
    assert r["observables"][0, 0] == 20.000000000000004
    assert r["numSamples"] == 1001 and r["err"] <= 1e-8


def test_lucky_breakdown_golden_gpu(gpu):
    r = _bh_gpu(gpu, 1, 1, 1, 1000.0, 1.0)
    g = np.loadtxt(os.path.join(GOLD, "blackhole_N1_Nm1_K1.txt"))
    assert np.abs(r["observables"] - g).max() <= 1e-8
    assert r["statusCode"] == 1 and r["krylovDim"] <= 4 and r["n_steps"] == 1


def test_simple_example_csv_gpu(gpu):
    args, terms, init = two_oscillators()
    b, H, obs = _gpu_model(gpu, "basis", args, terms, number_operator_terms(2))
    psi = np.zeros(b.D, complex)
    psi[b.find(init)[0]] = 1
    r = capi.Evolver(gpu, H, psi, 10.0, 0.01, obs).run()     # default constructor: tol 1e-6, m 40
    for k in range(2):
        txt = ", ".join("%g" % v for v in r["observables"][k])
        with open(os.path.join(GOLD, "simple_example_mode%d.csv" % k)) as f:
            assert txt == f.read()


def _ring_pair(gpu, oracle, L, N, psi_fn, T, step, m=40, tol=1e-8, **kw):
    terms = as_tuples(_ring_terms(L))
    ob = oracle.Basis.basis(N, L)
    H = oracle.build_operator(ob, terms)
    oobs = [("sparse", oracle.build_operator(ob, as_tuples([createNumberOperator(k, 1)]))) for k in range(L)]
    psi = psi_fn(ob.D)
    ro = oracle.evolve(ob.D, H, psi, T, step, oobs, tol=tol, m=m, **kw)
    b, Hg, gobs = _gpu_model(gpu, "basis", (N, L), terms, [[createNumberOperator(k, 1)] for k in range(L)])
    gkw = {"fast_integration": kw["fastIntegration"]} if "fastIntegration" in kw else {}
    if "expFactor" in kw:
        gkw["exp_factor"] = kw["expFactor"]
    rg = capi.Evolver(gpu, Hg, psi, T, step, gobs, tol=tol, m=m, **gkw).run()
    return ro, rg


def _check_pair(ro, rg, tol=1e-8):
    """States within the sum of the two error bounds; observables within the requested tol."""
    bound = ro["err"] + rg["err"]
    dpsi = np.linalg.norm(rg["state"] - ro["state"])
    dobs = np.abs(rg["observables"] - ro["observables"]).max()
    print("steps gpu/oracle %d/%d  err %.3e/%.3e  |dpsi| %.3e  max|dobs| %.3e  status %d/%d" %
          (rg["n_steps"], ro["n_steps"], rg["err"], ro["err"], dpsi, dobs, rg["statusCode"], ro["statusCode"]))
    assert dpsi <= bound + 1e-10
    assert dobs <= tol
    assert rg["numSamples"] == ro["numSamples"] == ro["n_samples"]
    assert rg["statusCode"] == ro["statusCode"]


def test_ring_small_config_vs_oracle(gpu, oracle_lib):
    """Ring Bose-Hubbard L = N = 8 (D = 6435), Fock |8,0,...>, T = 10, m = 40, tol = 1e-8, step 0.1."""
    ro, rg = _ring_pair(gpu, oracle_lib, 8, 8, fock_state, 10.0, 0.1)
    _check_pair(ro, rg)


def test_evolver_runs_are_bitwise_repeatable(gpu):
    """The Lanczos scalars come from block partials summed in block order by separate finish
    launches, and the observables likewise: two runs of one evolver, and a run of a second evolver
    on the same operator, give the same state, samples, error and step count bit for bit (ring
    L = N = 10, D = 92,378, sliced ELL over many blocks)."""
    L = N = 10
    gb = capi.Basis.basis(gpu, N, L)
    H = capi.operator_build(gpu, gb, as_tuples(_ring_terms(L)))
    assert H.storage()[0] == 4
    obs = [capi.operator_build(gpu, gb, as_tuples([createNumberOperator(k, 1)])) for k in range(3)]
    psi = random_state(gb.D, 8)
    ev = capi.Evolver(gpu, H, psi, 2.0, 0.5, obs, tol=1e-8)
    runs = [ev.run(), ev.run(), capi.Evolver(gpu, H, psi, 2.0, 0.5, obs, tol=1e-8).run()]
    for r in runs[1:]:
        assert r["n_steps"] == runs[0]["n_steps"] and r["statusCode"] == runs[0]["statusCode"]
        assert np.float64(r["err"]).view(np.uint64) == np.float64(runs[0]["err"]).view(np.uint64)
        assert np.array_equal(r["state"].view(np.uint64), runs[0]["state"].view(np.uint64))
        assert np.array_equal(r["observables"].view(np.uint64), runs[0]["observables"].view(np.uint64))


def test_eigen_solvers_agree(gpu, monkeypatch):
    """The step controller's eigen-decomposition by bisection and twisted factorisations (the
    default) against the QL iteration (TE_EIG=ql, read when the evolver is created), ring
    L = N = 10 from a random state over 12 Krylov steps: the same step count and status, states
    and observables far inside the error bound."""
    L = N = 10
    gb = capi.Basis.basis(gpu, N, L)
    H = capi.operator_build(gpu, gb, as_tuples(_ring_terms(L)))
    obs = [capi.operator_build(gpu, gb, as_tuples([createNumberOperator(k, 1)])) for k in range(L)]
    psi = random_state(gb.D, 12)
    rf = capi.Evolver(gpu, H, psi, 6.0, 0.5, obs, tol=1e-8).run()
    monkeypatch.setenv("TE_EIG", "ql")
    rq = capi.Evolver(gpu, H, psi, 6.0, 0.5, obs, tol=1e-8).run()
    dpsi = np.linalg.norm(rf["state"] - rq["state"])
    dobs = np.abs(rf["observables"] - rq["observables"]).max()
    print("steps %d/%d err %.6e/%.6e |dpsi| %.2e max|dobs| %.2e" %
          (rf["n_steps"], rq["n_steps"], rf["err"], rq["err"], dpsi, dobs))
    assert rf["n_steps"] == rq["n_steps"] and rf["statusCode"] == rq["statusCode"]
    assert abs(rf["err"] - rq["err"]) <= 1e-3 * rq["err"]
    assert dpsi <= 1e-10 and dobs <= 1e-10


def test_ring_medium_config_vs_oracle(gpu, oracle_lib):
    """Ring Bose-Hubbard L = N = 12 (D = 1,352,078), seeded random complex initial state, T = 10,
    m = 40, tol = 1e-8, step 0.1 (101 samples of the 12 n_k); the basis and CSR it runs on are
    the ones test_ring_operator_bitexact[12-12] checks bit for bit."""
    ro, rg = _ring_pair(gpu, oracle_lib, 12, 12, lambda D: random_state(D, 9), 10.0, 0.1)
    _check_pair(ro, rg)
    assert rg["statusCode"] == 10


def test_ring_random_state_vs_oracle(gpu, oracle_lib):
    ro, rg = _ring_pair(gpu, oracle_lib, 6, 6, lambda D: random_state(D, 9), 3.0, 0.05, m=30)
    _check_pair(ro, rg)


@pytest.mark.parametrize("f", [-1.0, 0.5, 2.5])
def test_exp_factor_vs_oracle(gpu, oracle_lib, f):
    """exp(-i f H t), the constructor's expFactor (core/krylovTimeEvolver.cpp:31, applied in every
    Lanczos product at :530): L = N = 6 ring, random complex state, T = 3, step 0.1, m = 40,
    against the oracle with the same factor (the same bars as every other evolution)."""
    ro, rg = _ring_pair(gpu, oracle_lib, 6, 6, lambda D: random_state(D, 21), 3.0, 0.1, expFactor=f)
    _check_pair(ro, rg)


@pytest.mark.parametrize("m,tol", [(4, 1e-6), (10, 1e-10), (64, 1e-8)])
def test_krylov_dimension_and_tol_vs_oracle(gpu, oracle_lib, m, tol):
    """The constructor's mm and tol (core/krylovTimeEvolver.cpp:26, :35): a short basis at the
    default tol (many Krylov steps), a tight tol, and the largest Krylov dimension the device
    takes, on the L = N = 6 ring from a random state to T = 3 (step 0.1) against the oracle."""
    ro, rg = _ring_pair(gpu, oracle_lib, 6, 6, lambda D: random_state(D, 23), 3.0, 0.1, m=m, tol=tol)
    _check_pair(ro, rg, tol=tol)
    assert rg["krylovDim"] == ro["krylovDim"] == m


def test_krylov_dimension_limit(gpu):
    """min(mm, dimension) above 64 is refused with TE_ERR_UNSUPPORTED; mm = 65 on a dimension-10
    basis runs with m = 10 (:35)."""
    L = N = 6
    b, H, _ = _gpu_model(gpu, "basis", (N, L), as_tuples(_ring_terms(L)), [])
    with pytest.raises(capi.TEError) as e:
        capi.Evolver(gpu, H, random_state(b.D, 1), 1.0, 0.5, [], tol=1e-8, m=65)
    assert e.value.code == capi.TE_ERR_UNSUPPORTED and "65" in str(e.value)
    b3, H3, _ = _gpu_model(gpu, "basis", (3, 3), as_tuples(_ring_terms(3)), [])
    assert b3.D == 10
    r = capi.Evolver(gpu, H3, random_state(b3.D, 1), 1.0, 0.5, [], tol=1e-8, m=65).run()
    assert r["krylovDim"] == 10 and r["err"] <= 1e-8


@pytest.mark.parametrize("T,step,ns", [(1.0, 0.3, 4), (0.95, 0.1, 10), (1.0, 2.0, 1)])
def test_sampling_grid_vs_oracle(gpu, oracle_lib, T, step, ns):
    """Sampling steps that do not divide t, and one longer than t: n_samples = floor(t/step) + 1
    (core/krylovTimeEvolver.cpp:57), samples taken where the reference takes them (:405-437), and
    the loop ends once every sample is taken (:332), so with one sample (t = 0) no Krylov step
    runs and the returned state is psi0 at evolvedTime 0; L = N = 5 ring from a random state,
    against the oracle."""
    psi0 = lambda D: random_state(D, 31)
    ro, rg = _ring_pair(gpu, oracle_lib, 5, 5, psi0, T, step)
    _check_pair(ro, rg)
    assert rg["n_samples"] == ro["n_samples"] == ns
    assert rg["n_steps"] == ro["n_steps"] and rg["evolvedTime"] == ro["evolvedTime"]
    if ns == 1:
        assert rg["n_steps"] == 0 and rg["evolvedTime"] == 0.0
        assert np.array_equal(rg["state"], psi0(len(rg["state"])))
    else:
        assert rg["evolvedTime"] >= (ns - 1) * step


def test_diagonal_hamiltonian_closed_form(gpu, oracle_lib):
    """No hopping (J = 0, the on-site and interaction terms only): H is diagonal, its sliced-ELL
    copy has no off-diagonal slot and no shared record. psi(T) = exp(-i d T) psi0 exactly and
    every n_k is constant in time; L = N = 6 from a random state to T = 3 (step 0.1), against
    that closed form within the returned error bound and against the oracle."""
    L = N = 6
    eps, J = ring_couplings(L, 11)
    terms = as_tuples(ring_bose_hubbard_terms(L, eps, np.zeros(L), 0.5))
    ob = oracle_lib.Basis.basis(N, L)
    Hc = oracle_lib.build_operator(ob, terms)
    assert np.array_equal(Hc[1], Hc[2])                      # diagonal only
    oobs = [("sparse", oracle_lib.build_operator(ob, as_tuples([createNumberOperator(k, 1)]))) for k in range(L)]
    psi0 = random_state(ob.D, 41)
    ro = oracle_lib.evolve(ob.D, Hc, psi0, 3.0, 0.1, oobs, tol=1e-8, m=40)
    b, H, gobs = _gpu_model(gpu, "basis", (N, L), terms, [[createNumberOperator(k, 1)] for k in range(L)])
    assert H.nnz == ob.D
    rg = capi.Evolver(gpu, H, psi0, 3.0, 0.1, gobs, tol=1e-8, m=40).run()
    _check_pair(ro, rg)
    d = np.zeros(ob.D)
    d[Hc[2].astype(np.int64)] = Hc[0].real
    exact = np.exp(-1j * d * 3.0) * psi0
    occ = ob.states().astype(float)
    n0 = (np.abs(psi0) ** 2) @ occ                           # <n_k> at t = 0, for all t
    print("diagonal H: |psi - exact| %.3e err %.3e max|n_k(t) - n_k(0)| %.3e" %
          (np.linalg.norm(rg["state"] - exact), rg["err"], np.abs(rg["observables"] - n0[:, None]).max()))
    assert np.linalg.norm(rg["state"] - exact) <= rg["err"] + 1e-12
    assert np.abs(rg["observables"] - n0[:, None]).max() <= 1e-12


def test_exp_factor_round_trip(gpu):
    """exp(+iHT) exp(-iHT) psi0 = psi0: the L = N = 8 ring from a random state forward to T = 2
    (expFactor 1) and back (expFactor -1), within the sum of the two returned error bounds."""
    L = N = 8
    b, H, _ = _gpu_model(gpu, "basis", (N, L), as_tuples(_ring_terms(L)), [])
    psi0 = random_state(b.D, 5)
    fw = capi.Evolver(gpu, H, psi0, 2.0, 2.0, [], tol=1e-8).run()
    bw = capi.Evolver(gpu, H, fw["state"], 2.0, 2.0, [], exp_factor=-1.0, tol=1e-8).run()
    d = np.linalg.norm(bw["state"] - psi0)
    print("round trip: |psi_back - psi0| %.3e, err %.3e + %.3e, |psi(T) - psi0| %.3e" %
          (d, fw["err"], bw["err"], np.linalg.norm(fw["state"] - psi0)))
    assert np.linalg.norm(fw["state"] - psi0) > 0.1          # the state did move
    assert d <= fw["err"] + bw["err"] + 1e-10


def test_complex_hermitian_ring_vs_oracle(gpu, oracle_lib):
    """A ring threaded by a flux: hops t_k a_{k+1}^dag a_k + conj(t_k) a_k^dag a_{k+1} with
    t_k = J_k exp(i phi_k), so H is Hermitian with complex off-diagonal values (its value table,
    the builder's storage of non-real terms and every product run on complex values). Built
    bit-exactly and evolved against the oracle from a random state."""
    from timeevolver_amd.hamiltonian import OpTerm, createFourPointB
    L = N = 6
    rng = np.random.default_rng(5)
    eps, J = ring_couplings(L, 13)
    phi = rng.uniform(-np.pi, np.pi, L)
    terms = [createNumberOperator(k, eps[k]) for k in range(L)]
    for k in range(L):
        t = J[k] * np.exp(1j * phi[k])
        terms.append(OpTerm(complex(t), [-1, 1], [k, (k + 1) % L], [0, 0]))
        terms.append(OpTerm(complex(np.conj(t)), [-1, 1], [(k + 1) % L, k], [0, 0]))
    terms += [createFourPointB(k, k, k, k, 0, 0, 0, 0, 0.25) for k in range(L)]
    terms = as_tuples(terms)
    ob = oracle_lib.Basis.basis(N, L)
    H = oracle_lib.build_operator(ob, terms)
    assert np.abs(H[0].imag).max() > 0.1                       # genuinely complex entries
    Hd = np.zeros((ob.D, ob.D), complex)
    np.add.at(Hd, (H[2].astype(np.int64), H[1].astype(np.int64)), H[0])
    assert np.abs(Hd - Hd.conj().T).max() == 0.0              # Hermitian as stored
    oobs = [("sparse", oracle_lib.build_operator(ob, as_tuples([createNumberOperator(k, 1)]))) for k in range(L)]
    psi = random_state(ob.D, 12)
    ro = oracle_lib.evolve(ob.D, H, psi, 4.0, 0.1, oobs, tol=1e-8, m=40)
    b, Hg, gobs = _gpu_model(gpu, "basis", (N, L), terms, [[createNumberOperator(k, 1)] for k in range(L)])
    v, c, r = Hg.to_coo()
    assert np.array_equal(c, H[1]) and np.array_equal(v.view(np.uint64), H[0].view(np.uint64))
    rg = capi.Evolver(gpu, Hg, psi, 4.0, 0.1, gobs, tol=1e-8, m=40).run()
    _check_pair(ro, rg)


def test_fast_integration_vs_oracle(gpu, oracle_lib):
    ro, rg = _ring_pair(gpu, oracle_lib, 6, 6, fock_state, 2.0, 0.1, fastIntegration=True)
    _check_pair(ro, rg)


def test_generic_observables(gpu, oracle_lib):
    """Non-diagonal sparse (H itself: energy), vector and host-callback observables."""
    L = N = 5
    terms = as_tuples(_ring_terms(L))
    ob = oracle_lib.Basis.basis(N, L)
    Hc = oracle_lib.build_operator(ob, terms)
    psi = random_state(ob.D, 4)
    proj = random_state(ob.D, 8)
    n0 = oracle_lib.build_operator(ob, as_tuples([createNumberOperator(0, 1)]))
    ro = oracle_lib.evolve(ob.D, Hc, psi, 2.0, 0.1, [("sparse", Hc), ("vector", proj), ("sparse", n0)], tol=1e-8)
    H = capi.Matrix.from_coo(gpu, *Hc, ob.D)
    N0 = capi.Matrix.from_coo(gpu, *n0, ob.D)
    diag0 = np.zeros(ob.D)
    diag0[n0[2]] = n0[0].real
    host = ("host", lambda st: float(np.sum(diag0 * np.abs(st) ** 2)))
    rg = capi.Evolver(gpu, H, psi, 2.0, 0.1, [H, ("vector", proj), host, N0], tol=1e-8).run()
    bound = ro["err"] + rg["err"]
    assert np.linalg.norm(rg["state"] - ro["state"]) <= bound + 1e-10
    e0 = ro["observables"][0, 0]
    assert np.abs(rg["observables"][0] - ro["observables"][0]).max() <= 1e-9       # energy
    assert np.abs(rg["observables"][0] - e0).max() <= 1e-9                            # conserved
    assert np.abs(rg["observables"][1] - ro["observables"][1]).max() <= 1e-9
    assert np.abs(rg["observables"][2] - ro["observables"][2]).max() <= 1e-9         # host == diag
    assert np.abs(rg["observables"][3] - ro["observables"][2]).max() <= 1e-9


def test_stop_request(gpu, oracle_lib):
    """An observable throwing requestStopException (core/krylovTimeEvolver.cpp:423-436)."""
    L = N = 4
    ob = oracle_lib.Basis.basis(N, L)
    Hc = oracle_lib.build_operator(ob, as_tuples(_ring_terms(L)))
    H = capi.Matrix.from_coo(gpu, *Hc, ob.D)
    calls = {"n": 0}

    def obs(st):
        calls["n"] += 1
        if calls["n"] == 6:
            raise StopIteration
        return 0.0

    r = capi.Evolver(gpu, H, fock_state(ob.D), 2.0, 0.1, [("host", obs)], tol=1e-8).run()
    ro = oracle_lib.evolve(ob.D, Hc, fock_state(ob.D), 2.0, 0.1, [], tol=1e-8, stop_at_sample=5)
    assert r["statusCode"] == 3 and r["numSamples"] == 5
    assert np.linalg.norm(r["state"] - ro["state"]) <= r["err"] + ro["err"] + 1e-10


@pytest.mark.parametrize("at", [0, 5])
def test_stop_request_records_earlier_observables(gpu, oracle_lib, at):
    """sample() evaluates the observables in list order (core/krylovTimeEvolver.cpp:120-126), so
    at the stop sample the ones listed before the stopping observable have recorded a value and
    the later ones have not: a diagonal and a vector observable before a host stopper, a diagonal
    and a vector observable after it. Values and per-observable sample counts (sampleIndex)
    against the oracle with the stopper at the same place (at sample 0 and sample 5)."""
    L = N = 5
    ob = oracle_lib.Basis.basis(N, L)
    Hc = oracle_lib.build_operator(ob, as_tuples(_ring_terms(L)))
    n0 = oracle_lib.build_operator(ob, as_tuples([createNumberOperator(0, 1)]))
    n2 = oracle_lib.build_operator(ob, as_tuples([createNumberOperator(2, 1)]))
    p1, p2 = random_state(ob.D, 31), random_state(ob.D, 32)
    psi = random_state(ob.D, 33)
    H = capi.Matrix.from_coo(gpu, *Hc, ob.D)
    N0, N2 = capi.Matrix.from_coo(gpu, *n0, ob.D), capi.Matrix.from_coo(gpu, *n2, ob.D)
    calls = {"n": 0}

    def stopper(st):
        calls["n"] += 1
        if calls["n"] == at + 1:
            raise StopIteration
        return 0.0

    r = capi.Evolver(gpu, H, psi, 2.0, 0.1, [N0, ("vector", p1), ("host", stopper), N2, ("vector", p2)],
                     tol=1e-8).run()
    ro = oracle_lib.evolve(ob.D, Hc, psi, 2.0, 0.1, [("sparse", n0), ("vector", p1), ("sparse", n2), ("vector", p2)],
                           tol=1e-8, stop_at_sample=at, stop_obs=2)
    assert r["statusCode"] == ro["statusCode"] == (3 if at else 0)
    assert r["numSamples"] == ro["numSamples"] == at
    assert r["recorded"] == [at + 1, at + 1, at, at, at]
    for k, ko in ((0, 0), (1, 1), (3, 2), (4, 3)):
        assert np.abs(r["observables"][k] - ro["observables"][ko]).max() <= 1e-8
    assert np.all(r["observables"][3:, at:] == 0) and np.all(r["observables"][:2, at + 1:] == 0)
    assert np.linalg.norm(r["state"] - ro["state"]) <= r["err"] + ro["err"] + 1e-10


@pytest.mark.parametrize("bytes_form", [True, False])
def test_tiled_combine_many_samples(gpu, oracle_lib, monkeypatch, bytes_form):
    """Steps that cover more samples than one pass of combine_kernel takes (9 diagonal
    observables: 2 samples a pass) go through combine_tiled_kernel, 8 samples a pass. Against
    passes of combine_kernel only (TE_COMBINE_TILED=0): the same states, errors and step counts
    bit for bit (each sample's state is the same column-by-column sum), the observables to 1e-13
    (the same terms added in another order), and both within tol of the oracle; with the
    observables held as byte tables and as 8-byte diagonals (TE_DIAG_BYTES=0)."""
    L = N = 8
    ob = oracle_lib.Basis.basis(N, L)
    Hc = oracle_lib.build_operator(ob, as_tuples(_ring_terms(L)))
    rng = np.random.default_rng(77)
    d = rng.normal(size=ob.D)                                    # 6435 distinct values: 8 bytes
    d = d if not bytes_form else np.round(d * 4) / 4             # a few dozen values: one byte
    idx = np.arange(ob.D, dtype=np.int64)
    dcoo = (d.astype(np.complex128), idx, idx)
    nks = [oracle_lib.build_operator(ob, as_tuples([createNumberOperator(k, 1)])) for k in range(L)]
    psi = random_state(ob.D, 43)
    T, step = 2.0, 0.02                                          # many samples a Krylov step
    res = {}
    for tiled in ("1", "0"):
        monkeypatch.setenv("TE_COMBINE_TILED", tiled)
        gb = capi.Basis.basis(gpu, N, L)
        H = capi.operator_build(gpu, gb, as_tuples(_ring_terms(L)))
        obs = [capi.operator_build(gpu, gb, [createNumberOperator(k, 1)]) for k in range(L)]
        obs.append(capi.Matrix.from_coo(gpu, *dcoo, ob.D))
        res[tiled] = capi.Evolver(gpu, H, psi, T, step, obs, tol=1e-8).run()
    monkeypatch.delenv("TE_COMBINE_TILED", raising=False)
    a, b = res["1"], res["0"]
    print("steps %d samples %d max|dobs| tiled vs passes %.3e" %
          (a["n_steps"], a["numSamples"], np.abs(a["observables"] - b["observables"]).max()))
    assert a["numSamples"] > 4 * a["n_steps"]                    # steps with more than 2 samples
    assert a["n_steps"] == b["n_steps"] and a["statusCode"] == b["statusCode"]
    assert np.float64(a["err"]).view(np.uint64) == np.float64(b["err"]).view(np.uint64)
    assert np.array_equal(a["state"].view(np.uint64), b["state"].view(np.uint64))
    assert np.abs(a["observables"] - b["observables"]).max() <= 1e-13
    ro = oracle_lib.evolve(ob.D, Hc, psi, T, step, [("sparse", o) for o in nks] + [("sparse", dcoo)], tol=1e-8)
    assert np.abs(a["observables"] - ro["observables"]).max() <= 1e-8
    assert np.linalg.norm(a["state"] - ro["state"]) <= a["err"] + ro["err"] + 1e-10


@pytest.mark.parametrize("n_distinct", [2, 256, 257, 5000])
def test_diagonal_observables_byte_tables(gpu, oracle_lib, monkeypatch, n_distinct):
    """Diagonal observables with at most 256 distinct values are read by the combine as one-byte
    indices into a table of those values (the same doubles): the observables are bit for bit
    those of the 8-byte diagonals (TE_DIAG_BYTES=0), with 256 values and with more (257, 5000:
    every diagonal stays 8 bytes when one has too many), next to the n_k observables; and within
    tol of the oracle."""
    L = N = 8
    ob = oracle_lib.Basis.basis(N, L)
    Hc = oracle_lib.build_operator(ob, as_tuples(_ring_terms(L)))
    rng = np.random.default_rng(n_distinct)
    vals = rng.normal(size=n_distinct)
    d = vals[rng.integers(0, n_distinct, ob.D)]
    d[:n_distinct] = vals                                   # every value present
    idx = np.arange(ob.D, dtype=np.int64)
    dcoo = (d.astype(np.complex128), idx, idx)
    nks = [oracle_lib.build_operator(ob, as_tuples([createNumberOperator(k, 1)])) for k in range(L)]
    psi = random_state(ob.D, 41)
    res = {}
    for mode in ("bytes", "doubles"):
        if mode == "doubles":
            monkeypatch.setenv("TE_DIAG_BYTES", "0")
        gb = capi.Basis.basis(gpu, N, L)
        H = capi.operator_build(gpu, gb, as_tuples(_ring_terms(L)))
        obs = [capi.operator_build(gpu, gb, [createNumberOperator(k, 1)]) for k in range(L)]
        obs.append(capi.Matrix.from_coo(gpu, *dcoo, ob.D))
        res[mode] = capi.Evolver(gpu, H, psi, 2.0, 0.25, obs, tol=1e-8).run()
        monkeypatch.delenv("TE_DIAG_BYTES", raising=False)
    a, b = res["bytes"], res["doubles"]
    assert np.array_equal(a["observables"].view(np.uint64), b["observables"].view(np.uint64))
    assert np.array_equal(a["state"].view(np.uint64), b["state"].view(np.uint64))
    ro = oracle_lib.evolve(ob.D, Hc, psi, 2.0, 0.25, [("sparse", o) for o in nks] + [("sparse", dcoo)], tol=1e-8)
    assert np.abs(a["observables"] - ro["observables"]).max() <= 1e-8


def test_fatal_unnormalised(gpu, oracle_lib):
    L = N = 3
    ob = oracle_lib.Basis.basis(N, L)
    Hc = oracle_lib.build_operator(ob, as_tuples(_ring_terms(L)))
    H = capi.Matrix.from_coo(gpu, *Hc, ob.D)
    with pytest.raises(capi.TEError) as e:
        capi.Evolver(gpu, H, 2 * fock_state(ob.D), 1.0, 0.1, [], tol=1e-8)
    assert e.value.code == capi.TE_ERR_FATAL


@pytest.mark.parametrize("part_mb", ["0.04", "0.02"])
def test_column_split_spmv_and_evolution(gpu, oracle_lib, monkeypatch, part_mb):
    """The column-split SpMV (x slices of at most TE_SPMV_PART_MB) gives the unsplit results:
    3 and 5 slices of the L = N = 8 ring (x = 103 KB)."""
    L = N = 8
    terms = as_tuples(_ring_terms(L))
    ob = oracle_lib.Basis.basis(N, L)
    v, c, r = oracle_lib.build_operator(ob, terms)
    monkeypatch.setenv("TE_SPMV_PART_MB", part_mb)
    monkeypatch.setenv("TE_SPMV_VD", "0")            # the split applies to the plain CSR copy
    A = capi.Matrix.from_coo(gpu, v, c, r, ob.D)
    gb = capi.Basis.basis(gpu, N, L)
    Hg = capi.operator_build(gpu, gb, terms)
    monkeypatch.delenv("TE_SPMV_PART_MB")
    monkeypatch.delenv("TE_SPMV_VD")
    assert A.storage()[0] == 1 and Hg.storage()[0] == 1
    x = random_state(ob.D, 5)
    for M in (A, Hg):
        y = M.spmv(x, -0.3j)
        yr = oracle_lib.spmv(ob.D, v, c, r, x, -0.3j)
        assert np.linalg.norm(y - yr) <= 1e-14 * np.linalg.norm(yr)
    obs = [capi.operator_build(gpu, gb, [createNumberOperator(k, 1)]) for k in range(L)]
    rg = capi.Evolver(gpu, Hg, fock_state(ob.D), 3.0, 0.1, obs, tol=1e-8).run()
    oobs = [("sparse", oracle_lib.build_operator(ob, as_tuples([createNumberOperator(k, 1)]))) for k in range(L)]
    ro = oracle_lib.evolve(ob.D, (v, c, r), fock_state(ob.D), 3.0, 0.1, oobs, tol=1e-8)
    _check_pair(ro, rg)
