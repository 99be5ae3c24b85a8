"""GPU check of te_spmv_noalloc, the allocation-free variant of te_spmv: through the C-ABI, the
same products bit for bit as te_spmv (same kernels) on consecutive calls of different sizes (the
second smaller than the first, then larger again, so the shared staging array holds a larger
call's data when a smaller one runs), within 1e-14 of the oracle's CSR product, and a clear
error for matrices beyond the staging size."""
import numpy as np
import pytest

from timeevolver_amd import capi
from timeevolver_amd.hamiltonian import as_tuples
from timeevolver_amd.workloads import random_state, ring_bose_hubbard_terms, ring_couplings

pytestmark = pytest.mark.gpu


def _ring(gpu, oracle_lib, L, from_coo):
    eps, J = ring_couplings(L, 11)
    terms = as_tuples(ring_bose_hubbard_terms(L, eps, J, 0.5))
    ob = oracle_lib.Basis.basis(L, L)
    v, c, r = oracle_lib.build_operator(ob, terms)
    if from_coo:
        M = capi.Matrix.from_coo(gpu, v, c, r, ob.D)
    else:
        M = capi.operator_build(gpu, capi.Basis.basis(gpu, L, L), terms)
    return M, (v, c, r), ob.D


def test_spmv_noalloc_matches_spmv(gpu, oracle_lib):
    for L, from_coo, seed in ((10, False, 1), (8, True, 2), (9, True, 3), (10, True, 4)):
        M, (v, c, r), D = _ring(gpu, oracle_lib, L, from_coo)
        x = random_state(D, seed)
        for alpha in (1.0, 0.2 - 0.7j):
            y0 = M.spmv(x, alpha)
            y1 = M.spmv_noalloc(x, alpha)
            assert np.array_equal(y0.view(np.uint64), y1.view(np.uint64)), (L, from_coo, alpha)
            yr = oracle_lib.spmv(D, v, c, r, x, alpha)
            assert np.linalg.norm(y1 - yr) <= 1e-14 * np.linalg.norm(yr)
        M.close()


def test_spmv_noalloc_rejects_larger_matrices(gpu):
    n = (1 << 21) + 1
    idx = np.arange(n, dtype=np.uint64)
    M = capi.Matrix.from_coo(gpu, np.full(n, 0.5, dtype=np.complex128), idx, idx, n)
    with pytest.raises(capi.TEError) as e:
        M.spmv_noalloc(np.ones(n, dtype=np.complex128))
    assert e.value.code == capi.TE_ERR_UNSUPPORTED and "2097152" in str(e.value)
    y = M.spmv(np.ones(n, dtype=np.complex128))      # te_spmv still takes it
    assert np.all(y == 0.5)
    M.close()
