"""ORACLE — test infrastructure only (ctypes front-end of oracle/liboracle.so).

A CPU restatement of TimeEvolver's Krylov time-evolution path (core/krylovTimeEvolver.cpp,
core/matrixDataTypes.cpp, core/krylovObservables.cpp, helper/Basis.cpp,
helper/hamiltonian.cpp). Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg use it, and only as the checker or the timed CPU baseline; the product package
timeevolver_amd never imports it.

Pinning (tests/test_oracle_*.py): golden observables shipped with the reference
(output/*.h5 and output/*.csv, converted to tests/golden/), closed forms (two-oscillator
Rabi formula, free bosons), numpy/scipy dense linear algebra, mpmath quadrature, the basis
hand trace and closed-form basis/CSR counts.
Parity unpinned: the exact Krylov step schedule, n_steps and err to all digits (they
depend on Boost's tanh_sinh internals, which are not installed and are not reproduced).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

_d = ctypes.c_double
_i = ctypes.c_int
_l = ctypes.c_long
_p = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_long)


def build(force=False):
    """Compile liboracle.so from oracle.cpp with the committed Makefile."""
    if force or not os.path.exists(_LIB) or os.path.getmtime(_LIB) < os.path.getmtime(
            os.path.join(_HERE, "oracle.cpp")):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE, "liboracle.so"], check=True,
                       stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        sig = {
            "or_set_threads": (None, [_i]),
            "or_get_threads": (_i, []),
            "or_basis_new": (_p, [_i, _i, _i, _i, _i]),
            "or_tensor_basis_new": (_p, [_i, _i, _i, _i, _i]),
            "or_basis_size": (_l, [_p]),
            "or_basis_modes": (_i, [_p]),
            "or_basis_states": (None, [_p, _ip]),
            "or_basis_index": (None, [_p]),
            "or_basis_find": (_l, [_p, _ip]),
            "or_basis_free": (None, [_p]),
            "or_build_operator": (_p, [_p, _i, _dp, _ip, _ip, _ip, _ip]),
            "or_coo_nnz": (_l, [_p]),
            "or_coo_copy": (None, [_p, _dp, _lp, _lp]),
            "or_coo_free": (None, [_p]),
            "or_build_rows": (None, [_p, _i, _dp, _ip, _ip, _ip, _ip, _l, _lp, _i, _ip, _lp, _dp]),
            "or_norm1": (_d, [_l, _l, _dp, _lp]),
            "or_spmv": (None, [_l, _l, _dp, _lp, _lp, _d, _d, _dp, _dp]),
            "or_tridiag_eig": (_i, [_i, _dp, _dp, _dp, _dp]),
            "or_exp_krylov": (_i, [_i, _dp, _dp, _d, _dp]),
            "or_integrate_error": (_i, [_i, _dp, _dp, _d, _d, _d, _i, _d, _d, _dp]),
            "or_tanh_sinh_poly": (_d, [_i, _dp, _d, _d, _d, _dp, _dp]),
            "or_gauss_poly": (_d, [_i, _i, _dp, _d, _d]),
            "or_tanh_sinh_rows": (_i, [_i]),
            "or_evolve_csr": (_i, [_l, _lp, _ip, _dp, _dp, _d, _d, _d, _d, _i, _i, _l, _i, _dp, _dp, _dp, _dp]),
            "or_evolve": (_i, [_l, _l, _dp, _lp, _lp, _dp, _d, _d, _d, _d, _i, _i, _l, _l, _l, _i, _ip, _lp,
                               _dp, _lp, _lp, _dp, _dp, _dp, _dp, _dp, _l]),
            "or_mkl_enable": (_i, [ctypes.c_char_p, _i]),
            "or_mkl_version": (None, [ctypes.c_char_p, _i]),
            "or_mkl_disable": (None, []),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


def mkl_enable(path=None, threads=0):
    """Routes the evolvers' products and level-1/2 kernels through MKL (libmkl_rt opened with
    dlopen, ILP64, GNU OpenMP threads), as the reference's MKL build runs them. Returns 0 on
    success, else 1 (no library), 2 (missing entry point), 3 (ILP64 refused); the loop kernels
    stay in use then. On until mkl_disable()."""
    return lib().or_mkl_enable(path.encode() if path else None, int(threads))


def mkl_disable():
    lib().or_mkl_disable()


def mkl_version():
    buf = ctypes.create_string_buffer(256)
    lib().or_mkl_version(buf, 256)
    return buf.value.decode(errors="replace")


def set_threads(n):
    lib().or_set_threads(int(n))


def get_threads():
    return lib().or_get_threads()


class Basis:
    """helper/Basis.cpp:37-120 (basis) and :131-182 (tensorBasis)."""

    def __init__(self, handle):
        self.h = handle
        self.D = lib().or_basis_size(handle)
        self.K = lib().or_basis_modes(handle)

    @classmethod
    def basis(cls, N, K, nbQudits=0, capacity=0, quant=1):
        return cls(lib().or_basis_new(N, K, nbQudits, capacity, quant))

    @classmethod
    def tensor(cls, N1, K1, N2, K2, C2):
        return cls(lib().or_tensor_basis_new(N1, K1, N2, K2, C2))

    def states(self):
        out = np.zeros((self.D, self.K), dtype=np.int32)
        if self.D:
            lib().or_basis_states(self.h, _ptr(out, ctypes.c_int))
        return out

    def find(self, occ):
        e = np.ascontiguousarray(occ, dtype=np.int32)
        return lib().or_basis_find(self.h, _ptr(e, ctypes.c_int))

    def __del__(self):
        try:
            lib().or_basis_free(self.h)
        except Exception:
            pass


def _pack_terms(terms):
    """terms: list of (coef complex, ops list, modes list, caps list)."""
    coef = np.zeros(2 * len(terms))
    n_ops, ops, modes, caps = [], [], [], []
    for t, (c, o, md, cp) in enumerate(terms):
        coef[2 * t] = complex(c).real
        coef[2 * t + 1] = complex(c).imag
        n_ops.append(len(o))
        ops += list(o)
        modes += list(md)
        caps += list(cp)
    a = lambda x: np.ascontiguousarray(x, dtype=np.int32) if len(x) else np.zeros(1, np.int32)
    return coef, a(n_ops), a(ops), a(modes), a(caps)


def build_operator(basis, terms):
    """helper/hamiltonian.cpp:318-401 -> (values complex128, cols int64, rows int64)."""
    coef, n_ops, ops, modes, caps = _pack_terms(terms)
    L = lib()
    h = L.or_build_operator(basis.h, len(terms), _ptr(coef, ctypes.c_double), _ptr(n_ops, ctypes.c_int),
                            _ptr(ops, ctypes.c_int), _ptr(modes, ctypes.c_int), _ptr(caps, ctypes.c_int))
    nnz = L.or_coo_nnz(h)
    vals = np.zeros(2 * max(nnz, 1))
    cols = np.zeros(max(nnz, 1), dtype=np.int64)
    rows = np.zeros(max(nnz, 1), dtype=np.int64)
    L.or_coo_copy(h, _ptr(vals, ctypes.c_double), _ptr(cols, ctypes.c_long), _ptr(rows, ctypes.c_long))
    L.or_coo_free(h)
    return vals[:2 * nnz].view(np.complex128), cols[:nnz], rows[:nnz]


def build_rows(basis, terms, row_ids, cap=128):
    """Selected rows of createMatrix: list of (cols, values) per requested row."""
    coef, n_ops, ops, modes, caps = _pack_terms(terms)
    row_ids = np.ascontiguousarray(row_ids, dtype=np.int64)
    n = len(row_ids)
    counts = np.zeros(n, dtype=np.int32)
    cols = np.zeros(n * cap, dtype=np.int64)
    vals = np.zeros(2 * n * cap)
    lib().or_build_rows(basis.h, len(terms), _ptr(coef, ctypes.c_double), _ptr(n_ops, ctypes.c_int),
                        _ptr(ops, ctypes.c_int), _ptr(modes, ctypes.c_int), _ptr(caps, ctypes.c_int), n,
                        _ptr(row_ids, ctypes.c_long), cap, _ptr(counts, ctypes.c_int),
                        _ptr(cols, ctypes.c_long), _ptr(vals, ctypes.c_double))
    v = vals.view(np.complex128)
    out = []
    for i in range(n):
        c = min(int(counts[i]), cap)
        out.append((cols[i * cap:i * cap + c].copy(), v[i * cap:i * cap + c].copy()))
    return out


def norm1(n, vals, cols):
    v = np.ascontiguousarray(vals, dtype=np.complex128).view(np.float64)
    c = np.ascontiguousarray(cols, dtype=np.int64)
    return lib().or_norm1(n, len(c), _ptr(v, ctypes.c_double), _ptr(c, ctypes.c_long))


def spmv(n, vals, cols, rows, x, alpha=1.0):
    v = np.ascontiguousarray(vals, dtype=np.complex128).view(np.float64)
    c = np.ascontiguousarray(cols, dtype=np.int64)
    r = np.ascontiguousarray(rows, dtype=np.int64)
    xx = np.ascontiguousarray(x, dtype=np.complex128)
    y = np.zeros(n, dtype=np.complex128)
    alpha = complex(alpha)
    lib().or_spmv(n, len(c), _ptr(v, ctypes.c_double), _ptr(c, ctypes.c_long), _ptr(r, ctypes.c_long),
                  alpha.real, alpha.imag, _ptr(xx.view(np.float64), ctypes.c_double),
                  _ptr(y.view(np.float64), ctypes.c_double))
    return y


def tridiag_eig(a, b):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    m = len(a)
    aim = np.ascontiguousarray(a.imag)
    bb = np.zeros(max(m, 1))
    bb[:m - 1] = b[:m - 1]
    lam = np.zeros(m)
    Q = np.zeros((m, m))
    info = lib().or_tridiag_eig(m, _ptr(aim, ctypes.c_double), _ptr(bb, ctypes.c_double),
                                _ptr(lam, ctypes.c_double), _ptr(Q, ctypes.c_double))
    return info, lam, Q


def exp_krylov(a, b, x):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    m = len(a)
    bb = np.zeros(max(m, 1))
    bb[:m - 1] = b[:m - 1]
    out = np.zeros(m, dtype=np.complex128)
    info = lib().or_exp_krylov(m, _ptr(a.view(np.float64), ctypes.c_double), _ptr(bb, ctypes.c_double), x,
                               _ptr(out.view(np.float64), ctypes.c_double))
    return info, out


def integrate_error(a, b, h, lo, hi, method=2, tolRate=1e-9, tol=1e-8):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    m = len(a)
    bb = np.zeros(max(m, 1))
    bb[:m - 1] = b[:m - 1]
    out = np.zeros(3)
    ok = lib().or_integrate_error(m, _ptr(a.view(np.float64), ctypes.c_double), _ptr(bb, ctypes.c_double), h,
                                  lo, hi, method, tolRate, tol, _ptr(out, ctypes.c_double))
    return out, bool(ok)


def tanh_sinh_poly(coeffs, a, b, tol=1e-3):
    c = np.ascontiguousarray(coeffs, dtype=np.float64)
    err = ctypes.c_double()
    L1 = ctypes.c_double()
    v = lib().or_tanh_sinh_poly(len(c) - 1, _ptr(c, ctypes.c_double), a, b, tol, ctypes.byref(err),
                                ctypes.byref(L1))
    return v, err.value, L1.value


def gauss_poly(n, coeffs, a, b):
    c = np.ascontiguousarray(coeffs, dtype=np.float64)
    return lib().or_gauss_poly(n, len(c) - 1, _ptr(c, ctypes.c_double), a, b)


def evolve(n, H, psi0, t, samplingStep, observables=(), expFactor=1.0, tol=1e-6, m=40, fastIntegration=False,
           max_steps=0, stop_at_sample=-1, stop_obs=0, trace_cap=100000):
    """krylovTimeEvolver(t, psi0, samplingStep, observables, H, expFactor, tol, m, fastIntegration)
    followed by timeEvolve(). stop_at_sample >= 0: a host observable placed before observable
    `stop_obs` of the list throws requestStopException at that sample (the observables before it
    record the sample, the run returns with status 3).

    H: (values complex128, cols, rows) COO. observables: list of ("sparse", (vals, cols, rows)) or
    ("vector", ndarray). Returns dict with state, observables (n_obs x n_samples), err, n_steps,
    statusCode, ... or raises RuntimeError where the reference would exit(1)."""
    hv, hc, hr = H
    hv = np.ascontiguousarray(hv, dtype=np.complex128).view(np.float64)
    hc = np.ascontiguousarray(hc, dtype=np.int64)
    hr = np.ascontiguousarray(hr, dtype=np.int64)
    psi = np.ascontiguousarray(psi0, dtype=np.complex128)
    kinds, nnzs, ov, oc, orr, vecs = [], [], [], [], [], []
    for kind, data in observables:
        if kind == "sparse":
            v, c, r = data
            kinds.append(0)
            nnzs.append(len(c))
            ov.append(np.asarray(v, dtype=np.complex128))
            oc.append(np.asarray(c, dtype=np.int64))
            orr.append(np.asarray(r, dtype=np.int64))
        else:
            kinds.append(1)
            nnzs.append(0)
            vecs.append(np.asarray(data, dtype=np.complex128))
    cat = lambda xs, dt: np.ascontiguousarray(np.concatenate(xs) if xs else np.zeros(1, dt), dtype=dt)
    ov_ = cat(ov, np.complex128).view(np.float64)
    oc_ = cat(oc, np.int64)
    or_ = cat(orr, np.int64)
    vecs_ = cat(vecs, np.complex128).view(np.float64)
    kinds_ = np.ascontiguousarray(kinds if kinds else [0], dtype=np.int32)
    nnzs_ = np.ascontiguousarray(nnzs if nnzs else [0], dtype=np.int64)
    n_samples = int(np.floor(t / samplingStep)) + 1
    obs_out = np.zeros(max(len(kinds), 1) * n_samples)
    state = np.zeros(n, dtype=np.complex128)
    result = np.zeros(8)
    trace = np.zeros(trace_cap)
    rc = lib().or_evolve(n, len(hc), _ptr(hv, ctypes.c_double), _ptr(hc, ctypes.c_long), _ptr(hr, ctypes.c_long),
                         _ptr(psi.view(np.float64), ctypes.c_double), t, samplingStep, expFactor, tol, m,
                         int(fastIntegration), max_steps, stop_at_sample, stop_obs, len(kinds),
                         _ptr(kinds_, ctypes.c_int),
                         _ptr(nnzs_, ctypes.c_long), _ptr(ov_, ctypes.c_double), _ptr(oc_, ctypes.c_long),
                         _ptr(or_, ctypes.c_long), _ptr(vecs_, ctypes.c_double), _ptr(obs_out, ctypes.c_double),
                         _ptr(state.view(np.float64), ctypes.c_double), _ptr(result, ctypes.c_double),
                         _ptr(trace, ctypes.c_double), trace_cap)
    if rc != 0:
        raise RuntimeError("oracle evolution hit a fatal condition (the reference exits(1) here)")
    n_steps = int(result[3])
    return {
        "state": state,
        "observables": obs_out[:len(kinds) * n_samples].reshape(len(kinds), n_samples),
        "err": result[0],
        "evolvedTime": result[1],
        "numSamples": int(result[2]),
        "n_steps": n_steps,
        "dim": int(result[4]),
        "krylovDim": int(result[5]),
        "statusCode": int(result[6]),
        "n_samples": int(result[7]),
        "trace": trace[:min(n_steps, trace_cap)].copy(),
    }


def evolve_csr(ptr, col, vals, psi0, t, samplingStep, diags=(), expFactor=1.0, tol=1e-6, m=40,
               fastIntegration=False, max_steps=0, want_state=False):
    """timeEvolve on H given as CSR with diagonal observables (the CPU-baseline entry)."""
    n = len(ptr) - 1
    ptr = np.ascontiguousarray(ptr, dtype=np.int64)
    col = np.ascontiguousarray(col, dtype=np.int32)
    vals = np.ascontiguousarray(vals, dtype=np.complex128).view(np.float64)
    psi = np.ascontiguousarray(psi0, dtype=np.complex128)
    d = np.ascontiguousarray(np.array(diags, dtype=np.float64).reshape(len(diags), n) if len(diags)
                             else np.zeros((0, n)))
    n_samples = int(np.floor(t / samplingStep)) + 1
    obs_out = np.zeros(max(len(diags), 1) * n_samples)
    state = np.zeros(n, dtype=np.complex128) if want_state else None
    result = np.zeros(10)
    rc = lib().or_evolve_csr(n, _ptr(ptr, ctypes.c_long), _ptr(col, ctypes.c_int), _ptr(vals, ctypes.c_double),
                             _ptr(psi.view(np.float64), ctypes.c_double), t, samplingStep, expFactor, tol, m,
                             int(fastIntegration), max_steps, len(diags),
                             _ptr(d, ctypes.c_double) if len(diags) else None, _ptr(obs_out, ctypes.c_double),
                             _ptr(state.view(np.float64), ctypes.c_double) if want_state else None,
                             _ptr(result, ctypes.c_double))
    if rc != 0:
        raise RuntimeError("oracle evolution hit a fatal condition")
    return {"err": result[0], "evolvedTime": result[1], "numSamples": int(result[2]), "n_steps": int(result[3]),
            "statusCode": int(result[6]), "state": state, "run_s": result[8], "setup_s": result[9],
            "observables": obs_out[:len(diags) * n_samples].reshape(len(diags), n_samples)}
