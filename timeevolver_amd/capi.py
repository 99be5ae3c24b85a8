"""ctypes binding of the C-ABI (include/timeevolver.h) of libtimeevolver_hip.so.

This is the Python-side view of the drop-in boundary: the same entry points a cgo/JNI
binding would use. Loading fails loudly when the HIP library is not built; there is no
CPU fallback.
"""
import ctypes
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtimeevolver_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "timeevolver.h")

TE_OK, TE_ERR_ARG, TE_ERR_HIP, TE_ERR_ALLOC, TE_ERR_FATAL, TE_ERR_UNSUPPORTED, TE_ERR_STOP = range(7)
TE_OBS_SPARSE, TE_OBS_VECTOR, TE_OBS_HOST = 0, 1, 2

_c = ctypes
_p = _c.c_void_p
_pp = _c.POINTER(_c.c_void_p)
_dp = _c.POINTER(_c.c_double)
_szp = _c.POINTER(_c.c_size_t)
_ip = _c.POINTER(_c.c_int)
_i64p = _c.POINTER(_c.c_int64)

OBS_CALLBACK = _c.CFUNCTYPE(_c.c_int, _c.c_void_p, _dp, _c.c_size_t, _dp)


class te_op_term(_c.Structure):
    _fields_ = [("coef_re", _c.c_double), ("coef_im", _c.c_double), ("n_ops", _c.c_int),
                ("operations", _ip), ("modes", _ip), ("capacities", _ip)]


class te_evolve_params(_c.Structure):
    _fields_ = [("t", _c.c_double), ("sampling_step", _c.c_double), ("exp_factor", _c.c_double),
                ("tol", _c.c_double), ("m", _c.c_int), ("fast_integration", _c.c_int)]


class te_observable(_c.Structure):
    _fields_ = [("kind", _c.c_int), ("matrix", _p), ("vector", _dp), ("callback", OBS_CALLBACK), ("user", _p)]


class te_result(_c.Structure):
    _fields_ = [("err", _c.c_double), ("evolved_time", _c.c_double), ("num_samples", _c.c_size_t),
                ("n_steps", _c.c_size_t), ("dim", _c.c_size_t), ("krylov_dim", _c.c_size_t),
                ("status_code", _c.c_int)]


class te_adjoint_info(_c.Structure):
    _fields_ = [("forward_ms", _c.c_double), ("backward_ms", _c.c_double), ("n_steps", _c.c_int),
                ("nodes", _c.c_longlong), ("max_phase", _c.c_double), ("capped", _c.c_int),
                ("backward_err", _c.c_double), ("psi_err", _c.c_double), ("node_states", _c.c_longlong)]


_SIGS = {
    "te_version": (_c.c_int, []),
    "te_set_log_level": (None, [_c.c_int]),
    "te_ctx_create": (_c.c_int, [_c.c_int, _pp]),
    "te_ctx_destroy": (None, [_p]),
    "te_last_error": (_c.c_char_p, [_p]),
    "te_ctx_synchronize": (_c.c_int, [_p]),
    "te_matrix_from_coo": (_c.c_int, [_p, _dp, _szp, _szp, _c.c_size_t, _c.c_size_t, _c.c_size_t, _pp]),
    "te_matrix_destroy": (None, [_p]),
    "te_matrix_info": (_c.c_int, [_p, _szp, _szp, _szp]),
    "te_matrix_storage": (_c.c_int, [_p, _ip, _szp, _szp]),
    "te_matrix_norm1": (_c.c_int, [_p, _dp]),
    "te_matrix_to_coo": (_c.c_int, [_p, _dp, _szp, _szp]),
    "te_matrix_to_csr": (_c.c_int, [_p, _dp, _i64p, _ip]),
    "te_spmv": (_c.c_int, [_p, _c.c_double, _c.c_double, _dp, _dp]),
    "te_spmv_device": (_c.c_int, [_p, _c.c_double, _c.c_double, _p, _p]),
    "te_spmv_noalloc": (_c.c_int, [_p, _c.c_double, _c.c_double, _dp, _dp]),
    "te_basis_create": (_c.c_int, [_p, _c.c_int, _c.c_int, _c.c_int, _c.c_int, _c.c_int, _pp]),
    "te_tensor_basis_create": (_c.c_int, [_p, _c.c_int, _c.c_int, _c.c_int, _c.c_int, _c.c_int, _pp]),
    "te_basis_destroy": (None, [_p]),
    "te_basis_info": (_c.c_int, [_p, _szp, _ip]),
    "te_basis_states": (_c.c_int, [_p, _ip]),
    "te_basis_find": (_c.c_int, [_p, _ip, _c.c_size_t, _i64p]),
    "te_operator_build": (_c.c_int, [_p, _p, _c.POINTER(te_op_term), _c.c_int, _pp]),
    "te_evolver_create": (_c.c_int, [_p, _p, _dp, _c.POINTER(te_evolve_params), _c.POINTER(te_observable),
                                     _c.c_int, _pp]),
    "te_evolver_create_device": (_c.c_int, [_p, _p, _p, _c.POINTER(te_evolve_params), _c.POINTER(te_observable),
                                            _c.c_int, _pp]),
    "te_evolver_destroy": (None, [_p]),
    "te_evolver_num_samples": (_c.c_int, [_p, _szp]),
    "te_evolver_run": (_c.c_int, [_p, _dp, _dp, _c.POINTER(te_result)]),
    "te_evolver_state_device": (_c.c_int, [_p, _pp]),
    "te_evolver_obs_recorded": (_c.c_int, [_p, _szp]),
    "te_evolver_diag_bytes": (_c.c_int, [_p, _szp]),
    "te_evolver_step_trace": (_c.c_int, [_p, _dp, _c.c_size_t, _szp]),
    "te_evolver_set_profiling": (_c.c_int, [_p, _c.c_int]),
    "te_evolver_kernel_time": (_c.c_int, [_p, _c.c_int, _dp, _szp]),
    "te_batch_create": (_c.c_int, [_p, _c.c_int, _pp, _c.POINTER(_dp), _c.POINTER(te_evolve_params),
                                   _c.POINTER(te_observable), _c.c_int, _pp]),
    "te_batch_destroy": (None, [_p]),
    "te_batch_run": (_c.c_int, [_p, _dp, _c.POINTER(_dp), _c.POINTER(te_result)]),
    "te_batch_info": (_c.c_int, [_p, _szp, _szp, _c.POINTER(_c.c_longlong)]),
    "te_batch_set_graph": (_c.c_int, [_p, _c.c_int]),
    "te_batch_set_profiling": (_c.c_int, [_p, _c.c_int]),
    "te_batch_set_shape": (_c.c_int, [_p, _c.c_int, _c.c_int, _c.c_int, _c.c_int]),
    "te_batch_shape": (_c.c_int, [_p, _ip, _ip, _ip, _ip, _ip, _dp]),
    "te_batch_kernel_time": (_c.c_int, [_p, _c.c_int, _dp, _szp]),
    "te_batch_kernel_times": (_c.c_int, [_p, _c.c_int, _dp, _c.c_size_t, _szp]),
    "te_adjoint_create": (_c.c_int, [_p, _p, _c.c_int, _pp, _p, _c.POINTER(te_evolve_params), _c.c_int, _pp]),
    "te_adjoint_destroy": (None, [_p]),
    "te_adjoint_run": (_c.c_int, [_p, _dp, _dp, _dp, _c.POINTER(te_result), _c.POINTER(te_adjoint_info)]),
    "te_adjoint_set_hamiltonian": (_c.c_int, [_p, _p]),
    "te_adjoint_state": (_c.c_int, [_p, _dp]),
}

_lib = None


def header_symbols():
    """Function names declared in include/timeevolver.h."""
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(te_\w+)\s*\(", txt, re.M)))


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libtimeevolver_hip.so is not built (run __graft_entry__.build()); "
                           "there is no CPU fallback")
    L = _c.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        f = getattr(L, name, None)
        if f is None:
            continue
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


class TEError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (code %d)" % (msg, code))
        self.code = code


def _ptr(a, t):
    return a.ctypes.data_as(_c.POINTER(t))


class Context:
    def __init__(self, device=0):
        L = load()
        h = _p()
        rc = L.te_ctx_create(device, _c.byref(h))
        if rc != TE_OK:
            raise TEError(rc, "te_ctx_create failed")
        self.h = h
        self.L = L

    def check(self, rc, what=""):
        if rc != TE_OK:
            raise TEError(rc, "%s: %s" % (what, self.L.te_last_error(self.h).decode()))

    def synchronize(self):
        self.check(self.L.te_ctx_synchronize(self.h), "synchronize")

    def close(self):
        if self.h:
            self.L.te_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Matrix:
    def __init__(self, ctx, handle):
        self.ctx = ctx
        self.h = handle
        n, nc, nnz = _c.c_size_t(), _c.c_size_t(), _c.c_size_t()
        ctx.L.te_matrix_info(handle, _c.byref(n), _c.byref(nc), _c.byref(nnz))
        self.n_rows, self.n_cols, self.nnz = n.value, nc.value, nnz.value

    @classmethod
    def from_coo(cls, ctx, vals, cols, rows, n_rows, n_cols=None):
        v = np.ascontiguousarray(vals, dtype=np.complex128)
        c = np.ascontiguousarray(cols, dtype=np.uint64)
        r = np.ascontiguousarray(rows, dtype=np.uint64)
        h = _p()
        n_cols = n_rows if n_cols is None else n_cols
        ctx.check(ctx.L.te_matrix_from_coo(ctx.h, _ptr(v.view(np.float64), _c.c_double), _ptr(c, _c.c_size_t),
                                           _ptr(r, _c.c_size_t), len(c), n_rows, n_cols, _c.byref(h)),
                  "te_matrix_from_coo")
        return cls(ctx, h)

    def storage(self):
        """(format, n_values, stream_bytes): 0 CSR lanes-per-row, 1 CSR row blocks,
        2 value-dictionary row blocks, 3 the same in the locality order, 4 value-dictionary
        sliced ELL; bytes of matrix data one product reads."""
        f, nv, b = _c.c_int(), _c.c_size_t(), _c.c_size_t()
        self.ctx.check(self.ctx.L.te_matrix_storage(self.h, _c.byref(f), _c.byref(nv), _c.byref(b)), "storage")
        return f.value, nv.value, b.value

    def norm1(self):
        out = _c.c_double()
        self.ctx.check(self.ctx.L.te_matrix_norm1(self.h, _c.byref(out)), "norm1")
        return out.value

    def to_coo(self):
        nnz = self.nnz
        v = np.zeros(max(nnz, 1), dtype=np.complex128)
        c = np.zeros(max(nnz, 1), dtype=np.uint64)
        r = np.zeros(max(nnz, 1), dtype=np.uint64)
        self.ctx.check(self.ctx.L.te_matrix_to_coo(self.h, _ptr(v.view(np.float64), _c.c_double),
                                                   _ptr(c, _c.c_size_t), _ptr(r, _c.c_size_t)), "to_coo")
        return v[:nnz], c[:nnz].astype(np.int64), r[:nnz].astype(np.int64)

    def to_csr(self):
        ptr = np.zeros(self.n_rows + 1, dtype=np.int64)
        col = np.zeros(max(self.nnz, 1), dtype=np.int32)
        val = np.zeros(max(self.nnz, 1), dtype=np.complex128)
        self.ctx.check(self.ctx.L.te_matrix_to_csr(self.h, _ptr(val.view(np.float64), _c.c_double),
                                                   _ptr(ptr, _c.c_int64), _ptr(col, _c.c_int)), "to_csr")
        return ptr, col[:self.nnz], val[:self.nnz]

    def spmv(self, x, alpha=1.0):
        xx = np.ascontiguousarray(x, dtype=np.complex128)
        y = np.zeros(self.n_rows, dtype=np.complex128)
        alpha = complex(alpha)
        self.ctx.check(self.ctx.L.te_spmv(self.h, alpha.real, alpha.imag, _ptr(xx.view(np.float64), _c.c_double),
                                          _ptr(y.view(np.float64), _c.c_double)), "te_spmv")
        return y

    def spmv_noalloc(self, x, alpha=1.0):
        """te_spmv_noalloc: the product staged through the library's fixed device array."""
        xx = np.ascontiguousarray(x, dtype=np.complex128)
        y = np.zeros(self.n_rows, dtype=np.complex128)
        alpha = complex(alpha)
        self.ctx.check(self.ctx.L.te_spmv_noalloc(self.h, alpha.real, alpha.imag,
                                                  _ptr(xx.view(np.float64), _c.c_double),
                                                  _ptr(y.view(np.float64), _c.c_double)), "te_spmv_noalloc")
        return y

    def close(self):
        if self.h:
            self.ctx.L.te_matrix_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Basis:
    def __init__(self, ctx, handle):
        self.ctx = ctx
        self.h = handle
        n, k = _c.c_size_t(), _c.c_int()
        ctx.check(ctx.L.te_basis_info(handle, _c.byref(n), _c.byref(k)), "basis_info")
        self.D, self.K = n.value, k.value

    @classmethod
    def basis(cls, ctx, N, K, nbQudits=0, capacity=0, quant=1):
        h = _p()
        ctx.check(ctx.L.te_basis_create(ctx.h, N, K, nbQudits, capacity, quant, _c.byref(h)), "te_basis_create")
        return cls(ctx, h)

    @classmethod
    def tensor(cls, ctx, N1, K1, N2, K2, C2):
        h = _p()
        ctx.check(ctx.L.te_tensor_basis_create(ctx.h, N1, K1, N2, K2, C2, _c.byref(h)), "te_tensor_basis_create")
        return cls(ctx, h)

    def states(self):
        out = np.zeros((self.D, self.K), dtype=np.int32)
        if self.D:
            self.ctx.check(self.ctx.L.te_basis_states(self.h, _ptr(out, _c.c_int)), "basis_states")
        return out

    def find(self, occ):
        a = np.ascontiguousarray(np.atleast_2d(occ), dtype=np.int32)
        out = np.zeros(a.shape[0], dtype=np.int64)
        self.ctx.check(self.ctx.L.te_basis_find(self.h, _ptr(a, _c.c_int), a.shape[0], _ptr(out, _c.c_int64)),
                       "basis_find")
        return out

    def close(self):
        if self.h:
            self.ctx.L.te_basis_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def operator_build(ctx, basis, terms):
    """terms: (coef, ops, modes, caps) tuples or OpTerm objects (helper/hamiltonian.h:18-35)."""
    from .hamiltonian import as_tuples
    terms = as_tuples(terms)
    arr = (te_op_term * max(len(terms), 1))()
    keep = []
    for i, (coef, ops, modes, caps) in enumerate(terms):
        o = np.ascontiguousarray(ops if len(ops) else [0], dtype=np.int32)
        md = np.ascontiguousarray(modes if len(modes) else [0], dtype=np.int32)
        cp = np.ascontiguousarray(caps if len(caps) else [0], dtype=np.int32)
        keep += [o, md, cp]
        coef = complex(coef)
        arr[i] = te_op_term(coef.real, coef.imag, len(ops), _ptr(o, _c.c_int), _ptr(md, _c.c_int),
                            _ptr(cp, _c.c_int))
    h = _p()
    ctx.check(ctx.L.te_operator_build(ctx.h, basis.h, arr, len(terms), _c.byref(h)), "te_operator_build")
    return Matrix(ctx, h)


class Evolver:
    """krylovTimeEvolver on the device. observables: list of Matrix (sparse), ("vector", ndarray)
    or ("host", callable(state ndarray) -> float; raise StopIteration to request a stop)."""

    def __init__(self, ctx, H, psi0, t, sampling_step, observables=(), exp_factor=1.0, tol=1e-6, m=40,
                 fast_integration=False, psi0_device_ptr=None):
        self.ctx = ctx
        self.H = H
        self._keep = []
        obs = (te_observable * max(len(observables), 1))()
        self._cbs = []
        for i, o in enumerate(observables):
            if isinstance(o, Matrix):
                obs[i] = te_observable(TE_OBS_SPARSE, o.h, None, OBS_CALLBACK(0), None)
                self._keep.append(o)
            elif o[0] == "vector":
                v = np.ascontiguousarray(o[1], dtype=np.complex128)
                self._keep.append(v)
                obs[i] = te_observable(TE_OBS_VECTOR, None, _ptr(v.view(np.float64), _c.c_double), OBS_CALLBACK(0),
                                       None)
            elif o[0] == "host":
                fn = o[1]

                def cb(user, state, n, value, fn=fn):
                    arr = np.ctypeslib.as_array(state, shape=(2 * n,)).view(np.complex128)
                    try:
                        value[0] = float(fn(arr))
                        return 0
                    except StopIteration:
                        return 1
                c = OBS_CALLBACK(cb)
                self._cbs.append(c)
                obs[i] = te_observable(TE_OBS_HOST, None, None, c, None)
            else:
                raise ValueError("unknown observable %r" % (o,))
        params = te_evolve_params(t, sampling_step, exp_factor, tol, m, int(fast_integration))
        h = _p()
        if psi0_device_ptr is not None:
            ctx.check(ctx.L.te_evolver_create_device(ctx.h, H.h, _c.c_void_p(psi0_device_ptr), _c.byref(params), obs,
                                                     len(observables), _c.byref(h)), "te_evolver_create_device")
        else:
            psi = np.ascontiguousarray(psi0, dtype=np.complex128)
            ctx.check(ctx.L.te_evolver_create(ctx.h, H.h, _ptr(psi.view(np.float64), _c.c_double), _c.byref(params),
                                              obs, len(observables), _c.byref(h)), "te_evolver_create")
        self.h = h
        self.n_obs = len(observables)
        ns = _c.c_size_t()
        ctx.L.te_evolver_num_samples(h, _c.byref(ns))
        self.n_samples = ns.value
        self.dim = H.n_rows

    def run(self, want_state=True, want_obs=True):
        R = te_result()
        obs = np.zeros(max(self.n_obs, 1) * self.n_samples)
        state = np.zeros(self.dim, dtype=np.complex128) if want_state else None
        rc = self.ctx.L.te_evolver_run(self.h, _ptr(obs, _c.c_double) if want_obs else None,
                                       _ptr(state.view(np.float64), _c.c_double) if want_state else None,
                                       _c.byref(R))
        self.ctx.check(rc, "te_evolver_run")
        nt = _c.c_size_t()
        self.ctx.L.te_evolver_step_trace(self.h, None, 0, _c.byref(nt))
        trace = np.zeros(max(nt.value, 1))
        self.ctx.L.te_evolver_step_trace(self.h, _ptr(trace, _c.c_double), nt.value, _c.byref(nt))
        rec = (_c.c_size_t * max(self.n_obs, 1))()
        self.ctx.L.te_evolver_obs_recorded(self.h, rec)
        return {
            "state": state,
            "observables": obs[:self.n_obs * self.n_samples].reshape(self.n_obs, self.n_samples),
            "recorded": [rec[i] for i in range(self.n_obs)],
            "err": R.err, "evolvedTime": R.evolved_time, "numSamples": R.num_samples, "n_steps": R.n_steps,
            "dim": R.dim, "krylovDim": R.krylov_dim, "statusCode": R.status_code, "n_samples": self.n_samples,
            "trace": trace[:nt.value],
        }

    KINDS = ("apply", "project", "control", "combine")

    def set_profiling(self, kinds=KINDS):
        """Time the launches of the given kinds (names from KINDS) with HIP events in the next
        runs; False/None/() turns timing off, True times every kind."""
        if kinds is True:
            kinds = self.KINDS
        mask = 0
        for k in (kinds or ()):
            mask |= 1 << self.KINDS.index(k)
        self.ctx.check(self.ctx.L.te_evolver_set_profiling(self.h, mask), "set_profiling")

    def kernel_times(self):
        """{kind: (total_ms, launches)} of the last run (0 apply, 1 project, 2 control, 3 combine)."""
        out = {}
        for k, name in enumerate(("apply", "project", "control", "combine", "ctl_eigen", "ctl_search",
                                   "ctl_samples")):
            ms, n = _c.c_double(), _c.c_size_t()
            self.ctx.check(self.ctx.L.te_evolver_kernel_time(self.h, k, _c.byref(ms), _c.byref(n)), "kernel_time")
            out[name] = (ms.value, n.value)
        return out

    def diag_bytes_per_row(self):
        """Bytes of diagonal-observable data a combine pass reads per row."""
        b = _c.c_size_t()
        self.ctx.check(self.ctx.L.te_evolver_diag_bytes(self.h, _c.byref(b)), "te_evolver_diag_bytes")
        return b.value

    def state_device_ptr(self):
        p = _p()
        self.ctx.check(self.ctx.L.te_evolver_state_device(self.h, _c.byref(p)), "state_device")
        return p.value

    def close(self):
        if self.h:
            self.ctx.L.te_evolver_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchEvolver:
    """Many independent krylovTimeEvolver runs advanced together (te_batch_*). Hs: list of
    Matrix; psi0s: list of complex vectors; observables: per instance a list of diagonal Matrix
    observables (same count for every instance)."""

    def __init__(self, ctx, Hs, psi0s, t, sampling_step, observables=None, exp_factor=1.0, tol=1e-6, m=40,
                 fast_integration=False):
        self.ctx = ctx
        n = len(Hs)
        if observables is None:
            observables = [[] for _ in range(n)]
        n_obs = len(observables[0]) if n else 0
        assert all(len(o) == n_obs for o in observables)
        self._keep = [list(Hs), [o for ol in observables for o in ol]]
        self.psis = [np.ascontiguousarray(p, dtype=np.complex128) for p in psi0s]
        Harr = (_p * n)(*[H.h.value if isinstance(H.h, _p) else H.h for H in Hs])
        Parr = (_dp * n)(*[_ptr(p.view(np.float64), _c.c_double) for p in self.psis])
        obs = (te_observable * max(n * n_obs, 1))()
        for i in range(n):
            for o in range(n_obs):
                obs[i * n_obs + o] = te_observable(TE_OBS_SPARSE, observables[i][o].h, None, OBS_CALLBACK(0), None)
        params = te_evolve_params(t, sampling_step, exp_factor, tol, m, int(fast_integration))
        h = _p()
        ctx.check(ctx.L.te_batch_create(ctx.h, n, Harr, Parr, _c.byref(params), obs, n_obs, _c.byref(h)),
                  "te_batch_create")
        self.h = h
        self.n = n
        self.n_obs = n_obs
        self.dims = [H.n_rows for H in Hs]
        rows, ns = _c.c_size_t(), _c.c_size_t()
        ctx.L.te_batch_info(h, _c.byref(rows), _c.byref(ns), None)
        self.packed_rows = rows.value
        self.n_samples = ns.value

    def set_graph(self, on):
        self.ctx.check(self.ctx.L.te_batch_set_graph(self.h, int(on)), "te_batch_set_graph")

    SHAPE_SOURCES = ("default", "tuned", "tuned (cached)", "set")

    def set_shape(self, unit_slices=0, vec_rows=0, spmv_blocks_per_cu=0, vec_blocks_per_cu=0):
        """Fix the launch shape (tuner off); all zero: tune again on the next run."""
        self.ctx.check(self.ctx.L.te_batch_set_shape(self.h, unit_slices, vec_rows, spmv_blocks_per_cu,
                                                     vec_blocks_per_cu), "te_batch_set_shape")

    def shape(self):
        v = [_c.c_int() for _ in range(5)]
        ms = _c.c_double()
        self.ctx.check(self.ctx.L.te_batch_shape(self.h, *[_c.byref(x) for x in v], _c.byref(ms)), "te_batch_shape")
        return {"unit_slices": v[0].value, "vec_rows": v[1].value, "spmv_blocks_per_cu": v[2].value,
                "vec_blocks_per_cu": v[3].value, "source": self.SHAPE_SOURCES[v[4].value], "tune_ms": ms.value}

    KINDS = ("apply", "project", "control", "combine")

    def set_profiling(self, kinds=KINDS):
        """Time the launches of the given kinds (names from KINDS) with HIP events in the next
        runs (rounds are then launched kernel by kernel); False/None/() turns timing off, True
        times every kind."""
        if kinds is True:
            kinds = self.KINDS
        mask = 0
        for k in (kinds or ()):
            mask |= 1 << self.KINDS.index(k)
        self.ctx.check(self.ctx.L.te_batch_set_profiling(self.h, mask), "te_batch_set_profiling")

    def kernel_times(self):
        """{kind: (total_ms, launches)} of the last run."""
        out = {}
        for k, name in enumerate(self.KINDS):
            ms, n = _c.c_double(), _c.c_size_t()
            self.ctx.check(self.ctx.L.te_batch_kernel_time(self.h, k, _c.byref(ms), _c.byref(n)), "kernel_time")
            out[name] = (ms.value, n.value)
        return out

    def kernel_launch_times(self):
        """{kind: array of the last run's launch times in ms, in launch order}."""
        out = {}
        for k, name in enumerate(self.KINDS):
            n = _c.c_size_t()
            self.ctx.check(self.ctx.L.te_batch_kernel_times(self.h, k, None, 0, _c.byref(n)), "kernel_times")
            ms = np.zeros(n.value)
            self.ctx.check(self.ctx.L.te_batch_kernel_times(self.h, k, _ptr(ms, _c.c_double), n.value, _c.byref(n)),
                           "kernel_times")
            out[name] = ms
        return out

    def rounds(self):
        r = _c.c_longlong()
        self.ctx.L.te_batch_info(self.h, None, None, _c.byref(r))
        return r.value

    def run(self, want_states=True, want_obs=True, check=True):
        """Returns a list of per-instance dicts (the keys of Evolver.run) and the return code."""
        R = (te_result * self.n)()
        obs = np.zeros(max(self.n * self.n_obs * self.n_samples, 1))
        states = [np.zeros(d, dtype=np.complex128) for d in self.dims] if want_states else None
        Sarr = (_dp * self.n)(*[_ptr(s.view(np.float64), _c.c_double) for s in states]) if want_states else None
        rc = self.ctx.L.te_batch_run(self.h, _ptr(obs, _c.c_double) if want_obs else None, Sarr, R)
        if check:
            self.ctx.check(rc, "te_batch_run")
        O = obs[:self.n * self.n_obs * self.n_samples].reshape(self.n, self.n_obs, self.n_samples)
        out = []
        for i in range(self.n):
            r = R[i]
            out.append({"state": states[i] if want_states else None, "observables": O[i], "err": r.err,
                        "evolvedTime": r.evolved_time, "numSamples": r.num_samples, "n_steps": r.n_steps,
                        "dim": r.dim, "krylovDim": r.krylov_dim, "statusCode": r.status_code,
                        "n_samples": self.n_samples})
        return (out, rc) if not check else out

    def close(self):
        if self.h:
            self.ctx.L.te_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Adjoint:
    """te_adjoint_*: L = <psi(t)|O|psi(t)> and dL/dtheta_p for H = H_0 + sum_p theta_p G_p."""

    def __init__(self, ctx, H, generators, observable, t, tol=1e-8, m=40, exp_factor=1.0, fast_integration=False,
                 checkpoint_interval=0):
        self.ctx = ctx
        self._keep = [H, list(generators), observable]
        self.P = len(generators)
        Garr = (_p * max(self.P, 1))(*[g.h.value if isinstance(g.h, _p) else g.h for g in generators])
        params = te_evolve_params(t, t, exp_factor, tol, m, int(fast_integration))
        h = _p()
        ctx.check(ctx.L.te_adjoint_create(ctx.h, H.h, self.P, Garr, observable.h, _c.byref(params),
                                          int(checkpoint_interval), _c.byref(h)), "te_adjoint_create")
        self.h = h
        self.dim = H.n_rows

    def set_hamiltonian(self, H):
        self.ctx.check(self.ctx.L.te_adjoint_set_hamiltonian(self.h, H.h), "te_adjoint_set_hamiltonian")
        self._keep[0] = H

    def run(self, psi0):
        """Returns (loss, grad ndarray, result dict, info dict)."""
        psi = np.ascontiguousarray(psi0, dtype=np.complex128)
        loss = _c.c_double()
        grad = np.zeros(max(self.P, 1))
        R = te_result()
        info = te_adjoint_info()
        self.ctx.check(self.ctx.L.te_adjoint_run(self.h, _ptr(psi.view(np.float64), _c.c_double), _c.byref(loss),
                                                 _ptr(grad, _c.c_double), _c.byref(R), _c.byref(info)),
                       "te_adjoint_run")
        res = {"err": R.err, "evolvedTime": R.evolved_time, "n_steps": R.n_steps, "statusCode": R.status_code,
               "krylovDim": R.krylov_dim}
        inf = {k: getattr(info, k) for k, _ in te_adjoint_info._fields_}
        return loss.value, grad[:self.P].copy(), res, inf

    def state(self):
        out = np.zeros(self.dim, dtype=np.complex128)
        self.ctx.check(self.ctx.L.te_adjoint_state(self.h, _ptr(out.view(np.float64), _c.c_double)), "state")
        return out

    def close(self):
        if self.h:
            self.ctx.L.te_adjoint_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
