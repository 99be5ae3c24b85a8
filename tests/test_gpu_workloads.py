"""GPU checks of the three large workloads at their full sizes, through the C-ABI.

* Ring Bose-Hubbard L = N = 14 (D = 20,058,300, nnz = 311,275,100), the time-to-T workload of
  core/krylovTimeEvolver.cpp:261-512 at m = 40, tol = 1e-8:
  - closed-form dimension and stored-entry count;
  - U = 0 (free bosons) from |14,0,...,0> to T = 10, sampling step 0.5: every sample of every
    n_k against N |[exp(-i h t)]_{k0}|^2 (h the 14 x 14 single-particle ring matrix), within
    the bound the returned error implies, 2 ||n_k|| err = 2 N err (|<psi|O|psi> - <phi|O|phi>|
    <= 2 ||O|| ||psi - phi|| for unit vectors);
  - sum_k <n_k> = N and <psi|psi> = 1 at every sample (an identity observable), within 1e-9;
  - the same two checks at L = N = 15 (D = 77,558,760), the largest ring one GPU holds;
  - U = 0.5 (the timed workload, seed 2024) over a short horizon (T = 0.5, sampling step 0.25)
    against the oracle on the CSR the device built: ||psi_gpu - psi_oracle|| <= err_gpu +
    err_oracle + 1e-10, every observable within tol, the same step count and status.
* The batched workload: 512 ring instances, instance i = batch_instance(2024, i) (L = N uniform
  in {6..10}), T = 5, m = 30, tol = 1e-8, sampling step 0.25, observable n_0: every instance
  against the plain single-instance evolver (same steps and status, states within the sum of
  the error bounds, observables within tol), and the first instance of each L against the
  oracle (the same bars).
* The adjoint workload: L = N = 12 (D = 1,352,078), Fock |12,0,...,0>, T = 5, m = 40,
  tol = 1e-8, loss <n_0>(T), gradient with respect to the 12 eps_k and 12 J_k for checkpoint
  intervals K in {0, 1, 4, 16}, against central differences (h = 1e-6) of the plain evolver's
  loss at tol 1e-12: within 1e-6 relative to the gradient norm.
"""
import numpy as np
import pytest
import scipy.linalg as sla

from timeevolver_amd import capi
from timeevolver_amd.adjoint import ring_problem
from timeevolver_amd.hamiltonian import as_tuples, createNumberOperator
from timeevolver_amd.workloads import (batch_instance, fock_state, random_state, ring_bose_hubbard_terms,
                                       ring_couplings, ring_dimension, ring_nnz)

pytestmark = pytest.mark.gpu

SEED = 2024


# ---------------------------------------------------------------------------------------- L = N = 14
@pytest.fixture(scope="module")
def ring14(gpu):
    L = N = 14
    b = capi.Basis.basis(gpu, N, L)
    assert b.D == ring_dimension(L, N) == 20_058_300
    start = np.zeros((1, L), dtype=np.int32)
    start[0, 0] = N
    assert b.find(start)[0] == 0                       # |14,0,...,0> is the first basis state
    obs = [capi.operator_build(gpu, b, as_tuples([createNumberOperator(k, 1)])) for k in range(L)]
    yield L, N, b, obs
    for o in obs:
        o.close()


def _free_bosons_closed_form(gpu, L, N, b, obs):
    """U = 0 from |N,0,...,0> to T = 10 (sampling step 0.5, m = 40, tol = 1e-8): every sample of
    every n_k against the single-particle closed form, sum_k n_k = N and <psi|psi> = 1."""
    D = b.D
    eps, J = ring_couplings(L, SEED)
    terms = ring_bose_hubbard_terms(L, eps, J, 0.0)[:3 * L]     # number terms and hops, no U term
    H = capi.operator_build(gpu, b, as_tuples(terms))
    assert H.nnz == ring_nnz(L, N)
    idx = np.arange(D, dtype=np.uint64)
    ident = capi.Matrix.from_coo(gpu, np.ones(D, dtype=np.complex128), idx, idx, D)
    del idx
    T, step, tol = 10.0, 0.5, 1e-8
    r = capi.Evolver(gpu, H, fock_state(D), T, step, obs + [ident], tol=tol, m=40).run(want_state=False)
    H.close()
    ident.close()
    h = np.diag(eps).astype(float)
    for k in range(L):
        h[k, (k + 1) % L] += J[k]
        h[(k + 1) % L, k] += J[k]
    ts = np.arange(r["n_samples"]) * step
    ref = np.array([N * np.abs(sla.expm(-1j * h * t)[:, 0]) ** 2 for t in ts]).T
    nk = r["observables"][:L]
    dev = np.abs(nk - ref).max()
    norm2 = r["observables"][L]
    print("L=N=%d free bosons: D %d steps %d status %d err %.3e max|n_k - closed form| %.3e max|sum n_k - N| %.3e "
          "max|<psi|psi> - 1| %.3e" % (L, D, r["n_steps"], r["statusCode"], r["err"], dev,
                                        np.abs(nk.sum(0) - N).max(), np.abs(norm2 - 1).max()))
    assert r["numSamples"] == r["n_samples"] == 21
    assert r["err"] <= tol
    assert dev <= 2 * N * r["err"] + 1e-12
    assert np.abs(nk.sum(0) - N).max() <= N * 1e-9
    assert np.abs(norm2 - 1).max() <= 1e-9


def test_ring14_free_bosons_closed_form(gpu, ring14):
    L, N, b, obs = ring14
    assert ring_nnz(L, N) == 311_275_100
    _free_bosons_closed_form(gpu, L, N, b, obs)


def test_ring14_bose_hubbard_short_horizon_vs_oracle(gpu, oracle_lib, ring14):
    L, N, b, obs = ring14
    D = b.D
    eps, J = ring_couplings(L, SEED)
    H = capi.operator_build(gpu, b, as_tuples(ring_bose_hubbard_terms(L, eps, J, 0.5)))
    assert H.nnz == ring_nnz(L, N)
    T, step, tol, m = 0.5, 0.25, 1e-8, 40
    rg = capi.Evolver(gpu, H, fock_state(D), T, step, obs, tol=tol, m=m).run()
    ptr, col, val = H.to_csr()
    H.close()
    diags = []
    for o in obs:
        _, c, v = o.to_csr()
        d = np.zeros(D)
        d[c] = v.real
        diags.append(d)
    ro = oracle_lib.evolve_csr(ptr, col, val, fock_state(D), T, step, diags, tol=tol, m=m, want_state=True)
    del ptr, col, val, diags
    dpsi = np.linalg.norm(rg["state"] - ro["state"])
    dobs = np.abs(rg["observables"] - ro["observables"]).max()
    print("L=N=14 U=0.5 T=0.5: steps gpu/oracle %d/%d err %.3e/%.3e |dpsi| %.3e max|dobs| %.3e status %d/%d" %
          (rg["n_steps"], ro["n_steps"], rg["err"], ro["err"], dpsi, dobs, rg["statusCode"], ro["statusCode"]))
    assert rg["n_steps"] == ro["n_steps"] and rg["statusCode"] == ro["statusCode"]
    assert rg["numSamples"] == ro["numSamples"] == 3
    assert dpsi <= rg["err"] + ro["err"] + 1e-10
    assert dobs <= tol


def test_ring15_free_bosons_closed_form(gpu):
    """The largest ring that fits one GPU: L = N = 15, D = 77,558,760, nnz = 1,281,056,760 (below
    2^31, the limit of 32-bit entry offsets; about 80 GB of HBM with the Krylov basis), the same
    closed-form and invariant checks as at L = N = 14."""
    L = N = 15
    b = capi.Basis.basis(gpu, N, L)
    assert b.D == ring_dimension(L, N) == 77_558_760
    assert ring_nnz(L, N) == 1_281_056_760
    obs = [capi.operator_build(gpu, b, as_tuples([createNumberOperator(k, 1)])) for k in range(L)]
    try:
        _free_bosons_closed_form(gpu, L, N, b, obs)
    finally:
        for o in obs:
            o.close()
        b.close()


# ---------------------------------------------------------------------------------------- 512 instances
def test_batch_workload_512(gpu, oracle_lib):
    T, step, tol, m = 5.0, 0.25, 1e-8, 30
    specs, bases, Hs, psis, obs = [], [], [], [], []
    for i in range(512):
        L, N, eps, J, U, sseed = batch_instance(SEED, i)
        terms = as_tuples(ring_bose_hubbard_terms(L, eps, J, U))
        b = capi.Basis.basis(gpu, N, L)
        Hs.append(capi.operator_build(gpu, b, terms))
        obs.append([capi.operator_build(gpu, b, as_tuples([createNumberOperator(0, 1)]))])
        psis.append(random_state(b.D, sseed))
        specs.append((L, N, terms))
        bases.append(b)
    dims = [H.n_rows for H in Hs]
    assert set(specs[i][0] for i in range(512)) == {6, 7, 8, 9, 10}
    be = capi.BatchEvolver(gpu, Hs, psis, T, step, obs, tol=tol, m=m)
    rb = be.run()
    be.close()
    worst = 0.0
    for i in range(512):
        ev = capi.Evolver(gpu, Hs[i], psis[i], T, step, obs[i], tol=tol, m=m)
        rp = ev.run()
        ev.close()
        r = rb[i]
        assert r["n_steps"] == rp["n_steps"] and r["statusCode"] == rp["statusCode"], i
        assert r["numSamples"] == rp["numSamples"] == 21, i
        assert np.linalg.norm(r["state"] - rp["state"]) <= r["err"] + rp["err"] + 1e-10, i
        d = np.abs(r["observables"] - rp["observables"]).max()
        worst = max(worst, d)
        assert d <= tol, (i, d)
    print("512 instances (sum D %d): batch = plain in steps and status; max|dobs| %.3e" % (sum(dims), worst))
    seen = set()
    for i, (L, N, terms) in enumerate(specs):
        if L in seen:
            continue
        seen.add(L)
        ob = oracle_lib.Basis.basis(N, L)
        Hc = oracle_lib.build_operator(ob, terms)
        n0 = oracle_lib.build_operator(ob, as_tuples([createNumberOperator(0, 1)]))
        ro = oracle_lib.evolve(ob.D, Hc, psis[i], T, step, [("sparse", n0)], tol=tol, m=m)
        r = rb[i]
        dpsi = np.linalg.norm(r["state"] - ro["state"])
        dobs = np.abs(r["observables"] - ro["observables"]).max()
        print("instance %d (L=%d, D=%d): steps %d/%d |dpsi| %.3e max|dobs| %.3e" %
              (i, L, ob.D, r["n_steps"], ro["n_steps"], dpsi, dobs))
        assert r["statusCode"] == ro["statusCode"] and r["numSamples"] == ro["numSamples"]
        assert dpsi <= r["err"] + ro["err"] + 1e-10
        assert dobs <= tol
    assert seen == {6, 7, 8, 9, 10}
    for H, o in zip(Hs, obs):
        H.close()
        o[0].close()


# ---------------------------------------------------------------------------------------- adjoint, L = 12
@pytest.fixture(scope="module")
def adjoint12_fd(gpu):
    """Central differences (h = 1e-6) of <n_0>(T) from the plain evolver at tol 1e-12."""
    L = N = 12
    T, m, U = 5.0, 40, 0.5
    eps, J = ring_couplings(L, SEED)
    theta = np.array(eps + J)
    b = capi.Basis.basis(gpu, N, L)
    psi0 = fock_state(b.D)
    n0 = capi.operator_build(gpu, b, as_tuples([createNumberOperator(0, 1.0)]))

    def loss(th):
        H = capi.operator_build(gpu, b, as_tuples(ring_bose_hubbard_terms(L, list(th[:L]), list(th[L:]), U)))
        ev = capi.Evolver(gpu, H, psi0, T, T, [n0], tol=1e-12, m=m)
        v = ev.run(want_state=False)["observables"][0, 1]
        ev.close()
        H.close()
        return v

    h = 1e-6
    fd = np.zeros_like(theta)
    for p in range(len(theta)):
        tp, tm = theta.copy(), theta.copy()
        tp[p] += h
        tm[p] -= h
        fd[p] = (loss(tp) - loss(tm)) / (2 * h)
    return theta, fd


@pytest.mark.parametrize("K", [0, 1, 4, 16])
def test_adjoint_L12_vs_finite_differences(gpu, adjoint12_fd, K):
    L = N = 12
    theta, fd = adjoint12_fd
    prob = ring_problem(gpu, L, N, 0.5, 5.0, tol=1e-8, m=40, checkpoint_interval=K, theta0=theta)
    assert prob.basis.D == 1_352_078
    loss, grad, res, info = prob.loss_and_grad(theta, fock_state(prob.basis.D))
    rel = np.linalg.norm(grad - fd) / np.linalg.norm(fd)
    print("L=N=12 adjoint K=%d: steps %d nodes %d |grad| %.4e rel err vs central differences %.2e" %
          (K, info["n_steps"], info["nodes"], np.linalg.norm(grad), rel))
    assert info["capped"] == 0
    assert rel <= 1e-6
    prob.adj.close()
