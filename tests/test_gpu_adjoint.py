"""GPU checks of the adjoint gradient (te_adjoint_*) of L = <psi(T)|n_0|psi(T)> with respect to
the ring Bose-Hubbard couplings (eps_k, J_k).

References: the dense FP64 Frechet derivative of exp(-i H T) (scipy.linalg.expm_frechet on the
full matrix, L = N = 6, D = 462), and central finite differences (h = 1e-6) of the forward loss
evaluated at a tight tolerance. Bars: 1e-6 relative to the gradient norm,
the loss against the dense value within 2 N (err) + 1e-12, and the loss equal to the plain
evolver's observable at T.
"""
import numpy as np
import pytest
import scipy.linalg as sla

from timeevolver_amd import capi
from timeevolver_amd.adjoint import evolve_loss, ring_problem
from timeevolver_amd.hamiltonian import as_tuples, createNumberOperator
from timeevolver_amd.workloads import fock_state, random_state, ring_bose_hubbard_terms, ring_couplings

pytestmark = pytest.mark.gpu


def _dense(mat):
    v, c, r = mat.to_coo()
    D = mat.n_rows
    A = np.zeros((D, D), dtype=np.complex128)
    np.add.at(A, (r, c), v)
    return A


def _dense_grad(prob, theta, psi0, T):
    """2 Re <psi_T| O |d psi_T> with d psi_T = L(-iHT, -iG_p T) psi0 (augmented-matrix Frechet)."""
    H = _dense(prob.hamiltonian(theta))
    O = _dense(prob.observable)
    A = -1j * T * H
    E0 = sla.expm(A)
    psiT = E0 @ psi0
    g = []
    for G in prob.generators:
        _, dE = sla.expm_frechet(A, -1j * T * _dense(G))
        g.append(2.0 * np.real(np.vdot(psiT, O @ (dE @ psi0))))
    return float(np.real(np.vdot(psiT, O @ psiT))), np.array(g)


def _theta(L, seed):
    eps, J = ring_couplings(L, seed)
    return np.array(eps + J)


@pytest.mark.parametrize("K", [0, 1, 4, 16])
def test_adjoint_vs_dense_frechet_L6(gpu, K):
    L = N = 6
    T = 5.0
    theta = _theta(L, 17)
    prob = ring_problem(gpu, L, N, 0.5, T, tol=1e-8, m=40, checkpoint_interval=K, theta0=theta)
    psi0 = fock_state(prob.basis.D)
    loss, grad, res, info = prob.loss_and_grad(theta, psi0)
    l_ref, g_ref = _dense_grad(prob, theta, psi0, T)
    rel = np.linalg.norm(grad - g_ref) / np.linalg.norm(g_ref)
    print("K=%d steps=%d nodes=%d max phase %.1f capped %d |dL|=%.2e grad rel err %.2e bwd err %.2e psi err %.2e"
          % (K, info["n_steps"], info["nodes"], info["max_phase"], info["capped"], abs(loss - l_ref), rel,
             info["backward_err"], info["psi_err"]))
    assert info["capped"] == 0
    assert abs(loss - l_ref) <= 2 * N * res["err"] + 1e-12
    assert rel <= 1e-6


@pytest.mark.parametrize("K", [1, 2])
def test_adjoint_paired_recurrences_break_down_apart(gpu, K):
    """Uniform couplings on a 4-mode ring with 2 particles (D = 10) and the Fock start |2,0,0,0>
    keep psi in the subspace symmetric under the reflection about mode 0, while the loss n_1
    breaks that symmetry: psi's Lanczos recurrence breaks down after 6 iterations and lambda's
    after 7 (plain Lanczos on the dense matrix), so for one iteration the paired apply and
    projection go on with lambda alone. Gradient against the dense Frechet derivative."""
    L, N = 4, 2
    T = 2.0
    theta = np.array([0.3] * L + [0.7] * L)
    prob = ring_problem(gpu, L, N, 0.5, T, tol=1e-8, m=40, obs_mode=1, checkpoint_interval=K, theta0=theta)
    psi0 = fock_state(prob.basis.D)
    assert list(prob.basis.states()[0]) == [N, 0, 0, 0]
    loss, grad, res, info = prob.loss_and_grad(theta, psi0)
    l_ref, g_ref = _dense_grad(prob, theta, psi0, T)
    assert res["statusCode"] == 1                      # the forward broke down too (lucky)
    assert abs(loss - l_ref) <= 2 * N * res["err"] + 1e-12
    assert np.linalg.norm(grad - g_ref) <= 1e-6 * np.linalg.norm(g_ref)


def test_adjoint_random_state_and_loss_matches_plain(gpu):
    """A random initial state, a different observable mode; the loss is the plain evolver's <n_2>(T)."""
    L = N = 5
    T = 3.0
    theta = _theta(L, 23)
    prob = ring_problem(gpu, L, N, 0.7, T, tol=1e-8, m=30, obs_mode=2, theta0=theta)
    psi0 = random_state(prob.basis.D, 5)
    loss, grad, res, info = prob.loss_and_grad(theta, psi0)
    l_ref, g_ref = _dense_grad(prob, theta, psi0, T)
    assert np.linalg.norm(grad - g_ref) <= 1e-6 * np.linalg.norm(g_ref)
    H = prob.hamiltonian(theta)
    obs = capi.operator_build(gpu, prob.basis, as_tuples([createNumberOperator(2, 1.0)]))
    rp = capi.Evolver(gpu, H, psi0, T, T, [obs], tol=1e-8, m=30).run()
    assert rp["n_steps"] == res["n_steps"] and rp["statusCode"] == res["statusCode"]
    assert abs(rp["observables"][0, 1] - loss) <= 1e-13
    assert np.linalg.norm(prob.adj.state() - rp["state"]) <= 1e-13


def test_adjoint_vs_finite_differences_L8(gpu):
    """Central differences (h = 1e-6) of the loss evaluated at tol 1e-12, all 16 parameters."""
    L = N = 8
    T = 2.0
    theta = _theta(L, 31)
    prob = ring_problem(gpu, L, N, 0.5, T, tol=1e-8, m=40, theta0=theta)
    psi0 = fock_state(prob.basis.D)
    _, grad, _, _ = prob.loss_and_grad(theta, psi0)
    tight = ring_problem(gpu, L, N, 0.5, T, tol=1e-12, m=40, theta0=theta)
    h = 1e-6
    fd = np.zeros_like(theta)
    for p in range(len(theta)):
        tp, tm = theta.copy(), theta.copy()
        tp[p] += h
        tm[p] -= h
        fd[p] = (tight.loss_and_grad(tp, psi0)[0] - tight.loss_and_grad(tm, psi0)[0]) / (2 * h)
    rel = np.linalg.norm(grad - fd) / np.linalg.norm(fd)
    print("FD rel err %.2e" % rel)
    assert rel <= 1e-6


@pytest.mark.parametrize("K", [0, 1])
def test_adjoint_runs_are_bitwise_repeatable(gpu, K):
    """Every reduction of the forward and backward passes (the Lanczos scalars summed by their
    finish launches, the node passes' label sums, the observables) runs in a fixed order: two
    runs of one adjoint give the same loss and gradient bit for bit."""
    L = N = 8
    theta = _theta(L, 17)
    prob = ring_problem(gpu, L, N, 0.5, 1.5, tol=1e-8, m=40, checkpoint_interval=K, theta0=theta)
    psi0 = random_state(prob.basis.D, 6)
    l1, g1, _, _ = prob.loss_and_grad(theta, psi0)
    l2, g2, _, _ = prob.loss_and_grad(theta, psi0)
    assert np.float64(l1).view(np.uint64) == np.float64(l2).view(np.uint64)
    assert np.array_equal(np.asarray(g1).view(np.uint64), np.asarray(g2).view(np.uint64))


def test_torch_autograd_wrapper(gpu):
    import torch
    L = N = 4
    theta0 = _theta(L, 3)
    prob = ring_problem(gpu, L, N, 0.5, 1.5, tol=1e-8, m=20, theta0=theta0)
    psi0 = fock_state(prob.basis.D)
    theta = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
    loss = evolve_loss(theta, prob, psi0)
    (3.0 * loss).backward()
    l2, g2, _, _ = prob.loss_and_grad(theta0, psi0)
    assert abs(loss.item() - l2) == 0.0
    assert np.allclose(theta.grad.numpy(), 3.0 * g2, rtol=0, atol=1e-15)


@pytest.mark.parametrize("K", [0, 1])
def test_adjoint_rank_form_matches_quadrature_nodes(gpu, K, monkeypatch):
    """Each step's node sum in low-rank form (the default) against the quadrature nodes
    themselves (TE_ADJ_RANK_TOL=0, read at every run): the rank form is exact to 1e-14 of the
    step's coefficient matrix, so the gradients agree far below the 1e-6 bar, with fewer node
    states formed."""
    L = N = 8
    T = 2.0
    theta = _theta(L, 31)
    prob = ring_problem(gpu, L, N, 0.5, T, tol=1e-8, m=40, checkpoint_interval=K, theta0=theta)
    psi0 = fock_state(prob.basis.D)
    l1, g1, _, i1 = prob.loss_and_grad(theta, psi0)
    monkeypatch.setenv("TE_ADJ_RANK_TOL", "0")
    l0, g0, _, i0 = prob.loss_and_grad(theta, psi0)
    rel = np.linalg.norm(np.asarray(g1) - np.asarray(g0)) / np.linalg.norm(g0)
    print("K=%d quadrature nodes %d, node states %d (rank form) / %d; rel diff %.2e"
          % (K, i1["nodes"], i1["node_states"], i0["node_states"], rel))
    assert l1 == l0
    assert i0["node_states"] == i0["nodes"] == i1["nodes"]
    assert i1["node_states"] < i1["nodes"]
    assert rel <= 1e-11


@pytest.mark.parametrize("K", [0, 1])
def test_adjoint_eigen_solvers_agree(gpu, K, monkeypatch):
    """The coefficient kernel's eigen-decompositions by bisection and twisted factorisations
    (the default) against the QL iteration (TE_ADJ_EIG=ql, read at every run): both are
    accurate to about 1e-13 of ||T|| on these Krylov matrices, so the gradients agree far below
    the 1e-6 bar, and the loss (the forward pass) is the same bit for bit."""
    L = N = 8
    T = 2.0
    theta = _theta(L, 31)
    prob = ring_problem(gpu, L, N, 0.5, T, tol=1e-8, m=40, checkpoint_interval=K, theta0=theta)
    psi0 = fock_state(prob.basis.D)
    l1, g1, _, i1 = prob.loss_and_grad(theta, psi0)
    monkeypatch.setenv("TE_ADJ_EIG", "ql")
    l0, g0, _, i0 = prob.loss_and_grad(theta, psi0)
    rel = np.linalg.norm(np.asarray(g1) - np.asarray(g0)) / np.linalg.norm(g0)
    print("K=%d bisection/twisted vs QL: rel diff %.2e, node states %d / %d"
          % (K, rel, i1["node_states"], i0["node_states"]))
    assert l1 == l0
    assert i1["nodes"] == i0["nodes"]
    assert rel <= 1e-10


@pytest.mark.parametrize("m", [40, 48])
def test_adjoint_node_states_on_matrix_cores(gpu, monkeypatch, m):
    """The node states formed on the matrix cores (materialise_mfma_kernel, the default: 4-deep
    FP64 MFMA steps over the basis columns) against the vector form (TE_ADJ_MAT=valu, one FMA per
    column): the sums differ only in their order, so the gradients agree to rounding, and the
    loss (the forward pass) is the same bit for bit. m = 48 takes the wider coefficient table."""
    L = N = 8
    T = 2.0
    theta = _theta(L, 37)
    prob = ring_problem(gpu, L, N, 0.5, T, tol=1e-8, m=m, checkpoint_interval=1, theta0=theta)
    psi0 = fock_state(prob.basis.D)
    l1, g1, _, i1 = prob.loss_and_grad(theta, psi0)
    monkeypatch.setenv("TE_ADJ_MAT", "valu")
    l0, g0, _, i0 = prob.loss_and_grad(theta, psi0)
    rel = np.linalg.norm(np.asarray(g1) - np.asarray(g0)) / np.linalg.norm(g0)
    print("m=%d matrix cores vs vector form: rel diff %.2e" % (m, rel))
    assert l1 == l0
    assert i1["nodes"] == i0["nodes"]
    assert rel <= 1e-12


def test_adjoint_krylov_dimension_above_40(gpu):
    """m = 48 (the coefficient and rank kernels' wider form) against the dense Frechet gradient."""
    L = N = 6
    T = 3.0
    theta = _theta(L, 19)
    prob = ring_problem(gpu, L, N, 0.5, T, tol=1e-8, m=48, checkpoint_interval=1, theta0=theta)
    psi0 = fock_state(prob.basis.D)
    loss, grad, res, info = prob.loss_and_grad(theta, psi0)
    l_ref, g_ref = _dense_grad(prob, theta, psi0, T)
    rel = np.linalg.norm(grad - g_ref) / np.linalg.norm(g_ref)
    print("m=48: steps %d nodes %d node states %d rel err %.2e" % (info["n_steps"], info["nodes"],
                                                                   info["node_states"], rel))
    assert res["krylovDim"] == 48
    assert abs(loss - l_ref) <= 2 * N * res["err"] + 1e-12
    assert rel <= 1e-6
