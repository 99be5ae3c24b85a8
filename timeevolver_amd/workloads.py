"""Seeded synthetic workloads: ring Bose-Hubbard models and initial states.

Couplings come from a SplitMix64 stream with u = (x >> 11) * 2^-53, so the same seed gives
identical bits wherever the stream is regenerated. For a ring of L modes:
    eps_k = u,  J_k = 2u - 1  (bond k -> (k+1) mod L),  U = 0.5 unless given.
Term order (fixed):
  1. createNumberOperator(k, eps_k) for k = 0..L-1
  2. for k = 0..L-1: linInteraction(k, k+1, 0, 0, false, J_k), linInteraction(k, k+1, 0, 0, true, J_k)
  3. createFourPointB(k, k, k, k, 0, 0, 0, 0, U/2) for k = 0..L-1   (U/2 n_k (n_k - 1))
"""
import math

import numpy as np

from .hamiltonian import createFourPointB, createNumberOperator, linInteraction

_MASK = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.x = seed & _MASK

    def next_u64(self):
        self.x = (self.x + 0x9E3779B97F4A7C15) & _MASK
        z = self.x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
        return z ^ (z >> 31)

    def uniform(self):
        return (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)


def ring_couplings(L, seed):
    rng = SplitMix64(seed)
    eps = [rng.uniform() for _ in range(L)]
    J = [2.0 * rng.uniform() - 1.0 for _ in range(L)]
    return eps, J


def ring_bose_hubbard_terms(L, eps, J, U):
    terms = [createNumberOperator(k, eps[k]) for k in range(L)]
    for k in range(L):
        terms.append(linInteraction(k, (k + 1) % L, 0, 0, False, J[k]))
        terms.append(linInteraction(k, (k + 1) % L, 0, 0, True, J[k]))
    for k in range(L):
        terms.append(createFourPointB(k, k, k, k, 0, 0, 0, 0, U / 2.0))
    return terms


def ring_dimension(L, N):
    return math.comb(N + L - 1, L - 1)


def ring_nnz(L, N):
    """Closed form for L >= 3: every state has a non-zero diagonal (eps_k > 0 or U > 0 with
    N >= 1), and each directed hop a_a^dag a_b between ring neighbours is allowed from
    exactly C(N+L-2, L-1) states (those with n_b >= 1)."""
    return ring_dimension(L, N) + 2 * L * math.comb(N + L - 2, L - 1)


def random_state(D, seed):
    """Normalised complex Gaussian vector from SplitMix64 + Box-Muller, in FP64 on the host."""
    rng = SplitMix64(seed)
    u = np.array([rng.uniform() for _ in range(2 * D)]) if D <= 4096 else _fast_uniform(seed, 2 * D)
    u1 = np.maximum(u[0::2], 1e-300)
    u2 = u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    z = r * np.cos(2 * np.pi * u2) + 1j * r * np.sin(2 * np.pi * u2)
    return z / np.sqrt(np.sum(np.abs(z) ** 2))


def _fast_uniform(seed, n):
    """Vectorised SplitMix64 (same stream as SplitMix64.uniform)."""
    with np.errstate(over="ignore"):
        idx = np.arange(1, n + 1, dtype=np.uint64)
        x = np.uint64(seed & _MASK) + idx * np.uint64(0x9E3779B97F4A7C15)
        z = x
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
        return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def fock_state(D, index=0):
    v = np.zeros(D, dtype=np.complex128)
    v[index] = 1.0
    return v


def batch_instance(base_seed, i, L_choices=(6, 7, 8, 9, 10)):
    """Instance i of a seeded batch of ring Bose-Hubbard models: a SplitMix64 stream seeded with
    base_seed + i draws L uniformly from L_choices (N = L), then eps_k = u and J_k = 2u - 1 for
    k < L, then U = 0.2 + 0.8 u. The initial state is random_state(D, that seed x 7919 + 1).
    Returns (L, N, eps, J, U, state_seed)."""
    rng = SplitMix64(base_seed + i)
    nL = len(L_choices)
    L = L_choices[min(int(rng.uniform() * nL), nL - 1)]
    eps = [rng.uniform() for _ in range(L)]
    J = [2.0 * rng.uniform() - 1.0 for _ in range(L)]
    U = 0.2 + 0.8 * rng.uniform()
    return L, L, eps, J, U, (base_seed + i) * 7919 + 1
