"""Test models restated from the reference's example programs (data only).

* black hole: example/exampleHamiltonian.cpp:27-138 with the parameters example/main.cpp
  passes (:84, :88-89, :103); produces the golden .h5 observables.
* two coupled oscillators: example/simpleExample.cpp:21-67; produces the golden CSVs.
"""
import math

from timeevolver_amd.hamiltonian import createFourPointA, createNumberOperator, linInteraction, OpTerm


def _sym_break(a, b):
    """example/exampleHamiltonian.cpp:116-122."""
    coef = math.fmod(math.sqrt(2) * math.pow(a, 3) + math.sqrt(7) * math.pow(b, 5), 1.0)
    if coef < 0.5:
        coef -= 1.0
    return coef


def _spherical(Q1, Q2, gm, gms, C):
    """example/exampleHamiltonian.cpp:70-114."""
    terms = []
    for m in range(2, Q1 + Q2 + 2):
        for M in range(2, Q1 + Q2 + 2):
            coeff = 0.0
            if m <= 1 + Q1 and M > Q1 + 1:
                coeff = _sym_break(m, M) * gm
            elif ((m <= Q1 + 1 and M <= 1 + Q1) or (m > 1 + Q1 and M > 1 + Q1)) and m < M:
                coeff = _sym_break(m, M) * gms
            if coeff != 0:
                terms.append(OpTerm(complex(coeff, 0.0), [-1, 1], [m, M], [C, C]))
                terms.append(OpTerm(complex(coeff, 0.0), [-1, 1], [M, m], [C, C]))
    return terms


def blackhole(N0=20, Nm=2, K=4, C0=1.0, Cm=1.0, DeltaN=12.0, capacity=1):
    """Returns (tensor-basis args, term list, initial occupation vector)."""
    Q1 = Q2 = K
    C = capacity
    CGap1 = CGap2 = 1.0
    ham = []
    for i in range(2, Q1 + 2):
        ham.append(createNumberOperator(i, CGap1 * math.sqrt(N0)))
    for i in range(2 + Q1, Q1 + Q2 + 2):
        ham.append(createNumberOperator(i, CGap2 * math.sqrt(N0)))
    ham.append(linInteraction(0, 1, 0, 0, True, C0))
    ham.append(linInteraction(0, 1, 0, 0, False, C0))
    for i in range(2, 2 + Q1):
        ham.append(createFourPointA(0, 0, i, i, 0, 0, C, C, -math.sqrt(N0) * CGap1 / N0))
    for i in range(2 + Q1, 2 + Q1 + Q2):
        ham.append(createFourPointA(0, 0, i, i, 0, 0, C, C, -math.sqrt(N0) * CGap2 / (N0 - DeltaN)))
    ham += _spherical(Q1, Q2, Cm, Cm, C)
    n_modes = Q1 + Q2 + 2
    init = [0] * n_modes
    init[0] = N0
    for i in range(min(Q1, Nm)):
        init[i + 2] = 1
    return (N0, 2, Nm, 2 * K, capacity), ham, init


def two_oscillators(N0=200, E1=1.0, E2=2.0, lam=1.0):
    terms = [createNumberOperator(0, E1), createNumberOperator(1, E2),
             linInteraction(0, 1, 0, 0, False, lam), linInteraction(0, 1, 0, 0, True, lam)]
    return (N0, 2, 0, 0), terms, [N0, 0]


def number_operator_terms(n_modes):
    """helper/hamiltonian.cpp:289-301: one single-term list per mode."""
    return [[createNumberOperator(i, 1)] for i in range(n_modes)]
