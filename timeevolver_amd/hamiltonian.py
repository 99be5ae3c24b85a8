"""Operator-term factories: the Python mirror of the reference's Hamiltonian builder API.

An op term is (coef, operations, modes, capacities) exactly as helper/hamiltonian.h:18-35
stores it: operations are applied in list order to a basis state, +1 is a creation and -1
an annihilation operator, capacity 0 means "no maximal occupation". The factories keep the
reference's argument order and its reversed storage order (the operator written last in
the formula is applied first), so a term list built here produces the same matrix as the
reference's createMatrix (helper/hamiltonian.cpp:318-401) when handed to te_operator_build.
"""
from dataclasses import dataclass, field
from typing import List


@dataclass
class OpTerm:
    """helper/hamiltonian.h:18-35."""
    coef: complex = 0j
    operations: List[int] = field(default_factory=list)
    mode: List[int] = field(default_factory=list)
    capacity: List[int] = field(default_factory=list)

    def as_tuple(self):
        return (complex(self.coef), list(self.operations), list(self.mode), list(self.capacity))


def createKineticTerms(modes, gap):
    """helper/hamiltonian.cpp:11-31: gap * a_k^dag a_k for each k in modes."""
    return [OpTerm(complex(gap, 0.0), [-1, 1], [k, k], [0, 0]) for k in modes]


def linInteraction(a, b, ca, cb, conjugate, cof):
    """helper/hamiltonian.cpp:43-65: cof * a_a^dag a_b, or its conjugate a_b^dag a_a."""
    if not conjugate:
        return OpTerm(complex(cof, 0.0), [-1, 1], [b, a], [cb, ca])
    return OpTerm(complex(cof, 0.0), [-1, 1], [a, b], [ca, cb])


def createNumberOperator(mode, cof):
    """helper/hamiltonian.cpp:73-85: cof * a^dag a on one mode."""
    return OpTerm(complex(cof, 0.0), [-1, 1], [mode, mode], [0, 0])


def createFourPointA(a, b, c, d, ca, cb, cc, cd, cof):
    """helper/hamiltonian.cpp:100-119: cof * a_a^dag a_b a_c^dag a_d."""
    return OpTerm(complex(cof, 0.0), [-1, 1, -1, 1], [d, c, b, a], [cd, cc, cb, ca])


def createFourPointB(a, b, c, d, ca, cb, cc, cd, cof):
    """helper/hamiltonian.cpp:134-153: cof * a_a^dag a_b^dag a_c a_d."""
    return OpTerm(complex(cof, 0.0), [-1, -1, 1, 1], [d, c, b, a], [cd, cc, cb, ca])


def createFourPointC(a, b, c, d, ca, cb, cc, cd, cof):
    """helper/hamiltonian.cpp:168-187: cof * a_a a_b a_c^dag a_d^dag."""
    return OpTerm(complex(cof, 0.0), [1, 1, -1, -1], [d, c, b, a], [cd, cc, cb, ca])


def createSixPoint(a, b, c, d, e, f, ca, cb, cc, cd, ce, cf, cof):
    """helper/hamiltonian.cpp:206-231: alternating creation/annihilation on six modes."""
    return OpTerm(complex(cof, 0.0), [-1, 1, -1, 1, -1, 1], [f, e, d, c, b, a], [cf, ce, cd, cc, cb, ca])


def as_tuples(terms):
    return [t.as_tuple() if isinstance(t, OpTerm) else tuple(t) for t in terms]
