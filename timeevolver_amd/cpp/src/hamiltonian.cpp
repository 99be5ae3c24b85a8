#include "hamiltonian.h"

#include "runtime.h"

// Term factories (helper/hamiltonian.cpp:11-231: same argument meaning and the same storage
// order, the operator written last in the formula first) and the device matrix build that
// replaces createMatrix (:318-401).

Hamiltonian::Hamiltonian() {}

static opTerm make_term(double cof, std::vector<int> ops, std::vector<int> modes, std::vector<int> caps) {
    opTerm t;
    t.coef = std::complex<double>(0.0, 0.0);
    t.coef.real(cof);
    t.operations = std::move(ops);
    t.mode = std::move(modes);
    t.capacity = std::move(caps);
    return t;
}

std::vector<opTerm> Hamiltonian::createKineticTerms(std::vector<int> modes, double gap) {
    std::vector<opTerm> ret;
    for (int k : modes) {
        opTerm t = make_term(0.0, {-1, 1}, {k, k}, {0, 0});
        t.coef = std::complex<double>(gap, 0.0);
        ret.push_back(t);
    }
    return ret;
}

opTerm Hamiltonian::linInteraction(int a, int b, int ca, int cb, bool conjugate, double cof) {
    if (!conjugate) return make_term(cof, {-1, 1}, {b, a}, {cb, ca});
    return make_term(cof, {-1, 1}, {a, b}, {ca, cb});
}

opTerm Hamiltonian::createNumberOperator(int mode, double cof) {
    return make_term(cof, {-1, 1}, {mode, mode}, {0, 0});
}

opTerm Hamiltonian::createFourPointA(int a, int b, int c, int d, int ca, int cb, int cc, int cd, double cof) {
    return make_term(cof, {-1, 1, -1, 1}, {d, c, b, a}, {cd, cc, cb, ca});
}

opTerm Hamiltonian::createFourPointB(int a, int b, int c, int d, int ca, int cb, int cc, int cd, double cof) {
    return make_term(cof, {-1, -1, 1, 1}, {d, c, b, a}, {cd, cc, cb, ca});
}

opTerm Hamiltonian::createFourPointC(int a, int b, int c, int d, int ca, int cb, int cc, int cd, double cof) {
    return make_term(cof, {1, 1, -1, -1}, {d, c, b, a}, {cd, cc, cb, ca});
}

opTerm Hamiltonian::createSixPoint(int a, int b, int c, int d, int e, int f, int ca, int cb, int cc, int cd, int ce,
                                   int cf, double cof) {
    return make_term(cof, {-1, 1, -1, 1, -1, 1}, {f, e, d, c, b, a}, {cf, ce, cd, cc, cb, ca});
}

std::string Hamiltonian::toString() {
    std::string ret;
    for (const opTerm& t : hamiltonOperator) {
        ret += std::to_string(t.coef.real());
        for (size_t i = 0; i != t.capacity.size(); i++) {
            ret += t.capacity[i] == 0 ? "a" : "b";
            ret += std::to_string(t.mode[i]);
            if (t.operations[i] == 1) ret += "d";
        }
        ret += "\n+";
    }
    return ret;
}

std::unique_ptr<smatrix> Hamiltonian::createHamiltonMatrix(basicBasis* basis) { return createMatrix(hamiltonOperator, basis); }

std::vector<std::unique_ptr<smatrix>> Hamiltonian::createNumberOperatorObservables(basicBasis* basis) {
    std::vector<std::unique_ptr<smatrix>> out;
    for (int i = 0; i != basis->numberModes; i++) {
        std::vector<opTerm> ops;
        ops.push_back(createNumberOperator(i, 1));
        out.push_back(createMatrix(ops, basis));
    }
    return out;
}

std::unique_ptr<smatrix> Hamiltonian::createMatrix(std::vector<opTerm>& op, basicBasis* basis) {
    std::vector<te_op_term> terms(op.size());
    for (size_t i = 0; i < op.size(); i++) {
        terms[i].coef_re = op[i].coef.real();
        terms[i].coef_im = op[i].coef.imag();
        terms[i].n_ops = (int)op[i].operations.size();
        terms[i].operations = op[i].operations.data();
        terms[i].modes = op[i].mode.data();
        terms[i].capacities = op[i].capacity.data();
    }
    te_matrix* A = nullptr;
    tecpp::check(te_operator_build(tecpp::ctx(), basis->device(), terms.data(), (int)terms.size(), &A), "createMatrix");
    size_t nr = 0, nc = 0, nnz = 0;
    te_matrix_info(A, &nr, &nc, &nnz);
    std::vector<std::complex<double>> v(nnz);
    std::vector<size_t> col(nnz), row(nnz);
    tecpp::check(te_matrix_to_coo(A, reinterpret_cast<double*>(v.data()), col.data(), row.data()), "createMatrix");
    const unsigned int M = (unsigned int)basis->numberElements;
    auto S = std::make_unique<smatrix>(v.data(), col.data(), row.data(), nnz, M, M);
    S->adoptDevice(A);
    return S;
}
