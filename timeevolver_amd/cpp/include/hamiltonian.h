#pragma once
// Operator terms and their matrix representation (helper/hamiltonian.h:18-71). createMatrix
// runs on the device (te_operator_build) and returns an smatrix that already carries its
// device CSR, with the host COO arrays filled for callers that read them.
#include <complex>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "Basis.h"
#include "matrixDataTypes.h"

using namespace TE;

struct opTerm
{
    std::complex<double> coef;
    std::vector<int> operations;
    std::vector<int> mode;
    std::vector<int> capacity;

    void reset()
    {
        coef = std::complex<double>(0, 0);
        operations.clear();
        mode.clear();
        capacity.clear();
    }
};

class Hamiltonian
{
public:
    std::vector<opTerm> hamiltonOperator;

    Hamiltonian();

    std::string toString();
    std::unique_ptr<smatrix> createHamiltonMatrix(basicBasis* basis);
    std::vector<std::unique_ptr<smatrix>> createNumberOperatorObservables(basicBasis* basis);

    std::vector<opTerm> createKineticTerms(std::vector<int> modes, double gap);

    opTerm linInteraction(int a, int b, int ca, int cb, bool conjugate, double cof);
    opTerm createNumberOperator(int mode, double cof);
    opTerm createFourPointA(int a, int b, int c, int d, int ca, int cb, int cc, int cd, double cof);
    opTerm createFourPointB(int a, int b, int c, int d, int ca, int cb, int cc, int cd, double cof);
    opTerm createFourPointC(int a, int b, int c, int d, int ca, int cb, int cc, int cd, double cof);
    opTerm createSixPoint(int a, int b, int c, int d, int e, int f, int ca, int cb, int cc, int cd, int ce, int cf, double cof);

protected:
    std::unique_ptr<smatrix> createMatrix(std::vector<opTerm>& op, basicBasis* basis);
};
