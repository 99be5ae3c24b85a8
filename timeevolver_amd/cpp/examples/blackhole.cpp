// The black-hole model of the reference's example program, run through the reference-compatible
// C++ API and saved with saveResult (flow of example/main.cpp:80-163; the model of
// example/exampleHamiltonian.cpp:27-138 written out here for this layer).
//
// Two control modes (0, 1) holding N0 particles and 2K critical modes of the given capacity:
//   H = sqrt(N0) sum_{k critical} n_k + C0 (a_0^dag a_1 + h.c.)
//       - sqrt(N0)/N0 sum_{k in first K} n_0 n_k - sqrt(N0)/(N0 - DeltaN) sum_{k in last K} n_0 n_k
//       + sum_{m != M critical} c(m, M) a_m^dag a_M,
// with c(m, M) = Cm s(min, max) for each pair (m < M), s(a, b) = frac(sqrt2 a^3 + sqrt7 b^5) and
// s - 1 when that is below 1/2. The start state has N0 particles in mode 0 and one particle in
// each of the first min(K, Nm) critical modes. Observables: the occupation of every mode.
//
// Options (defaults of the reference's example): --N0 20 --Nm 2 --K 4 --C0 1 --Cm 1 --maxT 10
// --samplingStep 0.01 --tol 1e-8 --m 40 --DeltaN 12 --capacity 1 --fastIntegration 0
// --out <file name prefix, default "ResultBlackHole">.
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "Basis.h"
#include "hamiltonian.h"
#include "krylovHelper.h"
#include "krylovObservables.h"
#include "krylovTimeEvolver.h"
#include "matrixDataTypes.h"
#include "parameter.h"
#include "version.h"

namespace {

// Symmetry-breaking weight of the critical-mode pair (a, b).
double symBreak(int a, int b)
{
    double c = std::fmod(std::sqrt(2.0) * std::pow(a, 3) + std::sqrt(7.0) * std::pow(b, 5), 1.0);
    return c < 0.5 ? c - 1.0 : c;
}

class blackHoleHamiltonian : public Hamiltonian
{
public:
    blackHoleHamiltonian(int N0, int Nm, double DeltaN, int K, int capacity, double C0, double Cm)
        : N0_(N0), Nm_(Nm), K_(K), C_(capacity), DeltaN_(DeltaN), C0_(C0), Cm_(Cm) {}

    void createSimplifiedHamiltonian()
    {
        const double gap = std::sqrt(double(N0_));
        std::vector<opTerm>& h = hamiltonOperator;
        h.clear();
        for (int i = 2; i < 2 + 2 * K_; i++) h.push_back(createNumberOperator(i, gap));
        h.push_back(linInteraction(0, 1, 0, 0, true, C0_));
        h.push_back(linInteraction(0, 1, 0, 0, false, C0_));
        for (int i = 2; i < 2 + K_; i++) h.push_back(createFourPointA(0, 0, i, i, 0, 0, C_, C_, -gap / N0_));
        for (int i = 2 + K_; i < 2 + 2 * K_; i++)
            h.push_back(createFourPointA(0, 0, i, i, 0, 0, C_, C_, -gap / (N0_ - DeltaN_)));
        // hopping inside the critical sector: across the two halves, then within each half
        for (int m = 2; m < 2 + 2 * K_; m++) {
            for (int M = 2; M < 2 + 2 * K_; M++) {
                const bool mLow = m <= 1 + K_, MLow = M <= 1 + K_;
                double c = 0.0;
                if (mLow && !MLow) c = symBreak(m, M) * Cm_;
                else if (mLow == MLow && m < M) c = symBreak(m, M) * Cm_;
                if (c == 0.0) continue;
                h.push_back(hop(c, m, M));
                h.push_back(hop(c, M, m));
            }
        }
    }

    basisVector createInitState() const
    {
        basisVector v(2 + 2 * K_);
        v.e[0] = N0_;
        for (int i = 0; i < std::min(K_, Nm_); i++) v.e[2 + i] = 1;
        return v;
    }

private:
    opTerm hop(double c, int to, int from) const
    {
        opTerm t;
        t.coef = c;
        t.operations = {-1, 1};
        t.mode = {to, from};
        t.capacity = {C_, C_};
        return t;
    }

    int N0_, Nm_, K_, C_;
    double DeltaN_, C0_, Cm_;
};

}  // namespace

int main(int argc, char** argv)
{
    std::map<std::string, std::string> opt = {
        {"N0", "20"},   {"Nm", "2"},       {"K", "4"},         {"C0", "1"},         {"Cm", "1"},
        {"maxT", "10"}, {"samplingStep", "0.01"}, {"tol", "1e-8"}, {"m", "40"},  {"DeltaN", "12"},
        {"capacity", "1"}, {"fastIntegration", "0"}, {"out", "ResultBlackHole"}};
    for (int i = 1; i < argc; i += 2) {
        if (std::strncmp(argv[i], "--", 2) != 0 || !opt.count(argv[i] + 2) || i + 1 >= argc) {
            std::cerr << "usage: blackhole [--option value]...; options:";
            for (auto& o : opt) std::cerr << " --" << o.first;
            std::cerr << std::endl;
            return 1;
        }
        opt[argv[i] + 2] = argv[i + 1];
    }
    // not const: paraPush takes the parameter's type from the variable (int, double, bool)
    int N0 = std::stoi(opt["N0"]), Nm = std::stoi(opt["Nm"]), K = std::stoi(opt["K"]);
    double C0 = std::stod(opt["C0"]), Cm = std::stod(opt["Cm"]);
    double maxT = std::stod(opt["maxT"]), samplingStep = std::stod(opt["samplingStep"]);
    double tol = std::stod(opt["tol"]), DeltaN = std::stod(opt["DeltaN"]);
    int m = std::stoi(opt["m"]), capacity = std::stoi(opt["capacity"]);
    bool fastIntegration = std::stoi(opt["fastIntegration"]) != 0;

    tensorBasis basis(N0, 2, Nm, 2 * K, capacity);
    std::cout << "Created basis with " << basis.numberElements << " elements ..." << std::endl;

    blackHoleHamiltonian ham(N0, Nm, DeltaN, K, capacity, C0, Cm);
    ham.createSimplifiedHamiltonian();
    std::unique_ptr<smatrix> hamMatrix = ham.createHamiltonMatrix(&basis);
    std::vector<std::unique_ptr<smatrix>> observables = ham.createNumberOperatorObservables(&basis);

    std::vector<std::complex<double>> vec(basis.numberElements);
    vec[basis.hashTable.find(ham.createInitState())->second] = 1.0;

    std::vector<std::unique_ptr<krylovBasicObservable>> observableList;
    for (int i = 0; i != basis.numberModes; i++)
        observableList.push_back(
            std::make_unique<krylovSpMatrixObservable>("mode" + std::to_string(i), std::move(observables[i])));

    const auto begin = std::chrono::high_resolution_clock::now();
    krylovTimeEvolver timeEvolver(maxT, vec.data(), samplingStep, std::move(observableList), std::move(hamMatrix), 1.0,
                                  tol, m, fastIntegration, false);
    krylovReturn* results = timeEvolver.timeEvolve();
    const double elapsed =
        std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - begin).count();
    std::cout << "Time: " << elapsed << " sec" << std::endl;
    std::cout << "Number of steps: " << results->n_steps << "    error: " << results->err
              << "    status: " << results->statusCode << std::endl;

    parameter_list parameters;
    parameters.push_back(paraPush("N", true, N0));
    parameters.push_back(paraPush("Nm", true, Nm));
    parameters.push_back(paraPush("K", true, K));
    parameters.push_back(paraPush("C", true, capacity));
    parameters.push_back(paraPush("DeltaN", true, DeltaN));
    parameters.push_back(paraPush("C0", true, C0));
    parameters.push_back(paraPush("Cm", true, Cm));
    parameters.push_back(paraPush("maxT", true, maxT));
    parameters.push_back(paraPush("tol", true, tol));
    parameters.push_back(paraPush("samplingStep", true, samplingStep));
    parameters.push_back(paraPush("m", true, m));
    parameters.push_back(paraPush("fastIntegration", true, fastIntegration));
    parameters.push_back(paraPush("TE_MAJOR_VERSION", false, TIMEEVOLVER_VERSION / 100));
    parameters.push_back(paraPush("TE_MINOR_VERSION", false, TIMEEVOLVER_VERSION % 100));

    saveResult(results->observableList, parameters, opt["out"]);
    std::cout << "Results have been saved to file." << std::endl;
    delete results;
    return 0;
}
