// Writes known data through saveResult, saveSparseMatrix and saveMatrix into the directory
// argv[1]; the Python side reads the files back. Host data only: no device call is made.
//   save_output <dir> generic
//       a few observables and parameters of every kind, a sparse and a dense matrix
//   save_output <dir> result <values file> N Nm K C DeltaN C0 Cm maxT tol samplingStep m fastIntegration
//       observables "mode<k>" holding line k of the values file, saved under the name
//       "ResultBlackHole" with the parameter list of the reference's example program
#include <complex>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "krylovHelper.h"
#include "krylovObservables.h"
#include "matrixDataTypes.h"
#include "parameter.h"

namespace {

int saveExampleResult(const std::string& dir, char** a)
{
    std::vector<std::unique_ptr<krylovBasicObservable>> obs;
    std::ifstream in(a[0]);
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::vector<double> v;
        for (double x; ls >> x;) v.push_back(x);
        obs.push_back(std::make_unique<krylovOutputObservable>("mode" + std::to_string(obs.size()), v));
    }
    int N = std::stoi(a[1]), Nm = std::stoi(a[2]), K = std::stoi(a[3]), C = std::stoi(a[4]);
    double DeltaN = std::stod(a[5]), C0 = std::stod(a[6]), Cm = std::stod(a[7]), maxT = std::stod(a[8]);
    double tol = std::stod(a[9]), samplingStep = std::stod(a[10]);
    int m = std::stoi(a[11]);
    bool fastIntegration = std::stoi(a[12]) != 0;
    parameter_list p;
    p.push_back(paraPush("N", true, N));
    p.push_back(paraPush("Nm", true, Nm));
    p.push_back(paraPush("K", true, K));
    p.push_back(paraPush("C", true, C));
    p.push_back(paraPush("DeltaN", true, DeltaN));
    p.push_back(paraPush("C0", true, C0));
    p.push_back(paraPush("Cm", true, Cm));
    p.push_back(paraPush("maxT", true, maxT));
    p.push_back(paraPush("tol", true, tol));
    p.push_back(paraPush("samplingStep", true, samplingStep));
    p.push_back(paraPush("m", true, m));
    p.push_back(paraPush("fastIntegration", true, fastIntegration));
    saveResult(obs, p, dir + "/ResultBlackHole");
    return 0;
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc < 3) return 2;
    const std::string dir = argv[1], mode = argv[2];
    if (mode == "result") return argc == 16 ? saveExampleResult(dir, argv + 3) : 2;
    if (mode != "generic") return 2;

    std::vector<std::unique_ptr<krylovBasicObservable>> obs;
    obs.push_back(std::make_unique<krylovOutputObservable>("alpha", std::vector<double>{1.0, -2.5, 1e-300, 0.1, 20.000000000000004}));
    obs.push_back(std::make_unique<krylovOutputObservable>("beta", std::vector<double>{3.0}));

    int N = 20;
    double tol = 1e-8, dt = 0.01, C0 = 1.0;
    bool fast = false;
    unsigned int u = 7;         // neither int, bool nor double: in the file name, no attribute
    std::string label = "x";
    parameter_list p;
    p.push_back(paraPush("N", true, N));
    p.push_back(paraPush("tol", true, tol));
    p.push_back(paraPush("C0", true, C0));
    p.push_back(paraPush("fast", true, fast));
    p.push_back(paraPush("dt", false, dt));
    p.push_back(paraPush("u", true, u));
    p.push_back(paraPush("label", true, label));
    p.push_back(paraPush("pi", true, 3.14159265358979));
    saveResult(obs, p, dir + "/Res");

    std::complex<double> v[4] = {{1, 2}, {3, -4}, {0.5, 0}, {0, -1e-3}};
    size_t col[4] = {0, 2, 1, 2}, row[4] = {0, 0, 2, 1};
    smatrix A(v, col, row, 4, 3, 3);
    matrix B(2, 2, v);
    try {
        saveSparseMatrix(&A, dir + "/sparse.h5");
        saveMatrix(&B, dir + "/dense.h5");
        std::printf("matrices: written\n");
    } catch (const std::runtime_error& e) {
        std::printf("matrices: %s\n", e.what());
    }
    return 0;
}
