// Drop-in check of the reference-compatible C++ API: two coupled bosonic modes,
//   H = w0 n_0 + w1 n_1 + g (a_0^dag a_1 + a_1^dag a_0),   200 particles,
// evolved to t = 10 with the default constructor, writing <n_k(t)> as comma-separated CSV
// files <prefix><k>.csv; then the same run with a host observable that requests a stop.
#include <cstdio>
#include <fstream>
#include <memory>
#include <string>
#include <vector>

#include "hamiltonian.h"
#include "krylovObservables.h"
#include "krylovTimeEvolver.h"
#include "matrixDataTypes.h"

namespace {

// Records <psi|psi> on the host and asks for a stop at its `limit`-th call.
class StopAt : public krylovBasicObservable {
public:
    StopAt(const std::string& name, int limit) : krylovBasicObservable(name), limit_(limit) {}
    std::complex<double> expectation(std::complex<double>* vec, int len) override {
        if (calls_++ == limit_) throw requestStopException();
        double s = 0.0;
        for (int i = 0; i < len; i++) s += std::norm(vec[i]);
        expectationValues[sampleIndex++] = s;
        return {s, 0.0};
    }

private:
    int limit_, calls_ = 0;
};

std::unique_ptr<smatrix> model(Hamiltonian& h, basis& b) {
    std::vector<opTerm> ops;
    ops.push_back(h.createNumberOperator(0, 1.0));
    ops.push_back(h.createNumberOperator(1, 2.0));
    ops.push_back(h.linInteraction(0, 1, 0, 0, false, 1.0));
    ops.push_back(h.linInteraction(0, 1, 0, 0, true, 1.0));
    h.hamiltonOperator = ops;
    return h.createHamiltonMatrix(&b);
}

}  // namespace

int main(int argc, char** argv) {
    const std::string prefix = argc > 1 ? argv[1] : "oscillators_n";
    const int particles = 200, modes = 2;
    basis b(particles, modes, 0, 0);
    Hamiltonian h;
    std::unique_ptr<smatrix> H = model(h, b);
    std::vector<std::unique_ptr<smatrix>> nops = h.createNumberOperatorObservables(&b);
    std::vector<std::unique_ptr<krylovBasicObservable>> list;
    for (int k = 0; k < modes; k++) list.push_back(std::make_unique<krylovSpMatrixObservable>(std::to_string(k), std::move(nops[k])));

    basisVector start(modes);
    start.e[0] = particles;
    std::vector<std::complex<double>> psi(b.numberElements);
    psi[b.hashTable.find(start)->second] = 1.0;

    krylovTimeEvolver ev(10.0, psi.data(), 0.01, std::move(list), std::move(H));
    krylovReturn* r = ev.timeEvolve();
    for (auto& o : r->observableList) {
        std::ofstream f(prefix + o->retName() + ".csv");
        const double* v = o->retExpectationValues();
        const size_t n = o->retNumSamples();
        for (size_t i = 0; i + 1 < n; i++) f << v[i] << ", ";
        f << v[n - 1];
    }
    std::printf("run: dim=%zu n_steps=%zu err=%.6e status=%d samples=%zu\n", r->dim, r->n_steps, r->err, r->statusCode,
                r->numSamples);
    delete r;

    // stop request from a host observable at its 6th sample (status 3, 5 samples)
    Hamiltonian h2;
    std::unique_ptr<smatrix> H2 = model(h2, b);
    std::vector<std::unique_ptr<krylovBasicObservable>> list2;
    list2.push_back(std::make_unique<StopAt>("norm", 5));
    krylovTimeEvolver ev2(10.0, psi.data(), 0.01, std::move(list2), std::move(H2), 1.0, 1e-8, 40, false, false);
    krylovReturn* r2 = ev2.timeEvolve();
    std::printf("stop: status=%d samples=%zu norm0=%.15f\n", r2->statusCode, r2->numSamples,
                r2->observableList[0]->retExpectationValues()[0]);
    delete r2;
    return 0;
}
