#include "krylovObservables.h"

#include <cstdlib>

// Result-array bookkeeping of core/krylovObservables.cpp:11-80 and the direct (host) evaluation
// of each observable type (:93-259). Inside krylovTimeEvolver::timeEvolve the sparse and vector
// observables are evaluated on the device instead.

constexpr std::complex<double> krylovBasicObservable::one;
constexpr std::complex<double> krylovBasicObservable::zero;

std::string krylovBasicObservable::retName() { return obs_name; }
double* krylovBasicObservable::retExpectationValues() { return expectationValues; }
size_t krylovBasicObservable::retNumSamples() { return numSamples; }
obsType krylovBasicObservable::retType() { return type; }

void krylovBasicObservable::initializeResultArray(size_t size) {
    numSamples = size;
    expectationValues = new double[size];
}

size_t krylovBasicObservable::resetResultArray() {
    sampleIndex = 0;
    return numSamples;
}

krylovBasicObservable::krylovBasicObservable(const std::string& name, std::vector<double> values)
    : obs_name(name), dim(0), numSamples(values.size()), sampleIndex(values.size()), type(VOID_TYPE_OBS),
      expectationValues(nullptr) {
    expectationValues = new double[numSamples];
    for (size_t i = 0; i != numSamples; i++) expectationValues[i] = values[i];
}

krylovBasicObservable::~krylovBasicObservable() {
    if (numSamples > 0) delete[] expectationValues;
}

// ---- dense matrix observable (private constructor, as in the reference) ----
krylovMatrixObservable::krylovMatrixObservable(const std::string& name, std::unique_ptr<matrix> obser)
    : krylovBasicObservable(name) {
    dim = obser->m;
    type = MATRIX_TYPE_OBS;
    tmpBlasVec = dim > 0 ? new std::complex<double>[dim] : nullptr;
    obs = std::move(obser);
}

krylovMatrixObservable::~krylovMatrixObservable() {
    if (dim > 0) delete[] tmpBlasVec;
}

std::complex<double> krylovMatrixObservable::expectation(std::complex<double>* vec, int len) {
    if (sampleIndex >= numSamples) { std::cerr << "Too many samples." << std::endl; exit(1); }
    if ((size_t)len != dim) { std::cerr << "Incompatible dimensions" << std::endl; exit(1); }
    std::complex<double> s(0.0, 0.0);
    for (size_t r = 0; r < dim; r++) {
        std::complex<double> t(0.0, 0.0);
        for (size_t c = 0; c < dim; c++) t += obs->values[c * dim + r] * vec[c];
        s += std::conj(vec[r]) * t;
    }
    expectationValues[sampleIndex++] = s.real();
    return s;
}

// ---- sparse matrix observable ----
krylovSpMatrixObservable::krylovSpMatrixObservable(const std::string& name, std::unique_ptr<smatrix> obser)
    : krylovBasicObservable(name) {
    dim = obser->m;
    type = SPARSE_MATRIX_TYPE_OBS;
    obs = std::move(obser);
    obs->initialize();
    tmpBlasVec = new std::complex<double>[dim];
}

krylovSpMatrixObservable::~krylovSpMatrixObservable() {
    if (dim > 0) delete[] tmpBlasVec;
}

std::complex<double> krylovSpMatrixObservable::expectation(std::complex<double>* vec, int len) {
    if (sampleIndex >= numSamples) { std::cerr << "Too many samples." << std::endl; exit(1); }
    if ((size_t)len != dim) { std::cerr << "Incompatible dimensions" << std::endl; exit(1); }
    obs->spMV(one, vec, tmpBlasVec);
    std::complex<double> s(0.0, 0.0);
    for (int i = 0; i < len; i++) s += std::conj(vec[i]) * tmpBlasVec[i];
    expectationValues[sampleIndex++] = s.real();
    return s;
}

// ---- vector observable ----
krylovVectorObservable::krylovVectorObservable(const std::string& name, std::complex<double>* obser, size_t len)
    : krylovBasicObservable(name) {
    dim = len;
    type = VECTOR_TYPE_OBS;
    obs = std::make_unique<std::complex<double>[]>(len);
    for (size_t i = 0; i < len; i++) obs[i] = obser[i];
}

std::complex<double> krylovVectorObservable::expectation(std::complex<double>* vec, int len) {
    if (sampleIndex >= numSamples) { std::cerr << "Too many samples." << std::endl; exit(1); }
    if ((size_t)len != dim) { std::cerr << "Incompatible dimensions" << std::endl; exit(1); }
    std::complex<double> s(0.0, 0.0);
    for (int i = 0; i < len; i++) s += std::conj(vec[i]) * obs[i];
    std::complex<double> r(std::norm(s), 0.0);
    expectationValues[sampleIndex++] = r.real();
    return r;
}

// ---- output-only observable ----
krylovOutputObservable::krylovOutputObservable(const std::string& name, std::vector<double> values)
    : krylovBasicObservable(name) {
    this->numSamples = values.size();
    initializeResultArray(numSamples);
    for (size_t i = 0; i != numSamples; i++) expectationValues[i] = values[i];
}

std::complex<double> krylovOutputObservable::expectation(std::complex<double>*, int) {
    throw std::runtime_error("krylovOutputObservable should not call `expectation` function.");
}

const char* requestStopException::what() const throw() {
    return "Observable requested termination of time evolution.";
}
