#pragma once
// The time evolver of the TimeEvolver API (core/krylovTimeEvolver.h:32-133), running on the
// MI355X through the C-ABI of include/timeevolver.h.
#include <atomic>
#include <complex>
#include <limits>
#include <memory>
#include <thread>
#include <vector>

#include "krylovLogger.h"
#include "krylovObservables.h"
#include "mathHeader.h"
#include "matrixDataTypes.h"
#include "version.h"

struct te_evolver;

class krylovReturn
{
public:
    std::complex<double>* evolvedState;
    double err; double evolvedTime;
    size_t numSamples;
    size_t n_steps;
    size_t dim;
    size_t krylovDim;
    // 0 success, 1 lucky breakdown, 2 analytic error below the round-off estimate, 3 stopped by an
    // observable; 10 round-off may spoil the bound, 11 tolerance unreachable, 20 integration
    // inaccurate, 30 norm drift, 100 several of these (core/krylovTimeEvolver.h:41-45).
    int statusCode;
    std::vector<std::unique_ptr<krylovBasicObservable>> observableList;
    std::unique_ptr<smatrix> hamiltonianMatrix;

    krylovReturn(unsigned int Hsize, int status);
    ~krylovReturn();
};

class krylovTimeEvolver
{
public:
    krylovTimeEvolver(double t, std::complex<double>* v, double samplingStep, std::vector<std::unique_ptr<krylovBasicObservable>> observables, std::unique_ptr<smatrix> Ham, double expFactor, double tol, int mm, bool fastIntegration, bool progressBar);
    krylovTimeEvolver(double t, std::complex<double>* v, double samplingStep, std::vector<std::unique_ptr<krylovBasicObservable>> observables, std::unique_ptr<smatrix> Ham);
    krylovReturn* timeEvolve();
    ~krylovTimeEvolver();

    bool fastIntegration, progressBar;
    std::complex<double> expFactor;
    double tol; size_t m;

    void changeLogLevel(krylovLogger::loggingLevel level);

protected:
    double t; size_t Hsize;
    double samplingStep;
    int nbObservables;
    std::unique_ptr<smatrix> Ham;
    std::vector<std::unique_ptr<krylovBasicObservable>> obsVector;
    krylovLogger logger;
    size_t n_samples;
    std::vector<std::complex<double>> initialState;
    std::atomic<bool> stop_printing;
};
