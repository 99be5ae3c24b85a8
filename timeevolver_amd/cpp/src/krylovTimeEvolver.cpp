#include "krylovTimeEvolver.h"

#include <cmath>
#include <cstring>

#include "runtime.h"

// The reference's constructor and timeEvolve (core/krylovTimeEvolver.cpp:26-84, :261-512), with
// the whole evolution delegated to te_evolver_run on the device.

krylovTimeEvolver::krylovTimeEvolver(double t, std::complex<double>* v, double samplingStep,
                                     std::vector<std::unique_ptr<krylovBasicObservable>> observables,
                                     std::unique_ptr<smatrix> Ham, double expFactor, double tol, int mm,
                                     bool fastIntegration, bool progressBar) {
    this->t = t;
    this->samplingStep = samplingStep;
    this->tol = tol;
    this->progressBar = progressBar;
    this->Ham = std::move(Ham);
    this->expFactor = std::complex<double>(0, -1 * expFactor);
    this->fastIntegration = fastIntegration;
    this->nbObservables = (int)observables.size();
    this->Hsize = this->Ham->m;
    this->m = std::min<size_t>(mm, Hsize);
    stop_printing = false;
    if (Hsize == 0) {
        logger.log_message(krylovLogger::FATAL, "Invalid Hilbertspace dimension");
        exit(1);
    }
    double nrm = 0.0;
    for (size_t i = 0; i < Hsize; i++) nrm += std::norm(v[i]);
    if (std::abs(std::sqrt(nrm) - 1.0) > tol) {
        logger.log_message(krylovLogger::FATAL, "Initial vector is not normalized");
        exit(1);
    }
    this->Ham->initialize();
    n_samples = (size_t)std::floor(t / samplingStep) + 1;
    initialState.assign(v, v + Hsize);
    obsVector = std::move(observables);
    for (auto& o : obsVector) o->initializeResultArray(n_samples);
}

krylovTimeEvolver::krylovTimeEvolver(double t, std::complex<double>* v, double samplingStep,
                                     std::vector<std::unique_ptr<krylovBasicObservable>> observables,
                                     std::unique_ptr<smatrix> Ham)
    : krylovTimeEvolver(t, v, samplingStep, std::move(observables), std::move(Ham), 1.0, 1e-6, 40, false, false) {}

krylovTimeEvolver::~krylovTimeEvolver() { obsVector.clear(); }

void krylovTimeEvolver::changeLogLevel(krylovLogger::loggingLevel level) { logger.set_loggingLevel(level); }

namespace {
// Host observables: the plugin's own expectation() records the value (krylovObservables.cpp);
// a requestStopException becomes a stop request of the device loop.
int host_observable(void* user, const double* state, size_t n, double* value) {
    krylovBasicObservable* o = static_cast<krylovBasicObservable*>(user);
    try {
        std::complex<double> r =
            o->expectation(reinterpret_cast<std::complex<double>*>(const_cast<double*>(state)), (int)n);
        *value = r.real();
        return 0;
    } catch (requestStopException& e) {
        return 1;
    }
}
}  // namespace

krylovReturn* krylovTimeEvolver::timeEvolve() {
    te_ctx* ctx = tecpp::ctx();
    std::vector<te_observable> obs(obsVector.size());
    std::vector<bool> on_device(obsVector.size(), true);
    for (size_t i = 0; i < obsVector.size(); i++) {
        krylovBasicObservable* o = obsVector[i].get();
        std::memset(&obs[i], 0, sizeof(te_observable));
        if (auto* sp = dynamic_cast<krylovSpMatrixObservable*>(o)) {
            obs[i].kind = TE_OBS_SPARSE;
            obs[i].matrix = sp->matrixPtr()->device();
        } else if (auto* vo = dynamic_cast<krylovVectorObservable*>(o)) {
            obs[i].kind = TE_OBS_VECTOR;
            obs[i].vector = reinterpret_cast<const double*>(vo->data());
        } else {
            obs[i].kind = TE_OBS_HOST;
            obs[i].callback = host_observable;
            obs[i].user = o;
            on_device[i] = false;
        }
    }
    te_evolve_params P;
    P.t = t;
    P.sampling_step = samplingStep;
    P.exp_factor = -expFactor.imag();
    P.tol = tol;
    P.m = (int)m;
    P.fast_integration = fastIntegration ? 1 : 0;
    te_evolver* ev = nullptr;
    int rc = te_evolver_create(ctx, Ham->device(), reinterpret_cast<const double*>(initialState.data()), &P,
                               obs.data(), (int)obs.size(), &ev);
    tecpp::check(rc, "time evolver");
    std::vector<double> values(obs.size() * n_samples);
    krylovReturn* ret = new krylovReturn((unsigned int)Hsize, 0);
    te_result R;
    rc = te_evolver_run(ev, values.data(), reinterpret_cast<double*>(ret->evolvedState), &R);
    std::vector<size_t> recorded(obs.size(), 0);
    if (rc == TE_OK && !obs.empty()) rc = te_evolver_obs_recorded(ev, recorded.data());
    te_evolver_destroy(ev);
    tecpp::check(rc, "timeEvolve");
    // values and sampleIndex of the device-evaluated observables: the samples each recorded
    // (observables listed before one that requested a stop also hold the stop sample)
    for (size_t i = 0; i < obsVector.size(); i++) {
        if (!on_device[i]) continue;
        krylovBasicObservable* o = obsVector[i].get();
        for (size_t s = 0; s < recorded[i] && s < o->numSamples; s++) o->expectationValues[s] = values[i * n_samples + s];
        o->sampleIndex = recorded[i];
    }
    if (progressBar) std::printf("[%s] 100 %%\n", std::string(70, '|').c_str());
    ret->statusCode = R.status_code;
    ret->n_steps = R.n_steps;
    ret->err = R.err;
    ret->dim = R.dim;
    ret->krylovDim = R.krylov_dim;
    ret->evolvedTime = R.evolved_time;
    ret->numSamples = R.num_samples;
    m = R.krylov_dim;
    ret->observableList = std::move(obsVector);
    ret->hamiltonianMatrix = std::move(Ham);
    return ret;
}

krylovReturn::krylovReturn(unsigned int Hsize, int status) {
    err = 0;
    n_steps = 0;
    krylovDim = 0;
    dim = Hsize;
    statusCode = status;
    evolvedTime = 0;
    numSamples = 0;
    evolvedState = new std::complex<double>[Hsize];
}

krylovReturn::~krylovReturn() { delete[] evolvedState; }
