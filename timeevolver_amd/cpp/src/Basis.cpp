#include "Basis.h"

#include <cmath>

#include "runtime.h"

// basis / tensorBasis (helper/Basis.cpp:5-191), built on the device; the host copies of the
// elements and the hash table are the public members the reference exposes.

void basicBasis::loadFromDevice() {
    size_t D = 0;
    int K = 0;
    tecpp::check(te_basis_info(dev, &D, &K), "basis");
    numberElements = (int)D;
    numberModes = K;
    std::vector<int> occ(D * (size_t)K);
    if (D) tecpp::check(te_basis_states(dev, occ.data()), "basis");
    basisElements = new basisVector[D ? D : 1];
    for (size_t i = 0; i < D; i++) {
        basisElements[i].zeros(K);
        basisElements[i].initInt(&occ[i * K], K);
    }
}

basis::basis(int N, int K, int nbQudits, int capacity) : basis(N, K, nbQudits, capacity, 1) {}

basis::basis(int N, int K, int nbQudits, int capacity, int quant) {
    tecpp::check(te_basis_create(tecpp::ctx(), N, K, nbQudits, capacity, quant, &dev), "basis");
    loadFromDevice();
    numberModes = K;
    qudits = capacity;
    createHashTable();
}

basis::~basis() {
    delete[] basisElements;
    if (dev) te_basis_destroy(dev);
}

void basis::createHashTable() {
    for (int i = 0; i != numberElements; i++) hashTable.insert(std::make_pair(basisElements[i], i));
}

int basis::nchoosekSmart(int n, int k) {
    if (k > n) return 1;
    double r = 1.0;
    for (int i = 1; i <= k; i++) r = r * (n - k + i) / i;
    return (int)std::llround(r);
}

basis& basis::operator=(const basis& rhs) {
    if (numberElements > 0) delete[] basisElements;
    numberElements = rhs.numberElements;
    numberModes = rhs.numberModes;
    qudits = rhs.qudits;
    basisElements = new basisVector[numberElements];
    for (int i = 0; i != numberElements; i++) basisElements[i] = rhs.basisElements[i];
    return *this;
}

tensorBasis::tensorBasis(int N1, int K1, int N2, int K2, int C2) {
    tecpp::check(te_tensor_basis_create(tecpp::ctx(), N1, K1, N2, K2, C2, &dev), "tensorBasis");
    loadFromDevice();
    numberModes = K1 + K2;
    qudits = C2;
    if (numberElements > 0) createHashTable();
}

tensorBasis::~tensorBasis() {
    if (numberElements != 0) delete[] basisElements;
    if (dev) te_basis_destroy(dev);
}

void tensorBasis::createHashTable() {
    for (int i = 0; i != numberElements; i++) hashTable.insert(std::make_pair(basisElements[i], i));
}
