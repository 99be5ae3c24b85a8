#pragma once
// Dense and sparse matrix types of the TimeEvolver API (core/matrixDataTypes.h:18-99).
// smatrix keeps the reference's public COO arrays on the host; initialize() attaches a device
// CSR copy (te_matrix) that spMV, norm1 and the time evolver use.
#include <complex>
#include <cstddef>
#include <iostream>
#include <string>
#include <vector>

#include "mathHeader.h"

struct te_matrix;

namespace TE {

class matrix
{
public:
    size_t n, m;
    std::complex<double>* values;
    size_t numValues;
    matrix(size_t nn, size_t mm);
    matrix(size_t nn, size_t mm, std::complex<double>* vals);
    ~matrix();
};

class vector
{
public:
    std::complex<double>* values;
    size_t length;

    vector(unsigned int n)
    {
        if (n == 0) {
            std::cerr << "Empty vectors are not supported." << std::endl;
            exit(1);
        }
        length = n;
        values = new std::complex<double>[length];
    }
    ~vector() { delete[] values; }
    operator std::complex<double>*() const { return NULL; }
};

class smatrix
{
public:
    std::complex<double>* values;
    size_t* columns;
    size_t* rowIndex;
    size_t numValues;
    size_t n, m;
    bool sym, hermitian;
    bool upperTri;
    bool initialized;

    double norm1();
    double normInf();

    smatrix();
    smatrix(std::complex<double>* val, size_t* col, size_t* row, size_t nbV, unsigned int nn, unsigned int mm);
    smatrix(const smatrix& old_obj);
    ~smatrix();

    int spMV(std::complex<double> alpha, std::complex<double>* in, std::complex<double>* out);
    int initialize();

    // Device copy (created by initialize(), or adopted from the device builder).
    te_matrix* device();
    void adoptDevice(te_matrix* A);

    static constexpr std::complex<double> one = std::complex<double>(1.0, 0.0);
    static constexpr std::complex<double> zero = std::complex<double>(0.0, 0.0);

private:
    te_matrix* dev = nullptr;
};

}  // namespace TE

// Element-wise helpers of core/matrixDataTypes.h:105-130.
inline void expV(size_t len, std::complex<double>* x, std::complex<double>* y)
{
    for (size_t i = 0; i < len; i++) y[i] = std::exp(x[i]);
}

inline void mulV(size_t len, std::complex<double>* a, std::complex<double>* b, std::complex<double>* y)
{
    for (size_t i = 0; i < len; i++) y[i] = b[i] * a[i];
}
