#include "matrixDataTypes.h"

#include <algorithm>
#include <cmath>

#include "runtime.h"

using namespace TE;

constexpr std::complex<double> smatrix::one;
constexpr std::complex<double> smatrix::zero;

// ---- dense matrix (core/matrixDataTypes.cpp:10-42) ----
matrix::matrix(size_t nn, size_t mm) {
    n = nn;
    m = mm;
    numValues = n * m;
    values = (n * m > 0) ? new std::complex<double>[n * m] : nullptr;
}

matrix::matrix(size_t nn, size_t mm, std::complex<double>* vals) : n(nn), m(mm) {
    numValues = n * m;
    values = new std::complex<double>[numValues];
    std::copy(vals, vals + numValues, values);
}

matrix::~matrix() {
    if (n * m > 0) delete[] values;
}

// ---- sparse matrix (core/matrixDataTypes.cpp:47-261) ----
smatrix::smatrix() {
    sym = hermitian = upperTri = initialized = false;
    rowIndex = columns = nullptr;
    values = nullptr;
    m = n = 0;
    numValues = 0;
}

smatrix::smatrix(std::complex<double>* val, size_t* col, size_t* row, size_t nbV, unsigned int nn, unsigned int mm) {
    if (nn == 0 || mm == 0) {
        std::cerr << "Empty matrices are not supported." << std::endl;
        exit(1);
    }
    numValues = nbV;
    n = nn;
    m = mm;
    sym = hermitian = upperTri = false;
    columns = new size_t[nbV];
    rowIndex = new size_t[nbV];
    values = new std::complex<double>[nbV];
    initialized = false;
    std::copy(val, val + nbV, values);
    std::copy(col, col + nbV, columns);
    std::copy(row, row + nbV, rowIndex);
}

smatrix::smatrix(const smatrix& o) {
    numValues = o.numValues;
    n = o.n;
    m = o.m;
    sym = o.sym;
    hermitian = o.hermitian;
    upperTri = o.upperTri;
    columns = new size_t[numValues];
    rowIndex = new size_t[numValues];
    values = new std::complex<double>[numValues];
    initialized = false;
    std::copy(o.values, o.values + numValues, values);
    std::copy(o.columns, o.columns + numValues, columns);
    std::copy(o.rowIndex, o.rowIndex + numValues, rowIndex);
}

smatrix::~smatrix() {
    if (n > 1 || m > 1) {
        delete[] rowIndex;
        delete[] values;
        delete[] columns;
    }
    if (dev) te_matrix_destroy(dev);
}

// Maximal absolute column sum, accumulated in COO order (core/matrixDataTypes.cpp:67-79).
double smatrix::norm1() {
    if (!values && dev) {
        double v = 0.0;
        tecpp::check(te_matrix_norm1(dev, &v), "norm1");
        return v;
    }
    std::vector<double> colVal(n);
    for (size_t i = 0; i != numValues; i++) colVal[columns[i]] += std::abs(values[i]);
    return colVal.empty() ? 0.0 : *std::max_element(colVal.begin(), colVal.end());
}

// Maximal absolute row sum (core/matrixDataTypes.cpp:85-97).
double smatrix::normInf() {
    std::vector<double> rowVal(m);
    for (size_t i = 0; i != numValues; i++) rowVal[rowIndex[i]] += std::abs(values[i]);
    return rowVal.empty() ? 0.0 : *std::max_element(rowVal.begin(), rowVal.end());
}

// Uploads the COO arrays as a device CSR (replaces the MKL set-up of :196-240).
int smatrix::initialize() {
    if (initialized) return -1;
    if (!dev) {
        int rc = te_matrix_from_coo(tecpp::ctx(), reinterpret_cast<const double*>(values), columns, rowIndex,
                                    numValues, m, n, &dev);
        if (rc != TE_OK) {
            std::cerr << "Problem with device sparse matrix creation: " << te_last_error(tecpp::ctx()) << std::endl;
            return -1;
        }
    }
    initialized = true;
    return 0;
}

te_matrix* smatrix::device() {
    if (!dev) initialize();
    return dev;
}

void smatrix::adoptDevice(te_matrix* A) {
    if (dev && dev != A) te_matrix_destroy(dev);
    dev = A;
    initialized = true;
}

// out = alpha * this * in on the device (core/matrixDataTypes.cpp:169-190).
int smatrix::spMV(std::complex<double> alpha, std::complex<double>* in, std::complex<double>* out) {
    te_matrix* A = device();
    if (!A) return 1;
    return te_spmv(A, alpha.real(), alpha.imag(), reinterpret_cast<const double*>(in), reinterpret_cast<double*>(out));
}
