#pragma once
// Occupation-number basis of the TimeEvolver helper API (helper/Basis.h:15-357). The basis is
// built on the device (te_basis_create); basisElements and hashTable are the host copies the
// reference exposes, and the device handle feeds Hamiltonian::createMatrix.
#include <complex>
#include <cstdint>
#include <iostream>
#include <unordered_map>
#include <vector>

struct te_basis;

struct basisVector
{
    int* e;
    size_t length;

    basisVector() { length = 0; e = nullptr; }
    basisVector(unsigned int le)
    {
        if (le == 0) { e = nullptr; length = 0; return; }
        length = le;
        e = new int[le]();
    }
    basisVector(int le, int pos)
    {
        if (le <= pos) { std::cerr << "basisvector too small" << std::endl; exit(12); }
        if (le > 0) { length = le; e = new int[le](); e[pos] = 1; }
        else { length = 0; e = nullptr; }
    }
    void ones() { for (unsigned int i = 0; i < length; i++) e[i] = 1; }
    void zeros() { for (unsigned int i = 0; i < length; i++) e[i] = 0; }
    void zeros(int n)
    {
        if (n > 0) {
            if (length != 0) delete[] e;
            length = n;
            e = new int[n];
        }
        zeros();
    }
    void initInt(int* init, unsigned int le)
    {
        if (length == le)
            for (unsigned int i = 0; i != le; i++) e[i] = init[i];
    }
    int occuNumber()
    {
        int ret = 0;
        for (unsigned int i = 0; i < length; i++) ret += e[i];
        return ret;
    }
    basisVector(const basisVector& o)
    {
        if (o.length > 0) {
            length = o.length;
            e = new int[length];
            for (unsigned int i = 0; i != length; i++) e[i] = o.e[i];
        } else { e = nullptr; length = 0; }
    }
    ~basisVector() { if (length > 0) delete[] e; }
    bool operator==(const basisVector& b) const
    {
        if (b.length != length) return false;
        for (unsigned int i = 0; i != length; i++)
            if (e[i] != b.e[i]) return false;
        return true;
    }
    basisVector& operator=(const basisVector& rhs)
    {
        if (this == &rhs) return *this;
        if (rhs.length != length) {
            if (e != nullptr) delete[] e;
            e = rhs.length ? new int[rhs.length] : nullptr;
            length = rhs.length;
        }
        for (unsigned int i = 0; i != length; i++) e[i] = rhs.e[i];
        return *this;
    }
};

// Hash of the occupation vector (exact-match lookups only, helper/hamiltonian.cpp:363).
struct basisVectorHasher
{
    std::size_t operator()(const basisVector& b) const
    {
        uint64_t h = 1469598103934665603ull;
        for (size_t i = 0; i < b.length; i++) { h ^= (uint32_t)b.e[i]; h *= 1099511628211ull; }
        return (std::size_t)h;
    }
};

struct basisState
{
public:
    basisVector b;
    std::complex<double> coef;
    basisState(const basisState& rhs) { b = rhs.b; coef = rhs.coef; }
    basisState() { b = basisVector(); coef = 0; }
    basisState(basisVector& rhs, std::complex<double>& c) { b = basisVector(rhs); coef = c; }
};

class basicBasis
{
public:
    int numberModes;
    int qudits;
    int numberElements;
    basisVector* basisElements;
    std::unordered_map<basisVector, int, basisVectorHasher> hashTable;

    virtual void createHashTable() = 0;
    basicBasis() {}
    virtual ~basicBasis() {}

    // Device copy used by Hamiltonian::createMatrix.
    te_basis* device() const { return dev; }

protected:
    te_basis* dev = nullptr;
    void loadFromDevice();
};

class basis : public basicBasis
{
public:
    basis(int N, int K, int nbQudits, int capacity, int quant);
    basis(int N, int K, int nbQudits, int capacity);
    int nchoosekSmart(int n, int k);
    basis& operator=(const basis& rhs);
    ~basis();
    void createHashTable();
};

class tensorBasis : public basicBasis
{
public:
    tensorBasis(int N1, int K1, int N2, int K2, int C2);
    tensorBasis() { numberElements = 0; numberModes = 0; qudits = -1; basisElements = nullptr; }
    void createHashTable();
    ~tensorBasis();
};
