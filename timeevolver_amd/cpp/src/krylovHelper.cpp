// Result and matrix output (helper/krylovHelper.cpp:10-106 saveResult, :109-170 saveSparseMatrix,
// :173-207 saveMatrix). The HDF5 path uses the library's C API, so the layer links against
// libhdf5 alone; it is compiled in when the build defines USE_HDF.
#include "krylovHelper.h"

#include <cstdint>
#include <fstream>
#include <stdexcept>

#ifdef USE_HDF
#include <hdf5.h>
#endif

namespace {

std::string resultFileStem(parameter_list& para, const std::string& name)
{
    std::string stem = name;
    for (auto& p : para)
        if (p->getPrintFilename()) stem += "_" + p->getName() + p->getData();
    return stem;
}

#ifdef USE_HDF
// Owns one HDF5 identifier; close_ is the matching H5?close.
class H5Id
{
public:
    H5Id(hid_t id, herr_t (*close)(hid_t), const std::string& what) : id_(id), close_(close)
    {
        if (id_ < 0) throw std::runtime_error("HDF5: cannot create " + what);
    }
    ~H5Id() { close_(id_); }
    H5Id(const H5Id&) = delete;
    H5Id& operator=(const H5Id&) = delete;
    operator hid_t() const { return id_; }

private:
    hid_t id_;
    herr_t (*close_)(hid_t);
};

H5Id createFile(const std::string& path)
{
    return H5Id(H5Fcreate(path.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose, "file " + path);
}

// A 1-D dataset of n elements stored as file_type, written from memory of mem_type.
void writeVector(hid_t file, const char* dsname, hid_t file_type, hid_t mem_type, const void* data, hsize_t n)
{
    H5Id space(H5Screate_simple(1, &n, nullptr), H5Sclose, "dataspace");
    H5Id ds(H5Dcreate2(file, dsname, file_type, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose,
            std::string("dataset ") + dsname);
    if (n && H5Dwrite(ds, mem_type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)
        throw std::runtime_error(std::string("HDF5: cannot write dataset ") + dsname);
}

// One-element attribute of the dataset (a rank-1 dataspace of size 1, as the reference writes).
void writeAttribute(hid_t ds, const std::string& aname, hid_t type, const void* value)
{
    const hsize_t one = 1;
    H5Id space(H5Screate_simple(1, &one, nullptr), H5Sclose, "dataspace");
    H5Id at(H5Acreate2(ds, aname.c_str(), type, space, H5P_DEFAULT, H5P_DEFAULT), H5Aclose, "attribute " + aname);
    if (H5Awrite(at, type, value) < 0) throw std::runtime_error("HDF5: cannot write attribute " + aname);
}

void splitComplex(const std::complex<double>* v, size_t n, std::vector<double>& re, std::vector<double>& im)
{
    re.resize(n);
    im.resize(n);
    for (size_t i = 0; i < n; i++) {
        re[i] = v[i].real();
        im[i] = v[i].imag();
    }
}
#endif

}  // namespace

void saveResult(const std::vector<std::unique_ptr<krylovBasicObservable>>& obs_list, parameter_list& para,
                const std::string& name)
{
    const std::string stem = resultFileStem(para, name);
#ifdef USE_HDF
    H5Id file = createFile(stem + ".h5");
    for (const auto& obs : obs_list) {
        const hsize_t n = obs->retNumSamples();
        const std::string dsname = obs->retName();
        H5Id space(H5Screate_simple(1, &n, nullptr), H5Sclose, "dataspace");
        H5Id ds(H5Dcreate2(file, dsname.c_str(), H5T_IEEE_F64LE, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                H5Dclose, "dataset " + dsname);
        if (n && H5Dwrite(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, obs->retExpectationValues()) < 0)
            throw std::runtime_error("HDF5: cannot write dataset " + dsname);
        for (auto& p : para) {
            if (p->isDouble()) {
                const double v = dynamic_cast<typedParameter<double>&>(*p).getValue();
                writeAttribute(ds, p->getName(), H5T_NATIVE_DOUBLE, &v);
            } else if (p->isInt()) {
                const int v = dynamic_cast<typedParameter<int>&>(*p).getValue();
                writeAttribute(ds, p->getName(), H5T_NATIVE_INT, &v);
            } else if (p->isBool()) {
                const int v = dynamic_cast<typedParameter<bool>&>(*p).getValue();
                writeAttribute(ds, p->getName(), H5T_NATIVE_INT, &v);
            }
        }
    }
#else
    for (const auto& obs : obs_list) {
        const std::string path = stem + obs->retName() + ".csv";
        std::ofstream out(path);
        if (!out) throw std::runtime_error("cannot open " + path);
        const double* v = obs->retExpectationValues();
        const size_t n = obs->retNumSamples();
        for (size_t i = 0; i < n; i++) out << (i ? ", " : "") << v[i];
    }
#endif
}

void saveMatrix(const matrix* mat, const std::string& name)
{
#ifdef USE_HDF
    std::vector<double> re, im;
    splitComplex(mat->values, mat->numValues, re, im);
    H5Id file = createFile(name);
    writeVector(file, "valuesRealPart", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, re.data(), re.size());
    writeVector(file, "valuesImagPart", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, im.data(), im.size());
#else
    (void)mat;
    throw std::runtime_error("saveMatrix(" + name + "): this build of the library has no HDF5 support");
#endif
}

void saveSparseMatrix(const smatrix* mat, const std::string& name)
{
#ifdef USE_HDF
    static_assert(sizeof(size_t) == sizeof(uint64_t), "index datasets are written as 64-bit");
    std::vector<double> re, im;
    splitComplex(mat->values, mat->numValues, re, im);
    H5Id file = createFile(name);
    const hsize_t n = mat->numValues;
    writeVector(file, "valuesRealPart", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, re.data(), n);
    writeVector(file, "valuesImagPart", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, im.data(), n);
    writeVector(file, "columIndex", H5T_STD_U64LE, H5T_NATIVE_UINT64, mat->columns, n);
    writeVector(file, "rowIndex", H5T_STD_U64LE, H5T_NATIVE_UINT64, mat->rowIndex, n);
    const uint64_t dim = mat->m;
    writeVector(file, "dimension", H5T_STD_U64LE, H5T_NATIVE_UINT64, &dim, 1);
#else
    (void)mat;
    throw std::runtime_error("saveSparseMatrix(" + name + "): this build of the library has no HDF5 support");
#endif
}
