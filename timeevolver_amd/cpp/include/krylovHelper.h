#pragma once
// Result and matrix output (helper/krylovHelper.h:18-27).
//
// saveResult names the output  name + "_" + paramName + paramText  for every parameter pushed
// with include_in_file_name set. Built with HDF5 (USE_HDF), it writes  <that>.h5  holding one
// 1-D little-endian float64 dataset per observable, named after it, with every int, bool and
// double parameter attached to each dataset as a scalar attribute (int and bool as 32-bit int).
// Built without HDF5 it writes one file per observable,  <that><observable name>.csv, with the
// samples separated by ", " at the stream's default precision.
//
// saveMatrix / saveSparseMatrix write a matrix to an HDF5 file; in a build without HDF5 they
// throw std::runtime_error. The HDF5 headers are not needed to include this file.
#include <memory>
#include <string>
#include <vector>

#include "krylovObservables.h"
#include "matrixDataTypes.h"
#include "parameter.h"

using namespace TE;

void saveResult(const std::vector<std::unique_ptr<krylovBasicObservable>>& obs_list, parameter_list& para,
                const std::string& name);

// Datasets "valuesRealPart" and "valuesImagPart" (numValues float64 each).
void saveMatrix(const matrix* mat, const std::string& name);

// The matrix's COO arrays: "valuesRealPart", "valuesImagPart", "columIndex", "rowIndex"
// (numValues entries each) and "dimension" (one entry, m); index datasets are unsigned 64-bit.
void saveSparseMatrix(const smatrix* mat, const std::string& name);
