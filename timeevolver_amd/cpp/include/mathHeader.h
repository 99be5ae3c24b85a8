#pragma once
// Numeric backend selection. The reference chooses MKL, OpenBLAS, Armadillo or Accelerate here
// (core/mathHeader.h:5-33); in this build the backend is the MI355X device behind the C-ABI of
// include/timeevolver.h, so no BLAS header is needed.
#include <complex>
