// dense.hpp -- dense inverse of the coarsest-level matrix on the GPU.
#pragma once

#include <vector>

#include "device.hpp"

namespace unap_hip {

// inv = a^-1 for a symmetric positive definite n x n matrix (row-major),
// by blocked Gauss-Jordan on the GPU (FP64 MFMA updates); inv is allocated
// to n * n doubles in device memory.
void denseInverseSPD(const std::vector<double>& a, int n, DBuf<double>& inv, hipStream_t s);

}  // namespace unap_hip
