// device_kernels_begin.hpp -- opens a region of kernel classes compiled as
// host and device code, closed by device_kernels_end.hpp.
//
// Reference-shaped kernel classes (user_kernel.hpp:10-179) included between
// the two headers compile unchanged; their member functions become
// __host__ __device__, so Kernel_Fun can run in the entry kernels on the GPU
// while the constructor, get_points and getMatrixEntry stay host code (a
// member function is only built for the device when device code calls it).
// Needs hipcc in HIP mode (-x hip), with -ffp-contract=off so Kernel_Fun
// rounds on the device as it does on the host.  No include guard: a program
// may open more than one region.
#if !defined(__HIP__)
#error "device kernel classes are compiled by hipcc as HIP (-x hip)"
#endif
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <iostream>
#include <set>
#include <string>
#include <vector>

#include "HODLRnD.hpp"

#pragma clang force_cuda_host_device begin
