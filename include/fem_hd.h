// Host/device qualifier used by the element-primitive headers.
//
// The basis and quadrature headers are shared by plain C++ host code (g++) and by the
// HIP kernels (hipcc). Under hipcc the functions are usable on both sides; under a host
// compiler the qualifier is empty.
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define FEM_HD __host__ __device__
#else
#define FEM_HD
#endif
