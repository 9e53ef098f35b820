// Drop-in replacement for the host dispatchers of the reference's element/FEKernels.h.
//
//   double runResidualKernel(...)   replaces element/FEKernels.h:1030-1102
//   double runJacobianKernel(...)   replaces element/FEKernels.h:1125-1225
//
// Same names, argument lists, pointer meanings and accumulate (+=) semantics: every pointer is caller-
// owned device memory, the caller zeroes `residual` / `matEntries` before the call, `residual` may be
// nullptr for the Jacobian, and elementBCSRMap holds 4 * BSR-slot offsets (TACSHelpers.h:138-182).
// The call is synchronous on the default stream and returns the elapsed time in seconds, measured
// with HIP events, as the reference's Jacobian dispatcher does.
//
// Errors: the reference calls exit() on a CUDA error and silently does nothing for an unknown order
// (FEKernels.h:1092-1093, 1208-1209). These wrappers print the library's message to stderr and return
// a negative time instead; the C ABI in fem_amd.h returns status codes.
//
// Precision: FP64, as the reference. Link with -lfem_amd.
#pragma once

#include <cstdio>

#include "GaussQuadrature.h"
#include "LagrangeShapeFuncs.h"
#include "fem_amd.h"

inline double runResidualKernel(const int elementOrder, const int *const connPtr, const int *const conn,
                                const int numElements, const double *const states, const double *const nodeCoords,
                                const double E, const double nu, const double t, double *const residual) {
  double seconds = 0.0;
  const int rc = fem_residual(elementOrder, connPtr, conn, numElements, states, nodeCoords, E, nu, t, residual,
                              FEM_PREC_FP64, nullptr, &seconds);
  if (rc != FEM_OK) {
    std::fprintf(stderr, "runResidualKernel: %s (status %d)\n", fem_last_error(), rc);
    return -1.0;
  }
  return seconds;
}

inline double runJacobianKernel(const int elementOrder, const int *const connPtr, const int *const conn,
                                const int numElements, const double *const states, const double *const nodeCoords,
                                const double E, const double nu, const double t, int *elementBCSRMap,
                                double *const residual, double *const matEntries) {
  double seconds = 0.0;
  // fem_jacobian_dropin: at orders 1-2 a cached resident plan in accumulate mode when the map is this library's
  // map of the connectivity; otherwise the element-parallel atomic kernel (fem_jacobian)
  const int rc = fem_jacobian_dropin(elementOrder, connPtr, conn, numElements, states, nodeCoords, E, nu, t,
                                     elementBCSRMap, residual, matEntries, FEM_PREC_FP64, nullptr, &seconds);
  if (rc != FEM_OK) {
    std::fprintf(stderr, "runJacobianKernel: %s (status %d)\n", fem_last_error(), rc);
    return -1.0;
  }
  return seconds;
}

// TACS-free overload of the reference's generateElementBCSRMap (verification/TACSHelpers.h:138-182):
// the BCSRMatData argument is replaced by its rowp/cols/bsize fields; the map is host memory the caller
// allocated (numElements * numNodesPerElement^2 ints). Returns false (after printing the reason) if an
// element couples two nodes that have no block in the pattern.
inline bool generateElementBCSRMap(const int *const connPtr, const int *const conn, const int numElements,
                                   const int numNodesPerElement, const int numRows, const int *const rowp,
                                   const int *const cols, const int bsize, int *bcsrMap) {
  const int rc = fem_element_bcsr_map(connPtr, conn, numElements, numNodesPerElement, numRows, rowp, cols, bsize,
                                      bcsrMap);
  if (rc != FEM_OK) {
    std::fprintf(stderr, "generateElementBCSRMap: %s (status %d)\n", fem_last_error(), rc);
    return false;
  }
  return true;
}
