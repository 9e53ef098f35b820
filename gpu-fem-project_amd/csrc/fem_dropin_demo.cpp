// C++ host program using the drop-in header exactly as the reference's verification/TestKernels.cpp
// uses its dispatchers (TestKernels.cpp:190-253): build a mesh, its BSR pattern and element map on the
// host, copy everything to the device, zero the outputs, call runResidualKernel / runJacobianKernel,
// copy back, and check R = K u on the host. The Jacobian is called twice on zeroed outputs: the first call
// builds the library's cached plan for this connectivity and map (orders 1-2), the second is served by it; both
// results must agree.
//
// Usage: fem_dropin_demo [order=2] [n=32] [outPrefix]. Prints one line per check; exit status 0 on success.
// With outPrefix, writes the residual and BSR values in the reference's formats (outPrefix_residual.csv,
// outPrefix_jacobian.mtx) for comparison with a CPU assembly.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "FEKernels.h"

#define CHECK_HIP(x)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (x);                                                               \
    if (e_ != hipSuccess) {                                                            \
      std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                     \
      return 2;                                                                        \
    }                                                                                  \
  } while (0)

template <typename T>
static T *to_device(const std::vector<T> &h) {
  T *d = nullptr;
  if (hipMalloc(&d, std::max<size_t>(1, h.size()) * sizeof(T)) != hipSuccess) return nullptr;
  if (hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) return nullptr;
  return d;
}

int main(int argc, char **argv) {
  const int order = argc > 1 ? std::atoi(argv[1]) : 2;
  const int n = argc > 2 ? std::atoi(argv[2]) : 32;
  const char *outPrefix = argc > 3 ? argv[3] : nullptr;
  const int N = (order + 1) * (order + 1);
  int64_t nn64, ne64;
  if (fem_square_mesh_size(n, n, order, &nn64, &ne64)) {
    std::fprintf(stderr, "%s\n", fem_last_error());
    return 1;
  }
  const int numNodes = int(nn64), numElements = int(ne64);
  std::vector<double> coords(2 * size_t(numNodes)), states(2 * size_t(numNodes));
  std::vector<int> conn(size_t(numElements) * N), connPtr(numElements);
  for (int e = 0; e < numElements; e++) connPtr[e] = e * N;
  if (fem_square_mesh(n, n, order, 0.2, 2, coords.data(), conn.data()) ||
      fem_fill_uniform(2 * int64_t(numNodes), -1e-3, 1e-3, 3, states.data())) {
    std::fprintf(stderr, "%s\n", fem_last_error());
    return 1;
  }
  std::vector<int> rowp(size_t(numNodes) + 1);
  fem_bsr_pattern_rowp(numNodes, numElements, N, connPtr.data(), conn.data(), rowp.data());
  std::vector<int> cols(static_cast<size_t>(rowp[numNodes]));
  fem_bsr_pattern_cols(numNodes, numElements, N, connPtr.data(), conn.data(), rowp.data(), cols.data());
  std::vector<int> map(size_t(numElements) * N * N);
  if (!generateElementBCSRMap(connPtr.data(), conn.data(), numElements, N, numNodes, rowp.data(), cols.data(), 2,
                              map.data())) {
    return 1;
  }
  const size_t nvals = size_t(rowp[numNodes]) * 4;

  int *dConnPtr = to_device(connPtr), *dConn = to_device(conn), *dMap = to_device(map);
  double *dStates = to_device(states), *dCoords = to_device(coords);
  double *dRes = nullptr, *dVals = nullptr;
  CHECK_HIP(hipMalloc(&dRes, 2 * size_t(numNodes) * sizeof(double)));
  CHECK_HIP(hipMalloc(&dVals, nvals * sizeof(double)));
  if (!dConnPtr || !dConn || !dMap || !dStates || !dCoords) return 2;

  const double E = 70e9, nu = 0.3, t = 1.0;
  CHECK_HIP(hipMemset(dRes, 0, 2 * size_t(numNodes) * sizeof(double)));
  const double resTime = runResidualKernel(order, dConnPtr, dConn, numElements, dStates, dCoords, E, nu, t, dRes);
  CHECK_HIP(hipMemset(dVals, 0, nvals * sizeof(double)));
  const double jacFirst = runJacobianKernel(order, dConnPtr, dConn, numElements, dStates, dCoords, E, nu, t, dMap,
                                            nullptr, dVals);
  std::vector<double> first(nvals);
  CHECK_HIP(hipMemcpy(first.data(), dVals, first.size() * sizeof(double), hipMemcpyDeviceToHost));
  CHECK_HIP(hipMemset(dVals, 0, nvals * sizeof(double)));
  const double jacTime = runJacobianKernel(order, dConnPtr, dConn, numElements, dStates, dCoords, E, nu, t, dMap,
                                           nullptr, dVals);
  if (resTime < 0 || jacFirst < 0 || jacTime < 0) return 1;

  std::vector<double> res(2 * size_t(numNodes)), vals(nvals);
  CHECK_HIP(hipMemcpy(res.data(), dRes, res.size() * sizeof(double), hipMemcpyDeviceToHost));
  CHECK_HIP(hipMemcpy(vals.data(), dVals, vals.size() * sizeof(double), hipMemcpyDeviceToHost));
  double again = 0, vmax = 0;
  for (size_t k = 0; k < nvals; k++) {
    again = std::max(again, std::fabs(vals[k] - first[k]));
    vmax = std::max(vmax, std::fabs(vals[k]));
  }
  again /= vmax;
  if (outPrefix) {
    const std::string r = std::string(outPrefix) + "_residual.csv", j = std::string(outPrefix) + "_jacobian.mtx";
    if (fem_write_array(res.data(), int64_t(res.size()), r.c_str()) ||
        fem_write_bsr_matrix(numNodes, numNodes, 2, rowp.data(), cols.data(), vals.data(), 0, 0, j.c_str())) {
      std::fprintf(stderr, "%s\n", fem_last_error());
      return 1;
    }
  }

  // Linear strain: the residual equals K u.
  double maxDiff = 0, maxRes = 0;
  for (int r = 0; r < numNodes; r++) {
    double y0 = 0, y1 = 0;
    for (int k = rowp[r]; k < rowp[r + 1]; k++) {
      const double *A = &vals[4 * size_t(k)];
      const int c = cols[k];
      y0 += A[0] * states[2 * c] + A[1] * states[2 * c + 1];
      y1 += A[2] * states[2 * c] + A[3] * states[2 * c + 1];
    }
    maxDiff = std::max({maxDiff, std::fabs(y0 - res[2 * r]), std::fabs(y1 - res[2 * r + 1])});
    maxRes = std::max({maxRes, std::fabs(res[2 * r]), std::fabs(res[2 * r + 1])});
  }
  const double rel = maxDiff / maxRes;
  int cached = 0, usesPlan = 0;
  fem_dropin_cache_info(order, dConnPtr, dConn, numElements, dMap, &cached, &usesPlan);
  std::printf("order %d, %d elements: residual %.3e s, jacobian %.3e s (first call %.3e s, %s), |Ku - R|/|R| = "
              "%.3e, second call vs first %.3e\n", order, numElements, resTime, jacTime, jacFirst,
              usesPlan ? "cached plan" : "atomic kernels", rel, again);
  // An unknown order is an error, not a silent no-op.
  const double bad = runResidualKernel(7, dConnPtr, dConn, numElements, dStates, dCoords, E, nu, t, dRes);
  std::printf("order 7 rejected: %s\n", bad < 0 ? "yes" : "no");
  (void)hipFree(dConnPtr);
  (void)hipFree(dConn);
  (void)hipFree(dMap);
  (void)hipFree(dStates);
  (void)hipFree(dCoords);
  (void)hipFree(dRes);
  (void)hipFree(dVals);
  return (rel < 1e-10 && again < 1e-14 && bad < 0) ? 0 : 1;
}
