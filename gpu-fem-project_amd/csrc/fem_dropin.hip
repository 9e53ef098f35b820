// Drop-in Jacobian with the reference's arguments (runJacobianKernel, element/FEKernels.h:1125-1225) backed by a
// resident plan: fem_jacobian_dropin.
//
// The reference's interface hands over only connectivity, the element->BSR map and the outputs; its kernels add
// every element contribution into matEntries with atomics (fem_jacobian does the same here). This entry point
// builds, on first use of a (connPtr, conn, map, numElements, order) combination, the BSR pattern and element map
// of the connectivity on the device; if that map equals the caller's, the caller's matrix has this library's
// pattern, and a resident plan is built and cached. Calls then run the plan's kernels in accumulate mode (orders
// 1-2: owner-computes parts; 3-4: element-once traversal and gathers): every output is read and written once
// (+=, the reference's semantics), with no atomics. Every call first checks that the cached arrays are unchanged (a 64-bit checksum of conn and of
// a strided sample of the map); a change rebuilds the entry. Anything else (orders 3-4, a different pattern, a failed
// build) runs fem_jacobian.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <memory>
#include <mutex>
#include <vector>

#include "fem_amd.h"
#include "fem_internal.h"

namespace fem {
namespace {

#define FEM_DI_CHECK(expr)                                                                                \
  do {                                                                                                    \
    hipError_t err_ = (expr);                                                                             \
    if (err_ != hipSuccess) {                                                                             \
      return set_error(FEM_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(err_), __FILE__,    \
                       __LINE__);                                                                         \
    }                                                                                                     \
  } while (0)

constexpr int kThreads = 256;
constexpr int64_t kMapStride = 61;  // map entries sampled by the per-call check
constexpr size_t kMaxEntries = 4;   // cached plans (least recently used evicted)

// 32-bit finaliser of a (position, value) pair (MurmurHash3's fmix32 on a multiplicative combination).
__device__ __forceinline__ unsigned int mix32(unsigned int k, unsigned int v) {
  unsigned int x = k * 0x9E3779B9u ^ v * 0x85EBCA6Bu;
  x ^= x >> 16;
  x *= 0x85EBCA6Bu;
  x ^= x >> 13;
  x *= 0xC2B2AE35u;
  return x ^ (x >> 16);
}

// Order-independent 64-bit sum of mix32 over conn[connPtr[e] + a] (all e, a) and map[k * stride] (sampled), plus
// the largest node id and a flag for a negative one. One atomic of each kind per workgroup.
__global__ __launch_bounds__(kThreads) void checksum_kernel(const int *connPtr, const int *conn, int numElements,
                                                            int npe, const int *map, int64_t mapSamples,
                                                            unsigned long long *sum, int *maxNode, int *negative) {
  const int64_t refs = int64_t(numElements) * npe;
  unsigned long long h = 0;
  int mx = -1;
  bool neg = false;
  for (int64_t k = int64_t(blockIdx.x) * kThreads + threadIdx.x; k < refs + mapSamples;
       k += int64_t(gridDim.x) * kThreads) {
    if (k < refs) {
      const int e = int(k / npe), a = int(k - int64_t(e) * npe);
      const int n = conn[int64_t(connPtr[e]) + a];
      h += mix32(unsigned(k), unsigned(n));
      mx = max(mx, n);
      neg = neg || n < 0;
    } else {
      const int64_t s = k - refs;
      h += (unsigned long long)mix32(unsigned(s) ^ 0x5bd1e995u, unsigned(map[s * kMapStride])) << 20;
    }
  }
  __shared__ unsigned long long hs[kThreads / 64];
  __shared__ int ms[kThreads / 64], ns[kThreads / 64];
  for (int off = 32; off > 0; off >>= 1) {
    h += __shfl_down(h, off);
    mx = max(mx, __shfl_down(mx, off));
  }
  const bool wneg = __any(neg);
  const int w = threadIdx.x / 64;
  if ((threadIdx.x & 63) == 0) {
    hs[w] = h;
    ms[w] = mx;
    ns[w] = wneg;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long bh = 0;
    int bm = -1, bn = 0;
    for (int i = 0; i < kThreads / 64; i++) {
      bh += hs[i];
      bm = max(bm, ms[i]);
      bn |= ns[i];
    }
    atomicAdd(sum, bh);
    atomicMax(maxNode, bm);
    if (bn) atomicOr(negative, 1);
  }
}

__global__ __launch_bounds__(kThreads) void differ_kernel(const int *a, const int *b, int64_t n, int *flag) {
  bool d = false;
  for (int64_t k = int64_t(blockIdx.x) * kThreads + threadIdx.x; k < n; k += int64_t(gridDim.x) * kThreads)
    d = d || a[k] != b[k];
  if (__any(d) && (threadIdx.x & 63) == 0) atomicOr(flag, 1);
}

struct Entry {
  int device = 0, order = 0, numElements = 0;
  const int *connPtr = nullptr, *conn = nullptr, *map = nullptr;
  unsigned long long sum = 0;
  fem_plan *plan = nullptr;  // NULL: this combination runs fem_jacobian
  uint64_t lastUse = 0;
  unsigned long long *dSum = nullptr;  // the check's device results: checksum, then (largest node, negative)
  int *dInts = nullptr;
  ~Entry() {
    fem_plan_destroy(plan);
    if (dSum) (void)hipFree(dSum);
    if (dInts) (void)hipFree(dInts);
  }
};

std::mutex cacheMu;
std::vector<std::unique_ptr<Entry>> cache;
uint64_t useClock = 0;

// Device buffers of one build, freed on every return path.
struct Temp {
  std::vector<void *> p;
  ~Temp() {
    for (void *q : p) (void)hipFree(q);
  }
  template <typename T>
  int alloc(T **out, size_t n) {
    void *q = nullptr;
    if (hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) {
      return set_error(FEM_ERR_OUT_OF_MEMORY, "hipMalloc of %zu bytes failed", n * sizeof(T));
    }
    p.push_back(q);
    *out = static_cast<T *>(q);
    return FEM_OK;
  }
};

unsigned grid_for(int64_t n) { return unsigned(std::min<int64_t>((n + kThreads - 1) / kThreads, 2048)); }

int64_t map_samples(int numElements, int npe) {
  const int64_t n = int64_t(numElements) * npe * npe;
  return (n + kMapStride - 1) / kMapStride;
}

// Checksum (and largest node id, negative-id flag) of the entry's arrays, in the entry's device results;
// synchronises `stream`.
int checksum(Entry &e, hipStream_t stream, unsigned long long &sum, int &maxNode, bool &negative) {
  const int order = e.order, numElements = e.numElements;
  const int *connPtr = e.connPtr, *conn = e.conn, *map = e.map;
  const int npe = nodes_per_element(order);
  if (!e.dSum) FEM_DI_CHECK(hipMalloc(reinterpret_cast<void **>(&e.dSum), sizeof(unsigned long long)));
  if (!e.dInts) FEM_DI_CHECK(hipMalloc(reinterpret_cast<void **>(&e.dInts), 2 * sizeof(int)));
  unsigned long long *dSum = e.dSum;
  int *dInts = e.dInts;
  const int init[2] = {-1, 0};
  const unsigned long long zero = 0;
  FEM_DI_CHECK(hipMemcpyAsync(dSum, &zero, sizeof(zero), hipMemcpyHostToDevice, stream));
  FEM_DI_CHECK(hipMemcpyAsync(dInts, init, sizeof(init), hipMemcpyHostToDevice, stream));
  const int64_t work = int64_t(numElements) * npe + map_samples(numElements, npe);
  checksum_kernel<<<grid_for(work), kThreads, 0, stream>>>(connPtr, conn, numElements, npe, map,
                                                          map_samples(numElements, npe), dSum, dInts, dInts + 1);
  FEM_DI_CHECK(hipGetLastError());
  int h[2];
  FEM_DI_CHECK(hipMemcpyAsync(&sum, dSum, sizeof(sum), hipMemcpyDeviceToHost, stream));
  FEM_DI_CHECK(hipMemcpyAsync(h, dInts, sizeof(h), hipMemcpyDeviceToHost, stream));
  FEM_DI_CHECK(hipStreamSynchronize(stream));
  maxNode = h[0];
  negative = h[1] != 0;
  return FEM_OK;
}

// Builds the entry's plan if the caller's map is this library's map of the connectivity (else leaves it NULL).
int build(Entry &e, const double *nodeCoords, hipStream_t stream) {
  fem_plan_destroy(e.plan);
  e.plan = nullptr;
  const int npe = nodes_per_element(e.order);
  int maxNode = -1;
  bool negative = false;
  int rc = checksum(e, stream, e.sum, maxNode, negative);
  if (rc || negative || maxNode < 0) return rc;
  const int numNodes = maxNode + 1;
  Temp t;
  int *rowp, *cols, *mine, *flag;
  if ((rc = t.alloc(&rowp, size_t(numNodes) + 1)) || (rc = t.alloc(&flag, 1))) return rc;
  int64_t nb = 0;
  rc = fem_bsr_pattern_device(numNodes, e.numElements, npe, e.connPtr, e.conn, rowp, nullptr, 0, &nb, stream);
  if (rc) return rc;
  if ((rc = t.alloc(&cols, size_t(nb)))) return rc;
  rc = fem_bsr_pattern_device(numNodes, e.numElements, npe, e.connPtr, e.conn, rowp, cols, nb, &nb, stream);
  if (rc) return rc;
  const int64_t nmap = int64_t(e.numElements) * npe * npe;
  if ((rc = t.alloc(&mine, size_t(nmap)))) return rc;
  rc = fem_element_bcsr_map_device(e.connPtr, e.conn, e.numElements, npe, numNodes, rowp, cols, 2, mine, stream);
  if (rc) {
    clear_error();  // e.g. a pattern beyond the map's 32-bit offsets: fem_jacobian serves it
    return FEM_OK;
  }
  FEM_DI_CHECK(hipMemsetAsync(flag, 0, sizeof(int), stream));
  differ_kernel<<<grid_for(nmap), kThreads, 0, stream>>>(mine, e.map, nmap, flag);
  FEM_DI_CHECK(hipGetLastError());
  int differ = 1;
  FEM_DI_CHECK(hipMemcpyAsync(&differ, flag, sizeof(int), hipMemcpyDeviceToHost, stream));
  FEM_DI_CHECK(hipStreamSynchronize(stream));
  if (differ) return FEM_OK;  // the caller's matrix has another layout
  // host copies for the plan builder
  std::vector<int> hConnPtr(static_cast<size_t>(e.numElements)), hRowp(static_cast<size_t>(numNodes) + 1),
      hCols(static_cast<size_t>(nb));
  FEM_DI_CHECK(hipMemcpy(hConnPtr.data(), e.connPtr, hConnPtr.size() * sizeof(int), hipMemcpyDeviceToHost));
  int64_t connLen = 0;
  for (int v : hConnPtr) connLen = std::max<int64_t>(connLen, int64_t(v) + npe);
  std::vector<int> hConn(static_cast<size_t>(connLen));
  std::vector<double> hCoords(static_cast<size_t>(numNodes) * 2);
  FEM_DI_CHECK(hipMemcpy(hConn.data(), e.conn, hConn.size() * sizeof(int), hipMemcpyDeviceToHost));
  FEM_DI_CHECK(hipMemcpy(hRowp.data(), rowp, hRowp.size() * sizeof(int), hipMemcpyDeviceToHost));
  FEM_DI_CHECK(hipMemcpy(hCols.data(), cols, hCols.size() * sizeof(int), hipMemcpyDeviceToHost));
  FEM_DI_CHECK(hipMemcpy(hCoords.data(), nodeCoords, hCoords.size() * sizeof(double), hipMemcpyDeviceToHost));
  fem_plan *plan = nullptr;
  rc = fem_plan_create(e.order, numNodes, e.numElements, hConnPtr.data(), hConn.data(), hRowp.data(), hCols.data(),
                       hCoords.data(), 0, &plan);
  if (rc) {
    clear_error();  // e.g. a row beyond the plan's 256-block limit: fem_jacobian serves the mesh
    return FEM_OK;
  }
  e.plan = plan;
  return FEM_OK;
}

}  // namespace
}  // namespace fem

using namespace fem;

extern "C" {

int fem_jacobian_dropin(int order, const int *connPtr, const int *conn, int numElements, const double *states,
                        const double *nodeCoords, double E, double nu, double t, const int *elementBCSRMap,
                        double *residual, double *matEntries, int precision, void *stream, double *elapsedSeconds) {
  auto plain = [&]() {
    return fem_jacobian(order, connPtr, conn, numElements, states, nodeCoords, E, nu, t, elementBCSRMap, residual,
                        matEntries, precision, stream, elapsedSeconds);
  };
  if (order < 1 || order > 2 || numElements <= 0 || !connPtr || !conn || !elementBCSRMap || !matEntries || !states ||
      !nodeCoords || (precision != FEM_PREC_FP64 && precision != FEM_PREC_MIXED)) {
    // argument errors are reported by fem_jacobian; at orders 3-4 its atomic kernels are faster than the
    // element-once route in accumulate mode (annulus 200x1250 order 4: 4.3 against 6.0 ms)
    return plain();
  }
  if (int64_t(numElements) * nodes_per_element(order) * nodes_per_element(order) > INT32_MAX) return plain();
  hipStream_t s = static_cast<hipStream_t>(stream);
  int device = 0;
  FEM_DI_CHECK(hipGetDevice(&device));
  std::lock_guard<std::mutex> lk(cacheMu);
  Entry *e = nullptr;
  for (auto &c : cache) {
    if (c->device == device && c->order == order && c->numElements == numElements && c->connPtr == connPtr &&
        c->conn == conn && c->map == elementBCSRMap) {
      e = c.get();
    }
  }
  int rc = FEM_OK;
  if (!e) {
    if (cache.size() >= kMaxEntries) {
      auto lru = std::min_element(cache.begin(), cache.end(),
                                  [](const auto &a, const auto &b) { return a->lastUse < b->lastUse; });
      cache.erase(lru);
    }
    cache.push_back(std::make_unique<Entry>());
    e = cache.back().get();
    e->device = device;
    e->order = order;
    e->numElements = numElements;
    e->connPtr = connPtr;
    e->conn = conn;
    e->map = elementBCSRMap;
    rc = build(*e, nodeCoords, s);
  } else {
    unsigned long long sum = 0;
    int maxNode = 0;
    bool negative = false;
    rc = checksum(*e, s, sum, maxNode, negative);
    if (!rc && sum != e->sum) rc = build(*e, nodeCoords, s);  // arrays rewritten in place
  }
  e->lastUse = ++useClock;
  if (rc || !e->plan) {
    clear_error();
    return plain();
  }
  return fem_plan_assemble_accumulate(e->plan, states, nodeCoords, E, nu, t, residual, matEntries, precision, stream,
                                      elapsedSeconds);
}

int fem_dropin_cache_clear(void) {
  std::lock_guard<std::mutex> lk(cacheMu);
  cache.clear();
  return FEM_OK;
}

int fem_dropin_cache_info(int order, const int *connPtr, const int *conn, int numElements, const int *elementBCSRMap,
                          int *cached, int *usesPlan) {
  if (!cached || !usesPlan) return set_error(FEM_ERR_NULL_POINTER, "cached/usesPlan is NULL");
  int device = 0;
  FEM_DI_CHECK(hipGetDevice(&device));
  std::lock_guard<std::mutex> lk(cacheMu);
  *cached = *usesPlan = 0;
  for (auto &c : cache) {
    if (c->device == device && c->order == order && c->numElements == numElements && c->connPtr == connPtr &&
        c->conn == conn && c->map == elementBCSRMap) {
      *cached = 1;
      *usesPlan = c->plan != nullptr;
    }
  }
  return FEM_OK;
}

}  // extern "C"
