// Device-side BSR pattern and element->BSR map (the TACS createSchurMat pattern the reference builds on the
// host, TestKernels.cpp:105, and generateElementBCSRMap, verification/TACSHelpers.h:138-182). Results are
// identical to the host builders in fem_host.cpp: block (i, j) exists iff nodes i and j share an element,
// columns sorted ascending, an isolated node keeps its diagonal block, map[e*N*N + i*N + j] = slot * bsize^2.
//
// Steps (all on one stream):
//   1. node -> element adjacency: per-node counts (atomics), exclusive scan, fill;
//   2. one wavefront per row gathers the node ids of the row's elements into LDS, sorts them (bitonic
//      network) and counts the distinct ids -> row lengths -> exclusive scan -> rowp;
//   3. the same per-row sort again writes the distinct ids into cols;
//   4. the map: one thread per (element, i, j), binary search of node j in row i.
// Invalid node ids, rows with more candidates than a wavefront sorts, and map misses raise device flags
// that the host turns into status codes; no access depends on an unchecked id.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <mutex>
#include <vector>

#include "fem_amd.h"
#include "fem_internal.h"

namespace fem {
namespace {

#define FEM_HIP_CHECK(expr)                                                                               \
  do {                                                                                                    \
    hipError_t err_ = (expr);                                                                             \
    if (err_ != hipSuccess) {                                                                             \
      return set_error(FEM_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(err_), __FILE__,    \
                       __LINE__);                                                                         \
    }                                                                                                     \
  } while (0)

constexpr int kThreads = 256;
constexpr int kRowCap = 1024;  // candidates one wavefront sorts (node references of a row's elements)
constexpr int kScanItems = 4;  // items per thread in the block scan

// Device flags: [0] invalid node id, [1] row beyond kRowCap, [2] map miss, [3] pattern beyond 2^31 blocks.
enum Flag { kBadNode = 0, kBigRow = 1, kMiss = 2, kOverflow = 3 };

__global__ __launch_bounds__(kThreads) void count_refs(int numNodes, int numElements, int npe, const int *connPtr,
                                                       const int *conn, int *cnt, int *flags) {
  const int64_t k = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (k >= int64_t(numElements) * npe) return;
  const int e = int(k / npe), a = int(k - int64_t(e) * npe);
  const int n = conn[int64_t(connPtr[e]) + a];
  if (n < 0 || n >= numNodes) {
    flags[kBadNode] = 1;
    return;
  }
  atomicAdd(&cnt[n], 1);
}

// Exclusive scan of `in` (n entries) into `out` (n + 1 entries, out[n] = total), in three passes: block sums,
// a one-block scan of the block sums, block-local scans plus offsets. Totals are int64.
__global__ __launch_bounds__(kThreads) void scan_block_sums(const int *in, int64_t n, int64_t *blockSums) {
  __shared__ int64_t part[kThreads];
  const int64_t base = int64_t(blockIdx.x) * kThreads * kScanItems;
  int64_t s = 0;
  for (int i = 0; i < kScanItems; i++) {
    const int64_t k = base + int64_t(threadIdx.x) * kScanItems + i;
    if (k < n) s += in[k];
  }
  part[threadIdx.x] = s;
  __syncthreads();
  for (int off = kThreads / 2; off > 0; off /= 2) {
    if (int(threadIdx.x) < off) part[threadIdx.x] += part[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) blockSums[blockIdx.x] = part[0];
}

// One block: exclusive scan of the block sums in place (any count, in chunks of kThreads).
__global__ __launch_bounds__(kThreads) void scan_sums(int64_t *blockSums, int64_t numBlocks) {
  __shared__ int64_t buf[kThreads];
  int64_t carry = 0;
  for (int64_t c = 0; c < numBlocks; c += kThreads) {
    const int64_t k = c + threadIdx.x;
    const int64_t v = k < numBlocks ? blockSums[k] : 0;
    buf[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < kThreads; off *= 2) {  // inclusive Hillis-Steele scan
      const int64_t t = int(threadIdx.x) >= off ? buf[threadIdx.x - off] : 0;
      __syncthreads();
      buf[threadIdx.x] += t;
      __syncthreads();
    }
    if (k < numBlocks) blockSums[k] = carry + buf[threadIdx.x] - v;
    const int64_t total = buf[kThreads - 1];
    __syncthreads();
    carry += total;
  }
}

template <typename OutT>
__global__ __launch_bounds__(kThreads) void scan_apply(const int *in, int64_t n, const int64_t *blockOffsets, OutT *out,
                                                      int *flags) {
  __shared__ int64_t buf[kThreads];
  const int64_t base = int64_t(blockIdx.x) * kThreads * kScanItems;
  int64_t v[kScanItems];
  int64_t s = 0;
  for (int i = 0; i < kScanItems; i++) {
    const int64_t k = base + int64_t(threadIdx.x) * kScanItems + i;
    v[i] = k < n ? in[k] : 0;
    s += v[i];
  }
  buf[threadIdx.x] = s;
  __syncthreads();
  for (int off = 1; off < kThreads; off *= 2) {
    const int64_t t = int(threadIdx.x) >= off ? buf[threadIdx.x - off] : 0;
    __syncthreads();
    buf[threadIdx.x] += t;
    __syncthreads();
  }
  int64_t run = blockOffsets[blockIdx.x] + buf[threadIdx.x] - s;
  for (int i = 0; i < kScanItems; i++) {
    const int64_t k = base + int64_t(threadIdx.x) * kScanItems + i;
    if (k <= n) {
      if (run > INT32_MAX) flags[kOverflow] = 1;
      out[k] = OutT(run);  // k == n writes the total
    }
    run += v[i];
  }
}

__global__ __launch_bounds__(kThreads) void fill_adjacency(int numNodes, int numElements, int npe, const int *connPtr,
                                                           const int *conn, int *cursor, int *adj) {
  const int64_t k = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (k >= int64_t(numElements) * npe) return;
  const int e = int(k / npe), a = int(k - int64_t(e) * npe);
  const int n = conn[int64_t(connPtr[e]) + a];
  if (n < 0 || n >= numNodes) return;  // flagged by count_refs
  adj[atomicAdd(&cursor[n], 1)] = e;
}

// One wavefront per row: sorted distinct node ids of the row's elements. WRITE = false stores the count in
// len[row]; WRITE = true stores the ids at cols[rowp[row]...]. Returns through flags when the row has more
// than kRowCap candidates.
template <bool WRITE>
__global__ __launch_bounds__(kThreads) void row_columns(int numNodes, int npe, const int *connPtr, const int *conn,
                                                        const int *nptr, const int *adj, int *len, const int *rowp,
                                                        int *cols, int *flags) {
  constexpr int WAVES = kThreads / 64;
  __shared__ int keys[WAVES][kRowCap];
  const int wave = int(threadIdx.x / 64), lane = int(threadIdx.x % 64);
  const int row = int(blockIdx.x) * WAVES + wave;
  if (row >= numNodes) return;  // whole wavefront leaves together (row is wave-uniform)
  int *s = keys[wave];
  const int e0 = nptr[row], ne = nptr[row + 1] - e0;
  const int k = ne * npe;
  if (k > kRowCap) {
    if (lane == 0) flags[kBigRow] = 1;
    return;
  }
  if (k == 0) {  // isolated node: diagonal block only
    if (lane == 0) {
      if (WRITE) cols[rowp[row]] = row;
      else len[row] = 1;
    }
    return;
  }
  int size = 64;
  while (size < k) size *= 2;
  for (int i = lane; i < size; i += 64) {
    int v = INT32_MAX;
    if (i < k) {
      const int e = adj[e0 + i / npe];
      v = conn[int64_t(connPtr[e]) + i % npe];
    }
    s[i] = v;
  }
  __syncwarp();
  // Bitonic sort of s[0, size) ascending; each lane handles pairs (i, i ^ j) with i < i ^ j.
  for (int w = 2; w <= size; w *= 2) {
    for (int j = w / 2; j > 0; j /= 2) {
      for (int i = lane; i < size; i += 64) {
        const int p = i ^ j;
        if (p > i) {
          const int x = s[i], y = s[p];
          const bool up = (i & w) == 0;
          if ((x > y) == up) {
            s[i] = y;
            s[p] = x;
          }
        }
      }
      __syncwarp();
    }
  }
  // Distinct ids in order: entry i is kept iff it differs from entry i - 1 (entries past k are INT32_MAX).
  int count = 0;
  for (int c = 0; c < k; c += 64) {
    const int i = c + lane;
    const bool keep = i < k && (i == 0 || s[i] != s[i - 1]);
    const unsigned long long m = __ballot(keep);
    if (WRITE && keep) cols[rowp[row] + count + __popcll(m & ((1ull << lane) - 1))] = s[i];
    count += __popcll(m);
  }
  if (!WRITE && lane == 0) len[row] = count;
}

__device__ __forceinline__ int find_col(const int *cols, int b, int e, int col) {
  while (b < e) {
    const int m = (b + e) / 2;
    if (cols[m] < col) b = m + 1;
    else e = m;
  }
  return b;
}

__global__ __launch_bounds__(kThreads) void element_map(int numRows, int numElements, int npe, const int *connPtr,
                                                        const int *conn, const int *rowp, const int *cols, int bs2,
                                                        int *map, int *flags) {
  const int64_t k = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  const int64_t nn = int64_t(npe) * npe;
  if (k >= int64_t(numElements) * nn) return;
  const int e = int(k / nn), ij = int(k - int64_t(e) * nn), i = ij / npe, j = ij - i * npe;
  const int *en = conn + connPtr[e];
  const int r = en[i], c = en[j];
  if (r < 0 || r >= numRows || c < 0 || c >= numRows) {
    flags[kBadNode] = 1;
    return;
  }
  const int b = rowp[r], end = rowp[r + 1];
  const int pos = find_col(cols, b, end, c);
  if (pos >= end || cols[pos] != c) {
    flags[kMiss] = 1;
    return;
  }
  map[k] = pos * bs2;
}

__global__ __launch_bounds__(kThreads) void fill_zero(int *p, int64_t n) {
  const int64_t k = int64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (k < n) p[k] = 0;
}

// Scratch owned by one call; freed on every return path.
struct Scratch {
  void *p[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  ~Scratch() {
    for (void *q : p)
      if (q) (void)hipFree(q);
  }
};

unsigned grid_for(int64_t n) { return unsigned((n + kThreads - 1) / kThreads); }

// Block sums an exclusive scan of n entries (n + 1 outputs) needs.
int64_t scan_blocks(int64_t n) {
  const int64_t perBlock = int64_t(kThreads) * kScanItems;
  return (n + 1 + perBlock - 1) / perBlock;
}

// Exclusive scan of in[0, n) into out[0, n] (see the scan kernels), with scan_blocks(n) block sums in `sums`.
template <typename OutT>
int exclusive_scan(const int *in, int64_t n, OutT *out, int *flags, int64_t *sums, hipStream_t stream) {
  const int64_t blocks = scan_blocks(n);
  scan_block_sums<<<unsigned(blocks), kThreads, 0, stream>>>(in, n, sums);
  scan_sums<<<1, kThreads, 0, stream>>>(sums, blocks);
  scan_apply<OutT><<<unsigned(blocks), kThreads, 0, stream>>>(in, n, sums, out, flags);
  FEM_HIP_CHECK(hipGetLastError());
  return FEM_OK;
}

int check_flags(const int *dflags, hipStream_t stream) {
  int h[4] = {0, 0, 0, 0};
  FEM_HIP_CHECK(hipMemcpyAsync(h, dflags, sizeof(h), hipMemcpyDeviceToHost, stream));
  FEM_HIP_CHECK(hipStreamSynchronize(stream));
  if (h[kBadNode]) return set_error(FEM_ERR_INVALID_ARGUMENT, "a node id is outside [0, numNodes)");
  if (h[kBigRow]) {
    return set_error(FEM_ERR_UNSUPPORTED, "a node's elements hold more than %d node references (use the host "
                     "pattern builder)", kRowCap);
  }
  if (h[kMiss]) return set_error(FEM_ERR_MAP_MISS, "an element node pair has no block in the pattern");
  if (h[kOverflow]) return set_error(FEM_ERR_INVALID_ARGUMENT, "pattern exceeds 2^31 blocks");
  return FEM_OK;
}

}  // namespace
}  // namespace fem

using namespace fem;

// Scratch of fem_bsr_pattern_device_noalloc: one array per GPU, held by the process (64 MiB). Host code never
// uses the array's name as an address: each device's copy is reached through the address
// hipGetSymbolAddress gives on that device, looked up once per device (device_scratch).
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace fem {
namespace {

// Scratch of one pattern build: error flags, per-node reference counts (then the fill cursor), node->element
// offsets, the adjacency list, row lengths and the scans' block sums.
struct PatternScratch {
  int *flags, *cnt, *nptr, *adj, *len;
  int64_t *sums;
};

constexpr size_t kScratchAlign = 256;
size_t align_bytes(size_t n) { return (n + kScratchAlign - 1) / kScratchAlign * kScratchAlign; }

// Bytes of each scratch array of a build over numNodes nodes and refs node references, in PatternScratch order.
void pattern_scratch_sizes(int numNodes, int64_t refs, size_t (&bytes)[6]) {
  const size_t nodes1 = (size_t(numNodes) + 1) * sizeof(int);
  bytes[0] = 4 * sizeof(int);
  bytes[1] = bytes[2] = bytes[4] = nodes1;
  bytes[3] = size_t(std::max<int64_t>(refs, 1)) * sizeof(int);
  bytes[5] = size_t(scan_blocks(numNodes)) * sizeof(int64_t);
}

int check_pattern_args(int numNodes, int numElements, int nodesPerElement, const int *connPtr, const int *conn,
                       const int *rowp, const int64_t *numBlocks) {
  if (numNodes < 0 || numElements < 0 || nodesPerElement <= 0) {
    return set_error(FEM_ERR_INVALID_ARGUMENT, "negative size (numNodes=%d numElements=%d nodesPerElement=%d)",
                     numNodes, numElements, nodesPerElement);
  }
  if (!rowp || !numBlocks) return set_error(FEM_ERR_NULL_POINTER, "rowp/numBlocks is NULL");
  if (numElements > 0 && (!connPtr || !conn)) return set_error(FEM_ERR_NULL_POINTER, "connPtr/conn is NULL");
  if (int64_t(numElements) * nodesPerElement > INT32_MAX) {
    return set_error(FEM_ERR_UNSUPPORTED, "more than 2^31 element node references");
  }
  return FEM_OK;
}

// The pattern build of fem_bsr_pattern_device on caller-provided scratch (arguments already checked).
int pattern_build(int numNodes, int numElements, int nodesPerElement, const int *connPtr, const int *conn, int *rowp,
                  int *cols, int64_t colsCapacity, int64_t *numBlocks, hipStream_t stream, const PatternScratch &sc) {
  const int64_t refs = int64_t(numElements) * nodesPerElement;
  int *flags = sc.flags, *cnt = sc.cnt, *nptr = sc.nptr, *adj = sc.adj, *len = sc.len;
  // cleared with a fill kernel (scratch may be reused across calls; no memset nodes)
  fill_zero<<<1, kThreads, 0, stream>>>(flags, 4);
  fill_zero<<<grid_for(int64_t(numNodes) + 1), kThreads, 0, stream>>>(cnt, int64_t(numNodes) + 1);
  if (refs > 0) {
    count_refs<<<grid_for(refs), kThreads, 0, stream>>>(numNodes, numElements, nodesPerElement, connPtr, conn, cnt,
                                                        flags);
  }
  int rc = exclusive_scan<int>(cnt, numNodes, nptr, flags, sc.sums, stream);
  if (rc) return rc;
  rc = check_flags(flags, stream);  // invalid ids must not reach the fill
  if (rc) return rc;
  // cursor = nptr (cnt is reused as the fill cursor)
  FEM_HIP_CHECK(hipMemcpyAsync(cnt, nptr, size_t(numNodes) * sizeof(int), hipMemcpyDeviceToDevice, stream));
  if (refs > 0) {
    fill_adjacency<<<grid_for(refs), kThreads, 0, stream>>>(numNodes, numElements, nodesPerElement, connPtr, conn, cnt,
                                                            adj);
  }
  const unsigned rowBlocks = unsigned((int64_t(numNodes) + kThreads / 64 - 1) / (kThreads / 64));
  if (numNodes > 0) {
    row_columns<false><<<rowBlocks, kThreads, 0, stream>>>(numNodes, nodesPerElement, connPtr, conn, nptr, adj, len,
                                                          nullptr, nullptr, flags);
  }
  rc = exclusive_scan<int>(len, numNodes, rowp, flags, sc.sums, stream);
  if (rc) return rc;
  FEM_HIP_CHECK(hipGetLastError());
  int total = 0;
  FEM_HIP_CHECK(hipMemcpyAsync(&total, rowp + numNodes, sizeof(int), hipMemcpyDeviceToHost, stream));
  rc = check_flags(flags, stream);
  if (rc) return rc;
  *numBlocks = total;
  if (cols && total > colsCapacity) {
    return set_error(FEM_ERR_INVALID_ARGUMENT, "cols holds %lld entries, the pattern has %d blocks",
                     (long long)colsCapacity, total);
  }
  if (cols && numNodes > 0) {
    row_columns<true><<<rowBlocks, kThreads, 0, stream>>>(numNodes, nodesPerElement, connPtr, conn, nptr, adj, nullptr,
                                                         rowp, cols, flags);
    FEM_HIP_CHECK(hipGetLastError());
    rc = check_flags(flags, stream);
    if (rc) return rc;
  }
  return FEM_OK;
}

std::mutex scratchMu;
std::vector<void *> scratchAddr;  // per device ordinal; NULL until looked up

int device_scratch(void **out) {
  int dev = 0;
  FEM_HIP_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(scratchMu);
  if (dev >= int(scratchAddr.size())) scratchAddr.resize(size_t(dev) + 1, nullptr);
  if (!scratchAddr[size_t(dev)]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      return set_error(FEM_ERR_HIP, "hipGetSymbolAddress of the pattern scratch failed on device %d: %s", dev,
                       hipGetErrorString(e));
    }
    scratchAddr[size_t(dev)] = p;
  }
  *out = scratchAddr[size_t(dev)];
  return FEM_OK;
}

}  // namespace
}  // namespace fem

extern "C" int fem_bsr_pattern_device(int numNodes, int numElements, int nodesPerElement, const int *connPtr,
                                      const int *conn, int *rowp, int *cols, int64_t colsCapacity, int64_t *numBlocks,
                                      void *streamPtr) {
  int rc = check_pattern_args(numNodes, numElements, nodesPerElement, connPtr, conn, rowp, numBlocks);
  if (rc) return rc;
  hipStream_t stream = static_cast<hipStream_t>(streamPtr);
  size_t bytes[6];
  pattern_scratch_sizes(numNodes, int64_t(numElements) * nodesPerElement, bytes);
  Scratch sc;
  for (int i = 0; i < 6; i++) FEM_HIP_CHECK(hipMalloc(&sc.p[i], bytes[i]));
  const PatternScratch ps{static_cast<int *>(sc.p[0]), static_cast<int *>(sc.p[1]), static_cast<int *>(sc.p[2]),
                          static_cast<int *>(sc.p[3]), static_cast<int *>(sc.p[4]), static_cast<int64_t *>(sc.p[5])};
  return pattern_build(numNodes, numElements, nodesPerElement, connPtr, conn, rowp, cols, colsCapacity, numBlocks,
                       stream, ps);
}

extern "C" int fem_bsr_pattern_device_noalloc(int numNodes, int numElements, int nodesPerElement, const int *connPtr,
                                              const int *conn, int *rowp, int *cols, int64_t colsCapacity,
                                              int64_t *numBlocks, void *streamPtr) {
  int rc = check_pattern_args(numNodes, numElements, nodesPerElement, connPtr, conn, rowp, numBlocks);
  if (rc) return rc;
  size_t bytes[6];
  pattern_scratch_sizes(numNodes, int64_t(numElements) * nodesPerElement, bytes);
  size_t need = 0;
  for (size_t b : bytes) need += align_bytes(b);
  if (need > kNoAllocScratchBytes) {
    return set_error(FEM_ERR_INVALID_ARGUMENT, "the pattern of %d nodes and %lld node references needs %zu bytes of "
                     "scratch; fem_bsr_pattern_device_noalloc holds %zu (use fem_bsr_pattern_device)", numNodes,
                     (long long)(int64_t(numElements) * nodesPerElement), need, kNoAllocScratchBytes);
  }
  void *base = nullptr;
  rc = device_scratch(&base);
  if (rc) return rc;
  unsigned char *p = static_cast<unsigned char *>(base);
  void *arr[6];
  for (int i = 0; i < 6; i++) {
    arr[i] = p;
    p += align_bytes(bytes[i]);
  }
  const PatternScratch ps{static_cast<int *>(arr[0]), static_cast<int *>(arr[1]), static_cast<int *>(arr[2]),
                          static_cast<int *>(arr[3]), static_cast<int *>(arr[4]), static_cast<int64_t *>(arr[5])};
  return pattern_build(numNodes, numElements, nodesPerElement, connPtr, conn, rowp, cols, colsCapacity, numBlocks,
                       static_cast<hipStream_t>(streamPtr), ps);
}

extern "C" int fem_element_bcsr_map_device(const int *connPtr, const int *conn, int numElements, int nodesPerElement,
                                           int numRows, const int *rowp, const int *cols, int bsize, int *map,
                                           void *streamPtr) {
  if (numRows < 0 || numElements < 0 || nodesPerElement <= 0) {
    return set_error(FEM_ERR_INVALID_ARGUMENT, "negative size (numRows=%d numElements=%d nodesPerElement=%d)",
                     numRows, numElements, nodesPerElement);
  }
  if (bsize < 1) return set_error(FEM_ERR_INVALID_ARGUMENT, "bsize %d < 1", bsize);
  if (numElements == 0) return FEM_OK;
  if (!connPtr || !conn || !rowp || !cols || !map) {
    return set_error(FEM_ERR_NULL_POINTER, "connPtr/conn/rowp/cols/map is NULL");
  }
  hipStream_t stream = static_cast<hipStream_t>(streamPtr);
  int numBlocks = 0;
  FEM_HIP_CHECK(hipMemcpyAsync(&numBlocks, rowp + numRows, sizeof(int), hipMemcpyDeviceToHost, stream));
  FEM_HIP_CHECK(hipStreamSynchronize(stream));
  if (int64_t(numBlocks) * bsize * bsize > INT32_MAX) {
    return set_error(FEM_ERR_UNSUPPORTED, "%d blocks x %d entries overflow the map's 32-bit offsets", numBlocks,
                     bsize * bsize);
  }
  Scratch sc;
  FEM_HIP_CHECK(hipMalloc(&sc.p[0], 4 * sizeof(int)));
  int *flags = static_cast<int *>(sc.p[0]);
  FEM_HIP_CHECK(hipMemsetAsync(flags, 0, 4 * sizeof(int), stream));
  const int64_t n = int64_t(numElements) * nodesPerElement * nodesPerElement;
  element_map<<<grid_for(n), kThreads, 0, stream>>>(numRows, numElements, nodesPerElement, connPtr, conn, rowp, cols,
                                                    bsize * bsize, map, flags);
  FEM_HIP_CHECK(hipGetLastError());
  return check_flags(flags, stream);
}
