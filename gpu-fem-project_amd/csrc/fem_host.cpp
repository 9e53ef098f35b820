// Host-side part of libfem_amd: error reporting, element tables, synthetic meshes and fields,
// the TACS-style BSR pattern builder, the element->BSR map (TACSHelpers.h:138-182 of the
// reference), element colouring and the owner-computes partitioner behind fem_plan_create.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "GaussQuadrature.h"
#include "LagrangeShapeFuncs.h"
#include "fem_internal.h"

namespace fem {

namespace {
thread_local std::string g_lastError;

// SplitMix64 (Steele, Lea & Flood 2014).
struct SplitMix64 {
  uint64_t state;
  explicit SplitMix64(uint64_t seed) : state(seed) {}
  uint64_t next() {
    uint64_t z = (state += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  double u01() { return double(next() >> 11) * 0x1.0p-53; }
  double uniform(double lo, double hi) { return lo + (hi - lo) * u01(); }
};

int hardware_threads() {
  unsigned n = std::thread::hardware_concurrency();
  if (n == 0) n = 4;
  return int(std::min(n, 64u));
}

// Runs fn(begin, end) over [0, n) split across threads.
template <typename Fn>
void parallel_for(int64_t n, Fn fn) {
  const int nt = int(std::min<int64_t>(hardware_threads(), std::max<int64_t>(1, n / 4096)));
  if (nt <= 1) {
    fn(int64_t(0), n);
    return;
  }
  std::vector<std::thread> pool;
  const int64_t chunk = (n + nt - 1) / nt;
  for (int t = 0; t < nt; t++) {
    const int64_t b = t * chunk, e = std::min(n, b + chunk);
    if (b >= e) break;
    pool.emplace_back([=]() { fn(b, e); });
  }
  for (auto &th : pool) th.join();
}

bool valid_order(int order) { return order >= 1 && order <= 4; }

// 1-D knot positions of a structured axis: n elements on [0, 1], order p, cosine knots.
std::vector<double> axis_knots(int n, int p) {
  std::vector<double> x(size_t(p) * n + 1);
  for (int j = 0; j <= p * n; j++) {
    const int e = std::min(j / p, n - 1);
    const int k = j - e * p;
    x[j] = (e + 0.5 * (1.0 + getCosinePoint(k, p))) / n;
  }
  return x;
}

// Jitter half-width at knot j: frac times the smaller adjacent gap (periodic axes wrap).
double knot_gap(const std::vector<double> &x, int j, bool periodic, double period) {
  const int m = int(x.size());
  double left, right;
  if (periodic) {
    const int jl = (j - 1 + m) % m, jr = (j + 1) % m;
    left = x[j] - x[jl];
    if (left <= 0) left += period;
    right = x[jr] - x[j];
    if (right <= 0) right += period;
  } else {
    left = x[j] - x[j - 1];
    right = x[j + 1] - x[j];
  }
  return std::min(left, right);
}

// Node -> element adjacency (CSR).
void node_element_adjacency(int numNodes, int numElements, int npe, const int *connPtr, const int *conn,
                            std::vector<int> &ptr, std::vector<int> &adj) {
  ptr.assign(size_t(numNodes) + 1, 0);
  for (int e = 0; e < numElements; e++) {
    const int *en = conn + connPtr[e];
    for (int a = 0; a < npe; a++) ptr[en[a] + 1]++;
  }
  for (int n = 0; n < numNodes; n++) ptr[n + 1] += ptr[n];
  adj.resize(size_t(ptr[numNodes]));
  std::vector<int> fill(ptr.begin(), ptr.end() - 1);
  for (int e = 0; e < numElements; e++) {
    const int *en = conn + connPtr[e];
    for (int a = 0; a < npe; a++) adj[size_t(fill[en[a]]++)] = e;
  }
}

int check_connectivity(int numNodes, int numElements, int npe, const int *connPtr, const int *conn) {
  if (numNodes < 0 || numElements < 0 || npe <= 0) {
    return set_error(FEM_ERR_INVALID_ARGUMENT, "negative size (numNodes=%d numElements=%d nodesPerElement=%d)",
                     numNodes, numElements, npe);
  }
  if (numElements > 0 && (!connPtr || !conn)) return set_error(FEM_ERR_NULL_POINTER, "connPtr/conn is NULL");
  for (int e = 0; e < numElements; e++) {
    if (connPtr[e] < 0) return set_error(FEM_ERR_INVALID_ARGUMENT, "connPtr[%d] = %d is negative", e, connPtr[e]);
    const int *en = conn + connPtr[e];
    for (int a = 0; a < npe; a++) {
      if (en[a] < 0 || en[a] >= numNodes) {
        return set_error(FEM_ERR_INVALID_ARGUMENT, "element %d node %d has id %d outside [0, %d)", e, a, en[a],
                         numNodes);
      }
    }
  }
  return FEM_OK;
}

// Sorted unique node ids coupled to node n (nodes of every element containing n).
void row_columns(int n, int npe, const int *connPtr, const int *conn, const std::vector<int> &ptr,
                 const std::vector<int> &adj, std::vector<int> &tmp) {
  tmp.clear();
  for (int k = ptr[n]; k < ptr[n + 1]; k++) {
    const int *en = conn + connPtr[adj[size_t(k)]];
    tmp.insert(tmp.end(), en, en + npe);
  }
  if (tmp.empty()) tmp.push_back(n);  // isolated node: diagonal block only
  std::sort(tmp.begin(), tmp.end());
  tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
}

// Bijective 32-bit mix (xor-shift-multiply), used as a fixed pseudo-random sort key.
inline uint32_t scramble(uint32_t v) {
  v ^= v >> 16;
  v *= 0x7feb352du;
  v ^= v >> 15;
  v *= 0x846ca68bu;
  return v ^ (v >> 16);
}

// Position of `col` in cols[b, e) (sorted); -1 if absent.
inline int find_sorted(const int *cols, int b, int e, int col) {
  const int *p = std::lower_bound(cols + b, cols + e, col);
  return (p != cols + e && *p == col) ? int(p - cols) : -1;
}

uint32_t morton2(uint32_t x, uint32_t y) {
  auto spread = [](uint32_t v) {
    v &= 0xFFFF;
    v = (v | (v << 8)) & 0x00FF00FF;
    v = (v | (v << 4)) & 0x0F0F0F0F;
    v = (v | (v << 2)) & 0x33333333;
    v = (v | (v << 1)) & 0x55555555;
    return v;
  };
  return spread(x) | (spread(y) << 1);
}
}  // namespace

int set_error(int code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_lastError = buf;
  return code;
}

void clear_error() { g_lastError.clear(); }

// ---- owner-computes partitioner ------------------------------------------------------------------
int build_plan_host(int order, int numNodes, int numElements, const int *connPtr, const int *conn, const int *rowp,
                    const int *cols, const double *nodeCoords, int flags, PlanHost &out,
                    double partScale) {
  if (!valid_order(order)) return set_error(FEM_ERR_INVALID_ORDER, "element order %d outside 1..4", order);
  const int N = nodes_per_element(order);
  if (!rowp || !cols) return set_error(FEM_ERR_NULL_POINTER, "rowp/cols is NULL");
  int rc = check_connectivity(numNodes, numElements, N, connPtr, conn);
  if (rc) return rc;
  if (numNodes == 0 || numElements == 0) return set_error(FEM_ERR_INVALID_ARGUMENT, "empty mesh");

  out = PlanHost();
  out.order = order;
  out.numNodes = numNodes;
  out.numElements = numElements;
  out.numBlocks = rowp[numNodes];
  for (int n = 0; n < numNodes; n++) {
    const int len = rowp[n + 1] - rowp[n];
    if (len <= 0) return set_error(FEM_ERR_INVALID_ARGUMENT, "BSR row %d is empty", n);
    out.maxRowLength = std::max(out.maxRowLength, len);
  }
  if (out.maxRowLength > 256) {
    return set_error(FEM_ERR_UNSUPPORTED,
                     "a BSR row holds %d blocks; the plan's byte-wide column ranks allow at most 256",
                     out.maxRowLength);
  }
  const bool residualOnly = (flags & FEM_PLAN_RESIDUAL_ONLY) != 0;
  const OwnerConfig cfg = owner_config(order);
  const int maxBlocks = residualOnly ? (1 << 30) : cfg.maxBlocks;
  const int maxRows = residualOnly ? kResidualOnlyMaxRows : cfg.maxRows;
  if (!residualOnly && out.maxRowLength > maxBlocks) {
    return set_error(FEM_ERR_UNSUPPORTED, "a BSR row of %d blocks exceeds the %d-block part budget",
                     out.maxRowLength, maxBlocks);
  }

  // Column ranks of every element block, checked against the pattern, stored transposed and padded:
  // rank[(e*N + j)*RP + i] = slot of column conn[e,j] within row conn[e,i].
  const int RP = rank_pad(order);
  out.rank.assign(size_t(numElements) * N * RP, 0);
  std::atomic<int> missE{-1};
  parallel_for(numElements, [&](int64_t b, int64_t e) {
    for (int64_t el = b; el < e; el++) {
      const int *en = conn + connPtr[el];
      for (int i = 0; i < N; i++) {
        const int r = en[i];
        for (int j = 0; j < N; j++) {
          const int pos = find_sorted(cols, rowp[r], rowp[r + 1], en[j]);
          if (pos < 0) {
            missE.store(int(el));
            return;
          }
          out.rank[(size_t(el) * N + j) * RP + i] = uint8_t(pos - rowp[r]);
        }
      }
    }
  });
  if (missE.load() >= 0) {
    return set_error(FEM_ERR_MAP_MISS, "element %d couples nodes with no block in the BSR pattern", missE.load());
  }

  // Node order for partitioning: Morton curve over the coordinates, else natural.
  std::vector<int> order_nodes(numNodes);
  std::iota(order_nodes.begin(), order_nodes.end(), 0);
  if (nodeCoords) {
    double xmin = 1e300, xmax = -1e300, ymin = 1e300, ymax = -1e300;
    for (int n = 0; n < numNodes; n++) {
      xmin = std::min(xmin, nodeCoords[2 * n]);
      xmax = std::max(xmax, nodeCoords[2 * n]);
      ymin = std::min(ymin, nodeCoords[2 * n + 1]);
      ymax = std::max(ymax, nodeCoords[2 * n + 1]);
    }
    const double sx = xmax > xmin ? 65535.0 / (xmax - xmin) : 0.0;
    const double sy = ymax > ymin ? 65535.0 / (ymax - ymin) : 0.0;
    std::vector<uint64_t> key(numNodes);
    for (int n = 0; n < numNodes; n++) {
      const uint32_t qx = uint32_t(std::lround((nodeCoords[2 * n] - xmin) * sx));
      const uint32_t qy = uint32_t(std::lround((nodeCoords[2 * n + 1] - ymin) * sy));
      key[n] = (uint64_t(morton2(qx, qy)) << 32) | uint32_t(n);
    }
    std::sort(key.begin(), key.end());
    for (int n = 0; n < numNodes; n++) order_nodes[n] = int(key[n] & 0xFFFFFFFFu);
  }

  // Greedy cut along that order. A part closes when its rows would exceed the LDS block budget, the
  // row table, or the number of element visits one round of the workgroup's wavefronts covers.
  const double scale = std::min(1.0, std::max(partScale, 0.01));
  const int maxVisits = residualOnly ? (1 << 30) : std::max(1, int(cfg.maxVisits * scale));
  const int blockBudget = residualOnly ? maxBlocks : std::max(out.maxRowLength, int(maxBlocks * scale));
  const int rowBudget = residualOnly ? maxRows : std::max(1, int(maxRows * scale));
  std::vector<int> nptr, nadj;
  node_element_adjacency(numNodes, numElements, N, connPtr, conn, nptr, nadj);
  // Nodes that belong to no element keep zero blocks and zero residual entries. The cut runs over the other
  // nodes, so every part visits at least one element; each of them then joins the first part (in cut order)
  // with room for its row, whose write-out stores the zeros.
  std::vector<int> isolated;
  {
    std::vector<int> kept;
    kept.reserve(size_t(numNodes));
    for (int n : order_nodes) (nptr[n + 1] == nptr[n] ? isolated : kept).push_back(n);
    order_nodes.swap(kept);
  }
  const int numCut = int(order_nodes.size());
  std::vector<int> cutPtr{0}, partRowCount, partBlockCount;
  {
    std::vector<int> stamp(numElements, -1);
    int blocks = 0, rowsInPart = 0, visits = 0, part = 0;
    for (int k = 0; k < numCut; k++) {
      const int n = order_nodes[k];
      const int len = rowp[n + 1] - rowp[n];
      int newVisits = 0;
      for (int j = nptr[n]; j < nptr[n + 1]; j++) newVisits += stamp[nadj[size_t(j)]] != part;
      if (rowsInPart > 0 &&
          (blocks + len > blockBudget || rowsInPart + 1 > rowBudget || visits + newVisits > maxVisits)) {
        cutPtr.push_back(k);
        partRowCount.push_back(rowsInPart);
        partBlockCount.push_back(blocks);
        part++;
        blocks = 0;
        rowsInPart = 0;
        visits = 0;
        newVisits = nptr[n + 1] - nptr[n];
      }
      for (int j = nptr[n]; j < nptr[n + 1]; j++) stamp[nadj[size_t(j)]] = part;
      blocks += len;
      rowsInPart++;
      visits += newVisits;
    }
    cutPtr.push_back(numCut);
    partRowCount.push_back(rowsInPart);
    partBlockCount.push_back(blocks);
  }
  {
    const int cutParts = int(cutPtr.size()) - 1;
    std::vector<std::vector<int>> joined(static_cast<size_t>(cutParts));
    int p = 0;
    for (int n : isolated) {
      const int len = rowp[n + 1] - rowp[n];
      while (p < cutParts && (partRowCount[p] + 1 > rowBudget || partBlockCount[p] + len > blockBudget)) p++;
      if (p == cutParts) {
        return set_error(FEM_ERR_UNSUPPORTED, "no part has room for the row of node %d, which belongs to no element",
                         n);
      }
      joined[size_t(p)].push_back(n);
      partRowCount[p]++;
      partBlockCount[p] += len;
    }
    std::vector<int> placed;
    placed.reserve(size_t(numNodes));
    out.partRowPtr.assign(1, 0);
    for (int q = 0; q < cutParts; q++) {
      placed.insert(placed.end(), order_nodes.begin() + cutPtr[q], order_nodes.begin() + cutPtr[q + 1]);
      placed.insert(placed.end(), joined[size_t(q)].begin(), joined[size_t(q)].end());
      out.partRowPtr.push_back(int(placed.size()));
    }
    order_nodes.swap(placed);
  }
  out.numParts = int(out.partRowPtr.size()) - 1;
  const int P = out.numParts;
  out.rows = order_nodes;
  out.rowLbase.resize(numNodes);
  out.partBlocks.resize(P);
  out.nodeLrow.resize(numNodes);
  for (int p = 0; p < P; p++) {
    std::sort(out.rows.begin() + out.partRowPtr[p], out.rows.begin() + out.partRowPtr[p + 1]);
    int base = 0;
    for (int k = out.partRowPtr[p]; k < out.partRowPtr[p + 1]; k++) {
      const int n = out.rows[k];
      out.rowLbase[k] = base;
      out.nodeLrow[n] = k - out.partRowPtr[p];
      base += rowp[n + 1] - rowp[n];
    }
    out.partBlocks[p] = base;
  }
  // Runs of consecutive node ids: contiguous in both LDS and the BSR value array.
  out.partRunPtr.push_back(0);
  for (int p = 0; p < P; p++) {
    for (int k = out.partRowPtr[p]; k < out.partRowPtr[p + 1]; k++) {
      const int n = out.rows[k];
      const int len = rowp[n + 1] - rowp[n];
      const bool extend = k > out.partRowPtr[p] && out.rows[k - 1] == n - 1;
      if (extend) {
        out.runs[out.runs.size() - 2] += len;
      } else {
        out.runs.push_back(out.rowLbase[k]);
        out.runs.push_back(rowp[n]);
        out.runs.push_back(len);
        out.runs.push_back(0);
      }
    }
    out.partRunPtr.push_back(int(out.runs.size() / 4));
  }

  // Element colouring (used to order elements inside a part for deterministic LDS sums).
  std::vector<int> colour(numElements, 0);
  int numColours = 1;
  rc = fem_element_colouring(numNodes, numElements, N, connPtr, conn, colour.data(), &numColours);
  if (rc) return rc;
  out.numColours = numColours;

  // Elements touching each part, grouped by colour (deterministic plans sum colour by colour) and in a
  // fixed pseudo-random order within a colour. A wavefront takes consecutive elements of this order, and
  // their LDS adds then land on less regular bank patterns than for neighbouring ids of a structured mesh
  // (which repeat the same column ranks at regular row offsets).
  std::vector<std::vector<int>> partElems(P);
  parallel_for(P, [&](int64_t b, int64_t e) {
    std::vector<int> tmp;
    for (int64_t p = b; p < e; p++) {
      tmp.clear();
      for (int k = out.partRowPtr[p]; k < out.partRowPtr[p + 1]; k++) {
        const int n = out.rows[k];
        tmp.insert(tmp.end(), nadj.begin() + nptr[n], nadj.begin() + nptr[n + 1]);
      }
      std::sort(tmp.begin(), tmp.end());
      tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
      std::sort(tmp.begin(), tmp.end(), [&](int x, int y) {
        return colour[x] != colour[y] ? colour[x] < colour[y] : scramble(uint32_t(x)) < scramble(uint32_t(y));
      });
      partElems[p] = tmp;
    }
  });
  out.partElemPtr.assign(size_t(P) + 1, 0);
  out.colourOff.assign(size_t(P) * (numColours + 1), 0);
  for (int p = 0; p < P; p++) {
    out.partElemPtr[p + 1] = out.partElemPtr[p] + int(partElems[p].size());
    int *off = &out.colourOff[size_t(p) * (numColours + 1)];
    for (int el : partElems[p]) off[colour[el] + 1]++;
    for (int c = 0; c < numColours; c++) off[c + 1] += off[c];
  }
  out.partDesc.assign(size_t(P) * 8, 0);
  for (int p = 0; p < P; p++) {
    int *dsc = &out.partDesc[size_t(p) * 8];
    dsc[0] = out.partRowPtr[p];
    dsc[1] = out.partRowPtr[p + 1] - out.partRowPtr[p];
    dsc[2] = out.partBlocks[p];
    dsc[3] = out.partRunPtr[p];
    dsc[4] = out.partRunPtr[p + 1] - out.partRunPtr[p];
    dsc[5] = out.partElemPtr[p];
    dsc[6] = out.partElemPtr[p + 1] - out.partElemPtr[p];
    out.maxPartVisits = std::max(out.maxPartVisits, dsc[6]);
  }
  const size_t visits = size_t(out.partElemPtr[P]);
  out.visitEid.reserve(visits);
  for (int p = 0; p < P; p++) out.visitEid.insert(out.visitEid.end(), partElems[p].begin(), partElems[p].end());
  out.visitConn.resize(visits * N);
  parallel_for(int64_t(visits), [&](int64_t b, int64_t e) {
    for (int64_t v = b; v < e; v++) {
      std::memcpy(&out.visitConn[size_t(v) * N], conn + connPtr[out.visitEid[size_t(v)]], N * sizeof(int));
    }
  });
  return FEM_OK;
}

namespace {

// Shared by the matrix and residual halves of build_split_host: from the contributions' destinations
// (dest(i), i = 0..n-1, each in [0, outputs)), writes code[i] (the destination if it is its output's only
// contribution, else outputs + a slot) and the gather list of the other outputs. Slots are grouped by output,
// ascending, and within an output ordered by contribution index.
template <typename Dest>
int split_outputs(int64_t n, int64_t outputs, Dest &&dest, std::vector<int> &code, std::vector<int> &dst,
                  std::vector<int> &ptr, int64_t &slots) {
  std::vector<int> next(size_t(outputs), 0);  // first a count per output, then its next free slot
  for (int64_t i = 0; i < n; i++) next[size_t(dest(i))]++;
  int64_t running = 0;
  ptr.assign(1, 0);
  dst.clear();
  for (int64_t d = 0; d < outputs; d++) {
    const int c = next[size_t(d)];
    if (c == 1) {
      next[size_t(d)] = -1;  // exclusive
      continue;
    }
    next[size_t(d)] = int(running);
    running += c;
    if (outputs + running > INT32_MAX) {
      return set_error(FEM_ERR_UNSUPPORTED, "%lld outputs and %lld staged contributions exceed int codes",
                       (long long)outputs, (long long)running);
    }
    dst.push_back(int(d));
    ptr.push_back(int(running));
  }
  slots = running;
  code.resize(size_t(n));
  for (int64_t i = 0; i < n; i++) {
    const int d = dest(i);
    int &s = next[size_t(d)];
    code[size_t(i)] = s < 0 ? d : int(outputs) + s++;
  }
  return FEM_OK;
}

}  // namespace

int build_split_host(int order, int numNodes, int numElements, const int *conn, const int *rowp,
                     const std::vector<uint8_t> &rank, int64_t numBlocks, SplitHost &out) {
  const int N = nodes_per_element(order);
  const int RP = rank_pad(order);
  out = SplitHost();
  // block (e, a, b), row-major: rows of node conn[e,a], column rank rank[(e*N + b)*RP + a]
  auto blockDest = [&](int64_t i) {
    const int64_t e = i / (N * N);
    const int a = int(i / N % N), b = int(i % N);
    return rowp[conn[e * N + a]] + int(rank[size_t((e * N + b) * RP + a)]);
  };
  int rc = split_outputs(int64_t(numElements) * N * N, numBlocks, blockDest, out.code, out.dst, out.ptr, out.slots);
  if (rc) return rc;
  auto nodeDest = [&](int64_t i) { return conn[i]; };
  return split_outputs(int64_t(numElements) * N, numNodes, nodeDest, out.rcode, out.rdst, out.rptr, out.rslots);
}

}  // namespace fem

using namespace fem;

// ---- C ABI -----------------------------------------------------------------------------------------
extern "C" {

const char *fem_last_error(void) { return g_lastError.c_str(); }

int fem_version(void) { return 100; }

int fem_basis_table(int order, double *phi, double *dphi, double *gaussPoints, double *gaussWeights) {
  if (!valid_order(order)) return set_error(FEM_ERR_INVALID_ORDER, "element order %d outside 1..4", order);
  if (!phi || !dphi || !gaussPoints || !gaussWeights) return set_error(FEM_ERR_NULL_POINTER, "NULL output");
  const int np = order + 1;
  for (int i = 0; i < np; i++) {
    double x = 0, w = 0;
    switch (order) {
      case 1: x = getGaussQuadCoord<1>(i); w = getGaussQuadWeight<1>(i); break;
      case 2: x = getGaussQuadCoord<2>(i); w = getGaussQuadWeight<2>(i); break;
      case 3: x = getGaussQuadCoord<3>(i); w = getGaussQuadWeight<3>(i); break;
      default: x = getGaussQuadCoord<4>(i); w = getGaussQuadWeight<4>(i); break;
    }
    gaussPoints[i] = x;
    gaussWeights[i] = w;
    for (int k = 0; k < np; k++) {
      double N = 0, dN = 0;
      switch (order) {
        case 1: lagrangePoly1dDeriv<double, 1>(x, k, N, dN); break;
        case 2: lagrangePoly1dDeriv<double, 2>(x, k, N, dN); break;
        case 3: lagrangePoly1dDeriv<double, 3>(x, k, N, dN); break;
        default: lagrangePoly1dDeriv<double, 4>(x, k, N, dN); break;
      }
      phi[k * np + i] = N;
      dphi[k * np + i] = dN;
    }
  }
  return FEM_OK;
}

int fem_square_mesh_size(int nx, int ny, int order, int64_t *numNodes, int64_t *numElements) {
  if (!valid_order(order)) return set_error(FEM_ERR_INVALID_ORDER, "element order %d outside 1..4", order);
  if (nx < 1 || ny < 1) return set_error(FEM_ERR_INVALID_ARGUMENT, "nx=%d ny=%d must be >= 1", nx, ny);
  if (!numNodes || !numElements) return set_error(FEM_ERR_NULL_POINTER, "NULL output");
  *numNodes = int64_t(order * int64_t(nx) + 1) * (order * int64_t(ny) + 1);
  *numElements = int64_t(nx) * ny;
  if (*numNodes > INT32_MAX || *numElements * nodes_per_element(order) > INT32_MAX) {
    return set_error(FEM_ERR_INVALID_ARGUMENT, "mesh too large for 32-bit node/connectivity indices");
  }
  return FEM_OK;
}

int fem_square_mesh(int nx, int ny, int order, double jitter, uint64_t seed, double *nodeCoords, int *conn) {
  int64_t nn, ne;
  int rc = fem_square_mesh_size(nx, ny, order, &nn, &ne);
  if (rc) return rc;
  if (!nodeCoords || !conn) return set_error(FEM_ERR_NULL_POINTER, "NULL output");
  if (!(jitter >= 0.0 && jitter < 0.5)) return set_error(FEM_ERR_INVALID_ARGUMENT, "jitter %g outside [0, 0.5)", jitter);
  const int p = order, mx = p * nx + 1, my = p * ny + 1, NP = p + 1, N = NP * NP;
  const std::vector<double> kx = axis_knots(nx, p), ky = axis_knots(ny, p);
  SplitMix64 rng(seed);
  for (int jy = 0; jy < my; jy++) {
    for (int jx = 0; jx < mx; jx++) {
      double x = kx[jx], y = ky[jy];
      if (jitter > 0 && jx > 0 && jx < mx - 1 && jy > 0 && jy < my - 1) {
        const double dx = rng.uniform(-1.0, 1.0) * jitter * knot_gap(kx, jx, false, 1.0);
        const double dy = rng.uniform(-1.0, 1.0) * jitter * knot_gap(ky, jy, false, 1.0);
        x += dx;
        y += dy;
      }
      const int64_t n = int64_t(jy) * mx + jx;
      nodeCoords[2 * n] = x;
      nodeCoords[2 * n + 1] = y;
    }
  }
  for (int ey = 0; ey < ny; ey++) {
    for (int ex = 0; ex < nx; ex++) {
      const int64_t e = int64_t(ey) * nx + ex;
      for (int ay = 0; ay < NP; ay++) {
        for (int ax = 0; ax < NP; ax++) {
          conn[e * N + ay * NP + ax] = (ey * p + ay) * mx + (ex * p + ax);
        }
      }
    }
  }
  return FEM_OK;
}

int fem_annulus_mesh_size(int nr, int ntheta, int order, int64_t *numNodes, int64_t *numElements) {
  if (!valid_order(order)) return set_error(FEM_ERR_INVALID_ORDER, "element order %d outside 1..4", order);
  if (nr < 1 || ntheta < 3) return set_error(FEM_ERR_INVALID_ARGUMENT, "nr=%d must be >= 1, ntheta=%d >= 3", nr, ntheta);
  if (!numNodes || !numElements) return set_error(FEM_ERR_NULL_POINTER, "NULL output");
  *numNodes = int64_t(order * int64_t(nr) + 1) * (order * int64_t(ntheta));
  *numElements = int64_t(nr) * ntheta;
  if (*numNodes > INT32_MAX || *numElements * nodes_per_element(order) > INT32_MAX) {
    return set_error(FEM_ERR_INVALID_ARGUMENT, "mesh too large for 32-bit node/connectivity indices");
  }
  return FEM_OK;
}

int fem_annulus_mesh(int nr, int ntheta, int order, double rIn, double rOut, double jitter, uint64_t seed,
                     double *nodeCoords, int *conn) {
  int64_t nn, ne;
  int rc = fem_annulus_mesh_size(nr, ntheta, order, &nn, &ne);
  if (rc) return rc;
  if (!nodeCoords || !conn) return set_error(FEM_ERR_NULL_POINTER, "NULL output");
  if (!(rIn > 0.0 && rOut > rIn)) return set_error(FEM_ERR_INVALID_ARGUMENT, "need 0 < rIn < rOut");
  if (!(jitter >= 0.0 && jitter < 0.5)) return set_error(FEM_ERR_INVALID_ARGUMENT, "jitter %g outside [0, 0.5)", jitter);
  const int p = order, mr = p * nr + 1, mt = p * ntheta, NP = p + 1, N = NP * NP;
  const std::vector<double> kr = axis_knots(nr, p);
  std::vector<double> kt = axis_knots(ntheta, p);  // in [0, 1]; last knot (== 1) is the seam, dropped
  kt.pop_back();
  SplitMix64 rng(seed);
  for (int jt = 0; jt < mt; jt++) {
    for (int jr = 0; jr < mr; jr++) {
      double s = kr[jr], u = kt[jt];
      if (jitter > 0 && jr > 0 && jr < mr - 1) {
        const double ds = rng.uniform(-1.0, 1.0) * jitter * knot_gap(kr, jr, false, 1.0);
        const double du = rng.uniform(-1.0, 1.0) * jitter * knot_gap(kt, jt, true, 1.0);
        s += ds;
        u += du;
      }
      const double r = rIn + (rOut - rIn) * s;
      const double th = 2.0 * M_PI * u;
      const int64_t n = int64_t(jt) * mr + jr;
      nodeCoords[2 * n] = r * std::cos(th);
      nodeCoords[2 * n + 1] = r * std::sin(th);
    }
  }
  for (int et = 0; et < ntheta; et++) {
    for (int er = 0; er < nr; er++) {
      const int64_t e = int64_t(et) * nr + er;
      for (int ay = 0; ay < NP; ay++) {
        for (int ax = 0; ax < NP; ax++) {
          conn[e * N + ay * NP + ax] = ((et * p + ay) % mt) * mr + (er * p + ax);
        }
      }
    }
  }
  return FEM_OK;
}

int fem_fill_uniform(int64_t n, double lo, double hi, uint64_t seed, double *out) {
  if (n < 0) return set_error(FEM_ERR_INVALID_ARGUMENT, "n=%lld < 0", (long long)n);
  if (n > 0 && !out) return set_error(FEM_ERR_NULL_POINTER, "NULL output");
  SplitMix64 rng(seed);
  for (int64_t i = 0; i < n; i++) out[i] = rng.uniform(lo, hi);
  return FEM_OK;
}

int fem_fill_smooth_field(int64_t numNodes, const double *nodeCoords, double amplitude, uint64_t phaseSeed,
                          double pert, uint64_t pertSeed, double *states) {
  if (numNodes < 0) return set_error(FEM_ERR_INVALID_ARGUMENT, "numNodes < 0");
  if (numNodes > 0 && (!nodeCoords || !states)) return set_error(FEM_ERR_NULL_POINTER, "NULL pointer");
  double ph[4] = {0, 0, 0, 0};
  if (phaseSeed != 0) {
    SplitMix64 rng(phaseSeed);
    for (double &v : ph) v = rng.uniform(0.0, 2.0 * M_PI);
  }
  SplitMix64 prng(pertSeed);
  for (int64_t n = 0; n < numNodes; n++) {
    const double x = nodeCoords[2 * n], y = nodeCoords[2 * n + 1];
    double u = amplitude * (std::sin(2.0 * M_PI * x + ph[0]) + std::sin(2.0 * M_PI * y + ph[1]));
    double v = amplitude * (std::cos(2.0 * M_PI * x + ph[2]) + std::cos(2.0 * M_PI * y + ph[3]));
    if (pert > 0) {
      u += prng.uniform(-pert, pert);
      v += prng.uniform(-pert, pert);
    }
    states[2 * n] = u;
    states[2 * n + 1] = v;
  }
  return FEM_OK;
}

int fem_min_detj(int order, int numElements, const int *connPtr, const int *conn, const double *nodeCoords,
                 double *minDetJ) {
  if (!valid_order(order)) return set_error(FEM_ERR_INVALID_ORDER, "element order %d outside 1..4", order);
  if (!minDetJ || (numElements > 0 && (!connPtr || !conn || !nodeCoords))) {
    return set_error(FEM_ERR_NULL_POINTER, "NULL pointer");
  }
  const int np = order + 1;
  double phi[25], dphi[25], xg[5], wg[5];
  fem_basis_table(order, phi, dphi, xg, wg);
  std::vector<double> mins(size_t(hardware_threads()) + 1, 1e300);
  std::atomic<int> slot{0};
  parallel_for(numElements, [&](int64_t b, int64_t e) {
    double m = 1e300;
    for (int64_t el = b; el < e; el++) {
      const int *en = conn + connPtr[el];
      for (int qy = 0; qy < np; qy++) {
        for (int qx = 0; qx < np; qx++) {
          double J[4] = {0, 0, 0, 0};
          for (int ay = 0; ay < np; ay++) {
            for (int ax = 0; ax < np; ax++) {
              const double gx = phi[ay * np + qy] * dphi[ax * np + qx];
              const double gy = phi[ax * np + qx] * dphi[ay * np + qy];
              const int n = en[ay * np + ax];
              J[0] += nodeCoords[2 * n] * gx;
              J[1] += nodeCoords[2 * n] * gy;
              J[2] += nodeCoords[2 * n + 1] * gx;
              J[3] += nodeCoords[2 * n + 1] * gy;
            }
          }
          m = std::min(m, J[0] * J[3] - J[1] * J[2]);
        }
      }
    }
    mins[size_t(slot.fetch_add(1))] = m;
  });
  *minDetJ = *std::min_element(mins.begin(), mins.end());
  return FEM_OK;
}

int fem_bsr_pattern_rowp(int numNodes, int numElements, int nodesPerElement, const int *connPtr, const int *conn,
                         int *rowp) {
  int rc = check_connectivity(numNodes, numElements, nodesPerElement, connPtr, conn);
  if (rc) return rc;
  if (!rowp) return set_error(FEM_ERR_NULL_POINTER, "rowp is NULL");
  std::vector<int> ptr, adj;
  node_element_adjacency(numNodes, numElements, nodesPerElement, connPtr, conn, ptr, adj);
  std::vector<int64_t> len(size_t(numNodes) + 1, 0);
  parallel_for(numNodes, [&](int64_t b, int64_t e) {
    std::vector<int> tmp;
    for (int64_t n = b; n < e; n++) {
      row_columns(int(n), nodesPerElement, connPtr, conn, ptr, adj, tmp);
      len[size_t(n) + 1] = int64_t(tmp.size());
    }
  });
  for (int n = 0; n < numNodes; n++) len[n + 1] += len[n];
  if (len[numNodes] > INT32_MAX) return set_error(FEM_ERR_INVALID_ARGUMENT, "pattern exceeds 2^31 blocks");
  for (int n = 0; n <= numNodes; n++) rowp[n] = int(len[n]);
  return FEM_OK;
}

int fem_bsr_pattern_cols(int numNodes, int numElements, int nodesPerElement, const int *connPtr, const int *conn,
                         const int *rowp, int *cols) {
  int rc = check_connectivity(numNodes, numElements, nodesPerElement, connPtr, conn);
  if (rc) return rc;
  if (!rowp || !cols) return set_error(FEM_ERR_NULL_POINTER, "rowp/cols is NULL");
  std::vector<int> ptr, adj;
  node_element_adjacency(numNodes, numElements, nodesPerElement, connPtr, conn, ptr, adj);
  std::atomic<int> bad{-1};
  parallel_for(numNodes, [&](int64_t b, int64_t e) {
    std::vector<int> tmp;
    for (int64_t n = b; n < e; n++) {
      row_columns(int(n), nodesPerElement, connPtr, conn, ptr, adj, tmp);
      if (int(tmp.size()) != rowp[n + 1] - rowp[n]) {
        bad.store(int(n));
        return;
      }
      std::copy(tmp.begin(), tmp.end(), cols + rowp[n]);
    }
  });
  if (bad.load() >= 0) return set_error(FEM_ERR_INVALID_ARGUMENT, "rowp does not match the mesh at row %d", bad.load());
  return FEM_OK;
}

int fem_split_tables(int order, int numNodes, int numElements, const int *connPtr, const int *conn, const int *rowp,
                     const int *cols, int *code, int *rcode, int *gdst, int *gptr, int *rdst, int *rptr,
                     int64_t *numSlots, int64_t *numGather, int64_t *numRslots, int64_t *numRgather) {
  if (!valid_order(order)) return set_error(FEM_ERR_INVALID_ORDER, "element order %d outside 1..4", order);
  if (!rowp || !cols || !code || !rcode || !gdst || !gptr || !rdst || !rptr) {
    return set_error(FEM_ERR_NULL_POINTER, "an input or output array is NULL");
  }
  const int N = nodes_per_element(order), RP = rank_pad(order);
  int rc = check_connectivity(numNodes, numElements, N, connPtr, conn);
  if (rc) return rc;
  for (int n = 0; n < numNodes; n++) {
    if (rowp[n + 1] - rowp[n] > 256) return set_error(FEM_ERR_UNSUPPORTED, "BSR row %d exceeds 256 blocks", n);
  }
  // element-major connectivity and the plan's column ranks, as fem_plan_create forms them
  std::vector<int> c(size_t(numElements) * N);
  std::vector<uint8_t> rank(size_t(numElements) * N * RP, 0);
  for (int e = 0; e < numElements; e++) {
    const int *en = conn + connPtr[e];
    std::copy(en, en + N, c.begin() + size_t(e) * N);
    for (int i = 0; i < N; i++) {
      for (int j = 0; j < N; j++) {
        const int pos = find_sorted(cols, rowp[en[i]], rowp[en[i] + 1], en[j]);
        if (pos < 0) return set_error(FEM_ERR_MAP_MISS, "element %d couples nodes with no block in the pattern", e);
        rank[(size_t(e) * N + j) * RP + i] = uint8_t(pos - rowp[en[i]]);
      }
    }
  }
  SplitHost sp;
  if ((rc = build_split_host(order, numNodes, numElements, c.data(), rowp, rank, rowp[numNodes], sp))) return rc;
  std::copy(sp.code.begin(), sp.code.end(), code);
  std::copy(sp.rcode.begin(), sp.rcode.end(), rcode);
  std::copy(sp.dst.begin(), sp.dst.end(), gdst);
  std::copy(sp.ptr.begin(), sp.ptr.end(), gptr);
  std::copy(sp.rdst.begin(), sp.rdst.end(), rdst);
  std::copy(sp.rptr.begin(), sp.rptr.end(), rptr);
  if (numSlots) *numSlots = sp.slots;
  if (numGather) *numGather = int64_t(sp.dst.size());
  if (numRslots) *numRslots = sp.rslots;
  if (numRgather) *numRgather = int64_t(sp.rdst.size());
  return FEM_OK;
}

int fem_element_bcsr_map(const int *connPtr, const int *conn, int numElements, int nodesPerElement, int numRows,
                         const int *rowp, const int *cols, int bsize, int *map) {
  if (!rowp || !cols || !map) return set_error(FEM_ERR_NULL_POINTER, "rowp/cols/map is NULL");
  if (bsize < 1) return set_error(FEM_ERR_INVALID_ARGUMENT, "bsize %d < 1", bsize);
  int rc = check_connectivity(numRows, numElements, nodesPerElement, connPtr, conn);
  if (rc) return rc;
  const int N = nodesPerElement, bs2 = bsize * bsize;
  if (int64_t(rowp[numRows]) * bs2 > INT32_MAX) {
    return set_error(FEM_ERR_UNSUPPORTED, "%d blocks x %d entries overflow the map's 32-bit offsets", rowp[numRows],
                     bs2);
  }
  // Rows are sorted in TACS patterns; binary search when they are, linear search otherwise.
  bool sorted = true;
  for (int r = 0; r < numRows && sorted; r++) {
    for (int k = rowp[r] + 1; k < rowp[r + 1]; k++) {
      if (cols[k - 1] >= cols[k]) {
        sorted = false;
        break;
      }
    }
  }
  std::atomic<int64_t> miss{-1};
  parallel_for(numElements, [&](int64_t b, int64_t e) {
    for (int64_t el = b; el < e; el++) {
      const int *en = conn + connPtr[el];
      for (int i = 0; i < N; i++) {
        const int r = en[i];
        for (int j = 0; j < N; j++) {
          int pos = -1;
          if (sorted) {
            pos = find_sorted(cols, rowp[r], rowp[r + 1], en[j]);
          } else {
            for (int k = rowp[r]; k < rowp[r + 1]; k++) {
              if (cols[k] == en[j]) {
                pos = k;
                break;
              }
            }
          }
          if (pos < 0) {
            miss.store(el * int64_t(N * N) + i * N + j);
            return;
          }
          map[el * int64_t(N * N) + i * N + j] = pos * bs2;
        }
      }
    }
  });
  if (miss.load() >= 0) {
    const int64_t m = miss.load();
    const int el = int(m / (N * N)), i = int(m % (N * N)) / N, j = int(m % N);
    return set_error(FEM_ERR_MAP_MISS, "no block for column %d in row %d (element %d, nodes %d,%d)",
                     conn[connPtr[el] + j], conn[connPtr[el] + i], el, i, j);
  }
  return FEM_OK;
}

// One value per line, "% .17g" (writeArrayToFile, verification/TACSHelpers.h:234-246).
int fem_write_array(const double *values, int64_t n, const char *path) {
  if (n < 0) return set_error(FEM_ERR_INVALID_ARGUMENT, "n = %lld is negative", (long long)n);
  if (!path || (n > 0 && !values)) return set_error(FEM_ERR_NULL_POINTER, "values/path is NULL");
  FILE *fp = std::fopen(path, "w");
  if (!fp) return set_error(FEM_ERR_INVALID_ARGUMENT, "cannot open %s for writing", path);
  for (int64_t i = 0; i < n; i++) std::fprintf(fp, "% .17g\n", values[i]);
  const bool failed = std::ferror(fp) != 0;
  if (std::fclose(fp) != 0 || failed) return set_error(FEM_ERR_INVALID_ARGUMENT, "writing %s failed", path);
  return FEM_OK;
}

// The reference's matrix file (writeBCSRMatToFile, verification/TACSHelpers.h:248-307): a header
// "rows cols nnz" in scalar entries, then one "row col value" line per entry of each block, 0-based, blocks
// row-major, value "% .17g". maxRows / maxCols > 0 keep the leading block rows / block columns only (a row's
// columns are sorted, so its blocks stop at the first column beyond the limit, as in the reference).
int fem_write_bsr_matrix(int nrows, int ncols, int bsize, const int *rowp, const int *cols, const double *values,
                         int maxRows, int maxCols, const char *path) {
  if (nrows < 0 || ncols < 0 || bsize < 1) {
    return set_error(FEM_ERR_INVALID_ARGUMENT, "nrows=%d ncols=%d bsize=%d", nrows, ncols, bsize);
  }
  if (!path || !rowp || (rowp[nrows] > 0 && (!cols || !values))) {
    return set_error(FEM_ERR_NULL_POINTER, "rowp/cols/values/path is NULL");
  }
  const int64_t bs2 = int64_t(bsize) * bsize;
  int64_t nnz = int64_t(rowp[nrows]) * bs2;
  const bool limited = maxRows > 0 || maxCols > 0;
  if (maxRows > 0) nrows = std::min(maxRows, nrows);
  if (maxCols > 0) ncols = std::min(maxCols, ncols);
  if (limited) {
    nnz = 0;
    for (int i = 0; i < nrows; i++) {
      for (int k = rowp[i]; k < rowp[i + 1] && cols[k] < ncols; k++) nnz += bs2;
    }
  }
  FILE *fp = std::fopen(path, "w");
  if (!fp) return set_error(FEM_ERR_INVALID_ARGUMENT, "cannot open %s for writing", path);
  std::fprintf(fp, "%lld %lld %lld\n", (long long)nrows * bsize, (long long)ncols * bsize, (long long)nnz);
  for (int i = 0; i < nrows; i++) {
    for (int k = rowp[i]; k < rowp[i + 1] && cols[k] < ncols; k++) {
      const double *block = values + int64_t(k) * bs2;
      for (int r = 0; r < bsize; r++) {
        for (int c = 0; c < bsize; c++) {
          std::fprintf(fp, "%lld %lld % .17g\n", (long long)i * bsize + r, (long long)cols[k] * bsize + c,
                       block[r * bsize + c]);
        }
      }
    }
  }
  const bool failed = std::ferror(fp) != 0;
  if (std::fclose(fp) != 0 || failed) return set_error(FEM_ERR_INVALID_ARGUMENT, "writing %s failed", path);
  return FEM_OK;
}

int fem_element_colouring(int numNodes, int numElements, int nodesPerElement, const int *connPtr, const int *conn,
                          int *colour, int *numColours) {
  int rc = check_connectivity(numNodes, numElements, nodesPerElement, connPtr, conn);
  if (rc) return rc;
  if (!colour || !numColours) return set_error(FEM_ERR_NULL_POINTER, "NULL output");
  std::vector<int> ptr, adj;
  node_element_adjacency(numNodes, numElements, nodesPerElement, connPtr, conn, ptr, adj);
  std::fill(colour, colour + numElements, -1);
  int maxc = 0;
  std::vector<char> used;
  for (int e = 0; e < numElements; e++) {
    used.assign(size_t(maxc) + 2, 0);
    const int *en = conn + connPtr[e];
    for (int a = 0; a < nodesPerElement; a++) {
      const int n = en[a];
      for (int k = ptr[n]; k < ptr[n + 1]; k++) {
        const int c = colour[adj[size_t(k)]];
        if (c >= 0) used[size_t(c)] = 1;
      }
    }
    int c = 0;
    while (used[size_t(c)]) c++;
    colour[e] = c;
    maxc = std::max(maxc, c + 1);
  }
  *numColours = numElements > 0 ? maxc : 0;
  return FEM_OK;
}

}  // extern "C"
