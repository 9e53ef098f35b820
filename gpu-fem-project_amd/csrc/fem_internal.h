// Internal declarations shared by the host code (fem_host.cpp) and the HIP code (fem_device.hip).
#pragma once

#include <cstdint>
#include <vector>

#include "fem_amd.h"

namespace fem {

// Records `code` and a printf-style message as the calling thread's last error; returns `code`.
int set_error(int code, const char *fmt, ...);
void clear_error();

constexpr int nodes_per_element(int order) { return (order + 1) * (order + 1); }
// Bytes of column rank stored per (element, column node): N rounded up to whole 32-bit words.
constexpr int rank_pad(int order) { return (nodes_per_element(order) + 3) / 4 * 4; }

// Launch geometry of the owner-computes kernel, shared by the host partitioner (which sizes parts
// to fit) and the kernel templates (which size their LDS from it).
//   waves     wavefronts per workgroup
//   maxBlocks BSR blocks a part may own (FP64 LDS accumulator, 32 B per block)
//   maxRows   block rows a part may own
//   maxVisits element visits per part (waves x elements per wavefront: one batch per wave)
struct OwnerConfig {
  int waves;
  int maxBlocks;
  int maxRows;
  int maxVisits;
};
// Orders 1-2: one workgroup of 8 waves per CU (two per SIMD) sharing an accumulator of 2720 FP64 blocks,
// so parts are twice as large as with two 4-wave workgroups and fewer elements are visited twice
// (1.53 visits per element at 1000x1000 order 2, against 1.83). Orders 3-4: two workgroups of 4 waves
// per CU, each with half the LDS.
constexpr OwnerConfig owner_config(int order) {
  return order == 1   ? OwnerConfig{8, 2720, 352, 128}
         : order == 2 ? OwnerConfig{8, 2720, 320, 56}
         : order == 3 ? OwnerConfig{4, 1312, 96, 16}
                      : OwnerConfig{4, 1312, 64, 8};
}
// Residual-only parts: only the per-row residual lives in LDS.
constexpr int kResidualOnlyMaxRows = 1024;

// Host-side result of partitioning a mesh into owner parts.
struct PlanHost {
  int order = 0;
  int numNodes = 0;
  int numElements = 0;
  int numParts = 0;
  int numColours = 1;
  int maxRowLength = 0;
  int64_t numBlocks = 0;
  std::vector<int> visitConn;     // [visits*N] nodes of each element visit (parts' visits contiguous)
  std::vector<int> visitEid;      // [visits] element id of each visit
  std::vector<int> nodeLrow;      // [numNodes] index of the node within its part's row list
  std::vector<int> partRowPtr;    // [P+1]
  std::vector<int> rows;          // global row (node) ids, ascending within a part
  std::vector<int> rowLbase;      // LDS block offset of each owned row within its part
  std::vector<int> partBlocks;    // [P] blocks owned by the part
  std::vector<int> partRunPtr;    // [P+1]
  std::vector<int> runs;          // [R*4] (lds block start, global block start, blocks, 0)
  std::vector<int> partElemPtr;   // [P+1] visit range of each part; visits grouped by colour
  std::vector<int> colourOff;     // [P*(C+1)] offsets of each colour within a part's visit list
  std::vector<uint8_t> rank;      // [E][N][rank_pad] rank[e][j][i]: slot of column conn[e,j] in row conn[e,i]
  std::vector<int> partDesc;      // [P*8] (first row, rows, blocks, first run, runs, first visit, visits, 0)
  int maxPartVisits = 0;
};

// Element-once assembly tables (orders 3-4). Every output (a BSR block, or a node's residual pair) is either
// exclusive, formed by exactly one (element, node[, node]) contribution, or not. An exclusive output is
// stored by its one contributor; the contributions to any other output go to slots of a staging array,
// grouped by output, and a gather pass stores each such output as the sum of its slots in slot order (zero
// when no element touches it). Every output is written exactly once and the result is bitwise
// reproducible.
struct SplitHost {
  std::vector<int> code;   // [E][N][N], row-major (a, b): block of rows conn[e,a] x column conn[e,b]; below
                           // numBlocks the block's BSR index (exclusive), else numBlocks + staging slot
  std::vector<int> rcode;  // [E][N]: conn[e,a] if the node's residual is exclusive, else numNodes + slot
  std::vector<int> dst;    // BSR indices of the non-exclusive blocks, ascending
  std::vector<int> ptr;    // [dst.size() + 1] slot range of each
  std::vector<int> rdst;   // non-exclusive residual nodes, ascending
  std::vector<int> rptr;   // [rdst.size() + 1]
  int64_t slots = 0, rslots = 0;
};

// Builds the element-once tables from the connectivity ([E][N], element-major), the BSR row starts and a
// plan's column ranks. Returns FEM_ERR_UNSUPPORTED if the codes would not fit in an int.
int build_split_host(int order, int numNodes, int numElements, const int *conn, const int *rowp,
                     const std::vector<uint8_t> &rank, int64_t numBlocks, SplitHost &out);

// partScale (0, 1] scales the part budgets (blocks, rows, visits) below the kernel's maxima: smaller
// parts give small meshes more workgroups at the price of more elements recomputed on part boundaries.
int build_plan_host(int order, int numNodes, int numElements, const int *connPtr, const int *conn,
                    const int *rowp, const int *cols, const double *nodeCoords, int flags, PlanHost &out,
                    double partScale = 1.0);

// Batches (fem_device.hip). Doubles of element-once staging a Jacobian job of `plan` takes (0: the plan has no
// element-once tables); fem_batch_add with that staging owned by the caller (`stage`, NULL: the batch's own
// arena, grown at assembly); whether a plan's parts fit the streaming kernel (graph batches need it).
int64_t batch_stage_doubles(const fem_plan *plan, bool withResidual);
int batch_add_staged(fem_batch *batch, fem_plan *plan, const double *states, const double *nodeCoords, double E,
                     double nu, double t, double *residual, double *matEntries, double *stage);
bool plan_streams_public(const fem_plan *plan);
// Records a graph batch's launch graph for `precision` now (it does not depend on the batch's contents), so that
// no capture is in flight later, when other threads may be using the device, and (before the batch's first
// assembly) replays it once with no jobs, so the runtime's first-launch allocations happen now too. No-op without
// FEM_BATCH_GRAPH.
int batch_prepare_graph(fem_batch *batch, int precision);
// fem_plan_assemble (accumulate false) and fem_plan_assemble_accumulate (true).
int plan_assemble(fem_plan *plan, const double *states, const double *nodeCoords, double E, double nu, double t,
                  double *residual, double *matEntries, int precision, void *stream, double *elapsedSeconds,
                  bool accumulate);

}  // namespace fem

// Opaque plan handle (device-resident arrays built from a PlanHost).
struct fem_plan {
  int order = 0;
  int numNodes = 0;
  int numElements = 0;
  int numParts = 0;
  int numColours = 1;
  int flags = 0;
  int device = 0;
  int64_t numBlocks = 0;
  int64_t elementVisits = 0;
  int64_t deviceBytes = 0;
  int *dVisitConn = nullptr;
  int *dVisitEid = nullptr;
  int *dNodeLrow = nullptr;
  int *dRows = nullptr;
  int *dRowLbase = nullptr;
  int *dRuns = nullptr;
  int *dColourOff = nullptr;
  int *dPartDesc = nullptr;
  uint8_t *dRank = nullptr;
  int *dJobStart = nullptr;  // {0, numParts}
  // Element-once assembly of Jacobian plans of orders 3-4 (SplitHost): connectivity, codes, gather lists and
  // staging arrays. dConn is NULL when the plan has none.
  int *dConn = nullptr;      // [E][N]
  int *dCode = nullptr;      // [E][N][N]
  int *dRcode = nullptr;     // [E][N]
  int *dGdst = nullptr, *dGptr = nullptr, *dRdst = nullptr, *dRptr = nullptr;
  double *dStage = nullptr;  // [4 * slots]
  double *dRstage = nullptr; // [2 * rslots]
  int numGather = 0, numRgather = 0;
  int64_t numSlots = 0, numRslots = 0;  // staged contributions: blocks, residual entries
  double partScale = 1.0;    // part budgets relative to the kernel's (FEM_PLAN_AUTOTUNE picks it)
  double tunedSeconds = 0;   // measured assembly time of the chosen build (autotuned plans)
  int tiledWaves = 4;        // orders 3-4: wavefronts per workgroup of the element-once kernel (autotuned)
  void *dJob = nullptr;      // device job record of the last assembly
  alignas(8) unsigned char hostJob[256] = {};  // host copy of that record
  bool jobUploaded = false;
  int maxPartVisits = 0;
  int numCUs = 0;
};
