/*
 * C ABI of the MI355X plane-stress assembly library (libfem_amd.so).
 *
 * Element formulation: 2-D linear-elastic plane stress (LINEAR strain) on tensor-product Lagrange
 * quadrilaterals of order 1..4 (N = (order+1)^2 nodes, (order+1)^2 Gauss points), 2 states per node.
 *
 * Data layout (identical to the reference's, element/FEKernels.h and verification/TestKernels.cpp):
 *   nodeCoords  double [numNodes*2]   (x, y) interleaved        (TestKernels.cpp:175-179)
 *   states      double [numNodes*2]   (u, v) interleaved        (TestKernels.cpp:181-187)
 *   residual    double [numNodes*2]   node-interleaved          (FEKernels.h:44-61)
 *   connPtr     int    [numElements]  element e's nodes are conn[connPtr[e] + a], a = ay*(order+1)+ax
 *   conn        int    [numElements*N]
 *   BSR matrix  TACS BCSRMatData: rowp[numNodes+1], cols[nnzb] (sorted per row, diagonal present),
 *               values double[nnzb*4], 2x2 blocks stored row-major (TACSHelpers.h:248-307)
 *   elementBCSRMap int [numElements*N*N]: map[e*N*N + i*N + j] = 4 * (BSR slot of block
 *               (conn[e,i], conn[e,j]))  (TACSHelpers.h:132-182)
 *
 * Conventions: every function returns an fem_status code (0 = success, negative = error) and never
 * exits the process; fem_last_error() gives a message for the calling thread's last failure.
 * "device" pointers are HIP device memory owned by the caller; "host" pointers are host memory.
 * `stream` is a hipStream_t passed as void* (NULL = the default stream). Device entry points are
 * asynchronous on `stream` unless `elapsedSeconds` is non-NULL, in which case the call records HIP
 * events around its launches, waits for them and stores the elapsed time in seconds.
 */
#ifndef FEM_AMD_H
#define FEM_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes and modes ------------------------------------------------------------------ */
enum fem_status {
  FEM_OK = 0,
  FEM_ERR_INVALID_ORDER = -1,    /* element order outside 1..4 (the reference silently no-ops) */
  FEM_ERR_NULL_POINTER = -2,     /* a required pointer is NULL */
  FEM_ERR_INVALID_ARGUMENT = -3, /* sizes, node ids or parameters out of range */
  FEM_ERR_HIP = -4,              /* a HIP runtime call failed */
  FEM_ERR_MAP_MISS = -5,         /* an element node pair has no block in the BSR pattern */
  FEM_ERR_OUT_OF_MEMORY = -6,
  FEM_ERR_UNSUPPORTED = -7,      /* valid input that this build does not handle (message says why) */
  FEM_ERR_BAD_GEOMETRY = -8,     /* generated mesh has det(J) <= 0 at a quadrature point */
  FEM_ERR_BUDGET = -9            /* serving admission: request exceeds the device-memory budget */
};

enum fem_precision {
  FEM_PREC_FP64 = 0, /* all arithmetic in FP64 */
  FEM_PREC_MIXED = 1 /* basis, stress and B^T D B in FP32; det(J), J^-1 and all accumulation in FP64 */
};

/* Message describing the calling thread's most recent failure ("" if none). */
const char *fem_last_error(void);
/* Library version as major*10000 + minor*100 + patch. */
int fem_version(void);

/* ---- element primitives (host) ----------------------------------------------------------------
 * Tabulate the 1-D basis of `order` at the order+1 Gauss points: phi[k*(order+1)+i] = N_k(xi_i),
 * dphi[...] = dN_k/dxi(xi_i), and the Gauss points/weights. Same values as lagrangePoly1dDeriv and
 * getGaussQuad* in LagrangeShapeFuncs.h / GaussQuadrature.h (LagrangeShapeFuncs.h:129-216,
 * GaussQuadrature.h:46-99 of the reference). */
int fem_basis_table(int order, double *phi, double *dphi, double *gaussPoints, double *gaussWeights);

/* ---- synthetic meshes and fields (host) --------------------------------------------------------
 * Structured meshes built the way the reference's gmsh meshes are consumed (TACS tensor node order,
 * x fastest), with nodes at the cosine knots. Random draws use SplitMix64; u01 = (x >> 11) * 2^-53. */
int fem_square_mesh_size(int nx, int ny, int order, int64_t *numNodes, int64_t *numElements);
/* Unit square, nx*ny elements, node id = jy*(order*nx+1)+jx. Interior nodes are moved by
 * U(-jitter, jitter) times the smaller adjacent knot gap along each axis (draws in node-id order,
 * dx then dy). conn is element-major, element e = ey*nx+ex. */
int fem_square_mesh(int nx, int ny, int order, double jitter, uint64_t seed, double *nodeCoords, int *conn);
int fem_annulus_mesh_size(int nr, int ntheta, int order, int64_t *numNodes, int64_t *numElements);
/* Annulus rIn <= r <= rOut, nr radial x ntheta circumferential elements with a periodic seam.
 * xi runs radially, eta circumferentially; node id = jtheta*(order*nr+1)+jr; element e = etheta*nr+er.
 * Radially interior nodes are jittered in (r, theta) like the square mesh. */
int fem_annulus_mesh(int nr, int ntheta, int order, double rIn, double rOut, double jitter, uint64_t seed,
                     double *nodeCoords, int *conn);
/* out[i] = U(lo, hi), i = 0..n-1, in order. */
int fem_fill_uniform(int64_t n, double lo, double hi, uint64_t seed, double *out);
/* Smooth displacement field u = A(sin(2 pi x + p1) + sin(2 pi y + p2)), v = A(cos(2 pi x + p3) +
 * cos(2 pi y + p4)) with phases p1..p4 ~ U(0, 2 pi) drawn from phaseSeed (all zero when phaseSeed
 * is 0: the reference's field, TACSHelpers.h:192-196), plus U(-pert, pert) per state from pertSeed. */
int fem_fill_smooth_field(int64_t numNodes, const double *nodeCoords, double amplitude, uint64_t phaseSeed,
                          double pert, uint64_t pertSeed, double *states);
/* Minimum det(J) over all Gauss points of all elements (host, FP64). */
int fem_min_detj(int order, int numElements, const int *connPtr, const int *conn, const double *nodeCoords,
                 double *minDetJ);

/* ---- BSR pattern, element map, colouring (host) ------------------------------------------------
 * TACS-style block pattern of the mesh: block (i, j) exists iff nodes i and j share an element;
 * columns sorted ascending per row, diagonal included. Call _rowp first (fills numNodes+1 entries,
 * rowp[numNodes] = number of blocks), then _cols with a cols array of that length. */
int fem_bsr_pattern_rowp(int numNodes, int numElements, int nodesPerElement, const int *connPtr, const int *conn,
                         int *rowp);
int fem_bsr_pattern_cols(int numNodes, int numElements, int nodesPerElement, const int *connPtr, const int *conn,
                         const int *rowp, int *cols);
/* TACS-free overload of generateElementBCSRMap (verification/TACSHelpers.h:138-182): same output
 * layout, map[e*N*N + i*N + j] = slot * bsize * bsize. A missing block is an error
 * (FEM_ERR_MAP_MISS), not a printed warning with the entry left unset (TACSHelpers.h:176-178). */
int fem_element_bcsr_map(const int *connPtr, const int *conn, int numElements, int nodesPerElement, int numRows,
                         const int *rowp, const int *cols, int bsize, int *map);
/* Element-once tables of a mesh on the host: what fem_plan_create builds for Jacobian assemblies of orders
 * 3-4 (any order here). code[(e*N + a)*N + b] is the BSR index of block (conn[e,a], conn[e,b]) when element e
 * alone forms it, else numBlocks + a staging slot; rcode[e*N + a] likewise for node conn[e,a]'s residual,
 * against numNodes. gdst / gptr list the other blocks (ascending) and their slot ranges (gptr has
 * *numGather + 1 entries), rdst / rptr the other nodes. Each output is the sum of its slots in slot order
 * (zero if none). Array sizes: code E*N*N, rcode E*N, gdst numBlocks, gptr numBlocks + 1, rdst numNodes,
 * rptr numNodes + 1 (upper bounds). No counterpart in the reference (its kernels add with atomics); exposed
 * so that the tables behind the plans can be checked on the host. */
int fem_split_tables(int order, int numNodes, int numElements, const int *connPtr, const int *conn, const int *rowp,
                     const int *cols, int *code, int *rcode, int *gdst, int *gptr, int *rdst, int *rptr,
                     int64_t *numSlots, int64_t *numGather, int64_t *numRslots, int64_t *numRgather);
/* ---- Output files in the reference's formats (host arrays) -----------------------------------------
 * fem_write_array: one value per line, "% .17g" (writeArrayToFile, verification/TACSHelpers.h:234-246).
 * fem_write_bsr_matrix: header "rows cols nnz" in scalar entries, then "row col value" per entry, 0-based,
 * blocks in row order and each block row-major, "% .17g" (writeBCSRMatToFile, TACSHelpers.h:248-307);
 * maxRows / maxCols > 0 keep the leading block rows / columns, as the reference's optional limits do. A
 * file that cannot be opened or written is an error status (the reference prints and continues). */
int fem_write_array(const double *values, int64_t n, const char *path);
int fem_write_bsr_matrix(int nrows, int ncols, int bsize, const int *rowp, const int *cols, const double *values,
                         int maxRows, int maxCols, const char *path);

/* ---- BDF mesh files (host) ------------------------------------------------------------------------
 * The reference reads its meshes from Nastran bulk-data files through TACS (TACSMeshLoader::scanBDFFile,
 * verification/TACSHelpers.h:86-114; element by card: CQUAD4, CQUAD/CQUAD9, CQUAD16, CQUAD25), written by gmsh
 * and rewritten by fixGmshBDF (meshes/GmshUtils.py:118-196). fem_read_bdf reads GRID (small, large GRID* and
 * free field) and CQUADn cards with continuation lines; nodes are numbered 0..n-1 in GRID-card order (their ids
 * go to gridIds, optional), 2-D coordinates X1, X2 (CP and X3 must be 0). Element nodes are taken in card
 * order, or with FEM_BDF_NASTRAN_ORDER converted from Nastran's CQUAD4 / CQUAD8-9 order (corners
 * counter-clockwise, mid-sides, centre) to the tensor-product order a = ay (p+1) + ax. One element order per
 * file. fem_read_bdf_size gives the order and counts (parse once for the sizes, once for the arrays:
 * coords [numNodes*2], connPtr [numElements+1], conn [numElements*N]). fem_write_bdf writes GRID in free field
 * with "%.17g" reals (read back bit-exact) and CQUADn in fixed small field, nine fields per line, as
 * fixGmshBDF lays them out; nodes and elements numbered from 1. */
#define FEM_BDF_NASTRAN_ORDER 1
int fem_read_bdf_size(const char *path, int flags, int *order, int *numNodes, int *numElements);
int fem_read_bdf(const char *path, int flags, double *coords, int *connPtr, int *conn, long long *gridIds);
int fem_write_bdf(const char *path, int order, int numNodes, int numElements, const int *connPtr, const int *conn,
                  const double *nodeCoords);

/* ---- BSR pattern and element map on the device ---------------------------------------------------
 * The same pattern and map as fem_bsr_pattern_rowp/_cols and fem_element_bcsr_map (identical arrays), built
 * on the GPU from device arrays, in work ordered on `stream` (a hipStream_t; NULL = default stream). These
 * replace the host set-up the reference leaves to TACS (createSchurMat, verification/TestKernels.cpp:105)
 * and generateElementBCSRMap (verification/TACSHelpers.h:138-182). Each call allocates and frees its own
 * scratch and returns after its work has finished. connPtr entries must index valid conn entries, as for
 * every entry point here.
 * fem_bsr_pattern_device fills rowp (numNodes + 1 entries) and *numBlocks. Pass cols = NULL to size the
 * pattern, then call again with cols of colsCapacity >= *numBlocks entries. FEM_ERR_UNSUPPORTED if a node's
 * elements hold more than 1024 node references (the host builder has no such limit). */
int fem_bsr_pattern_device(int numNodes, int numElements, int nodesPerElement, const int *connPtr, const int *conn,
                           int *rowp, int *cols, int64_t colsCapacity, int64_t *numBlocks, void *stream);
/* fem_bsr_pattern_device without any device allocation: its scratch (per-node counts and offsets, the
 * node->element adjacency, row lengths, scan sums, error flags) lives in a fixed 64 MiB array the library holds
 * on each GPU it runs on until the process ends. No counterpart in the reference: it is for callers that
 * cannot allocate (tight serving loops, harnesses that forbid hipMalloc). The scratch is about 12 bytes per
 * node plus 4 per element node reference (a 700 x 700 order-2 mesh fits, a 1000 x 1000 one does not); a mesh
 * needing more than 64 MiB gets FEM_ERR_INVALID_ARGUMENT before any device work. Every call on a GPU shares that GPU's scratch:
 * calls that overlap on one GPU (other threads, other streams) race. Same results as fem_bsr_pattern_device. */
int fem_bsr_pattern_device_noalloc(int numNodes, int numElements, int nodesPerElement, const int *connPtr,
                                   const int *conn, int *rowp, int *cols, int64_t colsCapacity, int64_t *numBlocks,
                                   void *stream);
int fem_element_bcsr_map_device(const int *connPtr, const int *conn, int numElements, int nodesPerElement,
                                int numRows, const int *rowp, const int *cols, int bsize, int *map, void *stream);
/* Greedy element colouring: no two elements of one colour share a node. colour[e] in
 * 0..*numColours-1, assigned in element order. */
int fem_element_colouring(int numNodes, int numElements, int nodesPerElement, const int *connPtr, const int *conn,
                          int *colour, int *numColours);

/* ---- drop-in device entry points ---------------------------------------------------------------
 * Same arguments and accumulate (+=) semantics as the reference's dispatchers, plus precision,
 * stream and timing. The caller zeroes the outputs first (TestKernels.cpp:237,241).
 *
 * fem_residual replaces runResidualKernel (element/FEKernels.h:1030-1102): residual += R(states).
 * fem_jacobian replaces runJacobianKernel (element/FEKernels.h:1125-1225): matEntries += K through
 *   elementBCSRMap, and, when residual is non-NULL, residual += R. FP64 atomic scatter. */
int fem_residual(int order, const int *connPtr, const int *conn, int numElements, const double *states,
                 const double *nodeCoords, double E, double nu, double t, double *residual, int precision,
                 void *stream, double *elapsedSeconds);
int fem_jacobian(int order, const int *connPtr, const int *conn, int numElements, const double *states,
                 const double *nodeCoords, double E, double nu, double t, const int *elementBCSRMap,
                 double *residual, double *matEntries, int precision, void *stream, double *elapsedSeconds);
/* fem_jacobian with the same arguments and semantics (+= into caller-owned residual / matEntries), for callers
 * of the reference's interface (runJacobianKernel, element/FEKernels.h:1125-1225): at orders 1-2, on first use of
 * a (connPtr, conn, elementBCSRMap, numElements, order) combination the library builds the connectivity's BSR
 * pattern and element map on the device and, if that map equals the caller's, a resident plan (kept until
 * fem_dropin_cache_clear; at most four combinations, least recently used evicted). Calls then run the plan in
 * accumulate mode (fem_plan_assemble_accumulate): each output is read and written once, without atomics. Each
 * call first checks a 64-bit checksum of conn and of every 61st map entry (a change rebuilds the entry) and
 * synchronises the stream. Orders 3-4 (where the atomic kernels are faster than accumulate mode), another
 * matrix layout, or a failed build run fem_jacobian. elapsedSeconds covers the
 * assembly. Not for graph capture (it synchronises). */
int fem_jacobian_dropin(int order, const int *connPtr, const int *conn, int numElements, const double *states,
                        const double *nodeCoords, double E, double nu, double t, const int *elementBCSRMap,
                        double *residual, double *matEntries, int precision, void *stream, double *elapsedSeconds);
int fem_dropin_cache_clear(void);
/* Whether the combination is cached (*cached) and served by a plan (*usesPlan; else by fem_jacobian). */
int fem_dropin_cache_info(int order, const int *connPtr, const int *conn, int numElements, const int *elementBCSRMap,
                          int *cached, int *usesPlan);

/* NONLINEAR (Green) strain versions of fem_residual / fem_jacobian (same arguments and accumulate semantics):
 * the reference's kernels instantiated with GreenStrainType::NONLINEAR (element/FEKernels.h:286-315,
 * 585-1006; its dispatchers runResidualKernel / runJacobianKernel, :1030-1136, fix LINEAR). The stiffness is
 * the tangent at `states`. precision must be FEM_PREC_FP64 (FEM_ERR_UNSUPPORTED otherwise). */
int fem_residual_nonlinear(int order, const int *connPtr, const int *conn, int numElements, const double *states,
                           const double *nodeCoords, double E, double nu, double t, double *residual, int precision,
                           void *stream, double *elapsedSeconds);
int fem_jacobian_nonlinear(int order, const int *connPtr, const int *conn, int numElements, const double *states,
                           const double *nodeCoords, double E, double nu, double t, const int *elementBCSRMap,
                           double *residual, double *matEntries, int precision, void *stream, double *elapsedSeconds);
/* Deterministic element-colouring variant of fem_jacobian (no reference counterpart beyond the
 * element-colouring requirement of the same path): elements are processed colour by colour from
 * colourElements (device, element ids grouped by colour, colourPtr (host) gives numColours+1
 * offsets); sums use plain read-add-write, so results are bitwise reproducible. */
int fem_jacobian_coloured(int order, const int *connPtr, const int *conn, int numElements, const double *states,
                          const double *nodeCoords, double E, double nu, double t, const int *elementBCSRMap,
                          const int *colourElements, const int *colourPtr, int numColours, double *residual,
                          double *matEntries, int precision, void *stream, double *elapsedSeconds);

/* ---- resident assembly plan (owner-computes, write-once) ---------------------------------------
 * Built once per mesh from host arrays, kept on the device. Nodes are split into parts whose BSR
 * rows fit a workgroup's LDS; each workgroup assembles every element touching its rows and writes
 * its residual entries and BSR blocks exactly once (no atomics to HBM, no zero-fill). Results are
 * written (not accumulated) into the caller's residual / values in the layout of `rowp`/`cols`.
 * nodeCoordsHost (optional) orders nodes along a Morton curve so parts are compact.
 * At orders 3-4 a plan also holds element-once tables: a Jacobian assembly visits each element once,
 * stores the blocks (and residual entries) that element alone forms, stages the others in the plan's
 * staging buffer and sums them in a gather pass; every output is still written exactly once, in a fixed
 * order, so results are bitwise reproducible with or without FEM_PLAN_DETERMINISTIC. A plan's job record
 * and staging buffer are per plan: assemblies of one plan must not overlap on different streams. */
typedef struct fem_plan fem_plan;

int fem_plan_create(int order, int numNodes, int numElements, const int *connPtr, const int *conn,
                    const int *rowp, const int *cols, const double *nodeCoordsHost, int flags, fem_plan **plan);
/* flags for fem_plan_create */
#define FEM_PLAN_DETERMINISTIC 1 /* colour-phased LDS sums: bitwise reproducible results */
#define FEM_PLAN_RESIDUAL_ONLY 2 /* size parts for residual-only assembly */
#define FEM_PLAN_AUTOTUNE 4      /* build the plan at part sizes 1, 1/2 and 1/4 of the kernel's budget, time one
                                    mixed-precision residual + Jacobian assembly of each on the device (zero
                                    states, temporary outputs of the mesh's size) and keep the fastest; smaller
                                    parts give small meshes more workgroups. Not with the flags above. */
#define FEM_PLAN_OWNER_ONLY 8    /* orders 3-4: no element-once tables; Jacobians take the owner-computes parts
                                    (FP64: row-restricted VALU Gram stage; mixed: FP32 MFMA Gram stage), the
                                    route plans take when the element-once codes would overflow an int */

/* Part-size scale a plan was built with (1, 0.5 or 0.25; 1 without FEM_PLAN_AUTOTUNE) and, for an autotuned
 * plan, the measured seconds of the chosen build (0 otherwise). No counterpart in the reference: the
 * reference fixes elements per block at compile time (FEKernels.h:1015-1116). */
int fem_plan_tuning(const fem_plan *plan, double *partScale, double *seconds);
/* Launch shape of the plan's Jacobian kernel: wavefronts per workgroup (autotuned at orders 3-4 among 1, 2 and
 * 4 for the element-once kernel; the owner kernel's fixed count at orders 1-2), elements per wavefront and
 * lanes per element (64 / N and N, fixed by the order). The reference tunes elements per block by hand per
 * order (element/FEKernels.h:1009-1028, 1104-1123). */
int fem_plan_tuning_shape(const fem_plan *plan, int *wavesPerGroup, int *elementsPerWave, int *lanesPerElement);

/* What a plan holds: number of parts, BSR blocks, element visits (elements counted once per part
 * they touch), resident device bytes. */
int fem_plan_info(const fem_plan *plan, int64_t *numParts, int64_t *numBlocks, int64_t *elementVisits,
                  int64_t *deviceBytes);
/* residual (may be NULL) and matEntries (may be NULL) are overwritten. At least one must be given. */
int fem_plan_assemble(fem_plan *plan, const double *states, const double *nodeCoords, double E, double nu, double t,
                      double *residual, double *matEntries, int precision, void *stream, double *elapsedSeconds);
/* fem_plan_assemble that adds its sums to residual / matEntries (+=, the reference kernels' semantics) instead of
 * storing them: each output is still read and written once (element-once plans still write their own staging). */
int fem_plan_assemble_accumulate(fem_plan *plan, const double *states, const double *nodeCoords, double E, double nu,
                                 double t, double *residual, double *matEntries, int precision, void *stream,
                                 double *elapsedSeconds);
int fem_plan_destroy(fem_plan *plan);

/* ---- batches of resident plans ----------------------------------------------------------------
 * Many assemblies on many resident plans in one call: jobs are grouped by element order and outputs
 * and each group runs as ONE persistent launch over all of its meshes' parts (no per-mesh launches).
 * Jacobian jobs of orders 3-4 run element-once (see fem_plan_create): one persistent launch per order
 * and residual/no-residual group, then the gathers, with staging in the batch's own device arena (grown
 * at assembly when the batch's jobs need more, after a device synchronisation), so batches holding the
 * same plans may run concurrently on different streams.
 * Pointers are the same as fem_plan_assemble's; outputs are overwritten. Jobs of plans built with
 * FEM_PLAN_DETERMINISTIC or FEM_PLAN_RESIDUAL_ONLY run as their plan's own assembly on `stream` after the
 * groups (batches without FEM_BATCH_GRAPH only: graph batches reject them); they use their plan's staging,
 * so batches holding the same such plan must not run concurrently. fem_batch_assemble uploads the job
 * tables (asynchronously on `stream`) and launches; the batch may be cleared and refilled once the call
 * returns. */
typedef struct fem_batch fem_batch;
/* maxJobs: capacity per (order, outputs) group. flags: FEM_BATCH_GRAPH replays the twelve group
 * launches as one captured hipGraph whose kernels read their job counts from device memory. */
int fem_batch_create(int maxJobs, int flags, fem_batch **batch);
#define FEM_BATCH_GRAPH 1
int fem_batch_clear(fem_batch *batch);
int fem_batch_add(fem_batch *batch, fem_plan *plan, const double *states, const double *nodeCoords, double E,
                  double nu, double t, double *residual, double *matEntries);
int fem_batch_assemble(fem_batch *batch, int precision, void *stream, double *elapsedSeconds);
int fem_batch_destroy(fem_batch *batch);

/* ---- serving loop ------------------------------------------------------------------------------
 * Resident meshes and continuous batching. A worker thread admits queued requests in FIFO order
 * while their outputs and batch staging fit the device-memory budget (resident meshes + live outputs +
 * staging of the batches in flight), packs them into batches (fem_batch_*) and runs each batch on one of
 * numSlots streams, so batches formed while others run overlap them on the device. Outputs are carved
 * from the server's arena when a request is admitted and held until fem_server_release; staging is
 * returned when its batch completes. A request whose batch job is rejected fails alone. Request states are
 * caller device memory that must be complete when submitted and stay valid until the request completes.
 * fem_server_destroy fails the requests still queued and waits for the batches in flight. Times are
 * steady-clock seconds. */
typedef struct fem_server fem_server;
typedef struct fem_server_config {
  int64_t hbmBudgetBytes; /* resident meshes + request outputs + batch staging */
  int maxBatchJobs;       /* jobs per (order, outputs) group in one batch */
  int numSlots;           /* batches in flight (streams), 1..8 */
  int precision;          /* FEM_PREC_FP64 or FEM_PREC_MIXED */
  int useGraphs;          /* replay each batch's launches through a captured hipGraph */
} fem_server_config;

int fem_server_create(const fem_server_config *cfg, fem_server **server);
/* Builds the mesh's BSR pattern and resident plan (host arrays in, as fem_plan_create) and uploads its
 * coordinates. FEM_ERR_BUDGET if it does not fit; FEM_ERR_UNSUPPORTED (with useGraphs) if its parts do not
 * fit the graph-replayed kernel. */
int fem_server_add_mesh(fem_server *server, int order, int numNodes, int numElements, const int *connPtr,
                        const int *conn, const double *nodeCoordsHost, double E, double nu, double t, int *meshId);
int fem_server_mesh_info(fem_server *server, int meshId, int64_t *numNodes, int64_t *numBlocks,
                         int64_t *residentBytes);
/* Queues an assembly (withJacobian 0: residual only). FEM_ERR_BUDGET if its outputs can never fit. */
int fem_server_submit(fem_server *server, int meshId, const double *states, int withJacobian, int64_t *requestId);
int fem_server_poll(fem_server *server, int64_t requestId, int *done);
/* Blocks until done; device pointers to the outputs (values NULL for residual-only) and times
 * {submitted, launched, completed}. */
int fem_server_wait(fem_server *server, int64_t requestId, const double **residual, const double **values,
                    double *times);
/* Waits, then copies the outputs to host memory (either pointer may be NULL). */
int fem_server_read(fem_server *server, int64_t requestId, double *residualHost, double *valuesHost);
/* Returns the request's outputs to the pool. */
int fem_server_release(fem_server *server, int64_t requestId);
int fem_server_stats(fem_server *server, int64_t *batches, int64_t *completed, int64_t *bytesInUse);
/* Device memory of the server, all within hbmBudgetBytes: resident meshes (plans + coordinates), the arena
 * (created at the first admission with the budget left after the resident meshes; request outputs and the
 * element-once staging of their batch jobs are carved from it, so serving makes no allocator call), the arena
 * bytes held now and their high-water mark. Any pointer may be NULL. */
int fem_server_memory(fem_server *server, int64_t *residentBytes, int64_t *arenaBytes, int64_t *arenaInUse,
                      int64_t *arenaPeak);
/* Creates the arena now with arenaBytes (before the first admission; otherwise the first admission gives it the
 * whole budget left), so the rest of the budget stays open to meshes added while requests are served.
 * FEM_ERR_BUDGET if it does not fit next to the resident meshes. */
int fem_server_reserve(fem_server *server, int64_t arenaBytes);
int fem_server_destroy(fem_server *server);

/* ---- BSR utilities (device) -------------------------------------------------------------------
 * y = A x for the 2x2-block BSR matrix (used to check R = K u). */
int fem_bsr_spmv(int numRows, const int *rowp, const int *cols, const double *values, const double *x, double *y,
                 void *stream);

#ifdef __cplusplus
}
#endif

#endif /* FEM_AMD_H */
