// Serving loop of libfem_amd: resident meshes, HBM-budget admission and continuous batching.
//
// Meshes are registered once (BSR pattern + resident plan + device coordinates). Clients submit
// assembly requests (mesh, device states, residual-only or residual+Jacobian); the server allocates
// each request's outputs when the request is admitted under the device-memory budget and keeps them
// until the client releases them. A worker thread forms batches from the admitted requests and launches
// each batch as fem_batch_assemble (one persistent launch per element order and output set, optionally
// replayed through a captured hipGraph) on one of several streams ("slots"): requests that arrive while
// earlier batches run are collected into the next batch, which starts on a free slot and overlaps the
// running ones on the device.
//
// Every device byte the server uses while serving comes from the budget: the resident meshes, and one arena
// (sized at the first admission to what the budget leaves after them) holding each admitted request's outputs
// and the element-once staging of its batch job. Staging is returned when the batch completes, outputs when the
// client releases the request. Serving makes no allocator call.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "fem_internal.h"

namespace {

double now_seconds() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct ServerMesh {
  fem_plan *plan = nullptr;
  double *coords = nullptr;
  int order = 0;
  int numNodes = 0;
  int64_t numBlocks = 0;
  double E = 0, nu = 0, t = 0;
  int64_t residentBytes = 0;
};

enum RequestState { kQueued = 0, kRunning = 1, kDone = 2, kFailed = 3, kReleased = 4 };

struct Request {
  int64_t id = 0;
  int mesh = 0;
  ServerMesh m;  // copy of the mesh record, taken at submission under the lock (the table may grow meanwhile)
  const double *states = nullptr;
  bool withJacobian = false;
  double *residual = nullptr;
  double *values = nullptr;
  double *stage = nullptr;  // element-once staging of the batch job (arena; returned when the batch completes)
  int64_t stageBytes = 0;
  int64_t bytes = 0;
  RequestState state = kQueued;
  int status = FEM_OK;
  std::string error;
  double tSubmit = 0, tLaunch = 0, tDone = 0;
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  fem_batch *batch = nullptr;
  std::vector<Request *> requests;
  bool busy = false;
};

}  // namespace

struct fem_server {
  fem_server_config cfg{};
  int device = 0;
  std::vector<ServerMesh> meshes;
  int64_t residentBytes = 0;
  // Request outputs live in one device arena, allocated at the first admission with what the budget
  // leaves after the resident meshes, and sub-allocated first-fit (no hipMalloc/hipFree while serving:
  // hipFree would synchronise the device and stall the batches in flight).
  char *arena = nullptr;
  int64_t arenaBytes = 0;
  int64_t outputBytes = 0;                 // bytes of the arena held by admitted requests (outputs + staging)
  int64_t peakOutputBytes = 0;             // high-water mark of outputBytes
  std::map<int64_t, int64_t> freeRanges;   // arena offset -> length, coalesced
  std::unordered_map<int64_t, std::unique_ptr<Request>> requests;
  std::deque<Request *> queue;
  std::vector<Slot> slots;
  int64_t nextId = 1;
  int64_t batches = 0;
  int64_t completed = 0;
  std::mutex mu;
  std::condition_variable cv;      // worker wake-up
  std::condition_variable doneCv;  // request completion
  std::thread worker;
  bool stop = false;
};

namespace {

int64_t output_bytes(const ServerMesh &m, bool withJacobian) {
  return int64_t(m.numNodes) * 2 * int64_t(sizeof(double)) + (withJacobian ? m.numBlocks * 4 * 8 : 0);
}

constexpr int64_t kArenaAlign = 256;

int64_t align_up(int64_t n) { return (n + kArenaAlign - 1) / kArenaAlign * kArenaAlign; }

// Element-once staging of a request's batch job, in bytes (0: none).
int64_t stage_bytes(const ServerMesh &m, bool withJacobian) {
  return withJacobian ? fem::batch_stage_doubles(m.plan, true) * int64_t(sizeof(double)) : 0;
}

// Arena bytes a request holds while its batch runs (aligned as pool_alloc aligns them).
int64_t arena_need(const ServerMesh &m, bool withJacobian) {
  const int64_t st = stage_bytes(m, withJacobian);
  return align_up(int64_t(m.numNodes) * 16) + (withJacobian ? align_up(m.numBlocks * 32) : 0) +
         (st ? align_up(st) : 0);
}

// Creates the output arena of `bytes` bytes. Caller holds mu.
bool create_arena(fem_server *s, int64_t bytes) {
  bytes = bytes / kArenaAlign * kArenaAlign;
  if (bytes <= 0) return false;
  if (hipMalloc(reinterpret_cast<void **>(&s->arena), size_t(bytes)) != hipSuccess) {
    s->arena = nullptr;
    return false;
  }
  s->arenaBytes = bytes;
  s->freeRanges.clear();
  s->freeRanges.emplace(0, bytes);
  return true;
}

// Creates the output arena, unless fem_server_reserve has, with the budget left after the resident meshes.
// Caller holds mu.
bool ensure_arena(fem_server *s) {
  if (s->arena) return true;
  return create_arena(s, s->cfg.hbmBudgetBytes - s->residentBytes);
}

// First-fit allocation from the arena. Caller holds mu.
void *pool_alloc(fem_server *s, size_t bytes) {
  if (!ensure_arena(s)) return nullptr;
  const int64_t need = align_up(int64_t(bytes));
  for (auto it = s->freeRanges.begin(); it != s->freeRanges.end(); ++it) {
    if (it->second < need) continue;
    const int64_t off = it->first, len = it->second;
    s->freeRanges.erase(it);
    if (len > need) s->freeRanges.emplace(off + need, len - need);
    s->outputBytes += need;
    s->peakOutputBytes = std::max(s->peakOutputBytes, s->outputBytes);
    return s->arena + off;
  }
  return nullptr;
}

// Returns a range to the arena, merging it with free neighbours. Caller holds mu.
void pool_free(fem_server *s, void *p, size_t bytes) {
  if (!p) return;
  int64_t off = static_cast<char *>(p) - s->arena, len = align_up(int64_t(bytes));
  s->outputBytes -= len;
  auto next = s->freeRanges.lower_bound(off);
  if (next != s->freeRanges.end() && off + len == next->first) {
    len += next->second;
    next = s->freeRanges.erase(next);
  }
  if (next != s->freeRanges.begin()) {
    auto prev = std::prev(next);
    if (prev->first + prev->second == off) {
      prev->second += len;
      return;
    }
  }
  s->freeRanges.emplace(off, len);
}

void finish_request(fem_server *s, Request *r, double t, int status, const char *err);

// Admits queued requests in FIFO order while their outputs fit the budget. Caller holds mu.
void admit_into(fem_server *s, std::vector<Request *> &out, int capacity) {
  int perGroup[12] = {0};
  if (!s->queue.empty() && !ensure_arena(s)) {
    // The arena could not be created (device memory taken by others): nothing can ever be admitted, so the
    // queued requests fail instead of waiting.
    const double t = now_seconds();
    while (!s->queue.empty()) {
      finish_request(s, s->queue.front(), t, FEM_ERR_OUT_OF_MEMORY,
                     "the server's output arena (the budget left after resident meshes) could not be allocated");
      s->queue.pop_front();
    }
    s->doneCv.notify_all();
    return;
  }
  while (!s->queue.empty() && int(out.size()) < capacity) {
    Request *r = s->queue.front();
    const ServerMesh &m = r->m;
    const int g = (m.order - 1) * 3 + (r->withJacobian ? 2 : 0);
    if (perGroup[g] >= s->cfg.maxBatchJobs) break;
    const size_t rb = size_t(m.numNodes) * 2 * sizeof(double);
    const size_t vb = r->withJacobian ? size_t(m.numBlocks) * 4 * sizeof(double) : 0;
    const size_t sb = size_t(stage_bytes(m, r->withJacobian));
    double *res = static_cast<double *>(pool_alloc(s, rb));
    if (!res) break;
    double *vals = nullptr, *stage = nullptr;
    if (vb) {
      vals = static_cast<double *>(pool_alloc(s, vb));
      if (!vals) {
        pool_free(s, res, rb);
        break;
      }
    }
    if (sb) {
      stage = static_cast<double *>(pool_alloc(s, sb));
      if (!stage) {
        pool_free(s, vals, vb);
        pool_free(s, res, rb);
        break;
      }
    }
    r->residual = res;
    r->values = vals;
    r->stage = stage;
    r->stageBytes = int64_t(sb);
    s->queue.pop_front();
    perGroup[g]++;
    out.push_back(r);
  }
}

// Marks a request finished (status FEM_OK: done, else failed) and returns its staging to the arena. Caller holds mu.
void finish_request(fem_server *s, Request *r, double t, int status, const char *err) {
  if (r->stage) {
    pool_free(s, r->stage, size_t(r->stageBytes));
    r->stage = nullptr;
    r->stageBytes = 0;
  }
  r->tDone = t;
  r->status = status;
  if (status != FEM_OK) {
    r->state = kFailed;
    r->error = err ? err : "";
  } else {
    r->state = kDone;
  }
  s->completed++;
}

void complete_slot(fem_server *s, Slot &slot, int status, const char *err) {
  const double t = now_seconds();
  std::lock_guard<std::mutex> lk(s->mu);
  for (Request *r : slot.requests) finish_request(s, r, t, status, err);
  slot.requests.clear();
  slot.busy = false;
  s->doneCv.notify_all();
  s->cv.notify_one();  // staging returned: queued requests may now fit
}

// Fails every request still queued (the server is being destroyed). Caller holds mu.
void fail_queued(fem_server *s) {
  const double t = now_seconds();
  while (!s->queue.empty()) {
    finish_request(s, s->queue.front(), t, FEM_ERR_INVALID_ARGUMENT, "the server was destroyed before admission");
    s->queue.pop_front();
  }
  s->doneCv.notify_all();
}

void worker_loop(fem_server *s) {
  (void)hipSetDevice(s->device);
  for (;;) {
    // Retire finished batches.
    for (Slot &slot : s->slots) {
      if (!slot.busy) continue;
      const hipError_t q = hipEventQuery(slot.done);
      if (q == hipSuccess) {
        complete_slot(s, slot, FEM_OK, nullptr);
      } else if (q != hipErrorNotReady) {
        complete_slot(s, slot, FEM_ERR_HIP, hipGetErrorString(q));
      }
    }
    // Launch on free slots while requests wait.
    bool launched = false;
    for (Slot &slot : s->slots) {
      if (slot.busy) continue;
      std::vector<Request *> take;
      {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) {
          // Destroying: requests not yet admitted fail (they may wait for memory that only their client's
          // release would free); the batches in flight drain below.
          fail_queued(s);
          break;
        }
        admit_into(s, take, 12 * s->cfg.maxBatchJobs);
        const double t = now_seconds();
        for (Request *r : take) {
          r->state = kRunning;
          r->tLaunch = t;
        }
      }
      if (take.empty()) break;
      int rc = fem_batch_clear(slot.batch);
      std::vector<Request *> added;
      for (Request *r : take) {
        const ServerMesh &m = r->m;
        const int arc = rc ? rc
                           : fem::batch_add_staged(slot.batch, m.plan, r->states, m.coords, m.E, m.nu, m.t, r->residual,
                                              r->withJacobian ? r->values : nullptr, r->stage);
        if (arc) {
          // only this request fails; the rest of the batch runs
          std::lock_guard<std::mutex> lk(s->mu);
          finish_request(s, r, now_seconds(), arc, fem_last_error());
          s->doneCv.notify_all();
        } else {
          added.push_back(r);
        }
      }
      if (added.empty()) continue;
      rc = fem_batch_assemble(slot.batch, s->cfg.precision, slot.stream, nullptr);
      if (!rc && hipEventRecord(slot.done, slot.stream) != hipSuccess) rc = FEM_ERR_HIP;
      slot.requests = added;
      slot.busy = true;
      if (rc) {
        complete_slot(s, slot, rc, fem_last_error());
      } else {
        std::lock_guard<std::mutex> lk(s->mu);
        s->batches++;
      }
      launched = true;
    }
    bool anyBusy = false;
    for (Slot &slot : s->slots) anyBusy = anyBusy || slot.busy;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      if (s->stop && s->queue.empty() && !anyBusy) return;
      if (!launched) {
        if (anyBusy) {
          // Poll the device for completions at a fine grain.
          s->cv.wait_for(lk, std::chrono::microseconds(20));
        } else if (s->queue.empty()) {
          s->cv.wait(lk, [&] { return s->stop || !s->queue.empty(); });
        } else {
          // Requests wait for memory held by completed-but-unreleased results.
          s->cv.wait_for(lk, std::chrono::microseconds(100));
        }
      }
    }
  }
}

}  // namespace

using namespace fem;

extern "C" {

int fem_server_create(const fem_server_config *cfg, fem_server **server) {
  if (!cfg || !server) return set_error(FEM_ERR_NULL_POINTER, "cfg/server is NULL");
  if (cfg->hbmBudgetBytes <= 0 || cfg->maxBatchJobs < 1 || cfg->numSlots < 1 || cfg->numSlots > 8) {
    return set_error(FEM_ERR_INVALID_ARGUMENT, "budget > 0, maxBatchJobs >= 1 and 1 <= numSlots <= 8 required");
  }
  if (cfg->precision != FEM_PREC_FP64 && cfg->precision != FEM_PREC_MIXED) {
    return set_error(FEM_ERR_INVALID_ARGUMENT, "unknown precision mode %d", cfg->precision);
  }
  auto *s = new fem_server();
  s->cfg = *cfg;
  (void)hipGetDevice(&s->device);
  s->slots.resize(size_t(cfg->numSlots));
  for (Slot &slot : s->slots) {
    if (hipStreamCreateWithFlags(&slot.stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&slot.done, hipEventDisableTiming) != hipSuccess ||
        fem_batch_create(cfg->maxBatchJobs, cfg->useGraphs ? FEM_BATCH_GRAPH : 0, &slot.batch) != FEM_OK ||
        fem::batch_prepare_graph(slot.batch, cfg->precision) != FEM_OK) {
      // (graphs are recorded here, before the worker starts: a capture in flight would make other threads'
      // default-stream work, such as fem_server_add_mesh's uploads, fail)
      fem_server_destroy(s);
      return set_error(FEM_ERR_HIP, "creating server streams/events/batches failed");
    }
  }
  s->worker = std::thread(worker_loop, s);
  *server = s;
  return FEM_OK;
}

int fem_server_add_mesh(fem_server *server, int order, int numNodes, int numElements, const int *connPtr,
                        const int *conn, const double *nodeCoordsHost, double E, double nu, double t, int *meshId) {
  if (!server || !meshId || !nodeCoordsHost) return set_error(FEM_ERR_NULL_POINTER, "NULL argument");
  const int N = (order + 1) * (order + 1);
  if (order < 1 || order > 4) return set_error(FEM_ERR_INVALID_ORDER, "element order %d outside 1..4", order);
  std::vector<int> rowp(size_t(numNodes) + 1);
  int rc = fem_bsr_pattern_rowp(numNodes, numElements, N, connPtr, conn, rowp.data());
  if (rc) return rc;
  std::vector<int> cols(size_t(rowp[size_t(numNodes)]));
  rc = fem_bsr_pattern_cols(numNodes, numElements, N, connPtr, conn, rowp.data(), cols.data());
  if (rc) return rc;
  ServerMesh m;
  m.order = order;
  m.numNodes = numNodes;
  m.numBlocks = rowp[size_t(numNodes)];
  m.E = E;
  m.nu = nu;
  m.t = t;
  rc = fem_plan_create(order, numNodes, numElements, connPtr, conn, rowp.data(), cols.data(), nodeCoordsHost, 0,
                       &m.plan);
  if (rc) return rc;
  if (server->cfg.useGraphs && !fem::plan_streams_public(m.plan)) {
    // A graph batch replays the streaming kernel, whose parts are one batch per wavefront; this mesh's parts
    // are larger (a node shared by more elements than a workgroup takes at once).
    fem_plan_destroy(m.plan);
    return set_error(FEM_ERR_UNSUPPORTED, "the mesh's parts do not fit the graph-replayed batch kernel; "
                     "create the server with useGraphs = 0 to serve it");
  }
  int64_t planBytes = 0;
  fem_plan_info(m.plan, nullptr, nullptr, nullptr, &planBytes);
  m.residentBytes = planBytes + int64_t(numNodes) * 16;
  {
    std::lock_guard<std::mutex> lk(server->mu);
    const int64_t held = server->residentBytes + (server->arena ? server->arenaBytes : 0);
    if (held + m.residentBytes > server->cfg.hbmBudgetBytes) {
      fem_plan_destroy(m.plan);
      return set_error(FEM_ERR_BUDGET, "mesh needs %lld resident bytes; budget leaves %lld%s",
                       (long long)m.residentBytes, (long long)(server->cfg.hbmBudgetBytes - held),
                       server->arena ? " (the output arena was sized at the first admission)" : "");
    }
    server->residentBytes += m.residentBytes;
  }
  if (hipMalloc(reinterpret_cast<void **>(&m.coords), size_t(numNodes) * 16) != hipSuccess ||
      hipMemcpy(m.coords, nodeCoordsHost, size_t(numNodes) * 16, hipMemcpyHostToDevice) != hipSuccess) {
    fem_plan_destroy(m.plan);
    std::lock_guard<std::mutex> lk(server->mu);
    server->residentBytes -= m.residentBytes;
    return set_error(FEM_ERR_HIP, "uploading mesh coordinates failed");
  }
  std::lock_guard<std::mutex> lk(server->mu);
  *meshId = int(server->meshes.size());
  server->meshes.push_back(m);
  return FEM_OK;
}

int fem_server_mesh_info(fem_server *server, int meshId, int64_t *numNodes, int64_t *numBlocks,
                         int64_t *residentBytes) {
  if (!server) return set_error(FEM_ERR_NULL_POINTER, "server is NULL");
  std::lock_guard<std::mutex> lk(server->mu);
  if (meshId < 0 || meshId >= int(server->meshes.size())) return set_error(FEM_ERR_INVALID_ARGUMENT, "bad mesh id");
  const ServerMesh &m = server->meshes[size_t(meshId)];
  if (numNodes) *numNodes = m.numNodes;
  if (numBlocks) *numBlocks = m.numBlocks;
  if (residentBytes) *residentBytes = m.residentBytes;
  return FEM_OK;
}

int fem_server_submit(fem_server *server, int meshId, const double *states, int withJacobian, int64_t *requestId) {
  if (!server || !states || !requestId) return set_error(FEM_ERR_NULL_POINTER, "NULL argument");
  std::lock_guard<std::mutex> lk(server->mu);
  if (meshId < 0 || meshId >= int(server->meshes.size())) return set_error(FEM_ERR_INVALID_ARGUMENT, "bad mesh id");
  const ServerMesh &m = server->meshes[size_t(meshId)];
  const int64_t need = output_bytes(m, withJacobian != 0);
  const int64_t room = server->arena ? server->arenaBytes : server->cfg.hbmBudgetBytes - server->residentBytes;
  if (arena_need(m, withJacobian != 0) > room) {
    return set_error(FEM_ERR_BUDGET, "request outputs and staging need %lld bytes; the budget leaves %lld after "
                     "resident meshes", (long long)arena_need(m, withJacobian != 0), (long long)room);
  }
  auto r = std::make_unique<Request>();
  r->id = server->nextId++;
  r->mesh = meshId;
  r->m = m;
  r->states = states;
  r->withJacobian = withJacobian != 0;
  r->bytes = need;
  r->tSubmit = now_seconds();
  *requestId = r->id;
  server->queue.push_back(r.get());
  server->requests.emplace(r->id, std::move(r));
  server->cv.notify_one();
  return FEM_OK;
}

int fem_server_poll(fem_server *server, int64_t requestId, int *done) {
  if (!server || !done) return set_error(FEM_ERR_NULL_POINTER, "NULL argument");
  std::lock_guard<std::mutex> lk(server->mu);
  auto it = server->requests.find(requestId);
  if (it == server->requests.end()) return set_error(FEM_ERR_INVALID_ARGUMENT, "unknown request %lld", (long long)requestId);
  *done = it->second->state == kDone || it->second->state == kFailed;
  return FEM_OK;
}

int fem_server_wait(fem_server *server, int64_t requestId, const double **residual, const double **values,
                    double *times) {
  if (!server) return set_error(FEM_ERR_NULL_POINTER, "server is NULL");
  std::unique_lock<std::mutex> lk(server->mu);
  auto it = server->requests.find(requestId);
  if (it == server->requests.end()) return set_error(FEM_ERR_INVALID_ARGUMENT, "unknown request %lld", (long long)requestId);
  Request *r = it->second.get();
  server->doneCv.wait(lk, [&] { return r->state == kDone || r->state == kFailed || r->state == kReleased; });
  if (r->state == kReleased) return set_error(FEM_ERR_INVALID_ARGUMENT, "request %lld was released", (long long)requestId);
  if (residual) *residual = r->residual;
  if (values) *values = r->withJacobian ? r->values : nullptr;
  if (times) {
    times[0] = r->tSubmit;
    times[1] = r->tLaunch;
    times[2] = r->tDone;
  }
  if (r->state == kFailed) return set_error(r->status, "request failed: %s", r->error.c_str());
  return FEM_OK;
}

int fem_server_read(fem_server *server, int64_t requestId, double *residualHost, double *valuesHost) {
  const double *res = nullptr, *vals = nullptr;
  int rc = fem_server_wait(server, requestId, &res, &vals, nullptr);
  if (rc) return rc;
  int64_t nn = 0, nb = 0;
  {
    std::lock_guard<std::mutex> lk(server->mu);
    const Request &r = *server->requests[requestId];
    nn = r.m.numNodes;
    nb = r.m.numBlocks;
  }
  if (residualHost && hipMemcpy(residualHost, res, size_t(nn) * 16, hipMemcpyDeviceToHost) != hipSuccess) {
    return set_error(FEM_ERR_HIP, "copying the residual failed");
  }
  if (valuesHost && vals && hipMemcpy(valuesHost, vals, size_t(nb) * 32, hipMemcpyDeviceToHost) != hipSuccess) {
    return set_error(FEM_ERR_HIP, "copying the BSR values failed");
  }
  return FEM_OK;
}

int fem_server_release(fem_server *server, int64_t requestId) {
  if (!server) return set_error(FEM_ERR_NULL_POINTER, "server is NULL");
  std::lock_guard<std::mutex> lk(server->mu);
  auto it = server->requests.find(requestId);
  if (it == server->requests.end()) return set_error(FEM_ERR_INVALID_ARGUMENT, "unknown request %lld", (long long)requestId);
  Request *r = it->second.get();
  if (r->state != kDone && r->state != kFailed) {
    return set_error(FEM_ERR_INVALID_ARGUMENT, "request %lld has not completed", (long long)requestId);
  }
  const ServerMesh &m = r->m;
  if (r->residual) pool_free(server, r->residual, size_t(m.numNodes) * 16);
  if (r->values) pool_free(server, r->values, size_t(m.numBlocks) * 32);
  server->requests.erase(it);
  server->cv.notify_one();
  return FEM_OK;
}

int fem_server_stats(fem_server *server, int64_t *batches, int64_t *completed, int64_t *bytesInUse) {
  if (!server) return set_error(FEM_ERR_NULL_POINTER, "server is NULL");
  std::lock_guard<std::mutex> lk(server->mu);
  if (batches) *batches = server->batches;
  if (completed) *completed = server->completed;
  if (bytesInUse) *bytesInUse = server->residentBytes + server->outputBytes;  // arena bytes held by requests
  return FEM_OK;
}

int fem_server_reserve(fem_server *server, int64_t arenaBytes) {
  if (!server) return set_error(FEM_ERR_NULL_POINTER, "server is NULL");
  std::lock_guard<std::mutex> lk(server->mu);
  if (server->arena) return set_error(FEM_ERR_INVALID_ARGUMENT, "the output arena exists already");
  if (arenaBytes < kArenaAlign) return set_error(FEM_ERR_INVALID_ARGUMENT, "arena of %lld bytes", (long long)arenaBytes);
  if (server->residentBytes + arenaBytes > server->cfg.hbmBudgetBytes) {
    return set_error(FEM_ERR_BUDGET, "an arena of %lld bytes exceeds the budget left after resident meshes (%lld)",
                     (long long)arenaBytes, (long long)(server->cfg.hbmBudgetBytes - server->residentBytes));
  }
  if (!create_arena(server, arenaBytes)) return set_error(FEM_ERR_OUT_OF_MEMORY, "hipMalloc of the arena failed");
  return FEM_OK;
}

int fem_server_memory(fem_server *server, int64_t *residentBytes, int64_t *arenaBytes, int64_t *arenaInUse,
                      int64_t *arenaPeak) {
  if (!server) return set_error(FEM_ERR_NULL_POINTER, "server is NULL");
  std::lock_guard<std::mutex> lk(server->mu);
  if (residentBytes) *residentBytes = server->residentBytes;
  if (arenaBytes) *arenaBytes = server->arena ? server->arenaBytes : 0;
  if (arenaInUse) *arenaInUse = server->outputBytes;
  if (arenaPeak) *arenaPeak = server->peakOutputBytes;
  return FEM_OK;
}

int fem_server_destroy(fem_server *server) {
  if (!server) return FEM_OK;
  {
    std::lock_guard<std::mutex> lk(server->mu);
    server->stop = true;
    server->cv.notify_all();
  }
  if (server->worker.joinable()) server->worker.join();
  for (Slot &slot : server->slots) {
    if (slot.stream) (void)hipStreamSynchronize(slot.stream);
    if (slot.batch) fem_batch_destroy(slot.batch);
    if (slot.done) (void)hipEventDestroy(slot.done);
    if (slot.stream) (void)hipStreamDestroy(slot.stream);
  }
  if (server->arena) (void)hipFree(server->arena);
  for (ServerMesh &m : server->meshes) {
    fem_plan_destroy(m.plan);
    (void)hipFree(m.coords);
  }
  delete server;
  return FEM_OK;
}

}  // extern "C"
