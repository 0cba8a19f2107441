// Continuous-batching query server over a trained guiding field.
//
// Requests (n queries each: lookup + cosine/HG product + one sample + pdf,
// the batched form of Field::getRegion + SamplingDistribution::init/apply/
// samplePDF) are submitted from any thread with host pointers and queued as
// pieces of at most pieceQueries queries.  When every array of a request is
// pinned host memory mapped at the same address on the device, the query
// kernel reads and writes it in place over PCIe (no copies); otherwise the
// request is copied through the slot's arena.  The device-memory budget is split
// into one fixed arena per in-flight slot (allocated once).  A dispatcher
// thread takes a free slot, admits pending pieces in arrival order while the
// slot's arena has room, and issues copy-in, the query kernel and copy-out of
// that batch on the slot's stream, so batches overlap on the device and a
// large request's pieces pipeline across slots.  A completion thread polls
// the batches' events; a request is done when its last piece is.
// Every query's result depends only on its own inputs and the field, so a
// request gets bit-identical results alone or in any mix.
//
// The server serves a snapshot of the field's device state taken at creation
// and on refresh(), so the field can be updated while the server runs; the
// snapshot is double-buffered so that a refresh does not stall admission.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "field.h"

namespace pglgpu {

struct ServerConfig {
  size_t memoryBudget = size_t(512) << 20;  // device bytes of the batch arenas (split evenly over the slots)
  uint32_t slots = 4;                        // batches in flight (one stream each)
  uint32_t maxBatchQueries = 1u << 20;       // queries per batch
  uint32_t pieceQueries = 32768;             // requests are admitted in pieces of at most this many queries
};

struct QueryRequest {
  size_t n = 0;
  const float *pos = nullptr, *nrm = nullptr, *g = nullptr, *u2 = nullptr, *u1 = nullptr;
  const uint8_t* isVolume = nullptr;
  float *dir = nullptr, *pdf = nullptr, *integral = nullptr;
  uint32_t* id = nullptr;
};

// bytes of device arena per query: inputs pos 12, nrm 12, g 4, u2 8, u1 4,
// volume flag 4 (padded); outputs dir 12, pdf 4, id 4, integral 4
static const size_t kServerBytesPerQuery = 68;

class QueryServer {
 public:
  QueryServer(GuidingField* gf, const ServerConfig& cfg);
  ~QueryServer();
  QueryServer(const QueryServer&) = delete;
  QueryServer& operator=(const QueryServer&) = delete;

  // returns the request id (> 0); throws if the request can never be admitted
  uint64_t submit(const QueryRequest& r);
  bool poll(uint64_t id);
  bool wait(uint64_t id, double timeoutMs);
  void drain();
  // milliseconds since server start: submitted, admitted (batch issued), done
  bool timing(uint64_t id, double& submitMs, double& admitMs, double& doneMs);
  void forget(uint64_t id);
  // snapshot the field's current device state: waits for the field's queued
  // update, copies into the inactive snapshot (once no batch reads it) and
  // makes it active; batches keep being admitted meanwhile
  void refresh();
  // [0] requests done, [1] queries done, [2] batches, [3] peak in-flight bytes,
  // [4] peak pending requests, [5] largest batch (queries)
  void stats(double* out, size_t count);

 private:
  static void register_live(QueryServer* q);
  struct Req {
    uint64_t id;
    QueryRequest q;
    std::chrono::steady_clock::time_point tSubmit, tAdmit, tDone;
    size_t piecesLeft = 0;
    bool zeroCopy = false;   // every array is device-accessible pinned host memory
    bool admitted = false;
    bool done = false;
  };
  struct Piece {
    std::shared_ptr<Req> req;
    size_t off, n;
  };
  struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t event = nullptr;
    void* arena = nullptr;
    size_t arenaBytes = 0;
    QueryPiece* hostTable = nullptr;  // pinned staging of the batch's piece table
    QueryPiece* devTable = nullptr;
    bool busy = false;
    size_t bytes = 0;
    size_t queries = 0;
    bool issued = false;
    int snap = 0;  // snapshot the batch reads
    std::vector<Piece> pieces;
  };
  struct Snapshot {
    DevBuf nodes, states, nbIds, nbPts;
    QueryField view{};
  };

  void dispatchLoop();
  void completionLoop();
  void issue(Slot& s, std::vector<Piece>& batch, size_t queries);
  void snapshotKind(const KindField& k, Snapshot& snap, hipStream_t s);
  double msSince(std::chrono::steady_clock::time_point t) const;

  GuidingField* gf_;
  ServerConfig cfg_;
  std::vector<Slot> slots_;
  // two snapshots: batches read the active one while refresh() fills the other
  Snapshot surf_[2], vol_[2];
  int active_ = 0;
  size_t inUse_[2] = {0, 0};  // batches in flight per snapshot
  std::mutex refreshMu_;
  hipStream_t copyStream_ = nullptr;
  std::chrono::steady_clock::time_point t0_;

  std::mutex mu_;
  std::condition_variable cvDispatch_, cvDone_;
  std::deque<Piece> pending_;
  size_t slotQueries_ = 0;  // capacity of one slot's arena
  size_t maxPieces_ = 0;    // piece-table capacity of one slot
  std::unordered_map<uint64_t, std::shared_ptr<Req>> all_;
  uint64_t nextId_ = 1;
  size_t inFlightBytes_ = 0;
  size_t inFlightBatches_ = 0;
  bool stop_ = false;
  double statDoneReq_ = 0, statDoneQ_ = 0, statBatches_ = 0, statPeakBytes_ = 0, statPeakPending_ = 0,
         statMaxBatch_ = 0;
  std::thread dispatcher_, completer_;
  bool joined_ = false;  // threads stopped (by the destructor or at process exit)

 public:
  // Stops and joins the worker threads once the batches in flight are done
  // (requests not yet dispatched are dropped).  Called for servers still
  // alive at process exit, from a handler that runs before the HIP
  // runtime's own teardown, so no thread calls into a finalised runtime.
  void stopThreads();
};

}  // namespace pglgpu
