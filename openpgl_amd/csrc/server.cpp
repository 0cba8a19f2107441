// Continuous-batching query server (see server.h).
#include "server.h"

#include <algorithm>
#include <cstdlib>
#include <set>
#include <stdexcept>

namespace pglgpu {

namespace {

// true if p is null or device-accessible pinned host memory at the same address
bool zero_copy_ok(const void* p) {
  if (!p) return true;
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeHost && a.devicePointer == a.hostPointer && a.hostPointer != nullptr;
}

// batch arena layout for N queries (all segments 256-byte aligned)
struct Arena {
  float *pos, *nrm, *g, *u2, *u1, *dir, *pdf, *integral;
  uint32_t* id;
  uint8_t* vol;
};
size_t align256(size_t x) { return (x + 255) & ~size_t(255); }
size_t arena_bytes(size_t N) {
  return align256(12 * N) * 3 + align256(4 * N) * 6 + align256(8 * N) + align256(N);
}
Arena carve(void* base, size_t N) {
  char* p = (char*)base;
  Arena a;
  auto take = [&](size_t bytes) { char* q = p; p += align256(bytes); return q; };
  a.pos = (float*)take(12 * N);
  a.nrm = (float*)take(12 * N);
  a.dir = (float*)take(12 * N);
  a.g = (float*)take(4 * N);
  a.u1 = (float*)take(4 * N);
  a.pdf = (float*)take(4 * N);
  a.integral = (float*)take(4 * N);
  a.id = (uint32_t*)take(4 * N);
  a.u2 = (float*)take(8 * N);
  take(4 * N);  // spare segment (keeps the layout size of arena_bytes)
  a.vol = (uint8_t*)take(N);
  return a;
}

}  // namespace

QueryServer::QueryServer(GuidingField* gf, const ServerConfig& cfg) : gf_(gf), cfg_(cfg) {
  if (cfg_.slots == 0) cfg_.slots = 1;
  if (cfg_.maxBatchQueries == 0) cfg_.maxBatchQueries = 1u << 20;
  if (cfg_.pieceQueries == 0) cfg_.pieceQueries = 32768;
  t0_ = std::chrono::steady_clock::now();
  // every slot owns an arena of budget / slots bytes, allocated once
  const size_t slotBytes = cfg_.memoryBudget / cfg_.slots;
  size_t q = std::min<size_t>(cfg_.maxBatchQueries, slotBytes / kServerBytesPerQuery + 1);
  while (q > 0 && arena_bytes(q) > slotBytes) q--;
  if (q < 1024) throw std::runtime_error("server: memory budget too small for one batch of 1024 queries per slot");
  slotQueries_ = q;
  PGL_HIP_CHECK(hipStreamCreate(&copyStream_));
  slots_.resize(cfg_.slots);
  const size_t piece = std::min<size_t>(cfg_.pieceQueries, slotQueries_);
  maxPieces_ = (slotQueries_ + piece - 1) / piece + 1;
  for (auto& s : slots_) {
    PGL_HIP_CHECK(hipStreamCreate(&s.stream));
    PGL_HIP_CHECK(hipEventCreateWithFlags(&s.event, hipEventDisableTiming));
    PGL_HIP_CHECK(hipMalloc(&s.arena, arena_bytes(slotQueries_)));
    s.arenaBytes = arena_bytes(slotQueries_);
    PGL_HIP_CHECK(hipHostMalloc((void**)&s.hostTable, maxPieces_ * sizeof(QueryPiece), hipHostMallocDefault));
    PGL_HIP_CHECK(hipMalloc((void**)&s.devTable, maxPieces_ * sizeof(QueryPiece)));
  }
  refresh();
  dispatcher_ = std::thread([this] { dispatchLoop(); });
  completer_ = std::thread([this] { completionLoop(); });
  register_live(this);
}

// Servers whose owner never released them (a leaked handle) still have
// running threads when the process exits.  The handler registered with the
// first server runs before the HIP runtime's static teardown (it is
// registered later, after the runtime initialised) and stops them.
namespace {
std::mutex gLiveMu;
std::set<QueryServer*>* gLive = nullptr;  // intentionally never destroyed

void stop_live_servers() {
  std::vector<QueryServer*> v;
  {
    std::lock_guard<std::mutex> g(gLiveMu);
    if (gLive) v.assign(gLive->begin(), gLive->end());
  }
  for (QueryServer* q : v) q->stopThreads();
}
}  // namespace

void QueryServer::register_live(QueryServer* q) {
  std::lock_guard<std::mutex> g(gLiveMu);
  if (!gLive) {
    gLive = new std::set<QueryServer*>();
    std::atexit(stop_live_servers);
  }
  gLive->insert(q);
}

void QueryServer::stopThreads() {
  {
    std::unique_lock<std::mutex> lk(mu_);
    if (joined_) return;
    stop_ = true;
  }
  cvDispatch_.notify_all();
  cvDone_.notify_all();
  if (dispatcher_.joinable()) dispatcher_.join();
  if (completer_.joinable()) completer_.join();
  std::lock_guard<std::mutex> lk(mu_);
  joined_ = true;
}

QueryServer::~QueryServer() {
  {
    std::lock_guard<std::mutex> g(gLiveMu);
    if (gLive) gLive->erase(this);
  }
  {
    std::unique_lock<std::mutex> lk(mu_);
    if (!joined_) cvDone_.wait(lk, [&] { return pending_.empty() && inFlightBatches_ == 0; });
  }
  stopThreads();
  for (auto& s : slots_) {
    (void)hipStreamSynchronize(s.stream);
    (void)hipFree(s.arena);
    (void)hipHostFree(s.hostTable);
    (void)hipFree(s.devTable);
    (void)hipEventDestroy(s.event);
    (void)hipStreamDestroy(s.stream);
  }
  (void)hipStreamDestroy(copyStream_);
}

double QueryServer::msSince(std::chrono::steady_clock::time_point t) const {
  return std::chrono::duration<double, std::milli>(t - t0_).count();
}

void QueryServer::snapshotKind(const KindField& k, Snapshot& snap, hipStream_t s) {
  const QueryField src = k.queryView();
  QueryField v = src;
  const size_t nodes = k.nodes().size(), R = k.numRegions();
  if (src.trained && nodes && R) {
    snap.nodes.ensure(nodes * sizeof(KDNode));
    snap.states.ensure(R * sizeof(RegionState));
    PGL_HIP_CHECK(hipMemcpyAsync(snap.nodes.p, src.nodes, nodes * sizeof(KDNode), hipMemcpyDeviceToDevice, s));
    PGL_HIP_CHECK(hipMemcpyAsync(snap.states.p, src.regions, R * sizeof(RegionState), hipMemcpyDeviceToDevice, s));
    v.nodes = snap.nodes.as<KDNode>();
    v.regions = snap.states.as<RegionState>();
    if (src.knn) {
      snap.nbIds.ensure(R * 8 * 4);
      snap.nbPts.ensure(R * 8 * sizeof(float4));
      PGL_HIP_CHECK(hipMemcpyAsync(snap.nbIds.p, src.nbIds, R * 8 * 4, hipMemcpyDeviceToDevice, s));
      PGL_HIP_CHECK(hipMemcpyAsync(snap.nbPts.p, src.nbPts, R * 8 * sizeof(float4), hipMemcpyDeviceToDevice, s));
      v.nbIds = snap.nbIds.as<uint32_t>();
      v.nbPts = snap.nbPts.as<float4>();
    }
  } else {
    v.trained = 0;
  }
  snap.view = v;
}

void QueryServer::refresh() {
  std::lock_guard<std::mutex> serial(refreshMu_);
  // the field's queued update (its own streams), outside the admission lock
  gf_->surface.finish();
  gf_->volume.finish();
  int next;
  {
    std::unique_lock<std::mutex> lk(mu_);
    next = 1 - active_;
    cvDone_.wait(lk, [&] { return inUse_[next] == 0; });
  }
  snapshotKind(gf_->surface, surf_[next], copyStream_);
  snapshotKind(gf_->volume, vol_[next], copyStream_);
  PGL_HIP_CHECK(hipStreamSynchronize(copyStream_));
  {
    std::lock_guard<std::mutex> lk(mu_);
    active_ = next;
  }
}

uint64_t QueryServer::submit(const QueryRequest& r) {
  if (r.n == 0) throw std::runtime_error("server: empty request");
  if (!r.pos || !r.nrm || !r.u2 || !r.dir || !r.pdf) throw std::runtime_error("server: null query array");
  if (r.n * kServerBytesPerQuery > cfg_.memoryBudget)
    throw std::runtime_error("server: request larger than the memory budget");
  auto q = std::make_shared<Req>();
  q->q = r;
  q->tSubmit = std::chrono::steady_clock::now();
  q->zeroCopy = zero_copy_ok(r.pos) && zero_copy_ok(r.nrm) && zero_copy_ok(r.g) && zero_copy_ok(r.u2) &&
                zero_copy_ok(r.u1) && zero_copy_ok(r.isVolume) && zero_copy_ok(r.dir) && zero_copy_ok(r.pdf) &&
                zero_copy_ok(r.id) && zero_copy_ok(r.integral);
  const size_t piece = std::min<size_t>(cfg_.pieceQueries, slotQueries_);
  uint64_t id;
  {
    std::lock_guard<std::mutex> lk(mu_);
    if (stop_) throw std::runtime_error("server: stopped");
    id = nextId_++;
    q->id = id;
    q->piecesLeft = (r.n + piece - 1) / piece;
    for (size_t off = 0; off < r.n; off += piece) pending_.push_back(Piece{q, off, std::min(piece, r.n - off)});
    all_[id] = q;
    statPeakPending_ = std::max(statPeakPending_, (double)pending_.size());
  }
  cvDispatch_.notify_one();
  return id;
}

void QueryServer::dispatchLoop() {
  for (;;) {
    std::vector<Piece> batch;
    Slot* slot = nullptr;
    size_t queries = 0, bytes = 0;
    {
      std::unique_lock<std::mutex> lk(mu_);
      cvDispatch_.wait(lk, [&] {
        if (stop_) return true;
        if (pending_.empty()) return false;
        for (auto& s : slots_)
          if (!s.busy) return true;
        return false;
      });
      if (stop_) return;
      for (auto& s : slots_)
        if (!s.busy) { slot = &s; break; }
      // admit pieces in arrival order while the slot's arena has room
      const size_t cap = std::min<size_t>(slotQueries_, cfg_.maxBatchQueries);
      while (!pending_.empty() && queries + pending_.front().n <= cap && batch.size() + 1 < maxPieces_) {
        batch.push_back(pending_.front());
        pending_.pop_front();
        queries += batch.back().n;
      }
      bytes = queries * kServerBytesPerQuery;
      inFlightBytes_ += bytes;
      inFlightBatches_++;
      slot->busy = true;
      slot->issued = false;
      slot->snap = active_;
      inUse_[active_]++;
      slot->bytes = bytes;
      slot->queries = queries;
      statPeakBytes_ = std::max(statPeakBytes_, (double)inFlightBytes_);
      statMaxBatch_ = std::max(statMaxBatch_, (double)queries);
      statBatches_ += 1;
    }
    issue(*slot, batch, queries);
    {
      std::lock_guard<std::mutex> lk(mu_);
      slot->pieces = std::move(batch);
      slot->issued = true;
    }
    cvDone_.notify_all();
  }
}

void QueryServer::issue(Slot& s, std::vector<Piece>& batch, size_t N) {
  // N <= slotQueries_: the arena was carved for that many queries
  Arena a = carve(s.arena, slotQueries_);
  const auto tAdmit = std::chrono::steady_clock::now();
  size_t off = 0;
  for (size_t k = 0; k < batch.size(); k++) {
    Piece& pc = batch[k];
    const QueryRequest& q = pc.req->q;
    const size_t n = pc.n, o = pc.off;
    {
      std::lock_guard<std::mutex> lk(mu_);
      if (!pc.req->admitted) {
        pc.req->admitted = true;
        pc.req->tAdmit = tAdmit;
      }
    }
    QueryPiece& d = s.hostTable[k];
    d.begin = (uint32_t)off;
    d.n = (uint32_t)n;
    if (pc.req->zeroCopy) {
      d.pos = q.pos + 3 * o;
      d.nrm = q.nrm + 3 * o;
      d.u2 = q.u2 + 2 * o;
      d.g = q.g ? q.g + o : nullptr;
      d.u1 = q.u1 ? q.u1 + o : nullptr;
      d.vol = q.isVolume ? q.isVolume + o : nullptr;
      d.dir = q.dir + 3 * o;
      d.pdf = q.pdf + o;
      d.id = q.id ? q.id + o : nullptr;
      d.integral = q.integral ? q.integral + o : nullptr;
    } else {
      PGL_HIP_CHECK(hipMemcpyAsync(a.pos + 3 * off, q.pos + 3 * o, 12 * n, hipMemcpyHostToDevice, s.stream));
      PGL_HIP_CHECK(hipMemcpyAsync(a.nrm + 3 * off, q.nrm + 3 * o, 12 * n, hipMemcpyHostToDevice, s.stream));
      PGL_HIP_CHECK(hipMemcpyAsync(a.u2 + 2 * off, q.u2 + 2 * o, 8 * n, hipMemcpyHostToDevice, s.stream));
      if (q.g) PGL_HIP_CHECK(hipMemcpyAsync(a.g + off, q.g + o, 4 * n, hipMemcpyHostToDevice, s.stream));
      if (q.u1) PGL_HIP_CHECK(hipMemcpyAsync(a.u1 + off, q.u1 + o, 4 * n, hipMemcpyHostToDevice, s.stream));
      if (q.isVolume)
        PGL_HIP_CHECK(hipMemcpyAsync(a.vol + off, q.isVolume + o, n, hipMemcpyHostToDevice, s.stream));
      d.pos = a.pos + 3 * off;
      d.nrm = a.nrm + 3 * off;
      d.u2 = a.u2 + 2 * off;
      d.g = q.g ? a.g + off : nullptr;           // no HG mean cosine: 0
      d.u1 = q.u1 ? a.u1 + off : nullptr;        // no KNN random: -1 disables the KNN pick (Field.h:113)
      d.vol = q.isVolume ? a.vol + off : nullptr;  // no flags: surface queries
      d.dir = a.dir + 3 * off;
      d.pdf = a.pdf + off;
      d.id = q.id ? a.id + off : nullptr;
      d.integral = q.integral ? a.integral + off : nullptr;
    }
    off += n;
  }
  PGL_HIP_CHECK(hipMemcpyAsync(s.devTable, s.hostTable, batch.size() * sizeof(QueryPiece), hipMemcpyHostToDevice,
                               s.stream));
  launch_query_pieces(surf_[s.snap].view, vol_[s.snap].view, gf_->dHG.as<HGRep>(), s.devTable, (uint32_t)batch.size(),
                      (uint32_t)N, s.stream);
  PGL_HIP_CHECK(hipGetLastError());
  off = 0;
  for (auto& pc : batch) {
    const QueryRequest& q = pc.req->q;
    const size_t n = pc.n, o = pc.off;
    if (!pc.req->zeroCopy) {
      PGL_HIP_CHECK(hipMemcpyAsync(q.dir + 3 * o, a.dir + 3 * off, 12 * n, hipMemcpyDeviceToHost, s.stream));
      PGL_HIP_CHECK(hipMemcpyAsync(q.pdf + o, a.pdf + off, 4 * n, hipMemcpyDeviceToHost, s.stream));
      if (q.id) PGL_HIP_CHECK(hipMemcpyAsync(q.id + o, a.id + off, 4 * n, hipMemcpyDeviceToHost, s.stream));
      if (q.integral)
        PGL_HIP_CHECK(hipMemcpyAsync(q.integral + o, a.integral + off, 4 * n, hipMemcpyDeviceToHost, s.stream));
    }
    off += n;
  }
  PGL_HIP_CHECK(hipEventRecord(s.event, s.stream));
}

void QueryServer::completionLoop() {
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(mu_);
      cvDone_.wait(lk, [&] {
        if (stop_) return true;
        for (auto& s : slots_)
          if (s.busy && s.issued) return true;
        return false;
      });
      if (stop_ && inFlightBatches_ == 0) return;
    }
    bool any = false;
    for (auto& s : slots_) {
      {
        std::lock_guard<std::mutex> lk(mu_);
        if (!s.busy || !s.issued) continue;
      }
      const hipError_t e = hipEventQuery(s.event);
      if (e == hipErrorNotReady) continue;
      if (e != hipSuccess) PGL_HIP_CHECK(e);
      const auto tDone = std::chrono::steady_clock::now();
      {
        std::lock_guard<std::mutex> lk(mu_);
        for (auto& pc : s.pieces) {
          Req& r = *pc.req;
          statDoneQ_ += (double)pc.n;
          if (--r.piecesLeft == 0) {
            r.tDone = tDone;
            r.done = true;
            statDoneReq_ += 1;
          }
        }
        s.pieces.clear();
        inFlightBytes_ -= s.bytes;
        inFlightBatches_--;
        inUse_[s.snap]--;
        s.busy = false;
        s.issued = false;
        s.bytes = 0;
        s.queries = 0;
      }
      any = true;
      cvDone_.notify_all();
      cvDispatch_.notify_all();
    }
    if (!any) std::this_thread::yield();
  }
}

bool QueryServer::poll(uint64_t id) {
  std::lock_guard<std::mutex> lk(mu_);
  auto it = all_.find(id);
  if (it == all_.end()) throw std::runtime_error("server: unknown request id");
  return it->second->done;
}

bool QueryServer::wait(uint64_t id, double timeoutMs) {
  std::unique_lock<std::mutex> lk(mu_);
  auto it = all_.find(id);
  if (it == all_.end()) throw std::runtime_error("server: unknown request id");
  auto req = it->second;
  if (timeoutMs < 0) {
    cvDone_.wait(lk, [&] { return req->done; });
    return true;
  }
  return cvDone_.wait_for(lk, std::chrono::duration<double, std::milli>(timeoutMs), [&] { return req->done; });
}

void QueryServer::drain() {
  std::unique_lock<std::mutex> lk(mu_);
  cvDone_.wait(lk, [&] { return pending_.empty() && inFlightBatches_ == 0; });
}

bool QueryServer::timing(uint64_t id, double& submitMs, double& admitMs, double& doneMs) {
  std::lock_guard<std::mutex> lk(mu_);
  auto it = all_.find(id);
  if (it == all_.end() || !it->second->done) return false;
  submitMs = msSince(it->second->tSubmit);
  admitMs = msSince(it->second->tAdmit);
  doneMs = msSince(it->second->tDone);
  return true;
}

void QueryServer::forget(uint64_t id) {
  std::lock_guard<std::mutex> lk(mu_);
  auto it = all_.find(id);
  if (it != all_.end() && it->second->done) all_.erase(it);
}

void QueryServer::stats(double* out, size_t count) {
  std::lock_guard<std::mutex> lk(mu_);
  const double o[6] = {statDoneReq_, statDoneQ_, statBatches_, statPeakBytes_, statPeakPending_, statMaxBatch_};
  for (size_t i = 0; i < count && i < 6; i++) out[i] = o[i];
}

}  // namespace pglgpu
