// Serving runtime: requests of mixed sizes / precisions / directions are
// admitted in arrival order under an HBM budget and run on a few internal
// streams; one waiter thread per stream blocks on the completion event of
// that stream's oldest request, stamps the completion and admits the next
// requests.
//
// A request's life (host-staged):
//   submit -> waiting FIFO (small or large) -> admitted (budget, a stream
//   with room, arena block) -> H2D copy, transform passes, D2H copy on that
//   stream -> its waiter returns from the last event -> block freed, result
//   queued.
// Device requests skip the copies and are transformed in place; with a
// producer stream named, the server's stream first waits for an event
// recorded on the producer at submission.
//
// Plans are built once per (Nx, Ny, Nz, prec) and shared by every request of
// that geometry; exec only reads a plan, so the same plan runs on several
// streams at once.  A request with its own output scale executes a shallow
// copy of the plan with that scale (same kernels, same twiddle tables).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdint>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <tuple>
#include <vector>

#include "gfft_arena.h"
#include "gfft_plan.h"
#include "gpufft/gpufft.h"

// The ABI structs are bound field by field from other languages.
static_assert(sizeof(gfft_request) == 88, "gfft_request layout");
static_assert(sizeof(gfft_result) == 80, "gfft_result layout");
static_assert(sizeof(gfft_server_stats_t) == 96, "gfft_server_stats_t layout");

namespace {

using Clock = std::chrono::steady_clock;

// default bound of the reserved stream's requests: 32 MiB (a 200^3 FP32 or a
// 160^3 FP64 array)
constexpr size_t kDefaultSmallBytes = size_t(32) << 20;
// default cap of admitted, unfinished requests per stream: one, so a
// request is bound to a stream only when that stream is free and every wait
// happens (and is stamped) in the server's FIFO.  Measured on MI355X (5000
// served requests, 80 % load): capacity and p99 equal to depth 2 or no cap
// within run-to-run noise.
constexpr size_t kDefaultDepth = 1;
// default sleep between completion queries of a stream's waiter, in
// microseconds (0: block in hipEventSynchronize; its interrupt-driven
// wake-up was measured ~1.5 ms late on MI355X, too slow to keep the
// depth-capped streams fed)
constexpr int kDefaultPollUs = 20;

double now_s() { return std::chrono::duration<double>(Clock::now().time_since_epoch()).count(); }

struct Events {
  // before copy-in, before transform, after transform, after copy-out,
  // producer's work done (recorded on the producer stream at submission)
  hipEvent_t e[5] = {};
};

struct Req {
  uint64_t id = 0;
  gfft_request r{};
  size_t bytes = 0;
  bool device = false;
  const gfft::PlanCore* plan = nullptr;
  size_t off = gfft::Arena::kNone;
  int queue = 0;   // admission FIFO: 1 = at or under the small-lane threshold
  bool after_producer = false;
  int stream = -1;
  int ev = -1;     // index into the event pool
  Events e;        // copy of the pool entry's handles (the pool may grow meanwhile)
  double t_adm = 0;
  hipEvent_t done_event() const { return e.e[device ? 2 : 3]; }
};

// Requests on one stream finish in submission order: the stream's waiter
// only ever waits for the oldest one.
struct Lane {
  std::deque<Req*> q;
  size_t bytes = 0;  // array bytes of the lane's unfinished requests
  std::condition_variable cv;  // waiter: a request was queued, or stop
  std::thread waiter;
};

using Key = std::tuple<int64_t, int64_t, int64_t, int>;

int request_bytes(int64_t nx, int64_t ny, int64_t nz, int prec, size_t* bytes) {
  const int st = gfft::validate_dims(nx, ny, nz, 1, 0, prec);
  if (st != GFFT_OK) return st;
  const double b = (double)nx * (double)ny * (double)(2 * (nz / 2 + 1)) * (prec == GFFT_F32 ? 4.0 : 8.0);
  if (b > 9.0e18) return GFFT_ERR_INVALID_DIMS;
  *bytes = (size_t)nx * (size_t)ny * (size_t)(2 * (nz / 2 + 1)) * (prec == GFFT_F32 ? 4u : 8u);
  return GFFT_OK;
}

}  // namespace

struct gfft_server_s {
  int device = 0;
  int flags = 0;
  char* arena_base = nullptr;
  gfft::Arena arena;
  gfft::Admission<Req*> adm;
  std::vector<hipStream_t> streams;
  std::vector<std::unique_ptr<Lane>> lanes;  // one per stream
  size_t n_running = 0;
  // requests of at most this many bytes may use the last stream, which no
  // larger request uses (0: no reserved stream)
  size_t small_bytes = 0;
  size_t max_depth = kDefaultDepth;  // unfinished requests per stream (0: no cap)
  int poll_us = kDefaultPollUs;      // waiter: sleep between queries (0: blocking synchronise)
  std::map<Key, std::unique_ptr<gfft::PlanCore>> plans;
  std::map<Key, int> checked;  // host-side plan status of geometries seen at submission
  size_t plan_bytes = 0;
  std::vector<Events> ev_pool;
  std::vector<int> ev_free;
  std::vector<gfft_result> done;
  size_t done_head = 0;
  uint64_t next_id = 1, submitted = 0, completed = 0, rejected = 0;
  std::mutex mu;
  std::condition_variable cv_idle;  // gfft_server_wait: nothing waiting or running
  bool stop = false;
};

namespace {

int get_plan(gfft_server_s* s, int64_t nx, int64_t ny, int64_t nz, int prec, const gfft::PlanCore** out) {
  const Key k{nx, ny, nz, prec};
  auto it = s->plans.find(k);
  if (it != s->plans.end()) {
    *out = it->second.get();
    return GFFT_OK;
  }
  std::unique_ptr<gfft::PlanCore> p(new (std::nothrow) gfft::PlanCore);
  if (!p) return GFFT_ERR_OUT_OF_MEMORY;
  int st = gfft::plan_host(*p, nx, ny, nz, 1, 0, prec);
  if (st == GFFT_OK) st = gfft::upload_twiddles(*p);
  if (st != GFFT_OK) {
    gfft::free_twiddles(*p);
    return st;
  }
  (void)gfft::apply_tune_cache(*p);  // tuned choices of this geometry, if any
  s->plan_bytes += gfft::plan_device_bytes(*p);
  *out = p.get();
  s->plans.emplace(k, std::move(p));
  return GFFT_OK;
}

int take_events(gfft_server_s* s) {
  if (!s->ev_free.empty()) {
    const int i = s->ev_free.back();
    s->ev_free.pop_back();
    return i;
  }
  Events e;
  // blocking events: a waiter sleeps in hipEventSynchronize instead of spinning
  for (int j = 0; j < 5; ++j)
    if (hipEventCreateWithFlags(&e.e[j], hipEventBlockingSync) != hipSuccess) {
      for (int q = 0; q < j; ++q) (void)hipEventDestroy(e.e[q]);
      return -1;
    }
  s->ev_pool.push_back(e);
  return (int)s->ev_pool.size() - 1;
}

void finish(gfft_server_s* s, Req* q, int status, double t_done) {
  gfft_result r{};
  r.id = q->id;
  r.tag = q->r.tag;
  r.status = status;
  r.stream = q->stream;
  r.arrival = q->r.arrival;
  r.admitted = q->t_adm;
  r.done = t_done;
  r.bytes = q->bytes;
  r.started = t_done;
  if (q->t_adm > 0 && q->ev >= 0 && status == GFFT_OK) {
    const Events& e = q->e;
    float a = 0, b = 0;
    if (hipEventElapsedTime(&a, e.e[1], e.e[2]) == hipSuccess) r.gpu_ms = a;
    if (q->device)
      r.staged_ms = r.gpu_ms;
    else if (hipEventElapsedTime(&b, e.e[0], e.e[3]) == hipSuccess)
      r.staged_ms = b;
    // the device reached the request's first event staged_ms before its last
    r.started = t_done - r.staged_ms * 1e-3;
    if (r.started < q->t_adm) r.started = q->t_adm;
  }
  if (q->ev >= 0) s->ev_free.push_back(q->ev);
  if (q->off != gfft::Arena::kNone) s->arena.release(q->off, q->bytes);
  if (q->stream >= 0) s->lanes[q->stream]->bytes -= q->bytes;
  if (q->t_adm > 0) s->adm.finished(q->bytes);
  s->done.push_back(r);
  ++s->completed;
  delete q;
}

// Enqueue copy-in, passes, copy-out of an admitted request on its stream.
int launch(gfft_server_s* s, Req* q) {
  const hipStream_t st = s->streams[q->stream];
  const Events& e = q->e;
  void* data = q->device ? q->r.data : s->arena_base + q->off;
  if (q->after_producer && hipStreamWaitEvent(st, e.e[4], 0) != hipSuccess) return GFFT_ERR_HIP;
  if (!q->device) {
    if (hipEventRecord(e.e[0], st) != hipSuccess) return GFFT_ERR_HIP;
    if (hipMemcpyAsync(data, q->r.data, q->bytes, hipMemcpyHostToDevice, st) != hipSuccess) return GFFT_ERR_HIP;
  }
  if (hipEventRecord(e.e[1], st) != hipSuccess) return GFFT_ERR_HIP;
  int rc;
  const double sc = q->r.scale == 0.0 ? 1.0 : q->r.scale;
  if (sc == 1.0 && q->plan->r2c_scale == 1.0 && q->plan->c2r_scale == 1.0) {
    rc = q->r.direction == 0 ? gfft::exec_r2c(*q->plan, data, st) : gfft::exec_c2r(*q->plan, data, st);
  } else {
    gfft::PlanCore c = *q->plan;  // shares the plan's device tables
    c.r2c_scale = q->r.direction == 0 ? sc : 1.0;
    c.c2r_scale = q->r.direction == 1 ? sc : 1.0;
    rc = q->r.direction == 0 ? gfft::exec_r2c(c, data, st) : gfft::exec_c2r(c, data, st);
  }
  if (rc != GFFT_OK) return rc;
  if (hipEventRecord(e.e[2], st) != hipSuccess) return GFFT_ERR_HIP;
  if (!q->device) {
    if (hipMemcpyAsync(q->r.host_out, data, q->bytes, hipMemcpyDeviceToHost, st) != hipSuccess) return GFFT_ERR_HIP;
    if (hipEventRecord(e.e[3], st) != hipSuccess) return GFFT_ERR_HIP;
  }
  return GFFT_OK;
}

// Admit FIFO heads while one fits: of the two FIFOs' heads the earlier
// arrival is tried first; a head that cannot start yet (budget, no stream
// with room, no arena block) does not block the other FIFO.  Caller holds mu.
void admit(gfft_server_s* s) {
  const int nl = (int)s->lanes.size();
  const bool reserve = s->small_bytes > 0 && nl >= 2;
  std::vector<size_t> depth(nl), lbytes(nl);
  for (;;) {
    int order[2] = {0, 1};
    if (!s->adm.empty(0) && !s->adm.empty(1) && s->adm.head(1)->id < s->adm.head(0)->id) std::swap(order[0], order[1]);
    bool admitted = false;
    for (int k = 0; k < 2 && !admitted; ++k) {
      const int fq = order[k];
      if (s->adm.empty(fq)) continue;
      Req* q = s->adm.head(fq);
      if (!s->adm.fits(q->bytes)) continue;
      for (int i = 0; i < nl; ++i) {
        depth[i] = s->lanes[i]->q.size();
        lbytes[i] = s->lanes[i]->bytes;
      }
      const int lane = gfft::pick_stream(depth.data(), lbytes.data(), nl, reserve, fq == 1, s->max_depth);
      if (lane < 0) continue;
      size_t off = gfft::Arena::kNone;
      if (!q->device) {
        off = s->arena.alloc(q->bytes);
        if (off == gfft::Arena::kNone) continue;
      }
      s->adm.pop_admitted(q->bytes, fq);
      admitted = true;
      q->off = off;
      q->t_adm = now_s();
      int st = q->plan ? GFFT_OK : get_plan(s, q->r.nx, q->r.ny, q->r.nz, q->r.prec, &q->plan);
      if (st == GFFT_OK) {
        q->stream = lane;
        s->lanes[lane]->bytes += q->bytes;
        st = launch(s, q);
      }
      if (st != GFFT_OK) {
        // nothing of it may still be in flight when its block is reused
        if (q->stream >= 0) (void)hipStreamSynchronize(s->streams[q->stream]);
        finish(s, q, st, now_s());
        continue;
      }
      s->lanes[lane]->q.push_back(q);
      ++s->n_running;
      s->lanes[lane]->cv.notify_one();
    }
    if (!admitted) break;
  }
  if (s->adm.empty() && s->n_running == 0) s->cv_idle.notify_all();
}

// Waiter of one stream: waits for the completion of the stream's oldest
// request with the server lock released, either sleeping poll_us between
// event queries or blocking in hipEventSynchronize (blocking-sync events);
// only this thread removes requests from its lane, so the oldest request
// stays valid.
void lane_loop(gfft_server_s* s, int li) {
  (void)hipSetDevice(s->device);
  Lane& L = *s->lanes[li];
  for (;;) {
    Req* q = nullptr;
    int poll_us = 0;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      L.cv.wait(lk, [&] { return s->stop || !L.q.empty(); });
      if (L.q.empty()) return;  // stop requested and nothing left
      q = L.q.front();
      poll_us = s->poll_us;
    }
    hipError_t e;
    if (poll_us <= 0) {
      e = hipEventSynchronize(q->done_event());
    } else {
      while ((e = hipEventQuery(q->done_event())) == hipErrorNotReady)
        std::this_thread::sleep_for(std::chrono::microseconds(poll_us));
    }
    const double t = now_s();
    std::lock_guard<std::mutex> lk(s->mu);
    L.q.pop_front();
    --s->n_running;
    finish(s, q, e == hipSuccess ? GFFT_OK : GFFT_ERR_HIP, t);
    admit(s);
  }
}

void teardown(gfft_server_s* s) {
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  for (auto& l : s->lanes) {
    l->cv.notify_all();
    if (l->waiter.joinable()) l->waiter.join();
  }
  for (hipStream_t st : s->streams) (void)hipStreamSynchronize(st);
  for (auto& l : s->lanes)
    for (Req* q : l->q) delete q;
  for (int fq = 0; fq < 2; ++fq)
    while (!s->adm.empty(fq)) {
      delete s->adm.head(fq);
      s->adm.pop_admitted(0, fq);
    }
  for (auto& kv : s->plans) gfft::free_twiddles(*kv.second);
  for (Events& e : s->ev_pool)
    for (hipEvent_t x : e.e)
      if (x) (void)hipEventDestroy(x);
  for (hipStream_t st : s->streams) (void)hipStreamDestroy(st);
  if (s->arena_base) (void)hipFree(s->arena_base);
}

}  // namespace

extern "C" {

double gfft_server_clock(void) { return now_s(); }

int gfft_request_bytes(int64_t nx, int64_t ny, int64_t nz, int prec, size_t* bytes) {
  if (!bytes) return GFFT_ERR_INVALID_ARG;
  return request_bytes(nx, ny, nz, prec, bytes);
}

int gfft_server_set_small_lane(gfft_server s, size_t max_bytes) {
  if (!s) return GFFT_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  s->small_bytes = max_bytes;
  return GFFT_OK;
}

int gfft_server_set_poll_interval(gfft_server s, int microseconds) {
  if (!s || microseconds < 0) return GFFT_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  s->poll_us = microseconds;
  return GFFT_OK;
}

int gfft_server_set_stream_depth(gfft_server s, int depth) {
  if (!s || depth < 0) return GFFT_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  s->max_depth = (size_t)depth;
  admit(s);  // a larger cap may let waiting requests start now
  return GFFT_OK;
}

int gfft_server_create(gfft_server* out, size_t budget_bytes, int nstreams, int flags) {
  if (!out || budget_bytes == 0 || nstreams < 1 || nstreams > 64 ||
      (flags & ~(GFFT_SERVER_STAGING | GFFT_SERVER_SHARED_QUEUES)))
    return GFFT_ERR_INVALID_ARG;
  *out = nullptr;
  gfft_server_s* s = new (std::nothrow) gfft_server_s;
  if (!s) return GFFT_ERR_OUT_OF_MEMORY;
  s->flags = flags;
  s->small_bytes = nstreams >= 2 ? kDefaultSmallBytes : 0;
  s->adm = gfft::Admission<Req*>(budget_bytes);
  int st = hipGetDevice(&s->device) == hipSuccess ? GFFT_OK : GFFT_ERR_HIP;
  if (st == GFFT_OK && (flags & GFFT_SERVER_STAGING)) {
    if (hipMalloc(&s->arena_base, budget_bytes) != hipSuccess) {
      (void)hipGetLastError();
      s->arena_base = nullptr;
      st = GFFT_ERR_OUT_OF_MEMORY;
    } else {
      s->arena.reset(budget_bytes);
    }
  }
  // Streams with a CU mask (every CU enabled) get a hardware queue each.
  // Plain streams come from the runtime's pool of a few queues; once the
  // process holds more streams than the pool has queues (or has created and
  // destroyed some), two server streams can land on one queue, and a small
  // request then waits behind the large one on the shared queue.
  std::vector<uint32_t> cu_mask;
  if (st == GFFT_OK && !(flags & GFFT_SERVER_SHARED_QUEUES)) {
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, s->device) != hipSuccess || ncu < 1) {
      st = GFFT_ERR_HIP;
    } else {
      cu_mask.assign((size_t)(ncu + 31) / 32, 0xffffffffu);
      if (ncu % 32) cu_mask.back() = (1u << (ncu % 32)) - 1u;
    }
  }
  for (int i = 0; i < nstreams && st == GFFT_OK; ++i) {
    hipStream_t h = nullptr;
    const hipError_t e = cu_mask.empty()
                             ? hipStreamCreateWithFlags(&h, hipStreamNonBlocking)
                             : hipExtStreamCreateWithCUMask(&h, (uint32_t)cu_mask.size(), cu_mask.data());
    if (e != hipSuccess) {
      st = GFFT_ERR_HIP;
      break;
    }
    s->streams.push_back(h);
    s->lanes.emplace_back(new (std::nothrow) Lane);
    if (!s->lanes.back()) {
      s->lanes.pop_back();
      st = GFFT_ERR_OUT_OF_MEMORY;
    }
  }
  for (int i = 0; i < (int)s->lanes.size() && st == GFFT_OK; ++i) {
    try {
      s->lanes[i]->waiter = std::thread(lane_loop, s, i);
    } catch (...) {
      st = GFFT_ERR_OUT_OF_MEMORY;
    }
  }
  if (st != GFFT_OK) {
    teardown(s);
    delete s;
    return st;
  }
  *out = s;
  return GFFT_OK;
}

int gfft_server_prepare(gfft_server s, int64_t nx, int64_t ny, int64_t nz, int prec) {
  if (!s) return GFFT_ERR_INVALID_ARG;
  size_t b = 0;
  int st = request_bytes(nx, ny, nz, prec, &b);
  if (st != GFFT_OK) return st;
  std::lock_guard<std::mutex> lk(s->mu);
  const gfft::PlanCore* p = nullptr;
  return get_plan(s, nx, ny, nz, prec, &p);
}

int gfft_server_submit(gfft_server s, const gfft_request* req, uint64_t* id) {
  if (!s || !req || !req->data || (req->direction != 0 && req->direction != 1) ||
      (req->flags & ~(GFFT_REQ_DEVICE | GFFT_REQ_PRODUCER)) || req->reserved != 0)
    return GFFT_ERR_INVALID_ARG;
  const bool device = (req->flags & GFFT_REQ_DEVICE) != 0;
  if (!device && (!req->host_out || !s->arena_base)) return GFFT_ERR_INVALID_ARG;
  size_t bytes = 0;
  int st = request_bytes(req->nx, req->ny, req->nz, req->prec, &bytes);
  if (st != GFFT_OK) return st;
  // device arrays are transformed in place: whole complex elements (8 / 16 B)
  if (device && (reinterpret_cast<uintptr_t>(req->data) & (req->prec == GFFT_F64 ? 15u : 7u)) != 0)
    return GFFT_ERR_INVALID_ARG;
  {
    // sizes the planner cannot serve (a prime factor > 13, ...) fail here,
    // not at admission; the host-only check is cached per geometry
    const Key k{req->nx, req->ny, req->nz, req->prec};
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->checked.find(k);
    if (it == s->checked.end()) {
      gfft::PlanCore tmp;
      it = s->checked.emplace(k, gfft::plan_host(tmp, req->nx, req->ny, req->nz, 1, 0, req->prec)).first;
    }
    if (it->second != GFFT_OK) return it->second;
  }
  // a host-staged request needs one arena block (256-byte multiples, arena
  // = the budget rounded down): one that can never exist would block its
  // FIFO for good
  if (bytes > s->adm.budget() || (!device && !s->arena.can_hold(bytes))) {
    std::lock_guard<std::mutex> lk(s->mu);
    ++s->rejected;
    return GFFT_ERR_OVER_BUDGET;
  }
  if (device) {
    hipPointerAttribute_t a{};
    if (hipPointerGetAttributes(&a, req->data) != hipSuccess || a.type != hipMemoryTypeDevice ||
        a.device != s->device) {
      (void)hipGetLastError();
      return GFFT_ERR_INVALID_ARG;
    }
  }
  int cur = -1;
  if (hipGetDevice(&cur) != hipSuccess) return GFFT_ERR_HIP;
  if (cur != s->device) return GFFT_ERR_DEVICE_MISMATCH;
  Req* q = new (std::nothrow) Req;
  if (!q) return GFFT_ERR_OUT_OF_MEMORY;
  q->r = *req;
  q->bytes = bytes;
  q->device = device;
  // The legacy null stream as producer needs no event with the default
  // (blocking) server streams: work queued on a blocking stream already
  // waits for the null stream's earlier work.  An event recorded on the
  // null stream would also wait for every blocking stream's queued work,
  // serialising the server's streams.
  q->after_producer = (req->flags & GFFT_REQ_PRODUCER) != 0 &&
                      !(req->producer == nullptr && !(s->flags & GFFT_SERVER_SHARED_QUEUES));
  if (q->r.arrival < 0) q->r.arrival = now_s();
  std::lock_guard<std::mutex> lk(s->mu);
  q->ev = take_events(s);
  if (q->ev < 0) {
    delete q;
    return GFFT_ERR_HIP;
  }
  q->e = s->ev_pool[q->ev];
  // the producer's work queued so far; the server's stream waits for it
  if (q->after_producer && hipEventRecord(q->e.e[4], reinterpret_cast<hipStream_t>(req->producer)) != hipSuccess) {
    s->ev_free.push_back(q->ev);
    delete q;
    return GFFT_ERR_HIP;
  }
  q->queue = (s->small_bytes > 0 && s->lanes.size() >= 2 && bytes <= s->small_bytes) ? 1 : 0;
  q->id = s->next_id++;
  ++s->submitted;
  if (id) *id = q->id;
  s->adm.push(q, q->queue);
  admit(s);
  return GFFT_OK;
}

int gfft_server_poll(gfft_server s, gfft_result* out, int max, int* count) {
  if (!s || !count || max < 0 || (max > 0 && !out)) return GFFT_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  int n = 0;
  while (n < max && s->done_head < s->done.size()) out[n++] = s->done[s->done_head++];
  if (s->done_head == s->done.size()) {
    s->done.clear();
    s->done_head = 0;
  }
  *count = n;
  return GFFT_OK;
}

int gfft_server_wait(gfft_server s, double timeout_s) {
  if (!s) return GFFT_ERR_INVALID_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  auto idle = [&] { return s->adm.empty() && s->n_running == 0; };
  if (timeout_s <= 0) {
    s->cv_idle.wait(lk, idle);
    return GFFT_OK;
  }
  return s->cv_idle.wait_for(lk, std::chrono::duration<double>(timeout_s), idle) ? GFFT_OK : GFFT_ERR_TIMEOUT;
}

int gfft_server_stats(gfft_server s, gfft_server_stats_t* st) {
  if (!s || !st) return GFFT_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  std::memset(st, 0, sizeof *st);
  st->submitted = s->submitted;
  st->completed = s->completed;
  st->rejected = s->rejected;
  st->waiting = s->adm.waiting();
  st->running = s->n_running;
  st->plans = s->plans.size();
  st->budget = s->adm.budget();
  st->in_flight_bytes = s->adm.in_flight();
  st->peak_in_flight_bytes = s->adm.peak();
  st->arena_bytes = s->arena.size();
  st->arena_used = s->arena.used();
  st->plan_bytes = s->plan_bytes;
  return GFFT_OK;
}

int gfft_server_destroy(gfft_server s) {
  if (!s) return GFFT_ERR_INVALID_ARG;
  (void)gfft_server_wait(s, 0);
  teardown(s);
  delete s;
  return GFFT_OK;
}

}  // extern "C"
