// tm_server.hip -- serving loop over a resident graph (tm_server_* in
// include/tm_abi.h).
//
// Requests are queued by any host thread.  One dispatcher thread owns the
// graph while the server lives: it polls the launch slots, and whenever a
// slot is free it forms the next launch from what is queued (continuous
// dispatch).  By default each request is one launch, chosen highest response
// ratio first ((wait + S) / S, S the query's last service time); with
// max_batch > 1, the waiting count requests of the oldest one's stack-depth
// class go together into one multi-query launch.
// An enumerate request runs alone once its
// row buffer fits in the enumeration arena (admission by device memory).
// Completion is detected with an event per launch, the counts and scheduler
// state come back through pinned buffers, and the request's latency is
// stamped when its result is on the host.
#include <hip/hip_runtime.h>

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

#include "tm_abi.h"
#include "tm_internal.h"

using namespace tmh;
using Clock = std::chrono::steady_clock;

namespace {

double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

// Depth class of a count request: launches are formed within one class, so
// no launch sizes its LDS stack for a deeper motif than its own.
int depth_class(int M, int n) { return std::max(1, M - 2) * 2 + (n <= 4 ? 0 : 1); }  // as tm_shape_class

struct Request {
  int M = 0;
  int32_t src[TM_MAX_MOTIF_EDGES] = {}, dst[TM_MAX_MOTIF_EDGES] = {};
  int32_t delta = 0;
  int64_t cap = -1;  // < 0: count only
  int cls = 0;
  Clock::time_point submit, start, done;
  bool finished = false;
  tm_status status = TM_OK;
  int batch = 0;
  std::string key;  // the query (motif, delta, cap): its service times so far are the dispatcher's estimate
  bool limited = false;  // running under the first-run time slice
  bool probed = false;   // its time slice ran out once: the next run is a full one
  uint64_t count = 0;
  std::vector<int32_t> rows;
  size_t arena_off = 0, arena_len = 0;
};

// First-fit allocator over the enumeration arena (offsets only).
struct Arena {
  char *base = nullptr;
  size_t size = 0;
  std::map<size_t, size_t> free_;  // offset -> length
  void init(char *b, size_t n) {
    base = b;
    size = n;
    free_.clear();
    if (n) free_[0] = n;
  }
  bool alloc(size_t n, size_t *off) {
    n = (n + 255) & ~size_t(255);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= n) {
        *off = it->first;
        const size_t rest = it->second - n, at = it->first + n;
        free_.erase(it);
        if (rest) free_[at] = rest;
        return true;
      }
    }
    return false;
  }
  bool empty() const { return size == 0 || (free_.size() == 1 && free_.begin()->second == size); }
  void release(size_t off, size_t n) {
    n = (n + 255) & ~size_t(255);
    auto it = free_.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
};

struct Lane {  // one launch slot as the server drives it
  bool busy = false;
  bool enumerate = false;
  std::vector<uint64_t> ids;
  hipEvent_t done = nullptr;
  tm_sched *h_sched = nullptr;   // pinned
  uint64_t *h_counts = nullptr;  // pinned, max_batch entries
};

}  // namespace

struct tm_server {
  tm_graph *g = nullptr;
  // two concurrent launches: on the 64M-edge set at 200 requests/s, 88.7 served/s
  // against 68.9 with one slot (no overlap) and 73-80 with four (the launches
  // contend for CUs and L2)
  // one request per launch by default: in a class batch every light request
  // waits for the batch's heaviest one (64M-edge set, 300 requests offered
  // at 500/s: p50 15.7 s batched by class against 4.1 s for a plain FIFO loop)
  int slots = 2, max_batch = 1;
  size_t arena_bytes = 0;
  tm_tuning tune{};
  bool own_tune = false;  // else the graph's per-class table
  std::mutex mu;
  std::condition_variable cv_done, cv_work;
  std::unordered_map<uint64_t, std::unique_ptr<Request>> reqs;
  std::deque<uint64_t> waiting;
  uint64_t next_id = 1;
  int running = 0;
  bool stop = false;
  bool owns_graph = false;
  std::thread thr;
  Arena arena;
  Lane lane[TM_SLOTS];
  uint64_t n_launch = 0, n_count = 0, n_enum = 0, max_seen = 0;
  // service time (launch to result, ms) last seen per query: one request per
  // launch is dispatched highest response ratio first, (wait + S) / S, with S
  // this estimate (1 ms for a query not seen yet, so it runs soon and is learnt)
  std::unordered_map<std::string, double> service_ms;
  // the first run of a count query not seen yet gets this time limit (the
  // kernel's own deadline); one that runs out is re-queued for a full run
  // with an estimate of kHeavyMs: a heavy query costs at most one slice more
  // and no longer holds a launch slot ahead of light ones it could not know
  static constexpr int kSliceMs = 25;
  static constexpr double kHeavyMs = 1000.0;
};

namespace {

tm_status lane_error(const char *what, hipError_t e) {
  return set_error(TM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// Launch the next batch on a free lane; called with s->mu held.
// Returns false when nothing could be launched.
bool launch_next(tm_server *s, int k) {
  Lane &ln = s->lane[k];
  Slot &sl = s->g->slot[k];
  if (s->waiting.empty()) return false;
  // one request per launch: the highest response ratio first; class
  // batches: the oldest waiting request picks the kind and class of the
  // launch.  An enumerate request that does not fit the arena yet is passed
  // over either way.
  std::vector<uint64_t> pick;
  bool enumerate = false;
  if (s->max_batch == 1) {
    // one request per launch: highest response ratio next, so light requests
    // do not queue behind heavy ones and heavy ones still age to the front
    const Clock::time_point now = Clock::now();
    std::vector<std::pair<double, std::deque<uint64_t>::iterator>> order;
    order.reserve(s->waiting.size());
    for (auto it = s->waiting.begin(); it != s->waiting.end(); ++it) {
      const Request &r = *s->reqs[*it];
      auto c = s->service_ms.find(r.key);
      const double S = c == s->service_ms.end() ? 1.0 : std::max(c->second, 0.01);
      order.push_back({(ms_between(r.submit, now) + S) / S, it});
    }
    std::stable_sort(order.begin(), order.end(), [](const auto &a, const auto &b) { return a.first > b.first; });
    for (auto &o : order) {
      Request &r = *s->reqs[*o.second];
      if (r.cap >= 0) {
        size_t off = 0;
        const size_t len = sizeof(int32_t) * (size_t)std::max<int64_t>(r.cap, 1) * r.M;
        if (!s->arena.alloc(len, &off)) continue;  // waits for running enumerations to free the arena
        r.arena_off = off;
        r.arena_len = len;
        enumerate = true;
      } else {
        r.limited = !r.probed && s->service_ms.find(r.key) == s->service_ms.end();
      }
      pick.push_back(*o.second);
      s->waiting.erase(o.second);
      break;
    }
  }
  for (auto it = s->waiting.begin(); pick.empty() && s->max_batch > 1 && it != s->waiting.end(); ++it) {
    Request &r = *s->reqs[*it];
    if (r.cap >= 0) {
      size_t off = 0;
      const size_t len = sizeof(int32_t) * (size_t)std::max<int64_t>(r.cap, 1) * r.M;
      if (!s->arena.alloc(len, &off)) {
        if (!s->arena.empty()) continue;  // waits for running enumerations to free the arena
        // cannot fit even an empty arena (submit rejects such requests): fail it
        r.status = set_error(TM_ERR_NOMEM, "enumeration rows exceed the server's arena");
        r.finished = true;
        r.done = Clock::now();
        s->waiting.erase(it);
        s->cv_done.notify_all();
        return true;
      }
      r.arena_off = off;
      r.arena_len = len;
      pick.push_back(*it);
      enumerate = true;
      s->waiting.erase(it);
      break;
    }
    const int cls = r.cls;
    for (auto jt = it; jt != s->waiting.end() && (int)pick.size() < s->max_batch;) {
      Request &q = *s->reqs[*jt];
      if (q.cap < 0 && q.cls == cls) {
        pick.push_back(*jt);
        jt = s->waiting.erase(jt);
      } else {
        ++jt;
      }
    }
    break;
  }
  if (pick.empty()) return false;
  const int n = (int)pick.size();
  std::vector<int32_t> ne(n), src((size_t)n * TM_MAX_MOTIF_EDGES), dst((size_t)n * TM_MAX_MOTIF_EDGES), dl(n);
  const Clock::time_point now = Clock::now();
  for (int i = 0; i < n; i++) {
    Request &r = *s->reqs[pick[i]];
    ne[i] = r.M;
    memcpy(&src[(size_t)i * TM_MAX_MOTIF_EDGES], r.src, sizeof(r.src));
    memcpy(&dst[(size_t)i * TM_MAX_MOTIF_EDGES], r.dst, sizeof(r.dst));
    dl[i] = r.delta;
    r.start = now;
    r.batch = n;
  }
  tm_status st = stage_queries(s->g, sl, 0, n, nullptr, ne.data(), src.data(), dst.data(), TM_MAX_MOTIF_EDGES,
                               dl.data(), nullptr, nullptr);
  if (st == TM_OK && enumerate) {
    Request &r = *s->reqs[pick[0]];
    sl.s.h_queries[0].rows = (int *)(s->arena.base + r.arena_off);
    sl.s.h_queries[0].cap = r.cap;
  }
  // launch parameters: the server's own, else the graph's entry for the class
  const int cls = s->reqs[pick[0]]->cls;
  const tm_tuning *tune = s->own_tune ? &s->tune
                                      : ((cls >= 0 && cls < TM_NCLASSES && s->g->class_set[cls]) ? &s->g->class_tune[cls]
                                                                                                : nullptr);
  tm_tuning sliced;
  if (n == 1 && s->reqs[pick[0]]->limited) {
    if (tune) sliced = *tune; else memset(&sliced, 0, sizeof(sliced));
    if (sliced.time_limit_ms <= 0 || sliced.time_limit_ms > tm_server::kSliceMs) sliced.time_limit_ms = tm_server::kSliceMs;
    tune = &sliced;
  }
  if (st == TM_OK) st = launch(s->g, sl, 0, n, enumerate, tune, false, !enumerate && n > 1);
  hipError_t e = hipSuccess;
  if (st == TM_OK) e = hipMemcpyAsync(ln.h_sched, sl.s.d_sched, sizeof(tm_sched), hipMemcpyDeviceToHost, sl.stream);
  if (st == TM_OK && e == hipSuccess)
    e = hipMemcpyAsync(ln.h_counts, sl.s.d_counts, sizeof(uint64_t) * n, hipMemcpyDeviceToHost, sl.stream);
  if (st == TM_OK && e == hipSuccess) e = hipEventRecord(ln.done, sl.stream);
  if (st == TM_OK && e != hipSuccess) st = lane_error("server launch", e);
  if (st != TM_OK) {  // the requests fail without running
    for (uint64_t id : pick) {
      Request &r = *s->reqs[id];
      if (r.arena_len) s->arena.release(r.arena_off, r.arena_len);
      r.arena_len = 0;
      r.status = st;
      r.finished = true;
      r.done = Clock::now();
    }
    s->cv_done.notify_all();
    return true;
  }
  ln.busy = true;
  ln.enumerate = enumerate;
  ln.ids = std::move(pick);
  s->running += n;
  s->n_launch++;
  s->max_seen = std::max<uint64_t>(s->max_seen, (uint64_t)n);
  return true;
}

// Collect a lane whose launch has completed; called with s->mu held.
void complete(tm_server *s, int k) {
  Lane &ln = s->lane[k];
  Slot &sl = s->g->slot[k];
  const tm_status st = collect(s->g, sl, *ln.h_sched);
  for (size_t i = 0; i < ln.ids.size(); i++) {
    Request &r = *s->reqs[ln.ids[i]];
    if (r.limited && st == TM_ERR_TIMEOUT) {
      // its first-run slice ran out: back to the queue for a full run
      r.limited = false;
      r.probed = true;
      s->service_ms[r.key] = tm_server::kHeavyMs;
      s->waiting.push_back(ln.ids[i]);
      continue;
    }
    r.limited = false;
    r.status = st;
    r.count = ln.h_counts[i];
    if (ln.enumerate) {
      if (st == TM_OK && r.cap > 0 && r.count >= (uint64_t)r.cap) {
        // at the cap: the deterministic rows, as tm_enum returns them (the
        // launches run on this lane's slot, the other lanes go on meanwhile)
        r.rows.resize((size_t)r.cap * r.M);
        int64_t nw = 0;
        const tm_status cs = enum_cut(s->g, sl, r.M, r.src, r.dst, r.delta, 0, s->g->E - r.M + 1, r.cap, nullptr,
                                      r.rows.data(), &r.count, &nw);
        if (cs != TM_OK) {
          r.status = cs;
          r.rows.clear();
        }
      } else {
        const int64_t nrows = st == TM_OK ? std::min<int64_t>((int64_t)r.count, r.cap) : 0;
        r.rows.resize((size_t)nrows * r.M);
        if (nrows > 0) {
          hipError_t e = hipMemcpy(r.rows.data(), s->arena.base + r.arena_off, sizeof(int32_t) * r.rows.size(),
                                   hipMemcpyDeviceToHost);
          if (e != hipSuccess) {
            r.status = lane_error("server rows", e);
            r.rows.clear();
          }
        }
      }
      s->arena.release(r.arena_off, r.arena_len);
      r.arena_len = 0;
      s->n_enum++;
    } else {
      s->n_count++;
    }
    r.done = Clock::now();
    r.finished = true;
    if (ln.ids.size() == 1 && r.status == TM_OK) s->service_ms[r.key] = ms_between(r.start, r.done);
  }
  s->running -= (int)ln.ids.size();
  ln.ids.clear();
  ln.busy = false;
  s->cv_done.notify_all();
}

void dispatcher(tm_server *s, std::mutex *ready_mu, std::condition_variable *ready_cv, bool *ready) {
  std::unique_lock<std::mutex> graph_lock(s->g->mtx);  // the graph is the server's while it lives
  (void)hipSetDevice(s->g->device);
  {
    std::lock_guard<std::mutex> lk(*ready_mu);
    *ready = true;
  }
  ready_cv->notify_all();
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    bool progress = false;
    int busy = 0;
    for (int k = 0; k < s->slots; k++) {
      Lane &ln = s->lane[k];
      if (!ln.busy) continue;
      const hipError_t q = hipEventQuery(ln.done);
      if (q == hipSuccess) {
        complete(s, k);
        progress = true;
      } else if (q != hipErrorNotReady) {
        (void)hipEventSynchronize(ln.done);  // an error: drain the lane, report per request
        complete(s, k);
        progress = true;
      } else {
        busy++;
      }
    }
    for (int k = 0; k < s->slots; k++)
      if (!s->lane[k].busy && launch_next(s, k)) {
        progress = true;
        if (s->lane[k].busy) busy++;
      }
    if (s->stop && s->waiting.empty() && busy == 0) break;
    if (!progress) {
      if (busy)
        s->cv_work.wait_for(lk, std::chrono::microseconds(50));  // poll the running launches
      else
        s->cv_work.wait(lk);  // idle until a request or the stop arrives
    }
  }
}

}  // namespace

extern "C" {

tm_status tm_server_create(tm_graph *g, const tm_server_config *cfg, tm_server **out) {
  if (!g || !out) return set_error(TM_ERR_ARG, "null argument");
  *out = nullptr;
  std::unique_ptr<tm_server> s(new tm_server());
  s->g = g;
  if (cfg) {
    if (cfg->slots < 0 || cfg->slots > TM_SLOTS) return set_error(TM_ERR_ARG, "slots must be 0 .. 4");
    if (cfg->max_batch < 0 || cfg->max_batch > TM_LAUNCH_MAXQ) return set_error(TM_ERR_ARG, "max_batch must be 0 .. 512");
    if (cfg->arena_bytes < 0) return set_error(TM_ERR_ARG, "arena_bytes must be >= 0");
    if (cfg->slots) s->slots = cfg->slots;
    if (cfg->max_batch) s->max_batch = cfg->max_batch;
    // the arena hands out 256-byte-aligned pieces: its size is rounded down to match
    s->arena_bytes = cfg->arena_bytes ? ((size_t)cfg->arena_bytes & ~(size_t)255) : (size_t)1 << 30;
    if (s->arena_bytes == 0) return set_error(TM_ERR_ARG, "arena_bytes must be 0 (default) or at least 256");
    s->tune = cfg->tune;
    const tm_tuning zero{};
    s->own_tune = memcmp(&cfg->tune, &zero, sizeof(zero)) != 0;
  } else {
    s->arena_bytes = (size_t)1 << 30;
  }
  {
    std::lock_guard<std::mutex> glk(g->mtx);  // no call on the graph is running
    hipError_t e = hipSetDevice(g->device);
    if (e != hipSuccess) return lane_error("hipSetDevice", e);
    for (int k = 0; k < s->slots; k++) {
      tm_status st = reserve(g, g->slot[k], s->max_batch);
      if (st != TM_OK) return st;
    }
    char *base = nullptr;
    e = hipMalloc((void **)&base, s->arena_bytes);
    if (e != hipSuccess) return set_error(TM_ERR_NOMEM, "enumeration arena: " + std::string(hipGetErrorString(e)));
    s->arena.init(base, s->arena_bytes);
    for (int k = 0; k < s->slots; k++) {
      Lane &ln = s->lane[k];
      if ((e = hipEventCreateWithFlags(&ln.done, hipEventDisableTiming)) != hipSuccess ||
          (e = hipHostMalloc((void **)&ln.h_sched, sizeof(tm_sched))) != hipSuccess ||
          (e = hipHostMalloc((void **)&ln.h_counts, sizeof(uint64_t) * s->max_batch)) != hipSuccess) {
        for (int j = 0; j <= k; j++) {
          Lane &l2 = s->lane[j];
          if (l2.done) (void)hipEventDestroy(l2.done);
          if (l2.h_sched) (void)hipHostFree(l2.h_sched);
          if (l2.h_counts) (void)hipHostFree(l2.h_counts);
        }
        (void)hipFree(base);
        return lane_error("server lanes", e);
      }
    }
  }
  std::mutex ready_mu;
  std::condition_variable ready_cv;
  bool ready = false;
  tm_server *sp = s.get();
  s->thr = std::thread(dispatcher, sp, &ready_mu, &ready_cv, &ready);
  {
    std::unique_lock<std::mutex> lk(ready_mu);
    ready_cv.wait(lk, [&] { return ready; });
  }
  *out = s.release();
  return TM_OK;
}

tm_status tm_server_submit(tm_server *s, int32_t num_edges, const int32_t *msrc, const int32_t *mdst, int32_t delta,
                           int64_t enum_cap, uint64_t *ticket) {
  if (!s || !msrc || !mdst || !ticket) return set_error(TM_ERR_ARG, "null argument");
  if (num_edges < 2 || num_edges > TM_MAX_MOTIF_EDGES) return set_error(TM_ERR_MOTIF, "motif needs 2..8 edges");
  if (delta < 0) return set_error(TM_ERR_ARG, "delta must be >= 0");
  tm_motif_dev md;
  tm_status st = build_motif(num_edges, msrc, mdst, &md);
  if (st != TM_OK) return st;
  std::unique_ptr<Request> r(new Request());
  r->M = num_edges;
  memcpy(r->src, msrc, sizeof(int32_t) * num_edges);
  memcpy(r->dst, mdst, sizeof(int32_t) * num_edges);
  r->delta = delta;
  r->cap = enum_cap;
  r->cls = depth_class(md.M, md.n);
  {
    int32_t k[2 + 2 * TM_MAX_MOTIF_EDGES] = {num_edges, delta};
    memcpy(k + 2, msrc, sizeof(int32_t) * num_edges);
    memcpy(k + 2 + num_edges, mdst, sizeof(int32_t) * num_edges);
    r->key.assign((const char *)k, sizeof(int32_t) * (2 + 2 * num_edges));
    r->key.append((const char *)&enum_cap, sizeof(enum_cap));
  }
  if (enum_cap >= 0) {
    // the piece the arena would hand out (rounded up to 256 bytes) must fit
    // an empty arena, and its size must not overflow
    const int64_t most = (int64_t)(s->arena_bytes / (sizeof(int32_t) * (size_t)num_edges));
    if (enum_cap > most) return set_error(TM_ERR_NOMEM, "enumeration rows exceed the server's arena");
    const size_t len = sizeof(int32_t) * (size_t)std::max<int64_t>(enum_cap, 1) * num_edges;
    if (((len + 255) & ~(size_t)255) > s->arena_bytes)
      return set_error(TM_ERR_NOMEM, "enumeration rows exceed the server's arena");
  }
  r->submit = Clock::now();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return set_error(TM_ERR_ARG, "server is stopping");
    const uint64_t id = s->next_id++;
    s->reqs[id] = std::move(r);
    s->waiting.push_back(id);
    *ticket = id;
  }
  s->cv_work.notify_one();
  return TM_OK;
}

tm_status tm_server_wait(tm_server *s, uint64_t ticket, int32_t timeout_ms, tm_server_result *res) {
  if (!s || !res) return set_error(TM_ERR_ARG, "null argument");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return set_error(TM_ERR_ARG, "unknown ticket");
  Request *r = it->second.get();
  auto ready = [&] { return r->finished; };
  if (timeout_ms < 0) {
    s->cv_done.wait(lk, ready);
  } else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready)) {
    return set_error(TM_ERR_TIMEOUT, "request still running");
  }
  res->status = r->status;
  res->batch = r->batch;
  res->count = r->count;
  res->rows = (int64_t)(r->M ? r->rows.size() / r->M : 0);
  res->latency_ms = ms_between(r->submit, r->done);
  res->queue_ms = r->batch ? ms_between(r->submit, r->start) : res->latency_ms;
  return TM_OK;
}

tm_status tm_server_take_rows(tm_server *s, uint64_t ticket, int32_t *rows, int64_t max_rows, int64_t *rows_written) {
  if (!s || (max_rows > 0 && !rows)) return set_error(TM_ERR_ARG, "bad argument");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return set_error(TM_ERR_ARG, "unknown ticket");
  Request &r = *it->second;
  if (!r.finished) return set_error(TM_ERR_ARG, "request still running");
  const int64_t n = std::min<int64_t>(max_rows, (int64_t)(r.rows.size() / std::max(r.M, 1)));
  if (n > 0) memcpy(rows, r.rows.data(), sizeof(int32_t) * (size_t)n * r.M);
  if (rows_written) *rows_written = n;
  s->reqs.erase(it);
  return TM_OK;
}

tm_status tm_server_release(tm_server *s, uint64_t ticket) {
  if (!s) return set_error(TM_ERR_ARG, "null server");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return set_error(TM_ERR_ARG, "unknown ticket");
  if (!it->second->finished) return set_error(TM_ERR_ARG, "request still running");
  s->reqs.erase(it);
  return TM_OK;
}

tm_status tm_server_stats(tm_server *s, uint64_t *stats, int32_t n) {
  if (!s || !stats || n < 0) return set_error(TM_ERR_ARG, "bad argument");
  std::lock_guard<std::mutex> lk(s->mu);
  const uint64_t v[6] = {s->n_launch, s->n_count, s->n_enum, s->max_seen, (uint64_t)s->waiting.size(),
                         (uint64_t)s->running};
  for (int i = 0; i < n && i < 6; i++) stats[i] = v[i];
  return TM_OK;
}

tm_status tm_server_destroy(tm_server *s) {
  if (!s) return TM_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  if (s->thr.joinable()) s->thr.join();
  for (int k = 0; k < s->slots; k++) {
    Lane &ln = s->lane[k];
    if (ln.done) (void)hipEventDestroy(ln.done);
    if (ln.h_sched) (void)hipHostFree(ln.h_sched);
    if (ln.h_counts) (void)hipHostFree(ln.h_counts);
  }
  if (s->arena.base) (void)hipFree(s->arena.base);
  delete s;
  return TM_OK;
}

}  // extern "C"
