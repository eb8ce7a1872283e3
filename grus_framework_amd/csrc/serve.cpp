// serve.cpp -- continuous-batching query server over the multi-source engines.
//
// No reference counterpart (the reference answers one --src per process).
// The graph stays resident; each query kind (BFS, SSSP) has one or more
// engines, each owning 1..64 lane slots of device state and its own HIP
// stream, driven by its own worker thread.  Between two steps of its engine
// a worker admits waiting queries into free lanes, so queries that arrive
// while others run join the batch in flight, and engines of different kinds
// overlap on the device.  A lane whose frontier empties is finished: at the
// end of that step its result is digested (and copied to the caller's
// buffer when requested), the query is reported, and the slot is freed.
//
// A BFS engine holds one query per lane, an SSSP engine two (16-bit
// distances: 128 queries per batch for the bytes of 64 in 32 bits).  Under
// light load an idle engine's worker runs waiting queries one at a time with
// the single-source kernels on its own view of the graph: a batch costs a
// step per level whatever it holds.  Load is light while the engines of a
// kind, each serving one query at a time, would empty the queue within the
// time of one batch (running means of both times, per kind; until a batch
// has run, a batch is taken to cost 32 single queries, about what one of 64
// lanes costs on RMAT graphs), or while at most solo_max queries wait when
// the caller sets it.  Measured on RMAT scale 22 at 1100 queries/s (50/50
// BFS/SSSP): p50 5 ms when every query is served alone, 99-120 ms when a
// queue of 8 already starts a batch (which then slows the single queries
// beside it).
//
// Memory.  Each worker owns a context (stream).  When memory is capped
// (state_budget_bytes or a device budget), its engine's device state (the
// per-vertex rows of its slots) is taken when a batch needs it and fits
// under both, and given back after idle_release_ms without a query;
// queries wait while there is no room.  The first engine of each kind
// holds its state from creation, so every kind always makes progress.
// Uncapped, every engine holds its state from creation.
//
// Admission.  An idle engine takes up to a full batch.  A running engine
// takes arrivals into its free lanes during the first admit_window steps
// after it started from idle (the batch is still young: the newcomers trail
// its first lanes by a step or two), and later only once at least
// admit_min_free lanes are free.  Lanes admitted together stay in step: on
// small-world graphs their frontiers cover mostly the same vertices at the
// same step, and a step of the engine costs about one sweep of their union.
// Lanes admitted one at a time drift apart, every step sweeps a large union,
// and the engine spends its bandwidth on few queries at a time (RMAT scale
// 22, 50/50 BFS/SSSP, saturated, 4 engines per kind: 2780 queries/s when
// only idle engines admit against 1760 when any free lane admits).
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

#include "apps.hpp"
#include "multi.hpp"
#include "serve.hpp"

namespace mgg {

namespace {

double now_us() {
  using namespace std::chrono;
  return (double)duration_cast<nanoseconds>(steady_clock::now().time_since_epoch()).count() /
         1e3;
}

struct Query {
  uint64_t id;
  int kind;
  uint32_t src;
  uint32_t *out;
  double submit_us, start_us;
};

class ServerImpl : public Server {
 public:
  ServerImpl(Graph &g, const grus_server_opts *o) : g_(g) {
    opts_.max_inflight_batches = 4;
    opts_.max_batch = 64;
    opts_.state_budget_bytes = 0;
    opts_.copy_results = 0;
    opts_.admit_min_free = 0;
    opts_.admit_window = 0;
    if (o) {
      if (o->max_inflight_batches > 0) opts_.max_inflight_batches = o->max_inflight_batches;
      if (o->max_batch > 0) opts_.max_batch = o->max_batch;
      opts_.state_budget_bytes = o->state_budget_bytes;
      opts_.copy_results = o->copy_results;
      opts_.admit_min_free = o->admit_min_free;
      opts_.admit_window = o->admit_window;
    }
    window_ = opts_.admit_window == 0 ? 2 : std::max(0, opts_.admit_window);
    opts_.solo_max = o ? o->solo_max : 0;
    solo_max_ = opts_.solo_max == 0 ? 8 : std::max(0, opts_.solo_max);
    opts_.idle_release_ms = o ? o->idle_release_ms : 0;
    if (opts_.max_batch > 64) raise(GRUS_E_INVALID, "max_batch is at most 64");
    if (opts_.admit_min_free < 0 || opts_.admit_min_free > opts_.max_batch)
      raise(GRUS_E_INVALID, "admit_min_free is between 0 and max_batch");
    if (opts_.max_inflight_batches > 8) raise(GRUS_E_INVALID, "at most 8 engines per kind");
    // lanes per engine: one engine of each kind must fit the state budget
    // (a BFS lane holds a V-entry level row, an SSSP lane a V-entry column
    // of two 16-bit distances; both have V-entry masks, queues and a degree
    // scan beside, at most 52 bytes per vertex)
    int lanes = opts_.max_batch;
    if (opts_.state_budget_bytes) {
      uint64_t per_lane = g.V * 4;
      uint64_t shared = g.V * 52;
      uint64_t per_engine = opts_.state_budget_bytes / 2;
      if (per_engine <= shared + per_lane)
        raise(GRUS_E_BUDGET, "state budget too small for one lane per engine of each kind");
      lanes = (int)std::min<uint64_t>((uint64_t)lanes, (per_engine - shared) / per_lane);
    }
    // admission gate, in lanes of an engine (capped by a budget-shrunk width);
    // an SSSP engine holds two queries per lane, so its gate in slots doubles
    min_free_ = opts_.admit_min_free ? opts_.admit_min_free : opts_.max_batch;
    min_free_ = std::min(min_free_, lanes);
    lanes_ = lanes;
    const bool capped = opts_.state_budget_bytes || devmem_budget(g.ctx->device);
    release_ms_ = opts_.idle_release_ms < 0   ? -1.0
                  : opts_.idle_release_ms > 0 ? (double)opts_.idle_release_ms
                  : capped                    ? 1000.0
                                              : -1.0;
    if (!g.symmetric) g.csc();  // build the CSC once, before workers share it
    per_kind_ = opts_.max_inflight_batches;
    const int n = 2 * per_kind_;
    for (int i = 0; i < n; i++) {
      auto c = std::make_unique<Context>(g.ctx->device);
      c->tuning = g.ctx->tuning;
      c->shared_device = true;  // engines of a server run side by side
      ctxs_.push_back(std::move(c));
    }
    engines_.resize(n);
    state_.assign(n, 0);
    // the first engine of each kind holds its state from the start; with
    // memory uncapped every engine does (nothing to admit by, and no
    // allocation in a worker thread while the others launch)
    for (int i = 0; i < n; i++) {
      if (capped && i % per_kind_ != 0) continue;
      if (!start_engine((size_t)i))
        raise(GRUS_E_BUDGET, "no device room for one engine of each query kind");
    }
    // each worker's view of the graph for single-source queries; an SSSP
    // view builds its pull-round layout now, not in its first query
    if (solo_max_ > 0) {
      for (int i = 0; i < n; i++) {
        views_.push_back(graph_view(g, ctxs_[i].get()));
        if (kind_of(i) == GRUS_Q_SSSP) (void)sssp_binned_ready(*views_.back());
      }
    }
    for (int i = 0; i < n; i++) workers_.emplace_back([this, i] { this->work((size_t)i); });
  }

  ~ServerImpl() override {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_work_.notify_all();
    for (auto &t : workers_) t.join();
  }

  uint64_t submit(int kind, uint32_t src, uint32_t *out) override {
    if (kind != GRUS_Q_BFS && kind != GRUS_Q_SSSP) raise(GRUS_E_INVALID, "query kind");
    if (src >= g_.V) raise(GRUS_E_INVALID, "source vertex out of range");
    std::lock_guard<std::mutex> lk(mu_);
    if (stop_) raise(GRUS_E_CLOSED, "server is shut down");
    if (failed_) raise(failed_code_, "server engine failed: " + failed_msg_);
    Query q{next_id_++, kind, src, out, now_us(), 0.0};
    pending_[kind].push_back(q);
    cv_work_.notify_all();
    return q.id;
  }

  void poll(grus_query_result *res, int timeout_ms) override {
    std::unique_lock<std::mutex> lk(mu_);
    auto ready = [&] { return !done_.empty() || failed_; };
    if (timeout_ms < 0) {
      cv_done_.wait(lk, ready);
    } else if (!cv_done_.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready)) {
      raise(GRUS_E_TIMEOUT, "no finished query within the timeout");
    }
    if (done_.empty()) raise(failed_code_, "server engine failed: " + failed_msg_);
    *res = done_.front();
    done_.pop_front();
  }

 private:
  void finish(const Query &q, int status, uint64_t reached, uint64_t checksum) {
    grus_query_result r{};
    r.id = q.id;
    r.status = status;
    r.reached = reached;
    r.checksum = checksum;
    r.submit_us = q.submit_us;
    r.start_us = q.start_us;
    r.done_us = now_us();
    std::lock_guard<std::mutex> lk(mu_);
    done_.push_back(r);
    cv_done_.notify_all();
  }

  void stats(grus_server_stats *out) override {
    std::lock_guard<std::mutex> lk(mu_);
    *out = grus_server_stats{};
    for (size_t i = 0; i < engines_.size(); i++)
      if (engines_[i]) out->engines_held[kind_of(i)]++;
    out->state_bytes = held_state();
    out->engines_started = started_;
    out->engines_released = released_;
    out->memory_waits = memory_waits_;
  }

  int kind_of(size_t i) const { return (int)i < per_kind_ ? GRUS_Q_BFS : GRUS_Q_SSSP; }
  uint64_t held_state() const {  // under mu_
    uint64_t t = reserved_;
    for (uint64_t b : state_) t += b;
    return t;
  }

  // gives engine i its device state; false when it does not fit (the state
  // budget, or an allocation refused by the device budget or the device)
  bool start_engine(size_t i) {
    const int kind = kind_of(i);
    uint64_t est;
    {
      std::lock_guard<std::mutex> lk(mu_);
      est = est_[kind];
      if (opts_.state_budget_bytes && est && held_state() + est > opts_.state_budget_bytes) {
        memory_waits_++;
        return false;
      }
      reserved_ += est;
    }
    std::unique_ptr<MsEngine> e;
    try {
      e = std::make_unique<MsEngine>(g_, (MsKind)kind, lanes_, ctxs_[i].get());
    } catch (const Error &x) {
      std::lock_guard<std::mutex> lk(mu_);
      reserved_ -= est;
      if (x.code != GRUS_E_BUDGET && x.code != GRUS_E_NOMEM) throw;
      memory_waits_++;
      return false;
    }
    std::lock_guard<std::mutex> lk(mu_);
    reserved_ -= est;
    state_[i] = e->state_bytes();
    if (!est_[kind]) est_[kind] = state_[i];
    engines_[i] = std::move(e);
    started_++;
    return true;
  }

  // engine i (idle) gives its device state back; waiting workers retry
  void release_engine(size_t i) {
    std::unique_ptr<MsEngine> e;
    {
      std::lock_guard<std::mutex> lk(mu_);
      e = std::move(engines_[i]);
      state_[i] = 0;
      released_++;
    }
    e.reset();  // frees the rows, outside the lock
    {
      std::lock_guard<std::mutex> lk(mu_);
      mem_epoch_++;
    }
    cv_work_.notify_all();
  }

  void work(size_t ei) {
    const int kind = kind_of(ei);
    const bool first = (int)ei % per_kind_ == 0;  // holds its state throughout
    std::vector<Query> slot_q(128);
    uint64_t cohort_start = 0;  // engine step at which the running batch started
    double batch_t0 = 0;        // and its clock
    double idle_since = now_us();
    uint64_t refused_at = ~0ull;  // memory epoch of the last refusal
    try {
      GRUS_HIP(hipSetDevice(g_.ctx->device));
      for (;;) {
        MsEngine *eng = engines_[ei].get();  // only this worker replaces it
        std::vector<Query> take;
        bool solo = false, start = false, release = false;
        {
          std::unique_lock<std::mutex> lk(mu_);
          if (eng) state_[ei] = eng->state_bytes();  // widened SSSP rows grow
          // after a refusal, a worker runs nothing that needs new memory
          // until some engine has given its state back
          auto room = [&] { return refused_at != mem_epoch_; };
          auto ready = [&] {
            if (stop_ || (eng && eng->active())) return true;
            if (pending_[kind].empty()) return false;
            return eng != nullptr || room();
          };
          if (eng && !eng->active() && !first && release_ms_ >= 0 && !stop_) {
            const double left = release_ms_ - (now_us() - idle_since) / 1e3;
            const bool woke = left > 0 && cv_work_.wait_for(
                lk, std::chrono::microseconds((int64_t)(left * 1e3)), ready);
            release = !woke && !ready();
          } else {
            cv_work_.wait(lk, ready);
          }
          if (!release) {
            if (stop_ && !(eng && eng->active())) break;
            if (!eng) {
              // no state: serve a short queue alone, else try to take state
              if (!views_.empty() && short_queue(kind)) {
                solo = true;
              } else if (!pending_[kind].empty()) {
                start = true;
              }
            } else {
              int free = stop_ ? 0 : eng->free_slots();
              // light load: an idle engine serves a short queue one query at
              // a time with the single-source kernels (a batch pays a step
              // per level whatever it holds; one query alone is several
              // times faster)
              if (!stop_ && !eng->active() && !views_.empty() && room() && short_queue(kind)) {
                solo = true;
                free = 0;
              }
              // a running batch takes newcomers while it is young, or once
              // enough of its lanes are free: lanes that start together keep
              // their frontiers in step, so a step of the engine pays for one
              // union of mostly shared vertices
              const int gate = min_free_ * (eng->slots() / lanes_);
              if (eng->active() && eng->steps() - cohort_start >= (uint64_t)window_ && free < gate)
                free = 0;
              while (free-- > 0 && !pending_[kind].empty()) {
                take.push_back(pending_[kind].front());
                pending_[kind].pop_front();
              }
            }
            if (solo) {
              take.push_back(pending_[kind].front());
              pending_[kind].pop_front();
            }
          }
        }
        if (release) {
          release_engine(ei);
          continue;
        }
        if (start) {
          if (!start_engine(ei)) {
            std::lock_guard<std::mutex> lk(mu_);
            refused_at = mem_epoch_;
          }
          idle_since = now_us();
          continue;
        }
        if (solo) {
          if (!serve_alone(ei, take[0])) {  // no room for the view's workspace
            std::lock_guard<std::mutex> lk(mu_);
            pending_[kind].push_front(take[0]);
            refused_at = mem_epoch_;
            memory_waits_++;
          }
          idle_since = now_us();
          continue;
        }
        if (!take.empty()) {
          if (!eng->active()) {
            cohort_start = eng->steps();
            batch_t0 = now_us();
          }
          std::vector<uint32_t> srcs;
          for (auto &q : take) srcs.push_back(q.src);
          std::vector<int> slots = eng->admit(srcs.data(), (int)srcs.size());
          double t = now_us();
          for (size_t k = 0; k < slots.size(); k++) {
            take[k].start_us = t;
            slot_q[slots[k]] = take[k];
          }
        }
        if (!eng || !eng->active()) continue;
        SlotMask done = 0;
        for (int sl : eng->step()) done |= slot_bit(sl);
        if (!done) continue;
        // the slots that finished in this step are digested together (one
        // pass over the engine's state) and reported at once
        uint64_t reached[128], checksum[128];
        eng->digest(done, reached, checksum);
        for (SlotMask m = done; m; m &= m - 1) {
          const int sl = slot_lowest(m);
          const Query &q = slot_q[sl];
          if (opts_.copy_results) eng->extract(sl, q.out, nullptr, nullptr);
          eng->release(sl);
          finish(q, GRUS_OK, reached[sl], checksum[sl]);
        }
        if (!eng->active()) {
          note_time(batch_ms_[kind], (now_us() - batch_t0) / 1e3);
          idle_since = now_us();
        }
      }
    } catch (const Error &e) {
      fail(e.code, e.what());
    } catch (const std::exception &e) {
      fail(GRUS_E_HIP, e.what());
    }
    // queries still held by this engine or waiting for its kind fail
    std::vector<Query> orphans;
    int orphan_status;
    {
      std::lock_guard<std::mutex> lk(mu_);
      if (engines_[ei])
        for (SlotMask m = engines_[ei]->busy_mask(); m; m &= m - 1)
          orphans.push_back(slot_q[slot_lowest(m)]);
      if (stop_) {
        for (auto &q : pending_[kind]) orphans.push_back(q);
        pending_[kind].clear();
      }
      // read under the lock: the destructor and fail() write them under it
      orphan_status = stop_ ? GRUS_E_CLOSED : failed_code_;
    }
    for (auto &q : orphans) finish(q, orphan_status, 0, 0);
  }

  // one query with the single-source kernels on the worker's graph view
  // (false: no room for the view's workspace; the query is not run)
  bool serve_alone(size_t ei, Query q) {
    Graph &v = *views_[ei];
    Context &ctx = *v.ctx;
    q.start_us = now_us();
    const double t0 = q.start_us;
    try {
      if (q.kind == GRUS_Q_BFS) run_bfs(v, q.src, GRUS_DIR_AUTO, nullptr);
      else run_sssp(v, q.src, nullptr);
    } catch (const Error &e) {
      if (e.code != GRUS_E_BUDGET && e.code != GRUS_E_NOMEM) throw;
      return false;
    }
    uint64_t reached = 0, checksum = 0;
    row_digest(ctx, v.result.as<uint32_t>(), v.V, &reached, &checksum);
    if (opts_.copy_results && q.out) {
      GRUS_HIP(hipMemcpyAsync(q.out, v.result.p, v.V * 4, hipMemcpyDeviceToHost, ctx.stream));
      ctx.sync();
    }
    note_time(solo_ms_[q.kind], (now_us() - t0) / 1e3);
    finish(q, GRUS_OK, reached, checksum);
    return true;
  }

  // whether the waiting queries of a kind are few enough to serve one at a
  // time (called under mu_)
  bool short_queue(int kind) const {
    const size_t n = pending_[kind].size();
    if (n == 0) return false;
    if (opts_.solo_max > 0 || solo_ms_[kind] <= 0) return (int)n <= solo_max_;
    const double batch = batch_ms_[kind] > 0 ? batch_ms_[kind] : 32 * solo_ms_[kind];
    return (double)n * solo_ms_[kind] <= opts_.max_inflight_batches * batch;
  }

  // running mean (weight 1/8 to the newest), 0 = nothing measured yet
  void note_time(double &mean, double ms) {
    std::lock_guard<std::mutex> lk(mu_);
    mean = mean > 0 ? mean + (ms - mean) / 8 : ms;
  }

  void fail(int code, const std::string &msg) {
    std::lock_guard<std::mutex> lk(mu_);
    failed_ = true;
    failed_code_ = code;
    failed_msg_ = msg;
    cv_done_.notify_all();
  }

  Graph &g_;
  grus_server_opts opts_{};
  // per worker: its context, its engine (null while it holds no state), the
  // device bytes of that state, and its view of the graph for single-source
  // queries (views_ is empty when the single-query path is off)
  std::vector<std::unique_ptr<Context>> ctxs_;
  std::vector<std::unique_ptr<MsEngine>> engines_;
  std::vector<uint64_t> state_;
  std::vector<std::unique_ptr<Graph>> views_;
  std::vector<std::thread> workers_;
  std::mutex mu_;
  std::condition_variable cv_work_, cv_done_;
  std::deque<Query> pending_[2];
  std::deque<grus_query_result> done_;
  int min_free_ = 1;
  int window_ = 2;  // steps after an idle start during which a batch admits
  int lanes_ = 64;  // lanes per engine
  int solo_max_ = 8;  // queue length up to which idle engines serve queries alone
  double solo_ms_[2] = {0, 0}, batch_ms_[2] = {0, 0};  // running means per kind
  int per_kind_ = 4;        // workers (engines) per kind
  double release_ms_ = -1;  // idle time after which state is given back (< 0: never)
  uint64_t est_[2] = {0, 0};  // state bytes of an engine of each kind
  uint64_t reserved_ = 0;     // state being allocated
  uint64_t mem_epoch_ = 0;    // engines released so far
  uint64_t started_ = 0, released_ = 0, memory_waits_ = 0;
  bool stop_ = false, failed_ = false;
  int failed_code_ = GRUS_OK;
  std::string failed_msg_;
  uint64_t next_id_ = 1;
};

}  // namespace

std::unique_ptr<Server> make_server(Graph &g, const grus_server_opts *opts) {
  g.ctx->activate();
  return std::make_unique<ServerImpl>(g, opts);
}

}  // namespace mgg
