// Continuous-batching request server.
//
// A scheduler thread owns the library stream.  Each loop iteration it
//   1. admits waiting requests while the device arena stays within the HBM
//      budget: validates them, records their schedules (device pointers into
//      freshly allocated input/output blocks), and enqueues the input upload;
//   2. runs a few levels of every admitted schedule as merged launches, so a
//      request that arrived a moment ago shares launches with requests that
//      are half-way through theirs;
//   3. enqueues the download of each finished schedule and releases its
//      device blocks (later work is stream-ordered after the download);
//   4. completes requests whose download event has fired.
// At most two iterations are in flight ahead of the device, so a new request
// joins the running mix within about one iteration's device time.
//
// Results are the single calls' bit for bit: a request's schedule is the same
// sequence of kernels whatever runs beside it, each record touches only that
// request's blocks, and the arithmetic is exact.
#pragma once

#include <set>

namespace fpxg {

// Page-locked host blocks in power-of-two classes, reused across requests.
class PinnedPool {
 public:
  ~PinnedPool() {
    for (void* p : all_) (void)hipHostFree(p);
  }
  void* alloc(size_t bytes, int* cls) {
    int k = 12;
    while (((size_t)1 << k) < bytes) ++k;
    *cls = k;
    std::lock_guard<std::mutex> g(mu_);
    auto& fl = free_[k];
    if (!fl.empty()) {
      void* p = fl.back();
      fl.pop_back();
      return p;
    }
    void* p = nullptr;
    if (hipHostMalloc(&p, (size_t)1 << k, hipHostMallocDefault) != hipSuccess || !p) {
      (void)hipGetLastError();
      throw Error(E_NOMEM, "hipHostMalloc failed for a pinned staging block");
    }
    all_.push_back(p);
    return p;
  }
  void free(void* p, int cls) {
    if (!p) return;
    std::lock_guard<std::mutex> g(mu_);
    free_[cls].push_back(p);
  }

 private:
  std::map<int, std::vector<void*>> free_;
  std::vector<void*> all_;
  std::mutex mu_;
};

struct SItem {
  uint64_t id = 0;
  fpxg_req* user = nullptr;  // the caller's request: host pointers, updated on completion
  fpxg_req q{};              // device form
  Prepared pr;
  ScanOut scan{};
  Chain chain;
  uint32_t* din = nullptr;
  uint32_t* dout = nullptr;
  void* pin_in = nullptr;
  void* pin_out = nullptr;
  int pin_in_cls = 0, pin_out_cls = 0;
  hipEvent_t ev = nullptr;
  double t_submit = 0, t_admit = 0, t_done = 0;
  int status = OK;
  fpxg_result result{};
  bool has_scan = false;  // input scanned at submit (series ops)
  size_t staged_a = 0;    // series ops: input words staged at submit
};

// Host scan of a series input (same fields as the device scan).
static ScanOut host_scan(const uint32_t* a, size_t n) {
  ScanOut o{};
  o.first = o.second = (uint32_t)n;
  o.last1 = 0;
  for (size_t i = 0; i < n; ++i)
    if (a[i]) {
      if (o.first == n) o.first = (uint32_t)i;
      else if (o.second == n) o.second = (uint32_t)i;
      o.last1 = (uint32_t)(i + 1);
    }
  o.a0 = n ? a[0] : 0u;
  o.afirst = o.first < n ? a[o.first] : 0u;
  return o;
}

struct Server {
  Context& c;
  size_t budget;
  int levels;
  std::chrono::steady_clock::time_point t0;

  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::unique_ptr<SItem>> incoming;
  std::map<uint64_t, std::unique_ptr<SItem>> finished;  // completed, result not yet collected
  uint64_t next_id = 1;
  bool stopping = false;
  fpxg_server_stats st{};  // submitted / completed under mu; the rest published by the scheduler
  std::set<uint64_t> outstanding;
  std::string fatal;
  fpxg_server_stats sched{};  // scheduler-thread counters
  // Host-time breakdown, printed at destruction when FPXG_TRACE_SERVER is set.
  double tr_admit = 0, tr_run = 0, tr_drain = 0, tr_idle = 0;
  uint64_t tr_attempts = 0, tr_rollbacks = 0;

  // Scheduler-thread state.
  std::deque<std::unique_ptr<SItem>> waiting;
  std::vector<std::unique_ptr<SItem>> running;
  std::deque<std::unique_ptr<SItem>> draining;
  std::deque<hipEvent_t> inflight;
  std::vector<hipEvent_t> ev_pool;
  PinnedPool pinned;
  std::thread th;

  Server(Context& cc, size_t b, int lv) : c(cc), budget(b), levels(lv), t0(std::chrono::steady_clock::now()) {
    th = std::thread([this] { loop(); });
  }
  ~Server() {
    {
      std::lock_guard<std::mutex> g(mu);
      stopping = true;
    }
    cv_work.notify_all();
    th.join();
    for (hipEvent_t e : ev_pool) (void)hipEventDestroy(e);
    if (getenv("FPXG_TRACE_SERVER"))
      fprintf(stderr, "fpxg server: host s admit %.3f run %.3f drain %.3f idle %.3f; admit attempts %llu rollbacks %llu\n",
              tr_admit, tr_run, tr_drain, tr_idle, (unsigned long long)tr_attempts, (unsigned long long)tr_rollbacks);
  }

  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }

  hipEvent_t take_event() {
    if (!ev_pool.empty()) {
      hipEvent_t e = ev_pool.back();
      ev_pool.pop_back();
      return e;
    }
    hipEvent_t e;
    FPXG_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    return e;
  }

  // The result (pin_out) is copied into the caller's buffer when collected.
  void complete(std::unique_ptr<SItem> it) {
    it->t_done = now();
    fpxg_result& r = it->result;
    r.status = it->status;
    r.out_len = it->status == OK ? it->q.out_len : 0;
    r.submit_s = it->t_submit;
    r.admit_s = it->t_admit;
    r.done_s = it->t_done;
    {
      std::lock_guard<std::mutex> g(mu);
      const uint64_t id = it->id;
      finished[id] = std::move(it);
      ++st.completed;
    }
    cv_done.notify_all();
  }

  // Caller's thread, at submit: the input goes into a pinned block and series
  // inputs are scanned, so the scheduler only records and launches.
  int stage(SItem& it) {
    const fpxg_req& q = *it.user;
    size_t xa = 0, xb = 0;
    if (q.op == FPXG_CONV) {
      if (q.la && q.lb && q.n) {
        xa = std::min(q.la, q.n);
        xb = std::min(q.lb, q.n);
      }
    } else {
      if (q.la && q.n && !q.a) return E_ARG;
      xa = q.n ? series_input_words(q.op, q.a, q.la, q.n) : 0;
      it.staged_a = xa;
    }
    if ((xa && !q.a) || (xb && !q.b)) return E_ARG;
    if (xa + xb > ((size_t)1 << 33)) return E_SIZE;
    if (xa + xb) {
      it.pin_in = pinned.alloc((xa + xb) * 4, &it.pin_in_cls);
      memcpy(it.pin_in, q.a, xa * 4);
      if (xb) memcpy((uint32_t*)it.pin_in + xa, q.b, xb * 4);
    }
    if (q.op != FPXG_CONV) {
      it.scan = host_scan(q.a, xa);
      it.has_scan = true;
    }
    return OK;
  }

  // Caller's thread, at collect: result into the caller's buffer, blocks back to the pool.
  void finish(SItem& it) {
    if (it.result.status == OK && it.result.out_len && it.pin_out)
      memcpy(it.user->out, it.pin_out, it.result.out_len * 4);
    it.user->status = it.result.status;
    it.user->out_len = it.result.out_len;
    pinned.free(it.pin_in, it.pin_in_cls);
    pinned.free(it.pin_out, it.pin_out_cls);
    it.pin_in = it.pin_out = nullptr;
  }

  void release_device(SItem& it) {
    c.arena->flush_deferred(it.chain.bufs);
    c.arena->free(it.din);
    c.arena->free(it.dout);
    it.din = it.dout = nullptr;
  }

  // Tries to admit one request.  Returns false if it must wait for memory.
  bool admit(std::unique_ptr<SItem>& it) {
    fpxg_req& q = it->q;
    q = *it->user;
    if (q.op != FPXG_CONV) q.la = it->staged_a;  // what stage() copied into pin_in
    try {
      it->pr = prepare(c, q);
    } catch (const Error& e) {
      it->status = e.code;
      complete(std::move(it));
      return true;
    }
    const Prepared& pr = it->pr;
    if (pr.outc == 0) {
      complete(std::move(it));
      return true;
    }
    const size_t in_words = pr.xa + pr.xb;
    try {
      it->din = (uint32_t*)c.arena->alloc(std::max<size_t>(in_words, 1) * 4);
      it->dout = (uint32_t*)c.arena->alloc(pr.outc * 4);
    } catch (const Error& e) {
      release_device(*it);
      if (e.code == E_NOMEM && !running.empty()) return false;
      it->status = e.code;
      complete(std::move(it));
      return true;
    }
    q.a = it->din;
    q.b = q.op == FPXG_CONV ? it->din + pr.xa : nullptr;
    q.la = pr.xa;
    q.lb = pr.xb;
    q.out = it->dout;
    q.out_cap = pr.outc;
    if (pr.char2) {  // characteristic two: immediate, with its own read-back
      if (pr.xa) FPXG_HIP(hipMemcpyAsync(it->din, it->pin_in, pr.xa * 4, hipMemcpyHostToDevice, c.stream));
      try {
        series_dev(c, c.field(q.p), FPXG_SQRT, it->din, pr.xa, q.n, it->dout);
        it->pin_out = pinned.alloc(pr.outc * 4, &it->pin_out_cls);
        FPXG_HIP(hipMemcpyAsync(it->pin_out, it->dout, pr.outc * 4, hipMemcpyDeviceToHost, c.stream));
        FPXG_HIP(hipStreamSynchronize(c.stream));
      } catch (const Error& e) {
        it->status = e.code;
      }
      release_device(*it);
      it->t_admit = now();
      complete(std::move(it));
      return true;
    }
    const int rs = record_request(c, q, pr, pr.scan ? &it->scan : nullptr, it->chain);
    if (rs == E_NOMEM && !running.empty()) {
      release_device(*it);
      it->chain = Chain();
      return false;
    }
    if (rs == OK && c.arena->in_use() > budget && !running.empty()) {
      release_device(*it);
      it->chain = Chain();
      return false;
    }
    if (rs != OK) {
      it->status = rs;
      release_device(*it);
      it->t_admit = now();
      complete(std::move(it));
      return true;
    }
    // Input upload from the pinned block staged at submit (ordered before the schedule).
    if (in_words) FPXG_HIP(hipMemcpyAsync(it->din, it->pin_in, in_words * 4, hipMemcpyHostToDevice, c.stream));
    it->t_admit = now();
    sched.peak_in_use_bytes = std::max(sched.peak_in_use_bytes, c.arena->in_use());
    running.push_back(std::move(it));
    return true;
  }

  void loop() {
    (void)hipSetDevice(c.device);
    for (;;) {
      {
        std::unique_lock<std::mutex> g(mu);
        if (waiting.empty() && running.empty() && draining.empty() && incoming.empty()) {
          if (stopping) return;
          cv_work.wait(g, [&] { return stopping || !incoming.empty(); });
          if (stopping && incoming.empty()) return;
        }
        while (!incoming.empty()) {
          waiting.push_back(std::move(incoming.front()));
          incoming.pop_front();
        }
        sched.max_queue = std::max<uint64_t>(sched.max_queue, waiting.size());
        st.iterations = sched.iterations;
        st.launches = sched.launches;
        st.max_active = sched.max_active;
        st.max_queue = sched.max_queue;
        st.peak_in_use_bytes = sched.peak_in_use_bytes;
      }
      bool progressed = false;
      try {
        std::lock_guard<std::recursive_mutex> g(c.mu);
        // Retire finished iterations.
        while (!inflight.empty() && hipEventQuery(inflight.front()) == hipSuccess) {
          ev_pool.push_back(inflight.front());
          inflight.pop_front();
        }
        // 1. Admission.
        double t0 = host_seconds();
        while (!waiting.empty()) {
          std::unique_ptr<SItem> it = std::move(waiting.front());
          waiting.pop_front();
          ++tr_attempts;
          if (!admit(it)) {
            ++tr_rollbacks;
            waiting.push_front(std::move(it));
            break;
          }
          progressed = true;
        }
        tr_admit += host_seconds() - t0;
        t0 = host_seconds();
        sched.max_active = std::max<uint64_t>(sched.max_active, running.size());
        // 2. A few merged levels, at most two iterations ahead of the device.
        if (!running.empty() && inflight.size() < 2) {
          std::vector<Chain*> chains;
          for (auto& it : running) chains.push_back(&it->chain);
          sched.launches += run_levels(c, chains, (size_t)levels);
          ++sched.iterations;
          hipEvent_t e = take_event();
          FPXG_HIP(hipEventRecord(e, c.stream));
          inflight.push_back(e);
          progressed = true;
        }
        tr_run += host_seconds() - t0;
        t0 = host_seconds();
        // 3. Finished schedules: download, release device blocks.
        for (size_t i = 0; i < running.size();) {
          SItem& it = *running[i];
          if (!it.chain.done()) {
            ++i;
            continue;
          }
          it.pin_out = pinned.alloc(it.pr.outc * 4, &it.pin_out_cls);
          FPXG_HIP(hipMemcpyAsync(it.pin_out, it.dout, it.pr.outc * 4, hipMemcpyDeviceToHost, c.stream));
          it.ev = take_event();
          FPXG_HIP(hipEventRecord(it.ev, c.stream));
          release_device(it);
          draining.push_back(std::move(running[i]));
          running[i] = std::move(running.back());
          running.pop_back();
        }
        // 4. Completions (stream order: stop at the first pending one).
        while (!draining.empty() && hipEventQuery(draining.front()->ev) == hipSuccess) {
          std::unique_ptr<SItem> it = std::move(draining.front());
          draining.pop_front();
          ev_pool.push_back(it->ev);
          it->ev = nullptr;
          complete(std::move(it));
          progressed = true;
        }
        tr_drain += host_seconds() - t0;
      } catch (const std::exception& e) {
        // A device failure ends the server: fail everything outstanding.
        std::lock_guard<std::mutex> g(mu);
        fatal = e.what();
        auto fail = [&](std::unique_ptr<SItem>& it) {
          it->result = fpxg_result{};
          it->result.status = E_DEVICE;
          it->result.submit_s = it->t_submit;
          it->result.done_s = now();
          const uint64_t id = it->id;
          finished[id] = std::move(it);
        };
        for (auto& it : waiting) fail(it);
        for (auto& it : running) fail(it);
        for (auto& it : draining) fail(it);
        for (auto& it : incoming) fail(it);
        waiting.clear();
        running.clear();
        draining.clear();
        incoming.clear();
        stopping = true;
        cv_done.notify_all();
        return;
      }
      if (!progressed) {
        const double t0 = host_seconds();
        std::this_thread::sleep_for(std::chrono::microseconds(20));
        tr_idle += host_seconds() - t0;
      }
    }
  }
};

}  // namespace fpxg
