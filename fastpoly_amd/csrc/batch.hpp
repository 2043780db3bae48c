// Merged executor: many independent requests, few launches.
//
// Every request (a conv or a series call) is first recorded: its drivers run
// with Context::rec set, so each kernel launch becomes a LaunchRec holding the
// job it would have launched (device pointers, sizes, field constants) and
// every arena block it frees stays parked on the request's chain until the
// schedule has been enqueued.  The schedules of different requests are
// independent, so they are executed level by level: level t takes the t-th
// record of every chain that is still running, groups the records by kernel
// shape, and issues ONE launch per group with a device job array (blockIdx.y
// selects the job) or a chunk table (elementwise / Garner records).  Within a
// chain, order is kept because levels run in stream order.
//
// Requests of the same op align naturally: a Newton step at size m has the
// same kernel shapes whatever the final n, and convolutions of one size class
// share shapes across moduli (the three auxiliary primes and 998244353 all run
// the lazy kernels).  Thousands of per-call launches become a few hundred.
#pragma once

#include <chrono>
#include <condition_variable>
#include <deque>
#include <thread>
#include <tuple>

namespace fpxg {

struct Chain {
  std::vector<LaunchRec> recs;
  std::vector<void*> bufs;  // arena blocks of this schedule, released once its last launch is enqueued
  size_t pos = 0;
  bool done() const { return pos >= recs.size(); }
};

struct LaunchKey {
  int kind, a, b, c, d;
  bool operator<(const LaunchKey& o) const {
    return std::tie(kind, a, b, c, d) < std::tie(o.kind, o.a, o.b, o.c, o.d);
  }
};

// Rows: (logM, lazy, mode, nrows); columns: (S, lazy, inverse, L); recurrences
// by workgroup size; one key each for elementwise and Garner records.
static LaunchKey key_of(const LaunchRec& r) {
  switch (r.kind) {
    case LK_ROWS: return {LK_ROWS, r.k0, r.k1, r.k2, (int)r.row.nrows};
    case LK_COLS: return {LK_COLS, r.k0, r.k1, r.k2, r.k3};
    case LK_RECUR: return {LK_RECUR, r.rec.hi <= 128 ? 64 : 1024, 0, 0, 0};
    default: return {r.kind, 0, 0, 0, 0};
  }
}

struct Staged {
  LaunchKey key;
  size_t jobs = 0, chunks = 0;  // byte offsets into the upload
  uint32_t count = 0, nchunks = 0;
  RowJob2 row0{};  // by-value kernel argument (unused when a job array is given)
  ColJob2 col0{};
};

constexpr uint32_t kMaxJobsPerLaunch = 65535;  // gridDim.y

struct Plan {
  std::vector<uint8_t> bytes;
  std::vector<Staged> launches;
  size_t put(const void* p, size_t n) {
    const size_t off = (bytes.size() + 255) & ~(size_t)255;
    bytes.resize(off + n);
    if (n) memcpy(bytes.data() + off, p, n);
    return off;
  }
  template <class T>
  size_t put_vec(const std::vector<T>& v) { return put(v.data(), v.size() * sizeof(T)); }
};

static void stage_group(Plan& plan, const LaunchKey& k, const std::vector<const LaunchRec*>& g) {
  for (size_t s = 0; s < g.size(); s += kMaxJobsPerLaunch) {
    const size_t e = std::min(g.size(), s + kMaxJobsPerLaunch);
    Staged st;
    st.key = k;
    st.count = (uint32_t)(e - s);
    switch (k.kind) {
      case LK_ROWS: {
        std::vector<RowJob2> v;
        for (size_t i = s; i < e; ++i) v.push_back(g[i]->row);
        st.row0 = v[0];
        st.jobs = plan.put_vec(v);
        break;
      }
      case LK_COLS: {
        std::vector<ColJob2> v;
        for (size_t i = s; i < e; ++i) v.push_back(g[i]->col);
        st.col0 = v[0];
        st.jobs = plan.put_vec(v);
        break;
      }
      case LK_RECUR: {
        std::vector<RecurRec> v;
        for (size_t i = s; i < e; ++i) v.push_back(g[i]->rec);
        st.jobs = plan.put_vec(v);
        break;
      }
      case LK_EW:
      case LK_GARNER: {
        std::vector<uint2> ch;
        std::vector<EwRec> ve;
        std::vector<GarnerRec> vg;
        for (size_t i = s; i < e; ++i) {
          const uint32_t n = k.kind == LK_EW ? g[i]->ew.n : g[i]->gar.n;
          const uint32_t idx = (uint32_t)(i - s);
          for (uint32_t o = 0; o < n; o += kChunk) ch.push_back(make_uint2(idx, o));
          if (k.kind == LK_EW) ve.push_back(g[i]->ew);
          else vg.push_back(g[i]->gar);
        }
        if (ch.empty()) continue;
        st.jobs = k.kind == LK_EW ? plan.put_vec(ve) : plan.put_vec(vg);
        st.chunks = plan.put_vec(ch);
        st.nchunks = (uint32_t)ch.size();
        break;
      }
      default:
        throw Error(E_DEVICE, "unknown launch record");
    }
    plan.launches.push_back(st);
  }
}

static void issue(Context& c, const Staged& s, const uint8_t* base) {
  const LaunchKey& k = s.key;
  switch (k.kind) {
    case LK_ROWS:
      launch_rows(c, k.b != 0, k.c, s.row0, k.a, (uint32_t)k.d, s.count, (const RowJob2*)(base + s.jobs));
      return;
    case LK_COLS: {
      const PassPlan P{k.d, k.d - k.a, k.a, tile_log_cols(k.a)};
      launch_cols(c, k.b != 0, k.c != 0, s.col0, P, s.count, (const ColJob2*)(base + s.jobs));
      return;
    }
    case LK_EW:
      hipLaunchKernelGGL(k_ew, dim3(s.nchunks), dim3(256), 0, c.stream, (const EwRec*)(base + s.jobs),
                         (const uint2*)(base + s.chunks), EwRec{});
      break;
    case LK_RECUR:
      if (k.a == 64)
        hipLaunchKernelGGL(k_recur<64>, dim3(s.count), dim3(64), 0, c.stream, (const RecurRec*)(base + s.jobs),
                           RecurRec{});
      else
        hipLaunchKernelGGL(k_recur<1024>, dim3(s.count), dim3(1024), 0, c.stream, (const RecurRec*)(base + s.jobs),
                           RecurRec{});
      break;
    case LK_GARNER:
      hipLaunchKernelGGL(k_garner_batch, dim3(s.nchunks), dim3(256), 0, c.stream, (const GarnerRec*)(base + s.jobs),
                         (const uint2*)(base + s.chunks));
      break;
  }
  FPXG_HIP(hipGetLastError());
}

// Copies the job arrays to the device through a pinned staging slot.  Slots
// rotate; a slot is rewritten only after its previous copy has completed.
static void upload_plan(Context& c, void* dev, const std::vector<uint8_t>& bytes) {
  Context::Staging& s = c.staging[c.staging_next];
  c.staging_next = (c.staging_next + 1) % Context::kStagingSlots;
  if (s.used) FPXG_HIP(hipEventSynchronize(s.ev));
  if (s.cap < bytes.size()) {
    if (s.host) FPXG_HIP(hipHostFree(s.host));
    s.host = nullptr;
    s.cap = std::max<size_t>(bytes.size() * 2, (size_t)1 << 20);
    FPXG_HIP(hipHostMalloc((void**)&s.host, s.cap, hipHostMallocDefault));
  }
  memcpy(s.host, bytes.data(), bytes.size());
  FPXG_HIP(hipMemcpyAsync(dev, s.host, bytes.size(), hipMemcpyHostToDevice, c.stream));
  FPXG_HIP(hipEventRecord(s.ev, c.stream));
  s.used = true;
}

// Runs up to max_levels levels of the given chains.  Returns the launches issued.
static size_t run_levels(Context& c, const std::vector<Chain*>& chains, size_t max_levels) {
  Plan plan;
  std::map<LaunchKey, std::vector<const LaunchRec*>> groups;
  for (size_t lvl = 0; lvl < max_levels; ++lvl) {
    groups.clear();
    for (Chain* ch : chains)
      if (!ch->done()) {
        const LaunchRec& r = ch->recs[ch->pos++];
        groups[key_of(r)].push_back(&r);
      }
    if (groups.empty()) break;
    for (auto& kv : groups) stage_group(plan, kv.first, kv.second);
  }
  if (plan.launches.empty()) return 0;
  DevBuf dev = c.buf(plan.bytes.size() / 4 + 1);
  upload_plan(c, dev.get(), plan.bytes);
  for (const Staged& s : plan.launches) issue(c, s, (const uint8_t*)dev.get());
  return plan.launches.size();
}

// ---------------------------------------------------------------------------
// Requests
// ---------------------------------------------------------------------------

// Validates one request the way the single call would (fpxg_conv_dev /
// fpxg_series_dev) and sets q.out_len.  Throws the single call's error.
struct Prepared {
  size_t xa = 0, xb = 0, outc = 0;
  bool scan = false;  // series op needing the input scan
  bool char2 = false; // sqrt in characteristic two (immediate path)
};

static void check_modulus_p(uint32_t p) {
  if (p < 2 || p >= (1u << 31) || !is_prime_u32(p)) throw Error(E_UNSUPPORTED, "modulus must be a prime in [2, 2^31)");
}

static size_t conv_len_checked(size_t la, size_t lb, size_t n, size_t* pla, size_t* plb) {
  if (la == 0 || lb == 0 || n == 0) return 0;
  la = std::min(la, n);
  lb = std::min(lb, n);
  if (la > SIZE_MAX - (lb - 1)) throw Error(E_SIZE, "poly::conv: result length overflows size_t");
  if (la > ((size_t)1 << 31) || lb > ((size_t)1 << 31))
    throw Error(E_SIZE, "operands above 2^31 coefficients are not supported");
  *pla = la;
  *plb = lb;
  return std::min(la + lb - 1, n);
}

static Prepared prepare(Context& c, fpxg_req& q) {
  Prepared r;
  q.out_len = 0;
  check_modulus_p(q.p);
  if (q.op == FPXG_CONV) {
    r.outc = conv_len_checked(q.la, q.lb, q.n, &r.xa, &r.xb);
    if (r.outc && (!q.a || !q.b || !q.out || q.out_cap < r.outc))
      throw Error(E_ARG, "null pointer or output capacity too small");
    q.out_len = r.outc;
    return r;
  }
  if (q.op < FPXG_INV || q.op > FPXG_SQRT) throw Error(E_ARG, "unknown series op");
  if (q.n == 0) return r;
  if (!q.out || (q.la && !q.a) || q.out_cap < q.n) throw Error(E_ARG, "null pointer or output capacity too small");
  r.xa = std::min(q.la, q.n);
  r.outc = q.n;
  q.out_len = q.n;
  if (q.op == FPXG_SQRT && c.field(q.p).p == 2) r.char2 = true;
  else r.scan = true;
  return r;
}

// Records one validated request into ch.  Returns its status; on failure
// nothing is kept and its blocks are released.
static int record_request(Context& c, const fpxg_req& q, const Prepared& pr, const ScanOut* scan, Chain& ch) {
  Recorder R;
  int st = OK;
  c.rec = &R;
  c.arena->set_deferral(&ch.bufs);
  try {
    const HostField& T = c.field(q.p);
    if (q.op == FPXG_CONV) {
      if (pr.outc) conv_dev(c, T, q.a, pr.xa, q.b, pr.xb, q.out, pr.outc);
    } else if (pr.outc) {
      series_emit(c, T, SeriesReq{q.op, q.a, pr.xa, q.n, q.out}, *scan);
    }
  } catch (const Error& e) {
    st = e.code;
  } catch (const std::bad_alloc&) {
    st = E_NOMEM;
  }
  c.rec = nullptr;
  c.arena->set_deferral(nullptr);
  if (st != OK) {
    c.arena->flush_deferred(ch.bufs);
  } else {
    ch.recs = std::move(R.recs);
  }
  return st;
}

static void release_chains(Context& c, std::vector<Chain>& chains) {
  for (Chain& ch : chains) c.arena->flush_deferred(ch.bufs);
  chains.clear();
}

struct BatchStats {
  size_t launches = 0, waves = 0, records = 0;
};

// One batch of device-pointer requests, executed level by level; large
// batches run in waves so that recorded intermediates fit the arena.
static BatchStats batch_dev(Context& c, fpxg_req* reqs, size_t k) {
  BatchStats bs;
  std::vector<Prepared> pr(k);
  std::vector<size_t> scan_idx(k, SIZE_MAX);
  std::vector<ScanRec> scans;
  for (size_t i = 0; i < k; ++i) {
    fpxg_req& q = reqs[i];
    q.status = OK;
    try {
      pr[i] = prepare(c, q);
    } catch (const Error& e) {
      q.status = e.code;
      q.out_len = 0;
      continue;
    }
    if (pr[i].scan) {
      scan_idx[i] = scans.size();
      scans.push_back({q.a, (uint32_t)pr[i].xa, 0});
    }
  }
  // Characteristic-two square roots are read back on the host: run them first.
  for (size_t i = 0; i < k; ++i)
    if (reqs[i].status == OK && pr[i].char2) {
      try {
        series_dev(c, c.field(reqs[i].p), FPXG_SQRT, reqs[i].a, pr[i].xa, reqs[i].n, reqs[i].out);
      } catch (const Error& e) {
        reqs[i].status = e.code;
      }
    }
  const std::vector<ScanOut> so = scan_inputs(c, scans);

  std::vector<Chain> wave;
  size_t base = c.arena->in_use();
  size_t budget = c.arena->capacity() > base ? (c.arena->capacity() - base) / 2 : 0;
  auto flush_wave = [&]() {
    if (wave.empty()) return;
    std::vector<Chain*> ptrs;
    for (Chain& ch : wave) {
      ptrs.push_back(&ch);
      bs.records += ch.recs.size();
    }
    bs.launches += run_levels(c, ptrs, SIZE_MAX);
    ++bs.waves;
    release_chains(c, wave);
    base = c.arena->in_use();
    budget = c.arena->capacity() > base ? (c.arena->capacity() - base) / 2 : 0;
  };
  wave.reserve(k);
  for (size_t i = 0; i < k; ++i) {
    fpxg_req& q = reqs[i];
    if (q.status != OK || pr[i].char2 || pr[i].outc == 0) continue;
    const ScanOut* s = scan_idx[i] == SIZE_MAX ? nullptr : &so[scan_idx[i]];
    for (int attempt = 0; attempt < 3; ++attempt) {
      wave.emplace_back();
      const int st = record_request(c, q, pr[i], s, wave.back());
      if (st == E_NOMEM && attempt < 2) {
        wave.pop_back();
        if (!wave.empty()) {
          flush_wave();
        } else {
          FPXG_HIP(hipStreamSynchronize(c.stream));
          c.arena->release_idle();
        }
        continue;
      }
      if (st != OK) {
        wave.pop_back();
        q.status = st;
        q.out_len = 0;
      }
      break;
    }
    if (c.arena->in_use() > base + budget) flush_wave();
  }
  flush_wave();
  return bs;
}

}  // namespace fpxg
