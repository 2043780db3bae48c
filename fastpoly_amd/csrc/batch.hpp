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

// Kernel shape of a record, packed: kind | a | b | c | d (d = row count of a
// row pass or L of a column pass).  Rows: (logM, lazy, mode, rows); columns:
// (S, lazy, inverse, L); recurrences by workgroup size; one key each for
// elementwise and Garner records.
typedef uint64_t LaunchKey;
static LaunchKey pack_key(int kind, int a, int b, int c, uint32_t d) {
  return ((uint64_t)kind << 56) | ((uint64_t)(a & 0xff) << 48) | ((uint64_t)(b & 0xff) << 40) |
         ((uint64_t)(c & 0xff) << 32) | d;
}
static int key_kind(LaunchKey k) { return (int)(k >> 56); }
static int key_a(LaunchKey k) { return (int)((k >> 48) & 0xff); }
static int key_b(LaunchKey k) { return (int)((k >> 40) & 0xff); }
static int key_c(LaunchKey k) { return (int)((k >> 32) & 0xff); }
static uint32_t key_d(LaunchKey k) { return (uint32_t)k; }

static LaunchKey key_of(const LaunchRec& r) {
  switch (r.kind) {
    case LK_ROWS: return pack_key(LK_ROWS, r.k0, r.k1, r.k2, r.row.nrows);
    case LK_COLS: return pack_key(LK_COLS, r.k0, r.k1, r.k2, (uint32_t)r.k3);
    case LK_COLG: return pack_key(LK_COLG, r.k0, 1, 0, (uint32_t)r.k3);
    case LK_RECUR: return pack_key(LK_RECUR, r.rec.hi <= 128 ? 0 : 1, 0, 0, 0);
    default: return pack_key(r.kind, 0, 0, 0, 0);
  }
}

struct Staged {
  LaunchKey key;
  size_t jobs = 0, chunks = 0;  // byte offsets into the upload
  uint32_t count = 0, nchunks = 0;
  const LaunchRec* first = nullptr;  // by-value kernel argument (unused when a job array is given)
};

constexpr uint32_t kMaxJobsPerLaunch = 65535;  // gridDim.y

// Job arrays and chunk tables of one upload, written straight into a pinned
// staging slot sized in advance.
struct Plan {
  uint8_t* data = nullptr;
  size_t size = 0, cap = 0;
  std::vector<Staged> launches;
  size_t alloc(size_t n) {
    const size_t off = (size + 255) & ~(size_t)255;
    if (off + n > cap) throw Error(E_DEVICE, "batch staging overflow");
    size = off + n;
    return off;
  }
};

// Bytes stage_group will take for a group (same alignment rule as Plan::alloc).
static size_t group_bytes(LaunchKey k, const std::pair<LaunchKey, const LaunchRec*>* g, size_t n, size_t at) {
  const int kind = key_kind(k);
  auto take = [&](size_t m) { at = ((at + 255) & ~(size_t)255) + m; };
  for (size_t s = 0; s < n; s += kMaxJobsPerLaunch) {
    const size_t cnt = std::min<size_t>(n - s, kMaxJobsPerLaunch);
    switch (kind) {
      case LK_ROWS: take(cnt * sizeof(RowJob2)); break;
      case LK_COLS: take(cnt * sizeof(ColJob2)); break;
      case LK_COLG: take(cnt * sizeof(CrtColJob)); break;
      case LK_RECUR: take(cnt * sizeof(RecurRec)); break;
      case LK_TAIL: take(cnt * sizeof(TailRec)); break;
      case LK_DIRECT: {
        size_t nch = 0;
        for (size_t i = 0; i < cnt; ++i) nch += (g[s + i].second->dir.outc + kDirectTile - 1) / kDirectTile;
        if (nch == 0) break;
        take(cnt * sizeof(DirectRec));
        take(nch * sizeof(uint2));
        break;
      }
      default: {
        const bool ew = kind == LK_EW;
        size_t nch = 0;
        for (size_t i = 0; i < cnt; ++i) {
          const uint32_t m = ew ? g[s + i].second->ew.n : g[s + i].second->gar.n;
          nch += (m + kChunk - 1) / kChunk;
        }
        if (nch == 0) break;
        take(cnt * (ew ? sizeof(EwRec) : sizeof(GarnerRec)));
        take(nch * sizeof(uint2));
      }
    }
  }
  return at;
}

static void stage_group(Plan& plan, LaunchKey k, const std::pair<LaunchKey, const LaunchRec*>* g, size_t n) {
  const int kind = key_kind(k);
  for (size_t s = 0; s < n; s += kMaxJobsPerLaunch) {
    const size_t cnt = std::min<size_t>(n - s, kMaxJobsPerLaunch);
    Staged st;
    st.key = k;
    st.count = (uint32_t)cnt;
    st.first = g[s].second;
    switch (kind) {
      case LK_ROWS: {
        st.jobs = plan.alloc(cnt * sizeof(RowJob2));
        RowJob2* d = (RowJob2*)(plan.data + st.jobs);
        for (size_t i = 0; i < cnt; ++i) d[i] = g[s + i].second->row;
        break;
      }
      case LK_COLS: {
        st.jobs = plan.alloc(cnt * sizeof(ColJob2));
        ColJob2* d = (ColJob2*)(plan.data + st.jobs);
        for (size_t i = 0; i < cnt; ++i) d[i] = g[s + i].second->col;
        break;
      }
      case LK_COLG: {
        st.jobs = plan.alloc(cnt * sizeof(CrtColJob));
        CrtColJob* d = (CrtColJob*)(plan.data + st.jobs);
        for (size_t i = 0; i < cnt; ++i) d[i] = g[s + i].second->crt;
        break;
      }
      case LK_RECUR: {
        st.jobs = plan.alloc(cnt * sizeof(RecurRec));
        RecurRec* d = (RecurRec*)(plan.data + st.jobs);
        for (size_t i = 0; i < cnt; ++i) d[i] = g[s + i].second->rec;
        break;
      }
      case LK_TAIL: {
        st.jobs = plan.alloc(cnt * sizeof(TailRec));
        TailRec* d = (TailRec*)(plan.data + st.jobs);
        for (size_t i = 0; i < cnt; ++i) d[i] = g[s + i].second->tail;
        break;
      }
      case LK_DIRECT: {
        // One chunk a tile of kDirectTile outputs.
        size_t nch = 0;
        for (size_t i = 0; i < cnt; ++i) nch += (g[s + i].second->dir.outc + kDirectTile - 1) / kDirectTile;
        if (nch == 0) continue;
        st.jobs = plan.alloc(cnt * sizeof(DirectRec));
        st.chunks = plan.alloc(nch * sizeof(uint2));
        DirectRec* d = (DirectRec*)(plan.data + st.jobs);
        uint2* ch = (uint2*)(plan.data + st.chunks);
        size_t c = 0;
        for (size_t i = 0; i < cnt; ++i) {
          d[i] = g[s + i].second->dir;
          for (uint32_t o = 0; o < d[i].outc; o += kDirectTile) ch[c++] = make_uint2((uint32_t)i, o);
        }
        st.nchunks = (uint32_t)nch;
        break;
      }
      case LK_EW:
      case LK_GARNER: {
        const bool ew = kind == LK_EW;
        size_t nch = 0;
        for (size_t i = 0; i < cnt; ++i) {
          const uint32_t m = ew ? g[s + i].second->ew.n : g[s + i].second->gar.n;
          nch += (m + kChunk - 1) / kChunk;
        }
        if (nch == 0) continue;
        st.jobs = plan.alloc(cnt * (ew ? sizeof(EwRec) : sizeof(GarnerRec)));
        st.chunks = plan.alloc(nch * sizeof(uint2));
        uint2* ch = (uint2*)(plan.data + st.chunks);
        size_t c = 0;
        for (size_t i = 0; i < cnt; ++i) {
          const LaunchRec& r = *g[s + i].second;
          uint32_t m;
          if (ew) {
            ((EwRec*)(plan.data + st.jobs))[i] = r.ew;
            m = r.ew.n;
          } else {
            ((GarnerRec*)(plan.data + st.jobs))[i] = r.gar;
            m = r.gar.n;
          }
          for (uint32_t o = 0; o < m; o += kChunk) ch[c++] = make_uint2((uint32_t)i, o);
        }
        st.nchunks = (uint32_t)nch;
        break;
      }
      default:
        throw Error(E_DEVICE, "unknown launch record");
    }
    plan.launches.push_back(st);
  }
}

static void issue(Context& c, const Staged& s, const uint8_t* base) {
  const LaunchKey k = s.key;
  switch (key_kind(k)) {
    case LK_ROWS:
      launch_rows(c, key_b(k) != 0, key_c(k), s.first->row, key_a(k), key_d(k), s.count,
                  (const RowJob2*)(base + s.jobs));
      return;
    case LK_COLS: {
      const int S = key_a(k), L = (int)key_d(k);
      const PassPlan P{L, L - S, S, tile_log_cols(S, L)};
      launch_cols(c, key_b(k) != 0, key_c(k), s.first->col, P, s.count, (const ColJob2*)(base + s.jobs));
      return;
    }
    case LK_COLG: {
      const int S = key_a(k), L = (int)key_d(k);
      const PassPlan P{L, L - S, S, tile_log_cols(S, L)};
      launch_cols_crt(c, &s.first->crt, P, s.count, (const CrtColJob*)(base + s.jobs));
      return;
    }
    case LK_EW:
      hipLaunchKernelGGL(k_ew, dim3(s.nchunks), dim3(256), 0, c.stream, (const EwRec*)(base + s.jobs),
                         (const uint2*)(base + s.chunks), EwRec{});
      break;
    case LK_RECUR:
      if (key_a(k) == 0)
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
    case LK_DIRECT:
      hipLaunchKernelGGL(k_conv_direct_batch, dim3(s.nchunks), dim3(256), 0, c.stream,
                         (const DirectRec*)(base + s.jobs), (const uint2*)(base + s.chunks));
      break;
    case LK_TAIL:
      hipLaunchKernelGGL(k_tail_fix_batch, dim3(s.count), dim3(64), 0, c.stream, (const TailRec*)(base + s.jobs));
      break;
  }
  FPXG_HIP(hipGetLastError());
}

// Takes the next pinned staging slot with at least `bytes` capacity; a slot is
// rewritten only after its previous copy has completed.
static Context::Staging& take_slot(Context& c, size_t bytes) {
  Context::Staging& s = c.staging[c.staging_next];
  c.staging_next = (c.staging_next + 1) % Context::kStagingSlots;
  if (s.used) FPXG_HIP(hipEventSynchronize(s.ev));
  s.used = false;
  if (s.cap < bytes) {
    if (s.host) FPXG_HIP(hipHostFree(s.host));
    s.host = nullptr;
    s.cap = std::max<size_t>(bytes + bytes / 2, (size_t)1 << 20);
    FPXG_HIP(hipHostMalloc((void**)&s.host, s.cap, hipHostMallocDefault));
  }
  return s;
}

// Runs up to max_levels levels of the given chains.  Returns the launches issued.
static size_t run_levels(Context& c, const std::vector<Chain*>& chains, size_t max_levels) {
  typedef std::pair<LaunchKey, const LaunchRec*> Head;
  // Pass 1: every level's heads, sorted by shape; exact staging size.
  std::vector<Head> heads;
  std::vector<size_t> level_end;
  size_t bytes = 0;
  for (size_t lvl = 0; lvl < max_levels; ++lvl) {
    const size_t start = heads.size();
    for (Chain* ch : chains)
      if (!ch->done()) {
        // The chain's next record, with the records that share its level.
        do {
          const LaunchRec& r = ch->recs[ch->pos++];
          heads.emplace_back(key_of(r), &r);
        } while (!ch->done() && ch->recs[ch->pos].par);
      }
    if (heads.size() == start) break;
    std::stable_sort(heads.begin() + (ptrdiff_t)start, heads.end(),
                     [](const Head& x, const Head& y) { return x.first < y.first; });
    for (size_t i = start; i < heads.size();) {
      size_t j = i + 1;
      while (j < heads.size() && heads[j].first == heads[i].first) ++j;
      bytes = group_bytes(heads[i].first, heads.data() + i, j - i, bytes);
      i = j;
    }
    level_end.push_back(heads.size());
  }
  if (heads.empty() || bytes == 0) return 0;
  // Pass 2: write the job arrays straight into a pinned slot, upload, launch.
  Context::Staging& slot = take_slot(c, bytes);
  Plan plan;
  plan.data = slot.host;
  plan.cap = slot.cap;
  size_t start = 0;
  for (size_t e : level_end) {
    for (size_t i = start; i < e;) {
      size_t j = i + 1;
      while (j < e && heads[j].first == heads[i].first) ++j;
      stage_group(plan, heads[i].first, heads.data() + i, j - i);
      i = j;
    }
    start = e;
  }
  if (plan.launches.empty()) return 0;
  DevBuf dev = c.buf(plan.size / 4 + 1);
  FPXG_HIP(hipMemcpyAsync(dev.get(), slot.host, plan.size, hipMemcpyHostToDevice, c.stream));
  FPXG_HIP(hipEventRecord(slot.ev, c.stream));
  slot.used = true;
  for (const Staged& s : plan.launches) issue(c, s, (const uint8_t*)dev.get());
  return plan.launches.size();
}

// Runs every level of the chains, uploading the job arrays of a few levels at
// a time (2, 4, 8, then 16 levels per upload) so the device starts on the
// first levels while the host stages the later ones.
static size_t run_all(Context& c, const std::vector<Chain*>& chains) {
  size_t launches = 0, step = 2;
  for (;;) {
    bool more = false;
    for (const Chain* ch : chains)
      if (!ch->done()) {
        more = true;
        break;
      }
    if (!more) break;
    launches += run_levels(c, chains, step);
    step = std::min<size_t>(step * 2, 16);
  }
  return launches;
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

static void check_modulus_p(Context& c, uint32_t p) {
  auto it = c.prime_cache.find(p);
  const bool ok = it != c.prime_cache.end() ? it->second
                                            : (c.prime_cache[p] = p >= 2 && p < (1u << 31) && is_prime_u32(p));
  if (!ok) throw Error(E_UNSUPPORTED, "modulus must be a prime in [2, 2^31)");
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
  check_modulus_p(c, q.p);
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
  r.xa = series_input_words_dev(q.op, q.la, q.n);
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
  double t_prepare = 0, t_scan = 0, t_record = 0, t_run = 0;  // host seconds
};

static double host_seconds() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// One batch of device-pointer requests, executed level by level; large
// batches run in waves so that recorded intermediates fit the arena.
static BatchStats batch_dev(Context& c, fpxg_req* reqs, size_t k) {
  BatchStats bs;
  double t0 = host_seconds();
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
  bs.t_prepare = host_seconds() - t0;
  t0 = host_seconds();
  const std::vector<ScanOut> so = scan_inputs(c, scans);
  bs.t_scan = host_seconds() - t0;

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
    const double tr = host_seconds();
    bs.launches += run_all(c, ptrs);
    bs.t_run += host_seconds() - tr;
    ++bs.waves;
    release_chains(c, wave);
    base = c.arena->in_use();
    budget = c.arena->capacity() > base ? (c.arena->capacity() - base) / 2 : 0;
  };
  wave.reserve(k);
  // Requests are recorded one shape class (op, modulus, log2 of the output
  // length) after another, and every kOverlapChunk requests the levels
  // recorded so far start on the device (kOverlapLevels of them) while the
  // host records the next requests: the device no longer waits for the whole
  // batch to be recorded, and the chains started together are of one class,
  // so they still share their launches.
  // (Uneven batch of 9216 requests: 25.4 ms -> 21.7 ms, 261 -> 292 launches;
  // chunks of 256 .. 2048 requests and 2 .. 8 levels all within 21.7-22.6 ms.)
  constexpr size_t kOverlapChunk = 1024, kOverlapLevels = 8;
  std::vector<size_t> order(k);
  for (size_t i = 0; i < k; ++i) order[i] = i;
  auto class_key = [&](size_t i) -> uint64_t {
    const fpxg_req& q = reqs[i];
    return ((uint64_t)(uint32_t)q.op << 40) | ((uint64_t)q.p << 8) | (uint64_t)ceil_log2(std::max<size_t>(pr[i].outc, 1));
  };
  std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return class_key(x) < class_key(y); });
  size_t since = 0;
  for (size_t oi = 0; oi < k; ++oi) {
    const size_t i = order[oi];
    fpxg_req& q = reqs[i];
    if (q.status != OK || pr[i].char2 || pr[i].outc == 0) continue;
    const ScanOut* s = scan_idx[i] == SIZE_MAX ? nullptr : &so[scan_idx[i]];
    for (int attempt = 0; attempt < 3; ++attempt) {
      wave.emplace_back();
      const double tr = host_seconds();
      const int st = record_request(c, q, pr[i], s, wave.back());
      bs.t_record += host_seconds() - tr;
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
    if (c.arena->in_use() > base + budget) {
      flush_wave();
      since = 0;
    } else if (++since >= kOverlapChunk && oi + 1 < k) {
      since = 0;
      std::vector<Chain*> ptrs;
      for (Chain& ch : wave)
        if (!ch.done()) ptrs.push_back(&ch);
      const double tr = host_seconds();
      bs.launches += run_levels(c, ptrs, kOverlapLevels);
      bs.t_run += host_seconds() - tr;
    }
  }
  flush_wave();
  return bs;
}

}  // namespace fpxg
