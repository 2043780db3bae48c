// fastpoly_amd host runtime: device context, arena, per-prime plans and the
// transform / convolution drivers.  fpxg.hip includes it with the drivers and
// the series code; the launch_*.hip units include it for the kernel tables.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <map>
#include <array>
#include <unordered_map>
#include <unordered_set>
#include <memory>
#include <atomic>
#include <chrono>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "fpxg.h"
#include "crt.hpp"
#include "ntt2.hpp"
#include "records.hpp"

namespace fpxg {

// ---------------------------------------------------------------------------
// Errors: every failure becomes an Error carrying the C-ABI status.
// ---------------------------------------------------------------------------
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
enum { OK = 0, E_DOMAIN = 1, E_SIZE = 2, E_NOMEM = 3, E_DEVICE = 4, E_UNSUPPORTED = 5, E_ARG = 6, E_BUSY = 7 };

#define FPXG_HIP(expr)                                                                    \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess)                                                                 \
      throw ::fpxg::Error(::fpxg::E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// ---------------------------------------------------------------------------
// Host Montgomery field (modint.hpp:57-180 semantics, runtime modulus).
// ---------------------------------------------------------------------------
struct HostField {
  uint32_t p = 0, root = 0, ninv = 0, one = 0, r2 = 0;
  int maxlog = 0;   // v2(p - 1)
  bool lazy = false;
  FieldDev dev{};   // device tables (when maxlog >= 1)
  uint32_t* tables = nullptr;
  int table_log = 0;  // transforms up to 2^table_log points have tables

  uint32_t reduce(uint64_t t) const {
    if (p == 2) return (uint32_t)(t & 1u);
    const uint32_t m = (uint32_t)t * ninv;
    const uint64_t u = (t + (uint64_t)m * p) >> 32;
    return (uint32_t)(u >= p ? u - p : u);
  }
  uint32_t mul(uint32_t a, uint32_t b) const { return reduce((uint64_t)a * b); }
  uint32_t add(uint32_t a, uint32_t b) const { uint32_t s = a + b; return s >= p ? s - p : s; }
  uint32_t sub(uint32_t a, uint32_t b) const { return a >= b ? a - b : a - b + p; }
  uint32_t neg(uint32_t a) const { return a == 0 ? 0 : p - a; }
  uint32_t from_int(uint64_t x) const { return (uint32_t)((uint64_t)(uint32_t)(x % p) * one % p); }
  uint32_t val(uint32_t limb) const { return reduce(limb); }
  uint32_t pow(uint32_t a, uint64_t e) const {
    uint32_t r = from_int(1), b = a;
    for (; e; e >>= 1) {
      if (e & 1) r = mul(r, b);
      b = mul(b, b);
    }
    return r;
  }
  uint32_t inv(uint32_t a) const { return p == 2 ? a : pow(a, p - 2); }
  // Mont::sqrt (modint.hpp:139-166): the same root choice.
  bool sqrt(uint32_t a, uint32_t* out) const;
};

uint32_t smallest_primitive_root(uint32_t p);
uint32_t alias_root(uint32_t p);
bool is_prime_u32(uint32_t p);

// ---------------------------------------------------------------------------
// Device arena: power-of-two size classes with per-class free lists, carved
// from large hipMalloc chunks (the device counterpart of memory.hpp:92-181).
// ---------------------------------------------------------------------------
class Arena {
 public:
  explicit Arena(size_t cap) : cap_(cap) {}
  ~Arena();
  void* alloc(size_t bytes);
  void free(void* p);
  void release_idle();
  fpxg_stats stats() const;
  size_t capacity() const { return cap_; }
  size_t in_use() const;
  // While a deferral list is installed, frees are parked on it (a recorded
  // schedule still references the memory); flush_deferred() releases them.
  void set_deferral(std::vector<void*>* list) { defer_ = list; }
  void flush_deferred(std::vector<void*>& list) {
    std::vector<void*>* saved = defer_;
    defer_ = nullptr;
    for (void* p : list) free(p);
    list.clear();
    defer_ = saved;
  }

 private:
  struct Chunk { char* base; size_t size, used; };
  void release_all_locked();
  void reclaim_idle_locked();
  size_t cap_;
  size_t reserved_ = 0, in_use_ = 0, hits_ = 0, sys_allocs_ = 0, idle_bytes_ = 0, idle_blocks_ = 0;
  std::vector<Chunk> chunks_;                  // blocks are carved from the last chunk only
  std::array<std::vector<void*>, 64> free_;    // class -> idle blocks
  std::unordered_map<void*, uint8_t> owner_;   // block -> class
  std::vector<void*>* defer_ = nullptr;
  mutable std::mutex mu_;
};

// RAII device buffer from the arena.
class DevBuf {
 public:
  DevBuf() = default;
  DevBuf(Arena* a, size_t words) : a_(a), n_(words) {
    p_ = static_cast<uint32_t*>(a->alloc(std::max<size_t>(words, 1) * 4));
  }
  // A view of memory the buffer does not own (the static scratch region).
  DevBuf(uint32_t* p, size_t words) : a_(nullptr), p_(p), n_(words) {}
  DevBuf(DevBuf&& o) noexcept : a_(o.a_), p_(o.p_), n_(o.n_) { o.p_ = nullptr; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    reset();
    a_ = o.a_; p_ = o.p_; n_ = o.n_; o.p_ = nullptr;
    return *this;
  }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { reset(); }
  void reset() {
    if (p_ && a_) a_->free(p_);
    p_ = nullptr;
  }
  uint32_t* get() const { return p_; }
  size_t size() const { return n_; }

 private:
  Arena* a_ = nullptr;
  uint32_t* p_ = nullptr;
  size_t n_ = 0;
};

// ---------------------------------------------------------------------------
// Transform plan: how an N = 2^L transform splits into passes.
// ---------------------------------------------------------------------------
struct PassPlan {
  int L, b, S, logC;  // rows of 2^b words; S column levels in tiles of 2^logC columns (S == 0: one row pass)
};

// One recorded kernel launch of a request's schedule.
enum { LK_ROWS = 0, LK_COLS = 1, LK_EW = 2, LK_RECUR = 3, LK_GARNER = 4, LK_COLG = 5, LK_DIRECT = 6, LK_TAIL = 7 };
struct LaunchRec {
  int kind;
  int k0, k1, k2, k3;  // shape: rows (logM, lazy, mode, -), cols (S, lazy, inv, L), CRT cols (S, -, -, L)
  int par;             // independent of the previous record: may run in the same launch level
  union {
    RowJob2 row;
    ColJob2 col;
    EwRec ew;
    RecurRec rec;
    GarnerRec gar;
    CrtColJob crt;
    DirectRec dir;
    TailRec tail;
  };
  LaunchRec() {}
};

struct Recorder {
  std::vector<LaunchRec> recs;
  LaunchRec& next(int kind) {
    recs.emplace_back();
    LaunchRec& r = recs.back();
    r.kind = kind;
    r.k0 = r.k1 = r.k2 = r.k3 = 0;
    r.par = 0;
    return r;
  }
  void push_rows(int logM, bool lazy, int mode, const RowJob2& j, bool par = false) {
    LaunchRec& r = next(LK_ROWS);
    r.k0 = logM; r.k1 = lazy; r.k2 = mode; r.row = j; r.par = par;
  }
  void push_cols(int S, bool lazy, int dir, int L, const ColJob2& j, bool par = false) {
    LaunchRec& r = next(LK_COLS);
    r.k0 = S; r.k1 = lazy; r.k2 = dir; r.k3 = L; r.col = j; r.par = par;
  }
  void push_ew(const EwRec& e) { next(LK_EW).ew = e; }
  void push_recur(const RecurRec& e) { next(LK_RECUR).rec = e; }
  void push_garner(const GarnerRec& g) { next(LK_GARNER).gar = g; }
  void push_direct(const DirectRec& d) { next(LK_DIRECT).dir = d; }
  void push_tail(const TailRec& t) { next(LK_TAIL).tail = t; }
  void push_colg(int S, int L, const CrtColJob& j, bool par = false) {
    LaunchRec& r = next(LK_COLG);
    r.k0 = S; r.k1 = 1; r.k3 = L; r.crt = j; r.par = par;
  }
};

struct Context {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t events[16] = {};
  size_t lds_bytes = 65536;  // usable dynamic LDS per workgroup
  int row_log_max = 14;      // largest row pass
  // Largest Newton step run as one workgroup's launch (2^12 points: at 2^13
  // the four-step products' launches, each on many CUs, finish first; inv /
  // exp / sqrt at 2^20: -2 / -9 / -7 us).
  int step_log_max = 12;
  bool row_dma = true;       // persistent prefetching forward pass for many 2^14-word rows
  int single_log_max = 12;   // largest transform done as one row pass
  // CRT products with a column pass from 2^crt_cols_log points run
  // k_cols2_crt; crt_cols_log_merged: the same in merged (batch, serving)
  // schedules.
  static constexpr int kCrtColsLog = 20, kCrtColsLogMerged = 1;
  int crt_cols_log = kCrtColsLog;
  int crt_cols_log_merged = kCrtColsLogMerged;
  int cus = 256;
  std::map<int, PassPlan> plan_override;  // autotuned splits by log2 n
  std::map<int, PassPlan> conv_override;  // autotuned splits of fused products by log2 n
  Recorder* rec = nullptr;                // non-null: launches are recorded, not issued
  bool rec_merged = true;                 // recorded for the merged executor (else for a graph of one call)
  // Captured single-call schedules (hipGraph) with the arena blocks they use.
  struct GraphEntry {
    hipGraphExec_t exec = nullptr;
    std::vector<void*> bufs;
    size_t bytes = 0;
    uint64_t stamp = 0;
  };
  std::map<std::array<uint64_t, 11>, GraphEntry> graphs;
  uint64_t graph_clock = 0;
  size_t graph_bytes = 0;
  bool graphs_enabled = true;
  std::unordered_set<uint64_t> graph_seen;  // keys called once (hashes)
  // Pinned host slots for job-array uploads of the merged executor.
  struct Staging {
    uint8_t* host = nullptr;
    size_t cap = 0;
    hipEvent_t ev = nullptr;
    bool used = false;
  };
  static constexpr int kStagingSlots = 4;
  Staging staging[kStagingSlots];
  int staging_next = 0;
  // Pinned host memory the input scan of one or two records writes its
  // results to (allocated on first use).
  void* scan_host = nullptr;
  uint32_t scan_token = 0;  // last token the scan wrote into scan_host
  std::unordered_map<uint64_t, GarnerConst> garner_cache;  // by (target p, primes, accumulate)
  CrtFields crt_f{};  // the auxiliary primes' tables, k_cols2_crt's kernel argument
  bool crt_f_ok = false;
  std::unordered_map<uint32_t, bool> prime_cache;          // modulus -> accepted
  // Reciprocal tables 1/i (entry 0 = 0) by modulus: the largest one and its
  // length.  A table is only ever replaced by a longer one, and replaced
  // tables stay allocated until shutdown (cached graphs and recorded
  // schedules may still read them).
  std::unordered_map<uint32_t, std::pair<uint32_t*, size_t>> recip_tabs;
  std::vector<uint32_t*> recip_old;
  // Outstanding batch tickets (completion events).
  std::map<uint64_t, hipEvent_t> tickets;
  uint64_t next_ticket = 1;
  std::unique_ptr<Arena> arena;
  std::map<uint64_t, std::unique_ptr<HostField>> fields;
  std::recursive_mutex mu;

  PassPlan plan(int L) const;
  PassPlan conv_plan(int L) const;
  PassPlan conv_plan_base(int L) const;
  PassPlan plan_base(int L) const;
  HostField& field(uint32_t p, uint32_t root = 0);
  // Static scratch (fpxg_conv_dev_noalloc): this device's address of the
  // file-scope scratch array, and, while a no-allocation call runs, the bump
  // region every buffer is carved from instead of the arena.
  uint32_t* scratch_dev = nullptr;
  struct Bump {
    uint32_t* base;
    size_t cap, used;
  };
  Bump* bump = nullptr;
  DevBuf buf(size_t words) {
    if (!bump) return DevBuf(arena.get(), words);
    const size_t w = (std::max<size_t>(words, 1) + 63) & ~(size_t)63;  // 256-byte aligned, as the arena
    if (bump->used + w > bump->cap) throw Error(E_SIZE, "static scratch exhausted");
    uint32_t* p = bump->base + bump->used;
    bump->used += w;
    return DevBuf(p, words);
  }
};

Context& ctx();  // initialises on first use

// Drivers (all asynchronous on ctx().stream).
void ntt_forward(Context& c, const HostField& F, uint32_t* dst, const uint32_t* src, uint32_t in_count,
                 int L, bool lift, bool canon);
void ntt_inverse(Context& c, const HostField& F, uint32_t* dst, const uint32_t* src, int L,
                 uint32_t out_lo, uint32_t out_hi, uint32_t out_shift);
// Exact product window [0, outc) of a*b over the integers reduced mod the
// field of T (Montgomery limbs in, limbs out), native NTT or three-prime CRT.
void conv_dev(Context& c, const HostField& T, const uint32_t* a, size_t la, const uint32_t* b, size_t lb,
              uint32_t* out, size_t outc);

}  // namespace fpxg
