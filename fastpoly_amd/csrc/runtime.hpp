// fastpoly_amd host runtime: device context, arena, per-prime plans and the
// transform / convolution drivers.  One translation unit (fpxg.hip) includes
// this header, the kernels and the series code.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "fpxg.h"
#include "crt.hpp"
#include "ntt2.hpp"

namespace fpxg {

// ---------------------------------------------------------------------------
// Errors: every failure becomes an Error carrying the C-ABI status.
// ---------------------------------------------------------------------------
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
enum { OK = 0, E_DOMAIN = 1, E_SIZE = 2, E_NOMEM = 3, E_DEVICE = 4, E_UNSUPPORTED = 5, E_ARG = 6 };

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

 private:
  struct Chunk { char* base; size_t size, used; };
  size_t cap_;
  size_t reserved_ = 0, in_use_ = 0, hits_ = 0, sys_allocs_ = 0, idle_bytes_ = 0, idle_blocks_ = 0;
  std::vector<Chunk> chunks_;
  std::map<int, std::vector<void*>> free_;  // class -> blocks
  std::map<void*, int> owner_;              // block -> class
  mutable std::mutex mu_;
};

// RAII device buffer from the arena.
class DevBuf {
 public:
  DevBuf() = default;
  DevBuf(Arena* a, size_t words) : a_(a), n_(words) {
    p_ = static_cast<uint32_t*>(a->alloc(std::max<size_t>(words, 1) * 4));
  }
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
    if (p_) a_->free(p_);
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

struct Context {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t events[16] = {};
  size_t lds_bytes = 65536;  // usable dynamic LDS per workgroup
  int row_log_max = 14;      // largest row pass
  int single_log_max = 12;   // largest transform done as one row pass
  int cus = 256;
  std::map<int, PassPlan> plan_override;  // autotuned splits by log2 n
  std::unique_ptr<Arena> arena;
  std::map<uint64_t, std::unique_ptr<HostField>> fields;
  std::recursive_mutex mu;

  PassPlan plan(int L) const;
  HostField& field(uint32_t p, uint32_t root = 0);
  DevBuf buf(size_t words) { return DevBuf(arena.get(), words); }
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
