// fastpoly_amd: the main translation unit of libfpxg.so (hipcc, gfx950); the
// kernel launch tables live in the launch_*.hip units (launch.hpp).
//
// Host runtime (context, arena, per-prime device plans), the transform and
// convolution drivers, the Newton series drivers (series.hpp) and the C ABI
// declared in include/fpxg.h.
#include "launch.hpp"

#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace fpxg {

// ===========================================================================
// Field helpers
// ===========================================================================

static uint64_t powmod_u64(uint64_t a, uint64_t e, uint64_t m) {
  uint64_t r = 1 % m;
  a %= m;
  for (; e; e >>= 1, a = (unsigned __int128)a * a % m)
    if (e & 1) r = (unsigned __int128)r * a % m;
  return r;
}

bool is_prime_u32(uint32_t p) {
  if (p < 2) return false;
  for (uint32_t q : {2u, 3u, 5u, 7u, 11u, 13u})
    if (p % q == 0) return p == q;
  uint32_t d = p - 1;
  int s = 0;
  while ((d & 1) == 0) { d >>= 1; ++s; }
  for (uint32_t a : {2u, 7u, 61u}) {  // deterministic below 2^32
    if (a % p == 0) continue;
    uint64_t x = powmod_u64(a, d, p);
    if (x == 1 || x == p - 1) continue;
    bool comp = true;
    for (int r = 1; r < s; ++r) {
      x = x * x % p;
      if (x == p - 1) { comp = false; break; }
    }
    if (comp) return false;
  }
  return true;
}

// primitive_root (modint.hpp:28-52): smallest g of full order.
uint32_t smallest_primitive_root(uint32_t p) {
  if (p == 2) return 1;
  uint32_t factors[16];
  int count = 0;
  uint32_t rem = p - 1;
  for (uint32_t d = 2; (uint64_t)d * d <= rem; ++d) {
    if (rem % d) continue;
    factors[count++] = d;
    do rem /= d; while (rem % d == 0);
  }
  if (rem > 1) factors[count++] = rem;
  for (uint32_t g = 2; g < p; ++g) {
    bool ok = true;
    for (int i = 0; i < count && ok; ++i) ok = powmod_u64(g, (p - 1) / factors[i], p) != 1;
    if (ok) return g;
  }
  return 0;
}

// Roots of the reference's named aliases (modint.hpp:195-205).
uint32_t alias_root(uint32_t p) {
  switch (p) {
    case 998244353u: case 1004535809u: case 469762049u: case 167772161u: case 1224736769u: return 3;
    case 754974721u: return 11;
    default: return 0;
  }
}

bool HostField::sqrt(uint32_t a, uint32_t* out) const {
  if (p == 2) { *out = a; return true; }
  const uint32_t o = from_int(1);
  if (a == 0) { *out = 0; return true; }
  if (pow(a, (p - 1) / 2) != o) return false;
  if (p % 4 == 3) { *out = pow(a, ((uint64_t)p + 1) / 4); return true; }
  uint64_t q = p - 1;
  int s = 0;
  while (q % 2 == 0) { q /= 2; ++s; }
  uint32_t z = from_int(2);
  while (pow(z, (p - 1) / 2) != neg(o)) z = add(z, o);
  uint32_t cc = pow(z, q), x = pow(a, (q + 1) / 2), t = pow(a, q);
  int m = s;
  while (t != o) {
    int i = 0;
    for (uint32_t t2 = t; t2 != o; t2 = mul(t2, t2)) ++i;
    const uint32_t bb = pow(cc, (uint64_t)1 << (m - i - 1));
    x = mul(x, bb);
    cc = mul(bb, bb);
    t = mul(t, cc);
    m = i;
  }
  *out = x;
  return true;
}

static void init_field(HostField& F, uint32_t p, uint32_t root) {
  F.p = p;
  uint32_t inv = 1;
  for (int i = 0; i < 5; ++i) inv *= 2u - p * inv;
  F.ninv = 0u - inv;
  F.one = p == 2 ? 1u : (uint32_t)((((uint64_t)1) << 32) % p);
  F.r2 = (uint32_t)((uint64_t)F.one * F.one % p);
  F.root = root ? root : (alias_root(p) ? alias_root(p) : smallest_primitive_root(p));
  uint32_t t = p - 1;
  F.maxlog = 0;
  while (t && !(t & 1)) { t >>= 1; ++F.maxlog; }
  F.lazy = p < (1u << 30);
  F.dev.p = p;
  F.dev.ninv = F.ninv;
  F.dev.r2 = F.r2;
  F.dev.one = F.one;
  if (p == 2 || F.maxlog < 1) return;
  // Device tables for transforms of up to 2^logmax points.
  const int logmax = std::min(F.maxlog, 26);
  const int split = std::min(logmax, 12);
  const int plane_log = std::min(logmax, kPlaneLog);
  const size_t lo_n = (size_t)1 << split, hi_n = (size_t)1 << (logmax - split);
  const size_t tw_words = (2 * lo_n + 2 * hi_n + 3) & ~(size_t)3;  // keep the planes 16-byte aligned
  // After the planes: powers of the coset generator (the primitive root) and
  // of its inverse for exponents below 2^logmax, split like the root tables.
  std::vector<uint32_t> h(tw_words + 2 * kPlaneWords + 2 * lo_n + 2 * hi_n, 0);
  uint32_t* lo_f = h.data();
  uint32_t* lo_i = lo_f + lo_n;
  uint32_t* hi_f = lo_i + lo_n;
  uint32_t* hi_i = hi_f + hi_n;
  uint32_t* pl_f = h.data() + tw_words;
  uint32_t* pl_i = pl_f + kPlaneWords;
  const uint32_t g = F.from_int(F.root);
  const uint32_t wmax = F.pow(g, (uint64_t)(p - 1) >> logmax), wmax_i = F.inv(wmax);
  {
    uint32_t cf = F.one, ci = F.one;
    for (size_t k = 0; k < lo_n; ++k) {
      lo_f[k] = cf; lo_i[k] = ci;
      cf = F.mul(cf, wmax); ci = F.mul(ci, wmax_i);
    }
    const uint32_t sf = F.pow(wmax, lo_n), si = F.pow(wmax_i, lo_n);
    cf = F.one; ci = F.one;
    for (size_t k = 0; k < hi_n; ++k) {
      hi_f[k] = cf; hi_i[k] = ci;
      cf = F.mul(cf, sf); ci = F.mul(ci, si);
    }
  }
  {
    uint32_t* co = h.data() + tw_words + 2 * kPlaneWords;
    const uint32_t gi = F.inv(g);
    uint32_t cf = F.one, ci = F.one;
    for (size_t k = 0; k < lo_n; ++k) {
      co[k] = cf; co[lo_n + k] = ci;
      cf = F.mul(cf, g); ci = F.mul(ci, gi);
    }
    const uint32_t sf = F.pow(g, lo_n), si = F.pow(gi, lo_n);
    cf = F.one; ci = F.one;
    for (size_t k = 0; k < hi_n; ++k) {
      co[2 * lo_n + k] = cf; co[2 * lo_n + hi_n + k] = ci;
      cf = F.mul(cf, sf); ci = F.mul(ci, si);
    }
  }
  // Powers of the roots of order 2^k, k <= plane_log, then the planes.
  std::vector<std::vector<uint32_t>> pw_f(plane_log + 1), pw_i(plane_log + 1);
  for (int k = 1; k <= plane_log; ++k) {
    const uint32_t w = F.pow(wmax, (uint64_t)1 << (logmax - k)), wi = F.inv(w);
    pw_f[k].resize((size_t)1 << (k - 1));
    pw_i[k].resize((size_t)1 << (k - 1));
    uint32_t cf = F.one, ci = F.one;
    for (size_t m = 0; m < pw_f[k].size(); ++m) {
      pw_f[k][m] = cf; pw_i[k][m] = ci;
      cf = F.mul(cf, w); ci = F.mul(ci, wi);
    }
  }
  for (int r = 1; r <= 4; ++r)
    for (int logB = r; logB <= plane_log; ++logB) {
      const int lo_bits = logB - r;
      uint32_t* Pf = pl_f + plane_off(logB, r);
      uint32_t* Pi = pl_i + plane_off(logB, r);
      for (uint32_t lo = 0; lo < (1u << lo_bits); ++lo) {
        int idx = 0;
        for (int q = 0; q < r; ++q) {
          const int Lg = logB - q, hh = 1 << (r - q - 1);
          for (int k = 0; k < hh; ++k) {
            const uint32_t m = ((uint32_t)k << lo_bits) | lo;
            Pf[(lo << r) + idx + k] = pw_f[Lg][m];
            Pi[(lo << r) + idx + k] = pw_i[Lg][m];
          }
          idx += hh;
        }
      }
    }
  FPXG_HIP(hipMalloc(&F.tables, h.size() * 4));
  FPXG_HIP(hipMemcpy(F.tables, h.data(), h.size() * 4, hipMemcpyHostToDevice));
  F.table_log = logmax;
  F.dev.logmax = logmax;
  F.dev.split = split;
  F.dev.tw_lo = F.tables;
  F.dev.tw_lo_inv = F.tables + lo_n;
  F.dev.tw_hi = F.tables + 2 * lo_n;
  F.dev.tw_hi_inv = F.tables + 2 * lo_n + hi_n;
  F.dev.planes_fwd = F.tables + tw_words;
  F.dev.planes_inv = F.tables + tw_words + kPlaneWords;
}

static uint32_t inv_pow2(const HostField& F, int L) {
  return F.pow(F.from_int(((uint64_t)F.p + 1) / 2), (uint64_t)L);
}

// ===========================================================================
// Arena
// ===========================================================================

Arena::~Arena() {
  for (auto& c : chunks_) (void)hipFree(c.base);
}

static int size_class(size_t bytes) {
  int k = 8;
  while (((size_t)1 << k) < bytes) ++k;
  return k;
}

void* Arena::alloc(size_t bytes) {
  std::lock_guard<std::mutex> g(mu_);
  const int k = size_class(bytes);
  const size_t sz = (size_t)1 << k;
  auto& fl = free_[(size_t)k];
  if (!fl.empty()) {
    void* p = fl.back();
    fl.pop_back();
    ++hits_;
    idle_bytes_ -= sz;
    --idle_blocks_;
    in_use_ += sz;
    return p;
  }
  if (!chunks_.empty()) {
    Chunk& c = chunks_.back();
    const size_t off = (c.used + 255) & ~(size_t)255;
    if (off + sz <= c.size) {
      c.used = off + sz;
      void* p = c.base + off;
      owner_[p] = (uint8_t)k;
      in_use_ += sz;
      return p;
    }
  }
  const size_t chunk = std::max(sz, (size_t)256 << 20);
  if (reserved_ + sz > cap_) reclaim_idle_locked();
  if (reserved_ + sz > cap_) {
    // At the cap: an idle block of a larger class serves the request before
    // the arena gives up (it returns to its own class when freed).
    for (size_t j = (size_t)k + 1; j < free_.size(); ++j) {
      auto& big = free_[j];
      if (big.empty()) continue;
      void* p = big.back();
      big.pop_back();
      ++hits_;
      idle_bytes_ -= (size_t)1 << j;
      --idle_blocks_;
      in_use_ += (size_t)1 << j;
      return p;
    }
    throw Error(E_NOMEM, "device arena exhausted");
  }
  const size_t take = reserved_ + chunk > cap_ ? sz : chunk;
  char* base = nullptr;
  if (hipMalloc(&base, take) != hipSuccess) {
    (void)hipGetLastError();
    throw Error(E_NOMEM, "hipMalloc failed for the device arena");
  }
  ++sys_allocs_;
  reserved_ += take;
  if (take == sz && !chunks_.empty()) {
    chunks_.insert(chunks_.end() - 1, {base, take, sz});  // keep carving the current chunk
  } else {
    chunks_.push_back({base, take, sz});
  }
  owner_[base] = (uint8_t)k;
  in_use_ += sz;
  return base;
}

size_t Arena::in_use() const {
  std::lock_guard<std::mutex> g(mu_);
  return in_use_;
}

void Arena::free(void* p) {
  if (!p) return;
  std::lock_guard<std::mutex> g(mu_);
  if (defer_) {
    defer_->push_back(p);
    return;
  }
  auto it = owner_.find(p);
  if (it == owner_.end()) return;
  const size_t sz = (size_t)1 << it->second;
  free_[it->second].push_back(p);
  in_use_ -= sz;
  idle_bytes_ += sz;
  ++idle_blocks_;
}

void Arena::release_idle() {
  std::lock_guard<std::mutex> g(mu_);
  if (in_use_ != 0) return;  // chunks are carved; only an idle arena can return them
  release_all_locked();
}

void Arena::release_all_locked() {
  for (auto& c : chunks_) (void)hipFree(c.base);
  chunks_.clear();
  for (auto& fl : free_) fl.clear();
  owner_.clear();
  reserved_ = idle_bytes_ = idle_blocks_ = 0;
}

// Under cap pressure: an arena with nothing in use starts over; otherwise
// every idle block that is a whole allocation of its own (taken at the cap,
// not carved from a chunk) goes back to the device.  hipFree waits for the
// device, so no queued kernel still reads the memory.
void Arena::reclaim_idle_locked() {
  if (in_use_ == 0) {
    release_all_locked();
    return;
  }
  for (size_t i = 0; i < chunks_.size();) {
    Chunk& ch = chunks_[i];
    auto ow = owner_.find(ch.base);
    bool freed = false;
    if (ow != owner_.end() && ((size_t)1 << ow->second) == ch.size && ch.used == ch.size) {
      auto& fl = free_[ow->second];
      auto f = std::find(fl.begin(), fl.end(), (void*)ch.base);
      if (f != fl.end()) {
        fl.erase(f);
        (void)hipFree(ch.base);
        reserved_ -= ch.size;
        idle_bytes_ -= ch.size;
        --idle_blocks_;
        owner_.erase(ow);
        chunks_.erase(chunks_.begin() + (ptrdiff_t)i);
        freed = true;
      }
    }
    if (!freed) ++i;
  }
}

fpxg_stats Arena::stats() const {
  std::lock_guard<std::mutex> g(mu_);
  fpxg_stats s{};
  s.cached_bytes = idle_bytes_;
  s.cached_blocks = idle_blocks_;
  s.system_allocations = sys_allocs_;
  s.cache_hits = hits_;
  s.in_use_bytes = in_use_;
  s.capacity_bytes = cap_;
  return s;
}

// ===========================================================================
// Context and launch helpers
// ===========================================================================

// Scratch of fpxg_conv_dev_noalloc: 2^23 words (32 MiB) of static device
// memory, one copy per device, held until the process ends.  Host code never
// uses this name as an address: context_init asks the runtime for the device's
// copy (Context::scratch_dev) and only that pointer reaches kernels.
constexpr size_t kStaticScratchWords = (size_t)1 << 23;
static __device__ __align__(256) uint32_t scratch[kStaticScratchWords];

static std::unique_ptr<Context> g_ctx;
static std::mutex g_init_mu;
static size_t g_arena_bytes = (size_t)64 << 30;
static int g_device = 0;

// ---- launch routing: recording, then the part that instantiates the shape ----

// A row launch through the part that instantiates its shape.
static void rows2_size(Context& c, bool lazy, int logM, int mode, const RowJob2* j, uint32_t nrows,
                       uint32_t njobs, const RowJob2* dj) {
  if (logM < 1 || logM > 14) throw Error(E_SIZE, "row pass size out of range");
  if (!lazy) rows2_part_canon(c, logM, mode, j, nrows, njobs, dj);
  else if (logM <= 11) rows2_part_lazy_lo(c, logM, mode, j, nrows, njobs, dj);
  else if (logM == 12) rows2_part_lazy_12(c, logM, mode, j, nrows, njobs, dj);
  else if (logM == 13) rows2_part_lazy_13(c, logM, mode, j, nrows, njobs, dj);
  else rows2_part_lazy_14(c, logM, mode, j, nrows, njobs, dj);
}

static void launch_rows(Context& c, bool lazy, int mode, const RowJob2& j, int logM, uint32_t nrows,
                        uint32_t njobs = 1, const RowJob2* dj = nullptr) {
  if (c.rec && !dj) {
    RowJob2 jj = j;
    jj.nrows = nrows;
    c.rec->push_rows(logM, lazy, mode, jj);
    return;
  }
  rows2_size(c, lazy, logM, mode, &j, nrows, njobs, dj);
  FPXG_HIP(hipGetLastError());
}

// k <= kRowPack independent row passes of one shape (nrows rows of 2^logM
// each) in one launch; recorded as k records sharing one launch level.
static void launch_rows_group(Context& c, bool lazy, int mode, const RowJob2* js, int k, int logM,
                              uint32_t nrows) {
  if (k < 1 || k > kRowPack) throw Error(E_DEVICE, "row launch group size");
  RowJob2 jj[kRowPack];
  for (int i = 0; i < k; ++i) {
    jj[i] = js[i];
    jj[i].nrows = nrows;
  }
  if (c.rec) {
    for (int i = 0; i < k; ++i) c.rec->push_rows(logM, lazy, mode, jj[i], i > 0);
    return;
  }
  rows2_size(c, lazy, logM, mode, jj, nrows, (uint32_t)k, nullptr);
  FPXG_HIP(hipGetLastError());
}

// njobs jobs from the device array dj, or (dj null) the njobs by-value jobs
// j[0 .. njobs); recorded as one record per job, all but the first marked as
// sharing the first's launch level.
// dir: kColFwd, kColInv, or kColChain (the chained inverse-forward pass, lazy
// fields only).
enum : int { kColFwd = 0, kColInv = 1, kColChain = 2 };
static void launch_cols_n(Context& c, bool lazy, int dir, const ColJob2* j, const PassPlan& P, uint32_t njobs,
                          const ColJob2* dj) {
  if (!dj && (njobs < 1 || njobs > (uint32_t)kColPack)) throw Error(E_DEVICE, "column launch group size");
  if (dir == kColChain && !lazy) throw Error(E_DEVICE, "chained column pass: lazy fields only");
  if (dir == kColChain && !dj) {
    // The window moves by a register move: 0, a quarter, half or minus a
    // quarter of the transform; a prefix lies below the moved window.
    const uint32_t N = 1u << P.L;
    for (uint32_t i = 0; i < njobs; ++i) {
      const uint32_t sh = j[i].out_shift;
      if (sh != 0 && sh != N / 2 && sh != N / 4 && sh != 0u - N / 4)
        throw Error(E_DEVICE, "chained column pass: shift is not a quarter multiple");
      const uint32_t epif = j[i].flags & (R2_PREFIX | R2_SUBFROM | R2_MULTAB);
      if (epif & (epif - 1)) throw Error(E_DEVICE, "chained column pass: one epilogue operand");
      if ((j[i].flags & R2_PREFIX) && (j[i].out_lo - sh < j[i].epi_avail || j[i].out_hi - sh > N))
        throw Error(E_DEVICE, "chained column pass: prefix overlaps the window");
    }
  }
  if (c.rec && !dj) {
    for (uint32_t i = 0; i < njobs; ++i) c.rec->push_cols(P.S, lazy, dir, P.L, j[i], i > 0);
    return;
  }
  // Signed residues when every by-value job's prime allows them (a device
  // job array may mix primes).
  bool sgn = lazy && dir == kColFwd && !dj;
  for (uint32_t i = 0; sgn && i < njobs; ++i) sgn = signed_ok(j[i].f.p);
  if (!lazy) cols2_part_canon(c, dir != kColFwd, j, P, njobs, dj);
  else if (dir == kColChain) cols2_part_lazy_chain(c, j, P, njobs, dj);
  else if (dir == kColInv) cols2_part_lazy_inv(c, j, P, njobs, dj);
  else cols2_part_lazy_fwd(c, sgn, j, P, njobs, dj);
  FPXG_HIP(hipGetLastError());
}

static void launch_cols(Context& c, bool lazy, int dir, const ColJob2& j, const PassPlan& P, uint32_t njobs = 1,
                        const ColJob2* dj = nullptr) {
  launch_cols_n(c, lazy, dir, &j, P, njobs, dj);
}

// Two independent column passes of one shape in one launch.
static void launch_cols_pair(Context& c, bool lazy, bool inv, const ColJob2& j0, const ColJob2& j1,
                             const PassPlan& P) {
  const ColJob2 js[2] = {j0, j1};
  launch_cols_n(c, lazy, inv, js, P, 2, nullptr);
}

static const CrtFields& crt_fields(Context& c);

// Whether a CRT product of 2^L points (with a column pass) runs k_cols2_crt:
// no residue arrays, no Garner launch, but each workgroup takes every prime's
// tile in turn.  Merged launches (batches, serving) are throughput-bound and
// gain at every size; a lone call's or a Newton step's pass has few
// workgroups below 2^20 points and is faster as one workgroup a prime tile
// plus a Garner launch.
static bool crt_cols_fused(Context& c, int L) {
  const int from = c.rec && c.rec_merged ? c.crt_cols_log_merged : c.crt_cols_log;
  return L >= from && c.conv_plan(L).S > 0;
}

// njobs CRT products' inverse column passes (by value, or the device array
// dj); recorded as one record per product.
static void launch_cols_crt(Context& c, const CrtColJob* j, const PassPlan& P, uint32_t njobs,
                            const CrtColJob* dj = nullptr) {
  if (!dj && (njobs < 1 || njobs > (uint32_t)kCrtPack)) throw Error(E_DEVICE, "CRT column launch group size");
  if (c.rec && !dj) {
    for (uint32_t i = 0; i < njobs; ++i) c.rec->push_colg(P.S, P.L, j[i], i > 0);
    return;
  }
  colcrt_part(c, j, P, njobs, dj, crt_fields(c));
  FPXG_HIP(hipGetLastError());
}

static void context_init(Context& c) {
  FPXG_HIP(hipSetDevice(c.device));
  FPXG_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
  for (auto& e : c.events) FPXG_HIP(hipEventCreate(&e));
  for (auto& s : c.staging) FPXG_HIP(hipEventCreateWithFlags(&s.ev, hipEventDisableTiming));
  hipDeviceProp_t prop;
  FPXG_HIP(hipGetDeviceProperties(&prop, c.device));
  c.cus = prop.multiProcessorCount;
  // Row passes of 2^14 words and column tiles of 2^14 words need 67.6 KB of
  // LDS per workgroup (gfx950: 160 KB per CU).  Without it, stop at 2^13.
  bool big = rows2_attrs_lazy_13();
  big = rows2_attrs_lazy_14() && big;
  big = rows2_attrs_canon() && big;
  big = cols2_attrs_lazy_fwd() && big;
  big = cols2_attrs_lazy_inv() && big;
  big = cols2_attrs_canon() && big;
  big = colcrt_attrs() && big;
  (void)hipGetLastError();
  c.lds_bytes = big ? (size_t)lds_words(1u << 14) * 4 : 65536;
  c.row_log_max = big ? 14 : 13;
  c.arena = std::make_unique<Arena>(g_arena_bytes);
  c.graphs_enabled = getenv("FPXG_NO_GRAPHS") == nullptr;
  // This device's copy of the static scratch array (the device is current).
  void* sp = nullptr;
  FPXG_HIP(hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)));
  c.scratch_dev = static_cast<uint32_t*>(sp);
  // Plans of the reference's named NTT modulus, 1e9+7 and the three auxiliary
  // primes, built at set-up so that the no-allocation entry point finds them.
  for (uint32_t p : {998244353u, 1000000007u, 167772161u, 469762049u, 754974721u}) (void)c.field(p);
}

Context& ctx() {
  std::lock_guard<std::mutex> g(g_init_mu);
  if (!g_ctx) {
    auto c = std::make_unique<Context>();
    c->device = g_device;
    context_init(*c);
    g_ctx = std::move(c);
  }
  return *g_ctx;
}

HostField& Context::field(uint32_t p, uint32_t root) {
  const uint64_t key = ((uint64_t)p << 32) | root;
  auto it = fields.find(key);
  if (it != fields.end()) return *it->second;
  auto F = std::make_unique<HostField>();
  init_field(*F, p, root);
  HostField& ref = *F;
  fields[key] = std::move(F);
  return ref;
}

// Pass split of a 2^L transform: one row pass up to 2^single_log_max; above,
// S = L - b column levels over rows of 2^b words with b <= row_log_max and
// S <= 12, balanced so both passes keep enough workgroups.
PassPlan Context::plan(int L) const {
  if (!(rec && rec_merged)) {  // autotuned splits are measured on lone transforms
    auto it = plan_override.find(L);
    if (it != plan_override.end()) return it->second;
  }
  return plan_base(L);
}

// Split of a fused product (both forwards, the pointwise product and the
// inverse; three transforms per row in the row pass), which needs its own
// choice: the forward's best split can leave the product's row pass with a
// few long rows.  A product with no stored spectrum is free to use it.
// Untuned, a lone product up to 2^13 points is one workgroup's launch.
PassPlan Context::conv_plan(int L) const {
  if (!(rec && rec_merged)) {
    auto it = conv_override.find(L);
    if (it != conv_override.end()) return it->second;
  }
  return conv_plan_base(L);
}

PassPlan Context::conv_plan_base(int L) const {
  if (!(rec && rec_merged) && L <= std::min(row_log_max, 13)) return {L, L, 0, 0};
  return plan_base(L);
}

PassPlan Context::plan_base(int L) const {
  // Recorded schedules run merged with others, so a transform that fits one
  // workgroup's LDS is one row pass there; a lone call splits it for parallelism.
  if (L <= ((rec && rec_merged) ? row_log_max : single_log_max)) return {L, L, 0, 0};
  int S = std::max(L - row_log_max, std::min(L / 2, 10));
  if (S > 12) throw Error(E_SIZE, "transform too large for two passes");
  const int b = L - S;
  return {L, b, S, tile_log_cols(S, L)};
}

static unsigned grid_for(size_t n) {
  size_t g = (n + 255) / 256;
  return (unsigned)std::max<size_t>(1, std::min<size_t>(g, 8192));
}

static int ceil_log2(uint64_t n) {
  int k = 0;
  while (((uint64_t)1 << k) < n) ++k;
  return k;
}

}  // namespace fpxg

#include "ew.hpp"

namespace fpxg {

static void fill(Context& c, uint32_t* p, size_t n, uint32_t v) {
  if (n == 0) return;
  EwRec r{};
  r.op = EW_FILL; r.f = {2u, 0u, 0u, 1u}; r.dst = p; r.n = (uint32_t)n; r.c = v;
  if (c.rec) {
    c.rec->push_ew(r);
    return;
  }
  hipLaunchKernelGGL(k_ew, dim3((unsigned)((n + kChunk - 1) / kChunk)), dim3(256), 0, c.stream, (const EwRec*)nullptr,
                     (const uint2*)nullptr, r);
  FPXG_HIP(hipGetLastError());
}

// ===========================================================================
// Transforms
// ===========================================================================

void ntt_forward(Context& c, const HostField& F, uint32_t* dst, const uint32_t* src, uint32_t in_count, int L,
                 bool lift, bool canon) {
  const PassPlan P = c.plan(L);
  const uint32_t N = 1u << L;
  RowJob2 rj{};
  rj.f = F.dev;
  rj.out_lo = 0; rj.out_hi = N;
  rj.flags = canon ? R2_CANON : 0;
  if (P.S == 0) {
    rj.src_a = src; rj.dst = dst; rj.nrows = 1; rj.in_a = in_count;
    if (lift) rj.flags |= R2_LIFT;
    launch_rows(c, F.lazy, R2M_FWD, rj, L, 1);
    return;
  }
  ColJob2 cj{};
  cj.src = src; cj.dst = dst; cj.in_count = in_count; cj.flags = lift ? R2_LIFT : 0; cj.f = F.dev;
  launch_cols(c, F.lazy, false, cj, P);
  rj.src_a = dst; rj.dst = dst; rj.nrows = 1u << P.S; rj.in_a = N;
  launch_rows(c, F.lazy, R2M_FWD, rj, P.b, rj.nrows);
}

// k forward transforms of one size and field (canonical outputs), their
// column passes kColPack jobs to a launch and their row passes kRowPack: the
// split product's spectra, a few launches that each fill the GPU several
// times over instead of one under-filled launch per pass and transform.
static void ntt_forward_multi(Context& c, const HostField& F, uint32_t* const* dst, const uint32_t* const* src,
                              const uint32_t* in_count, int k, int L, const uint32_t* col_flags = nullptr) {
  // Transforms of 2^22 points and more: rows of 2^14 words, so that the row
  // passes of all k transforms run as one persistent prefetching pass.
  const PassPlan P = (L >= 22 && L - kDmaRowLog <= 12 && c.row_log_max >= kDmaRowLog)
                         ? PassPlan{L, kDmaRowLog, L - kDmaRowLog, tile_log_cols(L - kDmaRowLog, L)}
                         : c.plan(L);
  if (P.S == 0 || c.rec) {
    if (col_flags) throw Error(E_DEVICE, "twisted transforms need a column pass and an immediate launch");
    for (int i = 0; i < k; ++i) ntt_forward(c, F, dst[i], src[i], in_count[i], L, false, true);
    return;
  }
  const uint32_t N = 1u << L;
  for (int i0 = 0; i0 < k; i0 += kColPack) {
    ColJob2 cj[kColPack];
    const int n = std::min(kColPack, k - i0);
    for (int i = 0; i < n; ++i) {
      cj[i] = ColJob2{};
      cj[i].src = src[i0 + i]; cj[i].dst = dst[i0 + i]; cj[i].in_count = in_count[i0 + i]; cj[i].f = F.dev;
      cj[i].flags = col_flags ? col_flags[i0 + i] : 0u;
    }
    launch_cols_n(c, F.lazy, false, cj, P, (uint32_t)n, nullptr);
  }
  // Rows of 2^14 words over a lazy prime: up to kDmaPack transforms' rows in
  // one persistent launch (rows numbered across them), so that each CU walks
  // several rows and the prefetch pipeline fills (four 2^23-point pieces:
  // 61 + 26 us in two launches before).
  if (P.b == kDmaRowLog && F.lazy) {
    std::vector<RowJob2> rj((size_t)k);
    for (int i = 0; i < k; ++i) {
      rj[i] = RowJob2{};
      rj[i].f = F.dev; rj[i].out_lo = 0; rj[i].out_hi = N; rj[i].flags = R2_CANON;
      rj[i].src_a = dst[i]; rj[i].dst = dst[i]; rj[i].in_a = N;
    }
    bool all = true;
    for (int i0 = 0; i0 < k && all; i0 += kDmaPack)
      all = rows_dma_ok(c, rj.data() + i0, 1u << P.S, (uint32_t)std::min(kDmaPack, k - i0));
    if (all) {
      for (int i0 = 0; i0 < k; i0 += kDmaPack)
        if (!rows_dma_part_lazy(c, signed_ok(F.p), rj.data() + i0, 1u << P.S, (uint32_t)std::min(kDmaPack, k - i0)))
          throw Error(E_DEVICE, "persistent row pass refused a checked launch");
      return;
    }
  }
  for (int i0 = 0; i0 < k; i0 += kRowPack) {
    RowJob2 rj[kRowPack];
    const int n = std::min(kRowPack, k - i0);
    for (int i = 0; i < n; ++i) {
      rj[i] = RowJob2{};
      rj[i].f = F.dev; rj[i].out_lo = 0; rj[i].out_hi = N; rj[i].flags = R2_CANON;
      rj[i].src_a = dst[i0 + i]; rj[i].dst = dst[i0 + i]; rj[i].in_a = N;
    }
    launch_rows_group(c, F.lazy, R2M_FWD, rj, n, P.b, 1u << P.S);
  }
}

void ntt_inverse(Context& c, const HostField& F, uint32_t* dst, const uint32_t* src, int L, uint32_t out_lo,
                 uint32_t out_hi, uint32_t out_shift) {
  // src is clobbered (row pass in place); dst may alias src only if out_shift == 0.
  const PassPlan P = c.plan(L);
  const uint32_t N = 1u << L;
  const uint32_t inv_n = inv_pow2(F, L);
  RowJob2 rj{};
  rj.f = F.dev;
  if (P.S == 0) {
    rj.src_a = src; rj.dst = dst; rj.nrows = 1; rj.out_lo = out_lo; rj.out_hi = out_hi; rj.out_shift = out_shift;
    rj.inv_n = inv_n; rj.flags = R2_SCALE | R2_CANON;
    launch_rows(c, F.lazy, R2M_INV, rj, L, 1);
    return;
  }
  uint32_t* work = const_cast<uint32_t*>(src);
  rj.src_a = work; rj.dst = work; rj.nrows = 1u << P.S; rj.out_lo = 0; rj.out_hi = N;
  launch_rows(c, F.lazy, R2M_INV, rj, P.b, rj.nrows);
  ColJob2 cj{};
  cj.src = work; cj.dst = dst; cj.out_lo = out_lo; cj.out_hi = out_hi; cj.out_shift = out_shift;
  cj.inv_n = inv_n; cj.f = F.dev;
  launch_cols(c, F.lazy, true, cj, P);
}

// Product window [0, outc) of the cyclic-free product over one NTT field F;
// with lift the inputs are raw limbs of another field (CRT auxiliary use).
// The same product computed in k <= 3 fields at once (the auxiliary primes of
// a CRT product: one launch per pass carries every field's jobs).  out[i] gets
// the window [0, outc) in field F[i]; lift reads a and b as limbs of another
// field (times r2).  All fields must share the lazy mode.
// Garner fused into the product's inverse column pass (CRT products): the
// target limbs go to out, reconstructed with g.
struct GarnerFuse {
  uint32_t* out;
  GarnerConst g;
};

static bool conv_core_fields(Context& c, const HostField* const* F, int k, const uint32_t* a, size_t la,
                             const uint32_t* b, size_t lb, uint32_t* const* out, size_t outc, bool lift,
                             int L_cyclic = 0, size_t lo = 0, const GarnerFuse* gf = nullptr) {
  // L_cyclic: the 2^L_cyclic-point cyclic product, whose window [lo, lo +
  // outc) is written (lo + outc <= 2^L_cyclic); else the whole product from 0.
  // With gf, where crt_cols_fused says so, the inverse column pass of every
  // prime runs in one launch with Garner in its epilogue and writes gf->out:
  // returns true; otherwise every prime's residues are in out[i] and the
  // caller runs Garner.
  const int L = L_cyclic ? L_cyclic : std::max(1, ceil_log2(la + lb - 1));
  const uint32_t N = 1u << L;
  const PassPlan P = c.conv_plan(L);
  const bool lazy = F[0]->lazy;
  for (int i = 1; i < k; ++i)
    if (F[i]->lazy != lazy) throw Error(E_DEVICE, "fields of one launch differ in lazy mode");
  RowJob2 rj[kRowPack];
  for (int i = 0; i < k; ++i) {
    rj[i] = RowJob2{};
    rj[i].f = F[i]->dev;
    rj[i].inv_n = inv_pow2(*F[i], L);
  }
  if (P.S == 0) {
    for (int i = 0; i < k; ++i) {
      rj[i].src_a = a; rj[i].src_b = b; rj[i].dst = out[i]; rj[i].in_a = (uint32_t)la; rj[i].in_b = (uint32_t)lb;
      rj[i].out_lo = (uint32_t)lo; rj[i].out_hi = (uint32_t)(lo + outc); rj[i].out_shift = (uint32_t)lo;
      rj[i].flags = R2_SCALE | R2_CANON | (lift ? R2_LIFT : 0);
    }
    launch_rows_group(c, lazy, R2M_CONV, rj, k, L, 1);
    return false;
  }
  DevBuf A[kRowPack], B[kRowPack];
  ColJob2 cj[2 * kRowPack];
  for (int i = 0; i < k; ++i) {
    A[i] = c.buf(N);
    B[i] = c.buf(N);
    ColJob2& ja = cj[2 * i];
    ja = ColJob2{};
    ja.src = a; ja.dst = A[i].get(); ja.in_count = (uint32_t)la; ja.flags = lift ? R2_LIFT : 0; ja.f = F[i]->dev;
    ColJob2& jb = cj[2 * i + 1];
    jb = ja;
    jb.src = b; jb.dst = B[i].get(); jb.in_count = (uint32_t)lb;
  }
  launch_cols_n(c, lazy, false, cj, P, (uint32_t)(2 * k), nullptr);
  for (int i = 0; i < k; ++i) {
    rj[i].src_a = A[i].get(); rj[i].src_b = B[i].get(); rj[i].dst = A[i].get(); rj[i].in_a = N; rj[i].in_b = N;
    rj[i].out_lo = 0; rj[i].out_hi = N;
  }
  launch_rows_group(c, lazy, R2M_CONV, rj, k, P.b, 1u << P.S);
  ColJob2 ij[kRowPack];
  for (int i = 0; i < k; ++i) {
    ij[i] = ColJob2{};
    ij[i].src = A[i].get(); ij[i].dst = out[i]; ij[i].out_lo = (uint32_t)lo; ij[i].out_hi = (uint32_t)(lo + outc);
    ij[i].out_shift = (uint32_t)lo; ij[i].inv_n = inv_pow2(*F[i], L); ij[i].f = F[i]->dev;
  }
  if (gf && crt_cols_fused(c, L)) {
    // Every prime's inverse column pass in one workgroup a tile, Garner in
    // its epilogue (also recorded): no residues are written.
    const CrtFields& CF = crt_fields(c);
    if (!lazy) throw Error(E_DEVICE, "CRT column pass: lazy fields only");
    CrtColJob cj{};
    for (int i = 0; i < 3; ++i) {
      const int s = i < k ? i : 0;
      if (F[s]->p != CF.f[s].p) throw Error(E_DEVICE, "CRT column pass: fields out of order");
      cj.src[i] = A[s].get();
      cj.inv_n[i] = inv_pow2(*F[s], L);
    }
    cj.dst = gf->out; cj.out_lo = (uint32_t)lo; cj.out_hi = (uint32_t)(lo + outc); cj.out_shift = (uint32_t)lo;
    cj.g = gf->g;
    launch_cols_crt(c, &cj, P, 1);
    return true;
  }
  launch_cols_n(c, lazy, true, ij, P, (uint32_t)k, nullptr);
  return false;
}

static void conv_core(Context& c, const HostField& F, const uint32_t* a, size_t la, const uint32_t* b, size_t lb,
                      uint32_t* out, size_t outc, bool lift, int L_cyclic = 0) {
  const HostField* Fs[1] = {&F};
  uint32_t* outs[1] = {out};
  conv_core_fields(c, Fs, 1, a, la, b, lb, outs, outc, lift, L_cyclic);
}

static const uint32_t kCrtP[3] = {167772161u, 469762049u, 754974721u};  // crt_convolution.hpp:14-19
static const size_t kCrtBlock = (size_t)1 << 23;                          // crt_convolution.hpp:19

// crt_prime_count (crt_convolution.hpp:78-83).
static const CrtFields& crt_fields(Context& c) {
  if (!c.crt_f_ok) {
    for (int k = 0; k < 3; ++k) c.crt_f.f[k] = c.field(kCrtP[k]).dev;
    c.crt_f_ok = true;
  }
  return c.crt_f;
}


static int crt_prime_count(uint32_t p, size_t terms) {
  const uint64_t sq = (uint64_t)(p - 1) * (p - 1);
  if (terms == 0) terms = 1;
  if (sq <= (kCrtP[0] - 1) / terms) return 1;
  if (sq <= ((uint64_t)kCrtP[0] * kCrtP[1] - 1) / terms) return 2;
  return 3;
}

static GarnerConst garner_const(Context& c, const HostField& T, int primes, bool accumulate) {
  const uint64_t key = ((uint64_t)T.p << 8) | ((uint64_t)primes << 1) | (accumulate ? 1u : 0u);
  auto hit = c.garner_cache.find(key);
  if (hit != c.garner_cache.end()) return hit->second;
  GarnerConst g{};
  const HostField* P[3];
  for (int k = 0; k < 3; ++k) {
    P[k] = &c.field(kCrtP[k]);
    g.P[k] = P[k]->p;
    g.Pninv[k] = P[k]->ninv;
  }
  const uint64_t p1 = kCrtP[0], p2 = kCrtP[1], p3 = kCrtP[2];
  g.inv12 = P[1]->inv(P[1]->from_int(p1));
  g.p1_in3 = P[2]->from_int(p1);
  g.inv123 = P[2]->inv(P[2]->from_int(p1 * p2 % p3));
  g.p = T.p;
  g.ninv = T.ninv;
  g.p1_t = (uint32_t)(p1 % T.p);
  g.p12_t = (uint32_t)((p1 * p2) % T.p);
  g.primes = (uint32_t)primes;
  g.accumulate = accumulate;
  g.neg = 0;
  c.garner_cache[key] = g;
  return g;
}

static void crt_window(Context& c, const HostField& T, const uint32_t* a, size_t la, const uint32_t* b, size_t lb,
                       uint32_t* out, size_t outc, bool accumulate) {
  const int primes = crt_prime_count(T.p, std::min(la, lb));
  const int L = std::max(1, ceil_log2(la + lb - 1));
  const bool fused = crt_cols_fused(c, L);
  DevBuf res[3];
  const HostField* Fs[3];
  uint32_t* outs[3];
  for (int k = 0; k < primes; ++k) {
    if (!fused) res[k] = c.buf(outc);
    Fs[k] = &c.field(kCrtP[k]);
    outs[k] = res[k].get();
  }
  if (fused) {
    GarnerFuse gf{out, garner_const(c, T, primes, accumulate)};
    if (!conv_core_fields(c, Fs, primes, a, la, b, lb, outs, outc, true, 0, 0, &gf))
      throw Error(E_DEVICE, "CRT product: fused column pass not taken");
    return;
  }
  conv_core_fields(c, Fs, primes, a, la, b, lb, outs, outc, true);
  GarnerRec g{};
  g.c1 = res[0].get();
  g.c2 = primes > 1 ? res[1].get() : res[0].get();
  g.c3 = primes > 2 ? res[2].get() : res[0].get();
  g.out = out;
  g.n = (uint32_t)outc;
  g.g = garner_const(c, T, primes, accumulate);
  emit_garner(c, g);
}

static bool native_ok(const HostField& T, size_t full) {
  if (T.p == 2 || T.maxlog < 6) return false;  // ntt_max_size <= naive_threshold: CRT (convolution.hpp:97-100)
  const int L = std::max(1, ceil_log2(full));
  return L <= std::min(T.maxlog, 26);
}

// Products over an NTT prime that exceed its largest transform 2^M (M =
// min(v2(p - 1), 26)): both operands are cut into pieces of h = 2^(M-1)
// coefficients, each piece is transformed once at 2^M points, and output block
// k (2h - 1 coefficients at offset k*h) is the inverse of the spectral sum of
// A_i * B_(k-i), accumulated into the output (neighbouring blocks overlap by
// h - 1).  For 2^23 x 2^23 mod 998244353 that is seven 2^23-point transforms
// against the three-prime route's nine 2^24-point ones (crt_convolution.hpp),
// and the product is unique, so the limbs are the same.
static bool split_ok(const HostField& T, size_t la, size_t lb) {
  if (T.p == 2 || T.maxlog < 16) return false;
  const size_t h = (size_t)1 << (std::min(T.maxlog, 26) - 1);
  return ((la + h - 1) / h) * ((lb + h - 1) / h) <= 64;
}

static void split_conv(Context& c, const HostField& T, const uint32_t* a, size_t la, const uint32_t* b, size_t lb,
                       uint32_t* out, size_t outc) {
  const int M = std::min(T.maxlog, 26);
  const size_t N = (size_t)1 << M, h = N / 2;
  const size_t pa = (la + h - 1) / h, pb = (lb + h - 1) / h;
  const bool same = a == b && la == lb;
  std::vector<DevBuf> SA(pa), SB(same ? 0 : pb);
  std::vector<uint32_t*> fd;
  std::vector<const uint32_t*> fs;
  std::vector<uint32_t> fc;
  for (size_t i = 0; i < pa; ++i) {
    SA[i] = c.buf(N);
    fd.push_back(SA[i].get()); fs.push_back(a + i * h); fc.push_back((uint32_t)std::min(h, la - i * h));
  }
  for (size_t j = 0; j < SB.size(); ++j) {
    SB[j] = c.buf(N);
    fd.push_back(SB[j].get()); fs.push_back(b + j * h); fc.push_back((uint32_t)std::min(h, lb - j * h));
  }
  ntt_forward_multi(c, T, fd.data(), fs.data(), fc.data(), (int)fd.size(), M);
  auto spec_b = [&](size_t j) -> const uint32_t* { return same ? SA[j].get() : SB[j].get(); };
  // Block k covers [k h, k h + 2h - 1); it overlaps block k - 1 on its first
  // h - 1 positions only: block 0 stores, every later block adds below h - 1
  // and stores above, so every output position is written and none is
  // filled beforehand.
  DevBuf W = c.buf(N);
  const PassPlan P = c.plan(M);
  for (size_t k = 0; k + 1 < pa + pb && k * h < outc; ++k) {
    const size_t lo = k * h, hi = std::min(outc, lo + 2 * h - 1);
    const size_t i0 = k + 1 > pb ? k + 1 - pb : 0, i1 = std::min(k, pa - 1);
    RowJob2 rj{};
    rj.f = T.dev; rj.dst = W.get(); rj.nrows = 1u << P.S; rj.in_a = (uint32_t)N; rj.out_lo = 0; rj.out_hi = (uint32_t)N;
    if (i0 == i1) {  // one pair: product and row inverse in one pass
      rj.src_a = SA[i0].get(); rj.src_b = spec_b(k - i0);
      launch_rows(c, T.lazy, R2M_SPEC_ONLY, rj, P.b, rj.nrows);
    } else if (i1 == i0 + 1) {  // two pairs: both products summed in the row pass's loads
      rj.src_a = SA[i0].get(); rj.src_b = spec_b(k - i0);
      rj.src_c = SA[i1].get(); rj.epi = spec_b(k - i1);
      launch_rows(c, T.lazy, R2M_SPEC_SUM2, rj, P.b, rj.nrows);
    } else {
      for (size_t i = i0; i <= i1; ++i)
        emit_ew(c, T, EW_MULADD, W.get(), SA[i].get(), spec_b(k - i), N, 0, i == i0 ? 0u : 1u);
      rj.src_a = W.get();
      launch_rows(c, T.lazy, R2M_INV, rj, P.b, rj.nrows);
    }
    ColJob2 cj{};
    cj.src = W.get(); cj.dst = out + lo; cj.out_lo = 0; cj.out_hi = (uint32_t)(hi - lo); cj.out_shift = 0;
    cj.inv_n = inv_pow2(T, M); cj.flags = k == 0 ? 0u : R2_ACCLO; cj.acc_end = (uint32_t)(h - 1); cj.f = T.dev;
    launch_cols(c, T.lazy, true, cj, P);
  }
}

// Products of two operands of at most N = 2^M coefficients (M = min(v2(p -
// 1), 26)) whose full length exceeds N, over a lazy NTT prime with N < p - 1:
// c = c_lo + x^N c_hi from two N-point cyclic products, one at the roots of
// unity w^k and one on the coset g w^k (g a primitive root, G = g^N != 1):
// the first gives c_lo + c_hi, the second, untwisted by g^-t, c_lo + G c_hi,
// so c_hi = (d - c1) / (G - 1) and c_lo = c1 - c_hi, in the second inverse
// column pass's epilogue.  Four forward transforms (a and b, plain and
// twisted by g^t in the column pass's loads) and two inverse ones, against
// the split product's four and three; the product is unique, so the limbs
// are the same.
static bool coset_ok(const Context& c, const HostField& T, size_t la, size_t lb) {
  if (T.p == 2 || T.maxlog < 16 || !T.lazy || c.rec) return false;
  const int M = std::min(T.maxlog, 26);
  const size_t N = (size_t)1 << M;
  return la <= N && lb <= N && la + lb - 1 > N && (uint64_t)T.p - 1 > N && c.plan(M).S > 0;
}

static void coset_conv(Context& c, const HostField& T, const uint32_t* a, size_t la, const uint32_t* b, size_t lb,
                       uint32_t* out, size_t outc) {
  const int M = std::min(T.maxlog, 26);
  const size_t N = (size_t)1 << M;
  const bool same = a == b && la == lb;
  DevBuf SA = c.buf(N), SA2 = c.buf(N), SB, SB2;
  uint32_t* fd[4] = {SA.get(), SA2.get(), nullptr, nullptr};
  const uint32_t* fs[4] = {a, a, b, b};
  uint32_t fc[4] = {(uint32_t)la, (uint32_t)la, (uint32_t)lb, (uint32_t)lb};
  const uint32_t ff[4] = {0u, R2_TWIST, 0u, R2_TWIST};
  if (!same) {
    SB = c.buf(N);
    SB2 = c.buf(N);
    fd[2] = SB.get();
    fd[3] = SB2.get();
  }
  ntt_forward_multi(c, T, fd, fs, fc, same ? 2 : 4, M, ff);
  const uint32_t* B1 = same ? SA.get() : SB.get();
  const uint32_t* B2 = same ? SA2.get() : SB2.get();
  const PassPlan P = c.plan(M);
  DevBuf W = c.buf(N), C1 = c.buf(N);
  RowJob2 rj{};
  rj.f = T.dev; rj.dst = W.get(); rj.nrows = 1u << P.S; rj.in_a = (uint32_t)N; rj.out_lo = 0; rj.out_hi = (uint32_t)N;
  rj.src_a = SA.get(); rj.src_b = B1;
  launch_rows(c, T.lazy, R2M_SPEC_ONLY, rj, P.b, rj.nrows);
  ColJob2 cj{};
  cj.src = W.get(); cj.dst = C1.get(); cj.out_lo = 0; cj.out_hi = (uint32_t)N; cj.out_shift = 0;
  cj.inv_n = inv_pow2(T, M); cj.f = T.dev;
  launch_cols(c, T.lazy, kColInv, cj, P);
  rj.src_a = SA2.get(); rj.src_b = B2;
  launch_rows(c, T.lazy, R2M_SPEC_ONLY, rj, P.b, rj.nrows);
  const uint32_t G = T.pow(T.from_int(T.root), (uint64_t)N);
  cj.dst = out; cj.out_hi = (uint32_t)std::min(outc, N);
  cj.flags = R2_COSET; cj.epi = C1.get(); cj.epi_avail = (uint32_t)(outc > N ? outc - N : 0);
  cj.acc_end = T.inv(T.sub(G, T.one));
  launch_cols(c, T.lazy, kColInv, cj, P);
}

// The path conv_dev takes for a product of la x lb coefficients (both already
// clamped to outc), decided from the sizes and the field alone.
enum ConvRoute { CR_DIRECT, CR_TAIL, CR_NATIVE, CR_SPLIT, CR_CRT, CR_BLOCKS };

static ConvRoute conv_route(const Context& c, const HostField& T, size_t la, size_t lb, size_t outc) {
  const size_t full = la + lb - 1;
  // A short side: direct product (lone calls, and recorded schedules, whose
  // direct products merge into one launch per level over a chunk table).
  // Measured crossover against the transforms (long side 2^20): a short side
  // of ~150 over an NTT prime, ~350 over the auxiliary primes.
  (void)c;
  const size_t direct_max = native_ok(T, full) || split_ok(T, la, lb) ? 128 : kDirectShort;
  if (T.p != 2 && std::min(la, lb) <= direct_max && outc <= ((size_t)1 << 31)) return CR_DIRECT;
  // A product a few coefficients past a power of two N: the N-point cyclic
  // product plus the high tail computed directly (convolution.hpp:102-137).
  if (T.p != 2 && full > 64) {
    const int Lf = ceil_log2(full) - 1;
    const size_t N = (size_t)1 << Lf, tail = full - N;
    if (tail >= 1 && tail <= 40 && la <= N && lb <= N && native_ok(T, N)) return CR_TAIL;
  }
  if (native_ok(T, full)) return CR_NATIVE;
  if (split_ok(T, la, lb)) return CR_SPLIT;
  if (la <= kCrtBlock && lb <= kCrtBlock) return CR_CRT;
  return CR_BLOCKS;
}

// Device words of scratch the product takes (every buffer rounded to 64 words,
// the arena's and the static region's 256-byte alignment), or SIZE_MAX for
// the paths that are only served from the arena (the split and block-pair
// products above 2^23 coefficients).  Mirrors the buffers of conv_core_fields
// (two 2^L-point spectra per field when the product has a column pass) and
// crt_window (one residue window per auxiliary prime).
static size_t words64(size_t w) { return (std::max<size_t>(w, 1) + 63) & ~(size_t)63; }
static size_t conv_scratch_words(const Context& c, const HostField& T, size_t la, size_t lb, size_t outc) {
  if (la == 0 || lb == 0 || outc == 0) return 0;
  la = std::min(la, outc);
  lb = std::min(lb, outc);
  const size_t full = la + lb - 1;
  auto spectra = [&](int L, int fields) -> size_t {
    return c.conv_plan(L).S == 0 ? 0 : 2 * (size_t)fields * words64((size_t)1 << L);
  };
  switch (conv_route(c, T, la, lb, outc)) {
    case CR_DIRECT: return 0;
    case CR_TAIL: return spectra(ceil_log2(full) - 1, 1);
    case CR_NATIVE: return spectra(std::max(1, ceil_log2(full)), 1);
    case CR_CRT: {
      const int primes = crt_prime_count(T.p, std::min(la, lb));
      return (size_t)primes * words64(outc) + spectra(std::max(1, ceil_log2(full)), primes);
    }
    default: return SIZE_MAX;
  }
}

void conv_dev(Context& c, const HostField& T, const uint32_t* a, size_t la, const uint32_t* b, size_t lb,
              uint32_t* out, size_t outc) {
  if (la == 0 || lb == 0 || outc == 0) return;
  la = std::min(la, outc);
  lb = std::min(lb, outc);
  const size_t full = la + lb - 1;
  switch (conv_route(c, T, la, lb, outc)) {
    case CR_DIRECT: {
      const bool a_short = la <= lb;
      const uint32_t* sh = a_short ? a : b;
      const uint32_t* lg = a_short ? b : a;
      const uint32_t ns = (uint32_t)std::min(la, lb), nl = (uint32_t)std::max(la, lb);
      if (c.rec) {
        c.rec->push_direct(DirectRec{sh, lg, out, ns, nl, (uint32_t)outc, T.p, T.ninv, 0u});
        return;
      }
      const unsigned blocks = (unsigned)((outc + kDirectTile - 1) / kDirectTile);
      hipLaunchKernelGGL(k_conv_direct, dim3(blocks), dim3(256), 0, c.stream, sh, ns, lg, nl, out, (uint32_t)outc,
                         T.p, T.ninv);
      FPXG_HIP(hipGetLastError());
      return;
    }
    case CR_TAIL: {
      const int Lf = ceil_log2(full) - 1;
      const size_t N = (size_t)1 << Lf, tail = full - N;
      conv_core(c, T, a, la, b, lb, out, std::min(outc, N), false, Lf);
      if (c.rec) {
        c.rec->push_tail(TailRec{a, b, out, (uint32_t)la, (uint32_t)lb, (uint32_t)outc, (uint32_t)N, (uint32_t)tail,
                                 T.p, T.ninv, 0u});
        return;
      }
      hipLaunchKernelGGL(k_tail_fix, dim3(1), dim3(64), 0, c.stream, a, (uint32_t)la, b, (uint32_t)lb, out,
                         (uint32_t)outc, (uint32_t)N, (uint32_t)tail, T.p, T.ninv);
      FPXG_HIP(hipGetLastError());
      return;
    }
    case CR_NATIVE:
      conv_core(c, T, a, la, b, lb, out, outc, false);
      return;
    case CR_SPLIT:
      if (coset_ok(c, T, la, lb)) coset_conv(c, T, a, la, b, lb, out, outc);
      else split_conv(c, T, a, la, b, lb, out, outc);
      return;
    case CR_CRT:
      crt_window(c, T, a, la, b, lb, out, outc, false);
      return;
    case CR_BLOCKS:
      break;
  }
  // crt_convolution block pairs (crt_convolution.hpp:187-205).
  fill(c, out, outc, 0);
  for (size_t ai = 0; ai < la;) {
    const size_t na = std::min(kCrtBlock, la - ai);
    for (size_t bi = 0; bi < lb && bi < outc - ai;) {
      const size_t nb = std::min(kCrtBlock, lb - bi);
      const size_t cnt = std::min(na + nb - 1, outc - ai - bi);
      crt_window(c, T, a + ai, na, b + bi, nb, out + ai + bi, cnt, true);
      bi += nb;
    }
    ai += na;
  }
}

}  // namespace fpxg

#include "series.hpp"
#include "batch.hpp"
#include "serve.hpp"
#include "graphs.hpp"
#include "ops.hpp"

// ===========================================================================
// C ABI
// ===========================================================================

using namespace fpxg;

namespace {
thread_local std::string t_last_error;

template <class F>
int guarded(F&& f) {
  try {
    Context& c = ctx();
    std::lock_guard<std::recursive_mutex> g(c.mu);
    f(c);
    return OK;
  } catch (const Error& e) {
    t_last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    t_last_error = "host allocation failed";
    return E_NOMEM;
  } catch (const std::exception& e) {
    t_last_error = e.what();
    return E_DEVICE;
  }
}

void check_modulus(uint32_t p) {
  if (p < 2 || p >= (1u << 31) || !is_prime_u32(p))
    throw Error(E_UNSUPPORTED, "modulus must be a prime in [2, 2^31)");
}

// conv_limbs length rules (convolution.hpp:17-23).
size_t conv_len(size_t la, size_t lb, size_t n, size_t* pla, size_t* plb) {
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

struct Upload {
  DevBuf buf;
  Upload(Context& c, const uint32_t* h, size_t n) : buf(c.buf(n)) {
    if (n) FPXG_HIP(hipMemcpyAsync(buf.get(), h, n * 4, hipMemcpyHostToDevice, c.stream));
  }
};
}  // namespace

namespace fpxg {

// Microseconds per run of fn on the device alone: `reps` runs captured into
// one hipGraph (after an untimed eager run that warms the arena, so nothing
// is allocated during capture), the least of three timed replays.  Eager
// timing would charge every launch its host overhead, which is as long as
// the short passes being compared and is not paid by graph replay.
template <class Fn>
static float time_us_graph(Context& c, int reps, Fn&& fn) {
  fn();
  FPXG_HIP(hipStreamSynchronize(c.stream));
  hipGraph_t g = nullptr;
  FPXG_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
  try {
    for (int r = 0; r < reps; ++r) fn();
  } catch (...) {
    (void)hipStreamEndCapture(c.stream, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  FPXG_HIP(hipStreamEndCapture(c.stream, &g));
  hipGraphExec_t ge = nullptr;
  const hipError_t ie = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ie != hipSuccess) throw Error(E_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
  float best = 1e30f;
  try {
    FPXG_HIP(hipGraphLaunch(ge, c.stream));
    for (int trial = 0; trial < 3; ++trial) {
      FPXG_HIP(hipEventRecord(c.events[14], c.stream));
      FPXG_HIP(hipGraphLaunch(ge, c.stream));
      FPXG_HIP(hipEventRecord(c.events[15], c.stream));
      FPXG_HIP(hipEventSynchronize(c.events[15]));
      float ms = 0;
      FPXG_HIP(hipEventElapsedTime(&ms, c.events[14], c.events[15]));
      best = std::min(best, 1e3f * ms / (float)reps);
    }
  } catch (...) {
    (void)hipGraphExecDestroy(ge);
    throw;
  }
  (void)hipGraphExecDestroy(ge);
  return best;
}

// The autotuner's choice among timed candidates: the fastest, unless the
// default split is within 5% of it (timings of a few repetitions are noisy,
// and a split that only wins by noise can lose badly in a dependent chain).
static PassPlan tuned_pick(const std::vector<PassPlan>& cand, const std::vector<float>& us, const PassPlan& def,
                           float* best_us) {
  size_t ib = 0;
  for (size_t i = 1; i < cand.size(); ++i)
    if (us[i] < us[ib]) ib = i;
  for (size_t i = 0; i < cand.size(); ++i)
    if (cand[i].S == def.S && cand[i].b == def.b && us[i] <= 1.05f * us[ib]) ib = i;
  if (best_us) *best_us = us[ib];
  return cand[ib];
}

}  // namespace fpxg

extern "C" {

int fpxg_init(int device, size_t arena_bytes) {
  {
    std::lock_guard<std::mutex> g(g_init_mu);
    if (g_ctx) return g_ctx->device == device ? OK : E_ARG;
    g_device = device;
    if (arena_bytes) g_arena_bytes = arena_bytes;
  }
  return guarded([](Context&) {});
}

void fpxg_shutdown(void) {
  std::lock_guard<std::mutex> g(g_init_mu);
  if (!g_ctx) return;
  (void)hipStreamSynchronize(g_ctx->stream);
  drop_graphs(*g_ctx);
  for (auto& kv : g_ctx->fields)
    if (kv.second->tables) (void)hipFree(kv.second->tables);
  for (auto& e : g_ctx->events) (void)hipEventDestroy(e);
  for (auto& s : g_ctx->staging) {
    if (s.host) (void)hipHostFree(s.host);
    (void)hipEventDestroy(s.ev);
  }
  for (auto& kv : g_ctx->tickets) (void)hipEventDestroy(kv.second);
  if (g_ctx->scan_host) (void)hipHostFree(g_ctx->scan_host);
  for (auto& kv : g_ctx->recip_tabs) (void)hipFree(kv.second.first);
  for (uint32_t* t : g_ctx->recip_old) (void)hipFree(t);
  (void)hipStreamDestroy(g_ctx->stream);
  g_ctx.reset();
}

const char* fpxg_strerror(int code) {
  switch (code) {
    case OK: return "ok";
    case E_DOMAIN: return "domain error (poly::domain_error)";
    case E_SIZE: return "size error (ntt_size_error / length_error)";
    case E_NOMEM: return "device memory exhausted (bad_alloc)";
    case E_DEVICE: return "HIP device error";
    case E_UNSUPPORTED: return "unsupported modulus";
    case E_ARG: return "invalid argument";
    case E_BUSY: return "request still pending";
    default: return "unknown error";
  }
}

const char* fpxg_last_error(void) { return t_last_error.c_str(); }

int fpxg_arena_stats(fpxg_stats* out) {
  if (!out) return E_ARG;
  return guarded([&](Context& c) { *out = c.arena->stats(); });
}

int fpxg_release_scratch(void) {
  return guarded([&](Context& c) {
    FPXG_HIP(hipStreamSynchronize(c.stream));
    drop_graphs(c);
    c.arena->release_idle();
  });
}

int fpxg_conv_u32(uint32_t p, const uint32_t* a, size_t la, const uint32_t* b, size_t lb, size_t n,
                  uint32_t* out, size_t out_cap, size_t* out_len) {
  return guarded([&](Context& c) {
    check_modulus(p);
    if (!out_len) throw Error(E_ARG, "out_len is null");
    *out_len = 0;
    size_t xa = 0, xb = 0;
    const size_t outc = conv_len(la, lb, n, &xa, &xb);
    if (outc == 0) return;
    if (!a || !b || !out || out_cap < outc) throw Error(E_ARG, "null pointer or output capacity too small");
    const HostField& T = c.field(p);
    Upload ua(c, a, xa);
    const bool same = a == b && xa == xb;
    Upload ub(c, same ? nullptr : b, same ? 0 : xb);
    DevBuf o = c.buf(outc);
    conv_dev(c, T, ua.buf.get(), xa, same ? ua.buf.get() : ub.buf.get(), xb, o.get(), outc);
    FPXG_HIP(hipMemcpyAsync(out, o.get(), outc * 4, hipMemcpyDeviceToHost, c.stream));
    FPXG_HIP(hipStreamSynchronize(c.stream));
    *out_len = outc;
  });
}

int fpxg_conv_dev(uint32_t p, const uint32_t* a, size_t la, const uint32_t* b, size_t lb, size_t n,
                  uint32_t* out, size_t out_cap, size_t* out_len) {
  return guarded([&](Context& c) {
    check_modulus(p);
    if (!out_len) throw Error(E_ARG, "out_len is null");
    *out_len = 0;
    size_t xa = 0, xb = 0;
    const size_t outc = conv_len(la, lb, n, &xa, &xb);
    if (outc == 0) return;
    if (!a || !b || !out || out_cap < outc) throw Error(E_ARG, "null pointer or output capacity too small");
    conv_dev(c, c.field(p), a, xa, b, xb, out, outc);
    *out_len = outc;
  });
}

int fpxg_conv_dev_noalloc(uint32_t p, const uint32_t* a, size_t la, const uint32_t* b, size_t lb, size_t n,
                          uint32_t* out, size_t out_cap, size_t* out_len) {
  return guarded([&](Context& c) {
    check_modulus(p);
    if (!out_len) throw Error(E_ARG, "out_len is null");
    *out_len = 0;
    size_t xa = 0, xb = 0;
    const size_t outc = conv_len(la, lb, n, &xa, &xb);
    if (outc == 0) return;
    if (!a || !b || !out || out_cap < outc) throw Error(E_ARG, "null pointer or output capacity too small");
    // Only plans built earlier (at set-up, or by an earlier call with this
    // modulus): building one allocates device tables.
    auto f = c.fields.find((uint64_t)p << 32);
    if (f == c.fields.end()) throw Error(E_ARG, "modulus not prepared: no earlier call on this device used it");
    const HostField& T = *f->second;
    const size_t need = conv_scratch_words(c, T, xa, xb, outc);
    if (need > kStaticScratchWords)
      throw Error(E_SIZE, "product needs more scratch than the 32 MiB static array (fpxg_conv_dev serves it)");
    if (!c.scratch_dev) throw Error(E_DEVICE, "static scratch address unavailable");
    Context::Bump bump{c.scratch_dev, kStaticScratchWords, 0};
    c.bump = &bump;
    try {
      conv_dev(c, T, a, xa, b, xb, out, outc);
    } catch (...) {
      c.bump = nullptr;
      throw;
    }
    c.bump = nullptr;
    *out_len = outc;
  });
}

static void ntt_check(const HostField& F, uint32_t n, int variant, size_t count) {
  if (n < 2 || (n & (n - 1))) throw Error(E_SIZE, "NTT size must be a power of two");
  if (F.p == 2 || ceil_log2(n) > F.maxlog) throw Error(E_SIZE, "NTT size exceeds 2^v2(mod-1) for this modulus");
  if (ceil_log2(n) > 26) throw Error(E_SIZE, "NTT sizes above 2^26 are not supported on the device");
  if ((variant == FPXG_NTT_FORWARD_PREFIX || variant == FPXG_NTT_INVERSE_PREFIX) && count > n)
    throw Error(E_SIZE, "NTT prefix exceeds plan size");
  if (variant < 0 || variant > FPXG_NTT_INVERSE_HIGHHALF) throw Error(E_ARG, "unknown NTT variant");
}

static void ntt_run(Context& c, const HostField& F, uint32_t n, int variant, size_t count, uint32_t* d) {
  const int L = ceil_log2(n);
  switch (variant) {
    case FPXG_NTT_FORWARD: ntt_forward(c, F, d, d, n, L, false, true); break;
    case FPXG_NTT_FORWARD_PREFIX: ntt_forward(c, F, d, d, (uint32_t)count, L, false, true); break;
    case FPXG_NTT_INVERSE: ntt_inverse(c, F, d, d, L, 0, n, 0); break;
    case FPXG_NTT_INVERSE_PREFIX: ntt_inverse(c, F, d, d, L, 0, (uint32_t)count, 0); break;
    case FPXG_NTT_INVERSE_HIGHHALF: ntt_inverse(c, F, d, d, L, n / 2, n, 0); break;
  }
}

int fpxg_ntt_u32(uint32_t p, uint32_t root, uint32_t n, int variant, size_t count, uint32_t* data) {
  return guarded([&](Context& c) {
    check_modulus(p);
    const HostField& F = c.field(p, root);
    ntt_check(F, n, variant, count);
    if (!data) throw Error(E_ARG, "data is null");
    Upload u(c, data, n);
    ntt_run(c, F, n, variant, count, u.buf.get());
    FPXG_HIP(hipMemcpyAsync(data, u.buf.get(), (size_t)n * 4, hipMemcpyDeviceToHost, c.stream));
    FPXG_HIP(hipStreamSynchronize(c.stream));
  });
}

int fpxg_ntt_dev(uint32_t p, uint32_t root, uint32_t n, int variant, size_t count, uint32_t* data) {
  return guarded([&](Context& c) {
    check_modulus(p);
    const HostField& F = c.field(p, root);
    ntt_check(F, n, variant, count);
    if (!data) throw Error(E_ARG, "data is null");
    ntt_run(c, F, n, variant, count, data);
  });
}

int fpxg_ntt_pass_dev(uint32_t p, uint32_t root, uint32_t n, int pass, uint32_t* data) {
  return guarded([&](Context& c) {
    check_modulus(p);
    const HostField& F = c.field(p, root);
    ntt_check(F, n, FPXG_NTT_FORWARD, 0);
    if (!data) throw Error(E_ARG, "data is null");
    const int L = ceil_log2(n);
    const PassPlan P = c.plan(L);
    if (pass == 0) {
      if (P.S == 0) throw Error(E_ARG, "single-pass size has no column pass");
      ColJob2 cj{};
      cj.src = data; cj.dst = data; cj.in_count = n; cj.f = F.dev;
      launch_cols(c, F.lazy, false, cj, P);
    } else if (pass == 1) {
      RowJob2 rj{};
      rj.src_a = data; rj.dst = data; rj.nrows = 1u << P.S; rj.in_a = n; rj.out_lo = 0; rj.out_hi = n;
      rj.flags = R2_CANON; rj.f = F.dev;
      launch_rows(c, F.lazy, R2M_FWD, rj, P.b, rj.nrows);
    } else {
      throw Error(E_ARG, "pass must be 0 or 1");
    }
  });
}

int fpxg_ntt_plan_info(uint32_t n, int* row_log, int* col_levels, int* tile_cols_log) {
  return guarded([&](Context& c) {
    if (n < 2 || (n & (n - 1))) throw Error(E_SIZE, "NTT size must be a power of two");
    const PassPlan P = c.plan(ceil_log2(n));
    if (row_log) *row_log = P.b;
    if (col_levels) *col_levels = P.S;
    if (tile_cols_log) *tile_cols_log = P.logC;
  });
}

// Autotuner: for each transform size, time the forward transform under every
// admissible pass split (one row pass while it fits a workgroup; otherwise
// S column levels with rows of 2^(L-S) words) and keep the fastest.
int fpxg_autotune(int min_log, int max_log, int reps, int* chosen_col_levels, float* best_us) {
  return guarded([&](Context& c) {
    if (min_log < 1 || max_log > 26 || min_log > max_log || reps < 1) throw Error(E_ARG, "bad autotune range");
    DevBuf buf = c.buf((size_t)1 << max_log);
    fill(c, buf.get(), (size_t)1 << max_log, 1u);
    for (int L = min_log; L <= max_log; ++L) {
      const HostField& F = c.field(L <= 25 ? kCrtP[0] : kCrtP[1]);
      std::vector<PassPlan> cand;
      if (L <= c.row_log_max) cand.push_back({L, L, 0, 0});
      for (int S = std::max(1, L - c.row_log_max); S <= std::min(12, L - 1); ++S)
        if (S >= 4 && L - S >= 4) cand.push_back({L, L - S, S, tile_log_cols(S, L)});
      const PassPlan def = c.plan_base(L);
      if (cand.empty()) cand.push_back(def);
      std::vector<float> times;
      for (const PassPlan& P : cand) {
        c.plan_override[L] = P;
        times.push_back(time_us_graph(c, std::max(reps, 8),
                                      [&] { ntt_forward(c, F, buf.get(), buf.get(), 1u << L, L, false, false); }));
      }
      float best = 0;
      const PassPlan pick = tuned_pick(cand, times, def, &best);
      c.plan_override[L] = pick;
      if (chosen_col_levels) chosen_col_levels[L - min_log] = pick.S;
      if (best_us) best_us[L - min_log] = best;
    }
    // Fused products (two half-length operands, the whole cyclic window):
    // one workgroup's launch up to 2^13 points, else every admissible split.
    DevBuf out = c.buf((size_t)1 << max_log);
    for (int L = std::max(min_log, 13); L <= max_log; ++L) {
      // Timed over 998244353 (the Newton chains' prime, unsigned forward
      // passes) where it has the transform, else an auxiliary prime.
      const HostField& F = c.field(L <= 23 ? 998244353u : (L <= 25 ? kCrtP[0] : kCrtP[1]));
      const size_t half = (size_t)1 << (L - 1);
      c.conv_override.erase(L);
      const PassPlan def = c.conv_plan_base(L);
      // Splits within two column levels of the default (and one workgroup
      // while the size fits it, up to 2^13): the product's passes are short,
      // latency-bound launches, where far-off splits only win by noise.
      const int s_ref = def.S ? def.S : std::max(1, L / 2);
      std::vector<PassPlan> cand;
      if (L <= std::min(c.row_log_max, 13)) cand.push_back({L, L, 0, 0});
      for (int S = std::max(1, L - c.row_log_max); S <= std::min(12, L - 1); ++S)
        if (S >= 4 && L - S >= 4 && std::abs(S - s_ref) <= 2) cand.push_back({L, L - S, S, tile_log_cols(S, L)});
      if (cand.empty()) continue;
      std::vector<float> times;
      for (const PassPlan& P : cand) {
        c.conv_override[L] = P;
        times.push_back(time_us_graph(c, std::max(reps, 8), [&] {
          native_conv(c, F, L, buf.get(), half, buf.get() + half, half, 0, 2 * half, out.get());
        }));
      }
      c.conv_override[L] = tuned_pick(cand, times, def, nullptr);
    }
  });
}

int fpxg_product_plan_info(uint32_t n, int* row_log, int* col_levels) {
  return guarded([&](Context& c) {
    if (n < 2 || (n & (n - 1))) throw Error(E_SIZE, "NTT size must be a power of two");
    const PassPlan P = c.conv_plan(ceil_log2(n));
    if (row_log) *row_log = P.b;
    if (col_levels) *col_levels = P.S;
  });
}

int fpxg_set_option(int option, int64_t value) {
  return guarded([&](Context& c) {
    switch (option) {
      case FPXG_OPT_GRAPHS:
        c.graphs_enabled = value != 0;
        if (!c.graphs_enabled) {
          FPXG_HIP(hipStreamSynchronize(c.stream));
          drop_graphs(c);
          c.graph_seen.clear();
        }
        return;
      case FPXG_OPT_ROW_DMA:
        // Captured graphs hold the old choice: drop them.
        FPXG_HIP(hipStreamSynchronize(c.stream));
        c.row_dma = value != 0;
        drop_graphs(c);
        c.graph_seen.clear();
        return;
      case FPXG_OPT_CRT_FUSED_LOG:
        if (value < 0 || value > 31) throw Error(E_ARG, "CRT fused pass size out of range");
        FPXG_HIP(hipStreamSynchronize(c.stream));
        c.crt_cols_log = value ? (int)value : Context::kCrtColsLog;
        c.crt_cols_log_merged = value ? (int)value : Context::kCrtColsLogMerged;
        drop_graphs(c);
        c.graph_seen.clear();
        return;
      case FPXG_OPT_STEP_LOG_MAX:
        if (value != 0 && (value < 9 || value > 13)) throw Error(E_ARG, "one-workgroup step size out of range");
        FPXG_HIP(hipStreamSynchronize(c.stream));
        c.step_log_max = value ? (int)value : 12;
        drop_graphs(c);
        c.graph_seen.clear();
        return;
      default:
        throw Error(E_ARG, "unknown option");
    }
  });
}

int fpxg_pow_u32(uint32_t p, const uint32_t* a, size_t la, uint64_t k, size_t n, uint32_t* out) {
  return guarded([&](Context& c) {
    check_modulus(p);
    if (n == 0) return;
    if (!out || (la && !a)) throw Error(E_ARG, "null pointer");
    const size_t use = std::min(la, n);
    Upload u(c, a, use);
    DevBuf o = c.buf(n);
    pow_dev(c, c.field(p), u.buf.get(), use, k, n, o.get());
    FPXG_HIP(hipMemcpyAsync(out, o.get(), n * 4, hipMemcpyDeviceToHost, c.stream));
    FPXG_HIP(hipStreamSynchronize(c.stream));
  });
}

int fpxg_pow_dev(uint32_t p, const uint32_t* a, size_t la, uint64_t k, size_t n, uint32_t* out) {
  return guarded([&](Context& c) {
    check_modulus(p);
    if (n == 0) return;
    if (!out || (la && !a)) throw Error(E_ARG, "null pointer");
    pow_dev(c, c.field(p), a, la, k, n, out);
  });
}

int fpxg_divmod_u32(uint32_t p, const uint32_t* a, size_t la, const uint32_t* b, size_t lb, uint32_t* q,
                    size_t q_cap, size_t* q_len, uint32_t* r, size_t r_cap, size_t* r_len) {
  return guarded([&](Context& c) {
    check_modulus(p);
    if (!q_len || !r_len || (la && !a) || (lb && !b)) throw Error(E_ARG, "null pointer");
    *q_len = *r_len = 0;
    Upload ua(c, a, la), ub(c, b, lb);
    DevBuf dq = c.buf(std::max<size_t>(la, 1)), dr = c.buf(std::max<size_t>(la, 1));
    size_t lq = 0, lr = 0;
    divmod_dev(c, c.field(p), ua.buf.get(), la, ub.buf.get(), lb, dq.get(), std::max<size_t>(la, 1), &lq,
               dr.get(), std::max<size_t>(la, 1), &lr);
    if ((q && q_cap < lq) || (r && r_cap < lr)) throw Error(E_ARG, "output capacity too small");
    if (q && lq) FPXG_HIP(hipMemcpyAsync(q, dq.get(), lq * 4, hipMemcpyDeviceToHost, c.stream));
    if (r && lr) FPXG_HIP(hipMemcpyAsync(r, dr.get(), lr * 4, hipMemcpyDeviceToHost, c.stream));
    FPXG_HIP(hipStreamSynchronize(c.stream));
    *q_len = lq;
    *r_len = lr;
  });
}

int fpxg_divmod_dev(uint32_t p, const uint32_t* a, size_t la, const uint32_t* b, size_t lb, uint32_t* q,
                    size_t q_cap, size_t* q_len, uint32_t* r, size_t r_cap, size_t* r_len) {
  return guarded([&](Context& c) {
    check_modulus(p);
    if (!q_len || !r_len || (la && !a) || (lb && !b)) throw Error(E_ARG, "null pointer");
    divmod_dev(c, c.field(p), a, la, b, lb, q, q_cap, q_len, r, r_cap, r_len);
  });
}

int fpxg_series_u32(int op, uint32_t p, const uint32_t* a, size_t la, size_t n, uint32_t* out) {
  return guarded([&](Context& c) {
    check_modulus(p);
    if (n == 0) return;
    if (!out || (la && !a)) throw Error(E_ARG, "null pointer");
    const HostField& T = c.field(p);
    const size_t use = series_input_words(op, a, la, n);
    Upload u(c, a, use);
    DevBuf o = c.buf(n);
    series_call(c, T, op, u.buf.get(), use, n, o.get());
    FPXG_HIP(hipMemcpyAsync(out, o.get(), n * 4, hipMemcpyDeviceToHost, c.stream));
    FPXG_HIP(hipStreamSynchronize(c.stream));
  });
}

int fpxg_series_dev(int op, uint32_t p, const uint32_t* a, size_t la, size_t n, uint32_t* out) {
  return guarded([&](Context& c) {
    check_modulus(p);
    if (n == 0) return;
    if (!out || (la && !a)) throw Error(E_ARG, "null pointer");
    series_call(c, c.field(p), op, a, series_input_words_dev(op, la, n), n, out);
  });
}

void* fpxg_dev_alloc(size_t bytes) {
  void* p = nullptr;
  int rc = guarded([&](Context& c) { p = c.arena->alloc(std::max<size_t>(bytes, 4)); });
  return rc == OK ? p : nullptr;
}

int fpxg_dev_free(void* ptr) {
  return guarded([&](Context& c) {
    FPXG_HIP(hipStreamSynchronize(c.stream));
    c.arena->free(ptr);
  });
}

int fpxg_memcpy_h2d(void* dst, const void* src, size_t bytes) {
  return guarded([&](Context& c) {
    FPXG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c.stream));
    FPXG_HIP(hipStreamSynchronize(c.stream));
  });
}

int fpxg_memcpy_d2h(void* dst, const void* src, size_t bytes) {
  return guarded([&](Context& c) {
    FPXG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c.stream));
    FPXG_HIP(hipStreamSynchronize(c.stream));
  });
}

int fpxg_synchronize(void) {
  return guarded([&](Context& c) { FPXG_HIP(hipStreamSynchronize(c.stream)); });
}

void* fpxg_stream(void) {
  void* s = nullptr;
  (void)guarded([&](Context& c) { s = (void*)c.stream; });
  return s;
}

int fpxg_event_record(int slot) {
  return guarded([&](Context& c) {
    if (slot < 0 || slot >= 16) throw Error(E_ARG, "event slot out of range");
    FPXG_HIP(hipEventRecord(c.events[slot], c.stream));
  });
}

int fpxg_event_elapsed(int a, int b, float* ms) {
  return guarded([&](Context& c) {
    if (a < 0 || a >= 16 || b < 0 || b >= 16 || !ms) throw Error(E_ARG, "bad event slot");
    FPXG_HIP(hipEventSynchronize(c.events[b]));
    FPXG_HIP(hipEventElapsedTime(ms, c.events[a], c.events[b]));
  });
}

// ---- merged batches ----

static void prune_tickets(Context& c) {
  for (auto it = c.tickets.begin(); it != c.tickets.end();) {
    if (hipEventQuery(it->second) == hipSuccess) {
      (void)hipEventDestroy(it->second);
      it = c.tickets.erase(it);
    } else {
      ++it;
    }
  }
}

int fpxg_batch_submit(fpxg_req* reqs, size_t count, uint64_t* ticket) {
  return guarded([&](Context& c) {
    if (!reqs && count) throw Error(E_ARG, "reqs is null");
    prune_tickets(c);
    const BatchStats bs = batch_dev(c, reqs, count);
    static const bool trace = getenv("FPXG_TRACE_BATCH") != nullptr;
    if (trace)
      fprintf(stderr, "fpxg batch: %zu requests, %zu records, %zu launches, %zu waves; host ms prepare %.2f scan %.2f "
              "record %.2f run %.2f\n", count, bs.records, bs.launches, bs.waves, 1e3 * bs.t_prepare,
              1e3 * bs.t_scan, 1e3 * bs.t_record, 1e3 * bs.t_run);
    hipEvent_t ev;
    FPXG_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    FPXG_HIP(hipEventRecord(ev, c.stream));
    const uint64_t t = c.next_ticket++;
    c.tickets[t] = ev;
    if (ticket) *ticket = t;
  });
}

int fpxg_wait(uint64_t ticket) {
  hipEvent_t ev = nullptr;
  int rc = guarded([&](Context& c) {
    if (ticket == 0 || ticket >= c.next_ticket) throw Error(E_ARG, "unknown ticket");
    auto it = c.tickets.find(ticket);
    if (it == c.tickets.end()) return;  // already complete and pruned
    ev = it->second;
    c.tickets.erase(it);
  });
  if (rc != OK || !ev) return rc;
  // Wait outside the library lock so other threads can keep submitting.
  const hipError_t e = hipEventSynchronize(ev);
  (void)hipEventDestroy(ev);
  if (e != hipSuccess) {
    t_last_error = hipGetErrorString(e);
    return E_DEVICE;
  }
  return OK;
}

int fpxg_batch_u32(fpxg_req* reqs, size_t count) {
  return guarded([&](Context& c) {
    if (!reqs && count) throw Error(E_ARG, "reqs is null");
    // One device block for every input and output, requests rewritten to it.
    std::vector<fpxg_req> dq(reqs, reqs + count);
    std::vector<size_t> in_a(count), in_b(count), out_n(count), off_a(count), off_b(count), off_o(count);
    size_t total = 0;
    for (size_t i = 0; i < count; ++i) {
      const fpxg_req& q = reqs[i];
      const bool conv = q.op == FPXG_CONV;
      in_a[i] = !q.a ? 0 : conv ? std::min(q.la, q.n) : (q.n ? series_input_words(q.op, q.a, q.la, q.n) : 0);
      in_b[i] = conv && q.b ? std::min(q.lb, q.n) : 0;
      out_n[i] = q.out ? std::min(q.out_cap, conv ? std::min(q.n, q.la + q.lb) : q.n) : 0;
      off_a[i] = total; total += in_a[i];
      off_b[i] = total; total += in_b[i];
      off_o[i] = total; total += out_n[i];
    }
    DevBuf dev = c.buf(std::max<size_t>(total, 1));
    uint32_t* d = dev.get();
    std::vector<uint32_t> host(total);
    for (size_t i = 0; i < count; ++i) {
      const fpxg_req& q = reqs[i];
      if (in_a[i]) memcpy(host.data() + off_a[i], q.a, in_a[i] * 4);
      if (in_b[i]) memcpy(host.data() + off_b[i], q.b, in_b[i] * 4);
      dq[i].a = q.a ? d + off_a[i] : nullptr;
      dq[i].la = q.a ? in_a[i] : q.la;
      dq[i].b = q.b ? d + off_b[i] : nullptr;
      dq[i].lb = q.b ? in_b[i] : q.lb;
      dq[i].out = q.out ? d + off_o[i] : nullptr;
      dq[i].out_cap = q.out ? out_n[i] : q.out_cap;
    }
    if (total) FPXG_HIP(hipMemcpyAsync(d, host.data(), total * 4, hipMemcpyHostToDevice, c.stream));
    batch_dev(c, dq.data(), count);
    if (total) FPXG_HIP(hipMemcpyAsync(host.data(), d, total * 4, hipMemcpyDeviceToHost, c.stream));
    FPXG_HIP(hipStreamSynchronize(c.stream));
    for (size_t i = 0; i < count; ++i) {
      reqs[i].status = dq[i].status;
      reqs[i].out_len = dq[i].status == OK ? dq[i].out_len : 0;
      if (reqs[i].out_len) memcpy(reqs[i].out, host.data() + off_o[i], reqs[i].out_len * 4);
    }
  });
}

// ---- serving ----

}  // extern "C"

struct fpxg_server {
  std::unique_ptr<fpxg::Server> s;
};

extern "C" {

fpxg_server* fpxg_server_create(size_t hbm_budget_bytes, int levels_per_iteration) {
  fpxg_server* out = nullptr;
  const int rc = guarded([&](Context& c) {
    const size_t budget = hbm_budget_bytes ? hbm_budget_bytes : c.arena->capacity() / 2;
    const int lv = levels_per_iteration > 0 ? levels_per_iteration : 16;
    auto h = std::make_unique<fpxg_server>();
    h->s = std::make_unique<Server>(c, budget, lv);
    out = h.release();
  });
  return rc == OK ? out : nullptr;
}

int fpxg_server_submit(fpxg_server* h, fpxg_req* req, uint64_t* id) {
  if (!h || !req || !id) return E_ARG;
  Server& S = *h->s;
  std::unique_ptr<SItem> it;
  // Staging allocates pinned host memory and may throw; nothing may escape
  // into a C caller, and the staged block goes back to the pool on every
  // early return.
  auto give_back = [&] {
    if (it) {
      S.pinned.free(it->pin_in, it->pin_in_cls);
      it->pin_in = nullptr;
    }
  };
  try {
    it = std::make_unique<SItem>();
    it->user = req;
    it->t_submit = S.now();
    const int rc = S.stage(*it);  // this thread copies and scans the input
    if (rc != OK) {
      give_back();
      t_last_error = "invalid request";
      return rc;
    }
    {
      std::lock_guard<std::mutex> g(S.mu);
      if (!S.fatal.empty() || S.stopping) {
        t_last_error = S.fatal.empty() ? "server is stopping" : S.fatal;
        give_back();
        return E_DEVICE;
      }
      it->id = S.next_id++;
      *id = it->id;
      S.outstanding.insert(it->id);
      ++S.st.submitted;
      S.incoming.push_back(std::move(it));
    }
  } catch (const Error& e) {
    give_back();
    t_last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    give_back();
    t_last_error = "host allocation failed";
    return E_NOMEM;
  } catch (const std::exception& e) {
    give_back();
    t_last_error = e.what();
    return E_DEVICE;
  }
  S.cv_work.notify_one();
  return OK;
}

static int server_collect(fpxg_server* h, uint64_t id, fpxg_result* res, bool block) {
  if (!h) return E_ARG;
  Server& S = *h->s;
  std::unique_ptr<SItem> it;
  {
    std::unique_lock<std::mutex> g(S.mu);
    if (!S.outstanding.count(id)) return E_ARG;
    if (block) S.cv_done.wait(g, [&] { return S.finished.count(id) > 0; });
    auto f = S.finished.find(id);
    if (f == S.finished.end()) return E_BUSY;
    it = std::move(f->second);
    S.finished.erase(f);
    S.outstanding.erase(id);
  }
  S.finish(*it);  // result copy in this thread
  if (res) *res = it->result;
  return OK;
}

int fpxg_server_wait(fpxg_server* h, uint64_t id, fpxg_result* res) { return server_collect(h, id, res, true); }
int fpxg_server_poll(fpxg_server* h, uint64_t id, fpxg_result* res) { return server_collect(h, id, res, false); }

double fpxg_server_now(fpxg_server* h) { return h ? h->s->now() : 0.0; }

int fpxg_server_stats_get(fpxg_server* h, fpxg_server_stats* out) {
  if (!h || !out) return E_ARG;
  std::lock_guard<std::mutex> g(h->s->mu);
  *out = h->s->st;
  return OK;
}

void fpxg_server_destroy(fpxg_server* h) { delete h; }

}  // extern "C"
