// fastpoly_amd Newton series on the device: inv, log, exp, sqrt.
//
// Control flow follows the reference (detail/poly/inverse.hpp:18-188,
// exponential.hpp:85-207, square_root.hpp:12-112, calculus.hpp:17-116): a
// 32-coefficient recurrence seed, doubling steps m -> min(2m, n) whose
// products are cyclic transforms of 2m points read only where no wraparound
// lands (the 2m middle-product trick), a persistent inverse prefix c for exp
// and sqrt, and <= 8-coefficient tails filled by the recurrence.  All series
// data stays in device memory; the host only sequences launches and reads back
// the few scalars a domain check needs.  Every output coefficient is unique,
// so the reference's special cases for short inputs (inverse_linear,
// inverse_short, exp_linear) are served by the same Newton path.
//
// Products go through an Engine: the target prime's own NTT when it supports
// the 2m-point transform, otherwise three auxiliary primes (167772161,
// 469762049, 754974721) with Garner reconstruction (crt_convolution.hpp).
#pragma once

namespace fpxg {

struct FieldScalar {
  uint32_t p, ninv, r2, one;
};

__device__ __forceinline__ uint32_t fs_mul(const FieldScalar& f, uint32_t a, uint32_t b) {
  if (f.p == 2) return a & b & 1u;
  return mulmod_full(a, b, f.p, f.ninv);
}
__device__ __forceinline__ uint32_t fs_add(const FieldScalar& f, uint32_t a, uint32_t b) {
  const uint32_t s = a + b;
  return s >= f.p ? s - f.p : s;
}
__device__ __forceinline__ uint32_t fs_sub(const FieldScalar& f, uint32_t a, uint32_t b) {
  return a >= b ? a - b : a - b + f.p;
}
__device__ __forceinline__ uint32_t fs_from_int(const FieldScalar& f, uint32_t x) {
  if (f.p == 2) return x & 1u;
  return mulmod_full(x % f.p, f.r2, f.p, f.ninv);
}
__device__ uint32_t fs_pow(const FieldScalar& f, uint32_t a, uint32_t e) {
  uint32_t r = f.one, b = a;
  for (; e; e >>= 1) {
    if (e & 1) r = fs_mul(f, r, b);
    b = fs_mul(f, b, b);
  }
  return r;
}
__device__ __forceinline__ uint32_t fs_inv(const FieldScalar& f, uint32_t a) {
  return f.p == 2 ? a : fs_pow(f, a, f.p - 2);
}

// ---------------------------------------------------------------------------
// Sequential recurrences (seeds and <= 8-coefficient tails), one workgroup:
// new values are sequential, each dot product is spread over the block.
//   INV : v[i] = -(sum_{j=1}^{min(i,lu-1)} u[j] v[i-j]) * v[0]     (inverse.hpp:68-77)
//   EXP : v[i] = (sum_{j=1}^{min(i,lu)} u[j-1] v[i-j]) / i         (exponential.hpp:158-163)
//   SQRT: v[i] = (w[i] - sum_{j=1}^{i-1} v[j] v[i-j]) * k          (square_root.hpp:46-52)
// For INV with lo == 0, v[0] = u[0]^-1; for EXP/SQRT v[0] is given in init.
// ---------------------------------------------------------------------------
enum { REC_INV = 0, REC_EXP = 1, REC_SQRT = 2 };
constexpr int kRecMax = 64;

// Block-wide sum of one uint64 per thread (blockDim = TPB, a multiple of 64).
template <int TPB>
__device__ __forceinline__ unsigned long long block_sum(unsigned long long v, unsigned long long* red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if constexpr (TPB == 64) {
    return v;
  } else {
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    unsigned long long t = 0;
#pragma unroll
    for (int w = 0; w < TPB / 64; ++w) t += red[w];
    __syncthreads();
    return t;
  }
}

template <int TPB>
__global__ void __launch_bounds__(TPB) k_recur(int kind, FieldScalar f, const uint32_t* __restrict__ u, uint32_t lu,
                                               const uint32_t* __restrict__ w, uint32_t lw, uint32_t* v, uint32_t lo,
                                               uint32_t hi, uint32_t init, uint32_t k) {
  // New values [lo, hi) (hi - lo <= kRecMax) live in LDS; older ones are read from v.
  __shared__ uint32_t nv[kRecMax];
  __shared__ uint32_t recip[kRecMax];
  __shared__ unsigned long long red[TPB / 64];
  __shared__ uint32_t v0s;
  const uint32_t base = lo;
  if (kind == REC_EXP && threadIdx.x < hi - lo) {
    const uint32_t i = lo + threadIdx.x;
    recip[threadIdx.x] = i == 0 ? 0u : fs_inv(f, fs_from_int(f, i));
  }
  if (threadIdx.x == 0) {
    uint32_t v0;
    if (lo == 0) {
      v0 = kind == REC_INV ? fs_inv(f, u[0]) : init;
      nv[0] = v0;
      v[0] = v0;
    } else {
      v0 = v[0];
    }
    v0s = v0;
  }
  __syncthreads();
  const uint32_t v0 = v0s;
  auto val = [&](uint32_t idx) -> uint32_t { return idx < base ? v[idx] : nv[idx - base]; };
  for (uint32_t i = lo == 0 ? 1 : lo; i < hi; ++i) {
    uint32_t jmax;
    if (kind == REC_INV) jmax = min(i, lu ? lu - 1 : 0);
    else if (kind == REC_EXP) jmax = min(i, lu);
    else jmax = i - 1;
    unsigned long long s = 0;
    for (uint32_t j = 1 + threadIdx.x; j <= jmax; j += TPB) {
      const uint32_t x = kind == REC_EXP ? u[j - 1] : (kind == REC_INV ? u[j] : val(j));
      s += fs_mul(f, x, val(i - j));
    }
    s = block_sum<TPB>(s, red);
    if (threadIdx.x == 0) {
      const uint32_t sum = (uint32_t)(s % f.p);
      uint32_t r;
      if (kind == REC_INV) r = fs_mul(f, sum == 0 ? 0u : f.p - sum, v0);
      else if (kind == REC_EXP) r = fs_mul(f, sum, recip[i - lo]);
      else r = fs_mul(f, fs_sub(f, i < lw ? w[i] : 0u, sum), k);
      nv[i - base] = r;
      v[i] = r;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Elementwise kernels.
// ---------------------------------------------------------------------------
#define FPXG_GRID_STRIDE(i, n) \
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < (n); i += (size_t)gridDim.x * blockDim.x)

__global__ void k_negate(uint32_t* x, size_t n, FieldScalar f) {
  FPXG_GRID_STRIDE(i, n) x[i] = x[i] == 0 ? 0 : f.p - x[i];
}
__global__ void k_scale(uint32_t* x, size_t n, uint32_t c, FieldScalar f) {
  FPXG_GRID_STRIDE(i, n) x[i] = fs_mul(f, x[i], c);
}
// derivative_prefix (calculus.hpp:17-39): d[i-1] = a[i] * i.
__global__ void k_derivative(const uint32_t* a, uint32_t* d, size_t cnt, FieldScalar f) {
  FPXG_GRID_STRIDE(i, cnt - 1) d[i] = fs_mul(f, a[i + 1], fs_from_int(f, (uint32_t)(i + 1)));
}
// t[i] *= 1/(off + i) (exponential.hpp:194-195, integral calculus.hpp:97-116).
__global__ void k_mul_recip(uint32_t* t, size_t n, uint32_t off, FieldScalar f) {
  FPXG_GRID_STRIDE(i, n) t[i] = fs_mul(f, t[i], fs_inv(f, fs_from_int(f, (uint32_t)(off + i))));
}
// integral: out[0] = 0, out[i+1] = pr[i]/(i+1) for i < lp, zero up to n.
__global__ void k_integral(const uint32_t* pr, size_t lp, uint32_t* out, size_t n, FieldScalar f) {
  FPXG_GRID_STRIDE(i, n) {
    uint32_t v = 0;
    if (i >= 1 && i - 1 < lp) v = fs_mul(f, pr[i - 1], fs_inv(f, fs_from_int(f, (uint32_t)i)));
    out[i] = v;
  }
}
// res[i] = (i < avail ? a[i] : 0) - sq[i]   (square_root.hpp:93-94)
__global__ void k_sub_from(const uint32_t* a, size_t avail, const uint32_t* sq, uint32_t* res, size_t n,
                           FieldScalar f) {
  FPXG_GRID_STRIDE(i, n) res[i] = fs_sub(f, i < avail ? a[i] : 0u, sq[i]);
}
__global__ void k_copy(uint32_t* dst, const uint32_t* src, size_t n) {
  FPXG_GRID_STRIDE(i, n) dst[i] = src[i];
}
// First and last nonzero index of a[lo, hi): res[0] = min (init hi), res[1] = max + 1 (init lo).
// Each block reduces in LDS and issues one atomic pair.
__global__ void __launch_bounds__(256) k_scan_nonzero(const uint32_t* a, size_t lo, size_t hi, unsigned long long* res) {
  __shared__ unsigned long long mn[256], mx[256];
  unsigned long long first = ~0ull, last = 0;
  for (size_t i = lo + blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < hi; i += (size_t)gridDim.x * blockDim.x)
    if (a[i] != 0) {
      if (first == ~0ull) first = i;
      last = i + 1;
    }
  mn[threadIdx.x] = first;
  mx[threadIdx.x] = last;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
      mn[threadIdx.x] = min(mn[threadIdx.x], mn[threadIdx.x + st]);
      mx[threadIdx.x] = max(mx[threadIdx.x], mx[threadIdx.x + st]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (mn[0] != ~0ull) atomicMin(&res[0], mn[0]);
    if (mx[0] != 0) atomicMax(&res[1], mx[0]);
  }
}

// Characteristic two (square_root.hpp:15-25): (sum b_i x^i)^2 = sum b_i x^(2i).
// out[i/2] = a[i] for even i; any nonzero odd coefficient sets *bad.
__global__ void k_frobenius_sqrt(const uint32_t* a, size_t cnt, uint32_t* out, unsigned int* bad) {
  FPXG_GRID_STRIDE(i, cnt) {
    if (i & 1) {
      if (a[i] != 0) atomicOr(bad, 1u);
    } else {
      out[i / 2] = a[i];
    }
  }
}

// ---------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------

static FieldScalar scalar_of(const HostField& F) { return {F.p, F.ninv, F.r2, F.one}; }

template <class K, class... Args>
static void launch_ew(Context& c, K kernel, size_t n, Args... args) {
  if (n == 0) return;
  hipLaunchKernelGGL(kernel, dim3(grid_for(n)), dim3(256), 0, c.stream, args...);
  FPXG_HIP(hipGetLastError());
}

// Seeds (dot products of at most ~128 terms) run on one wave; longer tails on
// a 1024-thread workgroup.
static void recur(Context& c, const HostField& T, int kind, const uint32_t* u, size_t lu, const uint32_t* w,
                  size_t lw, uint32_t* v, size_t lo, size_t hi, uint32_t init, uint32_t k) {
  if (hi <= lo) return;
  if (hi <= 128) {
    hipLaunchKernelGGL(k_recur<64>, dim3(1), dim3(64), 0, c.stream, kind, scalar_of(T), u, (uint32_t)lu, w,
                       (uint32_t)lw, v, (uint32_t)lo, (uint32_t)hi, init, k);
  } else {
    hipLaunchKernelGGL(k_recur<1024>, dim3(1), dim3(1024), 0, c.stream, kind, scalar_of(T), u, (uint32_t)lu, w,
                       (uint32_t)lw, v, (uint32_t)lo, (uint32_t)hi, init, k);
  }
  FPXG_HIP(hipGetLastError());
}

static uint32_t read_word(Context& c, const uint32_t* p) {
  uint32_t h = 0;
  FPXG_HIP(hipMemcpyAsync(&h, p, 4, hipMemcpyDeviceToHost, c.stream));
  FPXG_HIP(hipStreamSynchronize(c.stream));
  return h;
}

// (first nonzero index, one past the last nonzero) of a[lo, hi); (hi, lo) when all zero.
static std::pair<size_t, size_t> scan_nonzero(Context& c, const uint32_t* a, size_t lo, size_t hi) {
  if (hi <= lo) return {hi, lo};
  DevBuf r = c.buf(4);
  unsigned long long init[2] = {hi, lo}, out[2];
  FPXG_HIP(hipMemcpyAsync(r.get(), init, 16, hipMemcpyHostToDevice, c.stream));
  const unsigned blocks = (unsigned)std::min<size_t>(256, (hi - lo + 255) / 256);
  hipLaunchKernelGGL(k_scan_nonzero, dim3(blocks), dim3(256), 0, c.stream, a, lo, hi,
                     (unsigned long long*)r.get());
  FPXG_HIP(hipGetLastError());
  FPXG_HIP(hipMemcpyAsync(out, r.get(), 16, hipMemcpyDeviceToHost, c.stream));
  FPXG_HIP(hipStreamSynchronize(c.stream));
  return {(size_t)out[0], (size_t)out[1]};
}

static bool series_native(const HostField& T, int L) {
  return T.p != 2 && T.maxlog >= 6 && L <= std::min(T.maxlog, 26);
}

// Spectra of one operand in the engine's field(s).
struct Spec {
  DevBuf b[3];
  int L = 0;
};

struct Engine {
  Context& c;
  const HostField& T;
  int L;
  bool native;
  int k;
  const HostField* F[3];
  Engine(Context& cc, const HostField& t, int l) : c(cc), T(t), L(l), native(series_native(t, l)) {
    if (native) {
      k = 1;
      F[0] = &T;
    } else {
      if (L > 24) throw Error(E_SIZE, "series step beyond the auxiliary primes' 2^24-point transforms");
      k = 3;
      for (int i = 0; i < 3; ++i) F[i] = &c.field(kCrtP[i]);
    }
  }
  void forward(Spec& S, const uint32_t* src, size_t count) {
    const uint32_t N = 1u << L;
    S.L = L;
    for (int i = 0; i < k; ++i) {
      S.b[i] = c.buf(N);
      ntt_forward(c, *F[i], S.b[i].get(), src, (uint32_t)count, L, !native, false);
    }
  }
  // out[i - lo] = (x * y)[i] for i in [lo, hi) of the 2^L-point cyclic product,
  // x given as target limbs (count of them), y as a spectrum; neg negates.
  void product(const uint32_t* x, size_t count, const Spec& Y, size_t lo, size_t hi, uint32_t* out,
               bool neg = false) {
    const uint32_t N = 1u << L;
    const PassPlan P = c.plan(L);
    DevBuf res[3];
    const uint32_t negf = (neg && native) ? R2_NEG : 0u;
    for (int i = 0; i < k; ++i) {
      const HostField& Fi = *F[i];
      uint32_t* dst = out;
      if (!native) {
        res[i] = c.buf(hi - lo);
        dst = res[i].get();
      }
      const uint32_t inv_n = inv_pow2(Fi, L);
      RowJob2 rj{};
      rj.f = Fi.dev;
      if (P.S == 0) {
        rj.src_a = x; rj.src_b = Y.b[i].get(); rj.dst = dst; rj.nrows = 1; rj.in_a = (uint32_t)count;
        rj.out_lo = (uint32_t)lo; rj.out_hi = (uint32_t)hi; rj.out_shift = (uint32_t)lo; rj.inv_n = inv_n;
        rj.flags = R2_SCALE | R2_CANON | negf | (native ? 0u : R2_LIFT);
        launch_rows(c, Fi.lazy, R2M_SPEC, rj, L, 1);
      } else {
        DevBuf W = c.buf(N);
        ColJob2 cj{};
        cj.src = x; cj.dst = W.get(); cj.in_count = (uint32_t)count; cj.flags = native ? 0u : R2_LIFT;
        cj.f = Fi.dev;
        launch_cols(c, Fi.lazy, false, cj, P);
        rj.src_a = W.get(); rj.src_b = Y.b[i].get(); rj.dst = W.get(); rj.nrows = 1u << P.S; rj.in_a = N;
        rj.out_lo = 0; rj.out_hi = N;
        launch_rows(c, Fi.lazy, R2M_SPEC, rj, P.b, rj.nrows);
        ColJob2 ij{};
        ij.src = W.get(); ij.dst = dst; ij.out_lo = (uint32_t)lo; ij.out_hi = (uint32_t)hi;
        ij.out_shift = (uint32_t)lo; ij.inv_n = inv_n; ij.flags = negf; ij.f = Fi.dev;
        launch_cols(c, Fi.lazy, true, ij, P);
      }
    }
    if (!native) {
      const GarnerConst g = garner_const(c, T, 3, false);
      hipLaunchKernelGGL(k_garner, dim3(grid_for(hi - lo)), dim3(256), 0, c.stream, res[0].get(), res[1].get(),
                         res[2].get(), out, (uint32_t)(hi - lo), g);
      FPXG_HIP(hipGetLastError());
      if (neg) launch_ew(c, k_negate, hi - lo, out, hi - lo, scalar_of(T));
    }
  }
};

// extend_inverse (inverse.hpp:18-65): c[0, k) is 1/a mod x^k; fill c[k, wanted).
static void extend_inverse(Context& c, const HostField& T, const uint32_t* a, size_t la, uint32_t* cc, size_t k,
                           size_t wanted) {
  const size_t hlen = wanted - k;
  if (hlen <= 8) {
    recur(c, T, REC_INV, a, la, nullptr, 0, cc, k, wanted, 0, 0);
    return;
  }
  Engine E(c, T, ceil_log2(2 * k));
  Spec Sc;
  E.forward(Sc, cc, k);
  DevBuf e = c.buf(hlen);
  E.product(a, std::min(la, wanted), Sc, k, wanted, e.get(), true);
  E.product(e.get(), hlen, Sc, 0, hlen, cc + k);
}

static void inv_newton(Context& c, const HostField& T, const uint32_t* a, size_t la, size_t n, uint32_t* out) {
  const size_t seed = std::min<size_t>(n, 32);
  recur(c, T, REC_INV, a, la, nullptr, 0, out, 0, seed, 0, 0);
  for (size_t m = seed; m < n; m <<= 1) extend_inverse(c, T, a, la, out, m, std::min(2 * m, n));
}

// poly::inv (inverse.hpp:163-188).
static void series_inv(Context& c, const HostField& T, const uint32_t* a, size_t la, size_t n, uint32_t* out) {
  if (la == 0 || read_word(c, a) == 0) throw Error(E_DOMAIN, "poly::inv: constant term is zero");
  const auto nz = scan_nonzero(c, a, 1, la);
  const size_t cnt = nz.second > 1 ? nz.second : 1;
  if (cnt == 1) {
    fill(c, out, n, 0);
    recur(c, T, REC_INV, a, 1, nullptr, 0, out, 0, 1, 0, 0);
    return;
  }
  inv_newton(c, T, a, cnt, n, out);
}

// poly::log (exponential.hpp:85-98).
static void series_log(Context& c, const HostField& T, const uint32_t* a, size_t la, size_t n, uint32_t* out) {
  if (la == 0 || read_word(c, a) != T.from_int(1)) throw Error(E_DOMAIN, "poly::log: constant term must be 1");
  if (n > T.p) throw Error(E_DOMAIN, "poly::log: required denominators must be invertible (n <= mod)");
  if (n == 1) {
    fill(c, out, 1, 0);
    return;
  }
  const size_t cnt = std::min(la, n);
  const size_t ld = cnt > 1 ? cnt - 1 : 0;
  DevBuf d = c.buf(std::max<size_t>(ld, 1));
  launch_ew(c, k_derivative, ld, a, d.get(), cnt, scalar_of(T));
  DevBuf ia = c.buf(n - 1);
  series_inv(c, T, a, la, n - 1, ia.get());
  size_t lp = 0;
  DevBuf pr = c.buf(n - 1);
  if (ld) {
    lp = std::min(ld + (n - 1) - 1, n - 1);
    conv_dev(c, T, d.get(), ld, ia.get(), n - 1, pr.get(), lp);
  }
  launch_ew(c, k_integral, n, pr.get(), lp, out, n, scalar_of(T));
}

// poly::exp (exponential.hpp:107-207).
static void series_exp(Context& c, const HostField& T, const uint32_t* a, size_t la, size_t n, uint32_t* out) {
  if (la != 0 && read_word(c, a) != 0) throw Error(E_DOMAIN, "poly::exp: constant term must be 0");
  if (n - 1 >= T.p) throw Error(E_DOMAIN, "poly::inv_series: length must be < mod");
  const size_t cnt = std::min(la, n);
  size_t lap = cnt > 1 ? cnt - 1 : 0;
  DevBuf ap = c.buf(std::max<size_t>(lap, 1));
  if (lap) {
    launch_ew(c, k_derivative, lap, a, ap.get(), cnt, scalar_of(T));
    lap = scan_nonzero(c, ap.get(), 0, lap).second;  // trim (basic.hpp:15-17)
  }
  fill(c, out, n, 0);
  const uint32_t one = T.from_int(1);
  if (lap == 0) {
    FPXG_HIP(hipMemcpyAsync(out, &one, 4, hipMemcpyHostToDevice, c.stream));
    FPXG_HIP(hipStreamSynchronize(c.stream));
    return;
  }
  const size_t seed = std::min<size_t>(n, 32);
  uint32_t* b = out;
  recur(c, T, REC_EXP, ap.get(), lap, nullptr, 0, b, 0, seed, one, 0);
  if (n == seed) return;
  DevBuf cbuf = c.buf(n);
  recur(c, T, REC_INV, b, seed, nullptr, 0, cbuf.get(), 0, seed / 2, 0, 0);
  size_t csize = seed / 2;
  for (size_t m = seed; m < n; m *= 2) {
    const size_t m2 = std::min(2 * m, n), hlen = m2 - m;
    if (hlen <= 8) {
      recur(c, T, REC_EXP, ap.get(), lap, nullptr, 0, b, m, m2, one, 0);
      continue;
    }
    if (csize < hlen) {
      extend_inverse(c, T, b, m, cbuf.get(), csize, hlen);
      csize = hlen;
    }
    Engine E(c, T, ceil_log2(2 * m));
    Spec Sb, Sc;
    E.forward(Sb, b, m);
    DevBuf r = c.buf(hlen), t = c.buf(hlen);
    E.product(ap.get(), std::min(lap, m2 - 1), Sb, m - 1, m - 1 + hlen, r.get());
    E.forward(Sc, cbuf.get(), hlen);
    E.product(r.get(), hlen, Sc, 0, hlen, t.get());
    launch_ew(c, k_mul_recip, hlen, t.get(), hlen, (uint32_t)m, scalar_of(T));
    E.product(t.get(), hlen, Sb, 0, hlen, b + m);
  }
}

// poly::sqrt (square_root.hpp:12-112).
static void series_sqrt(Context& c, const HostField& T, const uint32_t* a, size_t la, size_t n, uint32_t* out) {
  fill(c, out, n, 0);
  if (T.p == 2) {
    const size_t cnt = std::min(la, n);
    DevBuf flag = c.buf(1);
    fill(c, flag.get(), 1, 0);
    launch_ew(c, k_frobenius_sqrt, cnt, a, cnt, out, (unsigned int*)flag.get());
    if (read_word(c, flag.get())) throw Error(E_DOMAIN, "poly::sqrt: nonzero odd coefficient in characteristic 2");
    return;
  }
  const auto nz = scan_nonzero(c, a, 0, la);
  const size_t v = nz.first;
  if (v >= la) return;
  if (v & 1) throw Error(E_DOMAIN, "poly::sqrt: odd valuation, no square root");
  uint32_t root0;
  if (!T.sqrt(read_word(c, a + v), &root0)) throw Error(E_DOMAIN, "poly::sqrt: non-residue constant term");
  const size_t half = v / 2;
  if (half >= n) return;
  const size_t inner = n - half, au = la - v;
  const size_t ucnt = std::min(au, inner);
  uint32_t* b = out + half;
  if (scan_nonzero(c, a + v, 1, ucnt).second <= 1) {
    FPXG_HIP(hipMemcpyAsync(b, &root0, 4, hipMemcpyHostToDevice, c.stream));
    FPXG_HIP(hipStreamSynchronize(c.stream));
    return;
  }
  const uint32_t half_inv = T.from_int(((uint64_t)T.p + 1) / 2);
  const uint32_t tri = T.mul(T.inv(root0), half_inv);
  const size_t seed = std::min<size_t>(inner, 32);
  const uint32_t* av = a + v;
  recur(c, T, REC_SQRT, nullptr, 0, av, au, b, 0, seed, root0, tri);
  if (inner == seed) return;
  DevBuf cbuf = c.buf(inner);
  recur(c, T, REC_INV, b, seed, nullptr, 0, cbuf.get(), 0, seed / 2, 0, 0);
  size_t csize = seed / 2;
  for (size_t m = seed; m < inner; m *= 2) {
    const size_t m2 = std::min(2 * m, inner), hlen = m2 - m;
    if (hlen <= 8) {
      recur(c, T, REC_SQRT, nullptr, 0, av, au, b, m, m2, root0, tri);
      continue;
    }
    if (csize < hlen) {
      extend_inverse(c, T, b, m, cbuf.get(), csize, hlen);
      csize = hlen;
    }
    Engine E(c, T, ceil_log2(2 * m));
    Spec Sb, Sc;
    E.forward(Sb, b, m);
    DevBuf sq = c.buf(hlen), res = c.buf(hlen);
    E.product(b, m, Sb, m, m2, sq.get());
    launch_ew(c, k_sub_from, hlen, av + std::min(m, au), au > m ? au - m : 0, (const uint32_t*)sq.get(), res.get(),
              hlen, scalar_of(T));
    E.forward(Sc, cbuf.get(), hlen);
    E.product(res.get(), hlen, Sc, 0, hlen, b + m);
    launch_ew(c, k_scale, hlen, b + m, hlen, half_inv, scalar_of(T));
  }
}

static void series_dev(Context& c, const HostField& T, int op, const uint32_t* a, size_t la, size_t n, uint32_t* out) {
  switch (op) {
    case FPXG_INV: series_inv(c, T, a, la, n, out); break;
    case FPXG_LOG: series_log(c, T, a, la, n, out); break;
    case FPXG_EXP: series_exp(c, T, a, la, n, out); break;
    case FPXG_SQRT: series_sqrt(c, T, a, la, n, out); break;
    default: throw Error(E_ARG, "unknown series op");
  }
}

}  // namespace fpxg
