/* fpxg — C ABI of the MI355X (gfx950) build of fastpoly's NTT convolution path.
 *
 * Every coefficient crosses the boundary as the raw Montgomery limb of the
 * reference representation (x * 2^32 mod p, include/fastpoly/modint.hpp:54-98),
 * so reinterpret_cast<const uint32_t*>(std::vector<fpx::Mont<P>>::data()) is
 * zero-copy.  Results are bit-identical to the reference CPU library.
 *
 * Host-pointer entry points copy in, compute on the GPU and copy out before
 * returning.  *_dev entry points take device pointers (library-allocated via
 * fpxg_dev_alloc or any HIP allocation on the same device), enqueue on the
 * library stream and return immediately; fpxg_synchronize() waits.
 *
 * Return codes map onto the reference's exceptions (see fpxg_strerror).
 * Thread safety: calls are serialised by an internal mutex.
 */
#ifndef FPXG_H
#define FPXG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  FPXG_OK = 0,
  FPXG_E_DOMAIN = 1,      /* poly::domain_error (detail/poly/base.hpp:26-30) */
  FPXG_E_SIZE = 2,        /* ntt_size_error (ntt.hpp:59-63), std::length_error (convolution.hpp:20-21) */
  FPXG_E_NOMEM = 3,       /* std::bad_alloc: device arena exhausted */
  FPXG_E_DEVICE = 4,      /* HIP runtime failure (std::runtime_error) */
  FPXG_E_UNSUPPORTED = 5, /* modulus not prime or outside [2, 2^31) */
  FPXG_E_ARG = 6,         /* null pointer / output capacity too small */
  FPXG_E_BUSY = 7         /* fpxg_server_poll: the request has not completed yet */
};

enum { FPXG_INV = 0, FPXG_LOG = 1, FPXG_EXP = 2, FPXG_SQRT = 3, FPXG_CONV = 4 };

/* NttPlan variants (ntt.hpp:142-192). */
enum {
  FPXG_NTT_FORWARD = 0,         /* forward(): canonical, bit-reversed order */
  FPXG_NTT_FORWARD_PREFIX = 1,  /* forward_prefix(a, count): a[count:n) treated as zero */
  FPXG_NTT_INVERSE = 2,         /* inverse(): normalised by 1/n */
  FPXG_NTT_INVERSE_PREFIX = 3,  /* inverse_prefix(a, count): only [0, count) valid */
  FPXG_NTT_INVERSE_HIGHHALF = 4 /* inverse_highhalf(): only [n/2, n) valid */
};

typedef struct {
  size_t cached_bytes;       /* idle bytes held by the arena (ScratchMemoryStats, memory.hpp:21-26) */
  size_t cached_blocks;
  size_t system_allocations; /* hipMalloc calls made by the arena */
  size_t cache_hits;
  size_t in_use_bytes;
  size_t capacity_bytes;
} fpxg_stats;

/* Library lifetime.  fpxg_init is implicit on first use (device 0, default
 * arena); call it explicitly to choose the device or cap the arena. */
int fpxg_init(int device, size_t arena_bytes);
void fpxg_shutdown(void);
const char* fpxg_strerror(int code);
const char* fpxg_last_error(void);
int fpxg_arena_stats(fpxg_stats* out);
/* release_scratch_memory (memory.hpp:251-258): drop idle arena blocks. */
int fpxg_release_scratch(void);

/* poly::conv(a, b, n) (detail/poly/convolution.hpp:145-147).  Writes
 * min(min(la,n)+min(lb,n)-1, n) limbs (0 if an input is empty or n == 0);
 * pass n = SIZE_MAX for the full product. */
int fpxg_conv_u32(uint32_t p, const uint32_t* a, size_t la, const uint32_t* b, size_t lb,
                  size_t n, uint32_t* out, size_t out_cap, size_t* out_len);

/* poly::inv / log / exp / sqrt (inverse.hpp:163, exponential.hpp:85 and :107,
 * square_root.hpp:12): n output limbs. */
int fpxg_series_u32(int op, uint32_t p, const uint32_t* a, size_t la, size_t n, uint32_t* out);

/* poly::pow(a, k, n) (detail/poly/power.hpp:17-87): n output limbs. */
int fpxg_pow_u32(uint32_t p, const uint32_t* a, size_t la, uint64_t k, size_t n, uint32_t* out);

/* poly::divmod(a, b) (detail/poly/division.hpp:12-104): a = b q + r with
 * deg r < deg b, q and r trimmed; *q_len / *r_len receive their lengths.
 * poly::div and poly::mod are the same call with r or q null.  Capacities of
 * la limbs always suffice.  FPXG_E_DOMAIN when b is zero (division.hpp:17,
 * :94). */
int fpxg_divmod_u32(uint32_t p, const uint32_t* a, size_t la, const uint32_t* b, size_t lb,
                    uint32_t* q, size_t q_cap, size_t* q_len, uint32_t* r, size_t r_cap, size_t* r_len);

/* NttPlan<Mont<p, root>>::get(n) transforms in place (ntt.hpp:130-192);
 * root == 0 selects the smallest primitive root (modint.hpp:28-52). */
int fpxg_ntt_u32(uint32_t p, uint32_t root, uint32_t n, int variant, size_t count, uint32_t* data);

/* Device-resident forms of conv, series, ntt, pow and divmod.  conv and ntt
 * are asynchronous (stream-ordered on the library stream).  series, pow and
 * divmod first read back a few words of a device scan of the input (the
 * domain checks and special cases are decided on the host before anything is
 * written, as the reference raises its errors), so they wait for the work
 * queued before them on the library stream; the Newton schedule they then
 * enqueue runs asynchronously. */
int fpxg_conv_dev(uint32_t p, const uint32_t* a, size_t la, const uint32_t* b, size_t lb,
                  size_t n, uint32_t* out, size_t out_cap, size_t* out_len);
int fpxg_series_dev(int op, uint32_t p, const uint32_t* a, size_t la, size_t n, uint32_t* out);
/* fpxg_conv_dev that allocates nothing.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, results and errors as fpxg_conv_dev; its scratch (spectra and
 * CRT residues) is a static 32 MiB device array (2^23 words) that each device
 * holds until the process ends.  Limits: products whose scratch exceeds the
 * array are refused with FPXG_E_SIZE (an NTT-prime product up to 2^22
 * coefficients fits; a three-prime product up to 2^19), and a modulus whose
 * plan no earlier call built is refused with FPXG_E_ARG (998244353,
 * 1000000007 and the auxiliary primes are built at set-up).  Every call
 * shares the array and leaves its contents behind; calls through this
 * library run in order on its one stream, but calls that overlap on one GPU
 * (e.g. from two copies of the library in one process) race. */
int fpxg_conv_dev_noalloc(uint32_t p, const uint32_t* a, size_t la, const uint32_t* b, size_t lb,
                          size_t n, uint32_t* out, size_t out_cap, size_t* out_len);
int fpxg_ntt_dev(uint32_t p, uint32_t root, uint32_t n, int variant, size_t count, uint32_t* data);
int fpxg_pow_dev(uint32_t p, const uint32_t* a, size_t la, uint64_t k, size_t n, uint32_t* out);
int fpxg_divmod_dev(uint32_t p, const uint32_t* a, size_t la, const uint32_t* b, size_t lb,
                    uint32_t* q, size_t q_cap, size_t* q_len, uint32_t* r, size_t r_cap, size_t* r_len);

/* One pass of a device-resident forward transform (0 = column pass with the
 * four-step twiddle, 1 = row pass; a single-pass size only has pass 1), for
 * per-kernel timing and tuning.  Returns FPXG_E_ARG for a pass the plan of
 * this size does not have. */
int fpxg_ntt_pass_dev(uint32_t p, uint32_t root, uint32_t n, int pass, uint32_t* data);
/* The plan chosen for an n-point transform: row words 2^*row_log, column
 * levels *col_levels in tiles of 2^*tile_cols_log columns. */
int fpxg_ntt_plan_info(uint32_t n, int* row_log, int* col_levels, int* tile_cols_log);

/* Autotuner (no counterpart in the reference, whose stage split is fixed):
 * for every size 2^min_log..2^max_log, times the forward transform under each
 * admissible pass split (`reps` runs each) and keeps the fastest for single
 * calls; then, from 2^13 up, does the same for the fused product (two
 * forwards, pointwise product, inverse) used by conv and the Newton steps.
 * Optional outputs per size: the forward's chosen column levels (0 = one row
 * pass) and its time in microseconds. */
int fpxg_autotune(int min_log, int max_log, int reps, int* chosen_col_levels, float* best_us);
/* The split chosen for an n-point fused product (no counterpart in the
 * reference): rows of 2^*row_log words after *col_levels column levels. */
int fpxg_product_plan_info(uint32_t n, int* row_log, int* col_levels);

/* Library options (no counterpart in the reference).
 *   FPXG_OPT_GRAPHS: 1 (default) replays repeated single series calls from
 *   captured hipGraphs; 0 drops the cached graphs and launches every call.
 *   FPXG_OPT_ROW_DMA: 1 (default) runs the forward row pass over many
 *   2^14-word rows as one persistent prefetching workgroup per CU; 0 runs one
 *   workgroup per row.
 *   FPXG_OPT_CRT_FUSED_LOG: v in [1, 31] makes every three-prime (CRT)
 *   product of at least 2^v points with a column pass reconstruct by Garner
 *   in its inverse column pass (one workgroup a tile takes every prime in
 *   turn); smaller ones write the primes' residues and run Garner after.
 *   0 restores the default (2^20 points for single calls, every size in
 *   batches and the server).
 *   FPXG_OPT_STEP_LOG_MAX: v in [9, 13]: Newton steps of 2^10 .. 2^v points
 *   over a prime with its own transform run as one workgroup's launch, larger
 *   ones as four-step products; 0 restores the default (12).
 * Results are identical under every option. */
enum { FPXG_OPT_GRAPHS = 0, FPXG_OPT_ROW_DMA = 1, FPXG_OPT_CRT_FUSED_LOG = 2, FPXG_OPT_STEP_LOG_MAX = 3 };
int fpxg_set_option(int option, int64_t value);

/* Device memory and stream plumbing for the *_dev entry points. */
void* fpxg_dev_alloc(size_t bytes);
int fpxg_dev_free(void* ptr);
int fpxg_memcpy_h2d(void* dst, const void* src, size_t bytes);
int fpxg_memcpy_d2h(void* dst, const void* src, size_t bytes);
int fpxg_synchronize(void);
void* fpxg_stream(void);
/* HIP events on the library stream: record slot (0..15), elapsed ms a -> b. */
int fpxg_event_record(int slot);
int fpxg_event_elapsed(int slot_a, int slot_b, float* ms);

/* ---- Batches ------------------------------------------------------------
 * One request = one poly::conv or poly::inv/log/exp/sqrt call, with that
 * call's arguments, result and error code (status).  A batch runs every
 * request's schedule merged with the others': one launch serves the same
 * stage of all requests of the same kernel shape, so thousands of uneven
 * requests over different moduli cost a few hundred launches.  Results are
 * bit-identical to the single calls; a request that fails (status != 0)
 * writes nothing and does not affect the others.
 */
typedef struct {
  int op;              /* FPXG_CONV or FPXG_INV / LOG / EXP / SQRT */
  uint32_t p;          /* modulus: a prime below 2^31 */
  const uint32_t* a;
  size_t la;
  const uint32_t* b;   /* FPXG_CONV only */
  size_t lb;
  size_t n;            /* conv: truncation (SIZE_MAX: full product); series: output length */
  uint32_t* out;
  size_t out_cap;      /* capacity of out in limbs */
  size_t out_len;      /* written: limbs produced */
  int status;          /* written: FPXG_OK or the single call's error code */
} fpxg_req;

/* Device-pointer batch (buffers from fpxg_dev_alloc or HIP on this device,
 * inputs complete in stream order).  Validation and status/out_len happen
 * before return; the computation is enqueued and *ticket names it. */
int fpxg_batch_submit(fpxg_req* reqs, size_t count, uint64_t* ticket);
/* Blocks until the batch named by ticket has completed. */
int fpxg_wait(uint64_t ticket);
/* Host-pointer batch: uploads, runs and downloads before returning. */
int fpxg_batch_u32(fpxg_req* reqs, size_t count);

/* ---- Serving -------------------------------------------------------------
 * A request server with continuous batching: a scheduler thread admits
 * submitted requests while the device arena stays within hbm_budget_bytes,
 * and runs every admitted request's schedule merged with the others', a few
 * levels (levels_per_iteration) per loop, so requests that arrive while others
 * run join their launches.  Requests carry HOST pointers: fpxg_server_submit
 * copies the inputs (a, b) before returning; the caller keeps out and the
 * fpxg_req itself valid until the request is collected by fpxg_server_wait or
 * fpxg_server_poll, which write status and out_len into it and the result into
 * out.  Results are bit-identical to the single calls.  Destroy servers before
 * fpxg_shutdown.
 */
typedef struct fpxg_server fpxg_server;
typedef struct {
  int status;
  size_t out_len;
  double submit_s, admit_s, done_s; /* fpxg_server_now() clock */
} fpxg_result;
typedef struct {
  uint64_t submitted, completed, iterations, launches, max_active, max_queue;
  size_t peak_in_use_bytes;
} fpxg_server_stats;

/* hbm_budget_bytes 0: half the arena; levels_per_iteration 0: 16. */
fpxg_server* fpxg_server_create(size_t hbm_budget_bytes, int levels_per_iteration);
/* FPXG_E_ARG for a null input with a nonzero length (nothing is queued). */
int fpxg_server_submit(fpxg_server* s, fpxg_req* req, uint64_t* id);
/* Blocks until request id completes; fills *res (may be null). */
int fpxg_server_wait(fpxg_server* s, uint64_t id, fpxg_result* res);
/* FPXG_E_BUSY while request id is pending; else as fpxg_server_wait. */
int fpxg_server_poll(fpxg_server* s, uint64_t id, fpxg_result* res);
/* Seconds since the server was created (the clock of fpxg_result). */
double fpxg_server_now(fpxg_server* s);
int fpxg_server_stats_get(fpxg_server* s, fpxg_server_stats* out);
/* Drains outstanding requests, stops the scheduler, frees the server. */
void fpxg_server_destroy(fpxg_server* s);

#ifdef __cplusplus
}
#endif

#endif /* FPXG_H */
