// Single calls replayed as hipGraphs.
//
// A Newton series call is a few hundred small launches whose shapes depend
// only on (op, p, n, la) and the five words of the input scan.  The first call
// with a given key (those, plus the input and output pointers) runs eagerly.
// The second records its schedule (the recorder of the merged executor, with
// the one-call pass plans), captures it on the library stream into a graph and
// keeps the graph with the arena blocks its kernels use; from then on the key
// replays the graph (the device reads the current input, so new data with the
// same scan signature gets its own result).  Least-recently-used graphs are
// dropped beyond 64 entries or 4 GiB of held blocks.
#pragma once

#include <array>

namespace fpxg {

static void issue_one(Context& c, const LaunchRec& r) {
  switch (r.kind) {
    case LK_ROWS:
      launch_rows(c, r.k1 != 0, r.k2, r.row, r.k0, r.row.nrows);
      return;
    case LK_COLS: {
      const PassPlan P{r.k3, r.k3 - r.k0, r.k0, tile_log_cols(r.k0, r.k3)};
      launch_cols(c, r.k1 != 0, r.k2, r.col, P);
      return;
    }
    case LK_EW:
      hipLaunchKernelGGL(k_ew, dim3((r.ew.n + kChunk - 1) / kChunk), dim3(256), 0, c.stream, (const EwRec*)nullptr,
                         (const uint2*)nullptr, r.ew);
      break;
    case LK_RECUR:
      if (r.rec.hi <= 128)
        hipLaunchKernelGGL(k_recur<64>, dim3(1), dim3(64), 0, c.stream, (const RecurRec*)nullptr, r.rec);
      else
        hipLaunchKernelGGL(k_recur<1024>, dim3(1), dim3(1024), 0, c.stream, (const RecurRec*)nullptr, r.rec);
      break;
    case LK_COLG: {
      const PassPlan P{r.k3, r.k3 - r.k0, r.k0, tile_log_cols(r.k0, r.k3)};
      launch_cols_crt(c, &r.crt, P, 1);
      return;
    }
    case LK_GARNER:
      hipLaunchKernelGGL(k_garner, dim3(grid_for(r.gar.n)), dim3(256), 0, c.stream, r.gar.c1, r.gar.c2, r.gar.c3,
                         r.gar.out, r.gar.n, r.gar.g);
      break;
    case LK_DIRECT:
      hipLaunchKernelGGL(k_conv_direct, dim3((r.dir.outc + kDirectTile - 1) / kDirectTile), dim3(256), 0, c.stream,
                         r.dir.sh, r.dir.ns, r.dir.lg, r.dir.nl, r.dir.out, r.dir.outc, r.dir.p, r.dir.ninv);
      break;
    case LK_TAIL:
      hipLaunchKernelGGL(k_tail_fix, dim3(1), dim3(64), 0, c.stream, r.tail.a, r.tail.la, r.tail.b, r.tail.lb,
                         r.tail.out, r.tail.outc, r.tail.N, r.tail.tail, r.tail.p, r.tail.ninv);
      break;
    default:
      throw Error(E_DEVICE, "unknown launch record");
  }
  FPXG_HIP(hipGetLastError());
}

// Records of one launch level that one launch can carry: the same pass shape
// (and for row passes the same row count), each after the first marked par.
static bool same_launch(const LaunchRec& a, const LaunchRec& b) {
  if (!b.par || a.kind != b.kind || a.k0 != b.k0 || a.k1 != b.k1 || a.k2 != b.k2 || a.k3 != b.k3) return false;
  if (a.kind == LK_COLS || a.kind == LK_COLG) return true;
  return a.kind == LK_ROWS && a.row.nrows == b.row.nrows;
}

// Issues recs[0 .. n) as one launch (n > 1 records of same_launch shape).
static void issue_group(Context& c, const LaunchRec* recs, size_t n) {
  const LaunchRec& r = recs[0];
  if (r.kind == LK_COLG) {
    CrtColJob js[kCrtPack];
    for (size_t i = 0; i < n; ++i) js[i] = recs[i].crt;
    const PassPlan P{r.k3, r.k3 - r.k0, r.k0, tile_log_cols(r.k0, r.k3)};
    launch_cols_crt(c, js, P, (uint32_t)n);
  } else if (r.kind == LK_COLS) {
    ColJob2 js[kColPack];
    for (size_t i = 0; i < n; ++i) js[i] = recs[i].col;
    const PassPlan P{r.k3, r.k3 - r.k0, r.k0, tile_log_cols(r.k0, r.k3)};
    launch_cols_n(c, r.k1 != 0, r.k2, js, P, (uint32_t)n, nullptr);
  } else {
    RowJob2 js[kRowPack];
    for (size_t i = 0; i < n; ++i) js[i] = recs[i].row;
    launch_rows_group(c, r.k1 != 0, r.k2, js, (int)n, r.k0, r.row.nrows);
  }
}

static void drop_graph(Context& c, Context::GraphEntry& e) {
  if (e.exec) (void)hipGraphExecDestroy(e.exec);
  e.exec = nullptr;
  c.arena->flush_deferred(e.bufs);  // later work is stream-ordered after any replay
  c.graph_bytes -= e.bytes;
  e.bytes = 0;
}

static void drop_graphs(Context& c) {
  for (auto& kv : c.graphs) drop_graph(c, kv.second);
  c.graphs.clear();
  c.graph_bytes = 0;
}

// Drops least-recently-used graphs beyond the limits, never the one launched
// last.  The stream is drained first, so no dropped graph is still running.
static void evict_graphs(Context& c) {
  constexpr size_t kMaxGraphs = 64, kMaxBytes = (size_t)4 << 30;
  bool synced = false;
  while (c.graphs.size() > 1 && (c.graphs.size() > kMaxGraphs || c.graph_bytes > kMaxBytes)) {
    auto lru = c.graphs.end();
    for (auto it = c.graphs.begin(); it != c.graphs.end(); ++it)
      if (it->second.stamp != c.graph_clock && (lru == c.graphs.end() || it->second.stamp < lru->second.stamp))
        lru = it;
    if (lru == c.graphs.end()) break;
    if (!synced) {
      FPXG_HIP(hipStreamSynchronize(c.stream));
      synced = true;
    }
    drop_graph(c, lru->second);
    c.graphs.erase(lru);
  }
}

// Records one series call's schedule with the one-call pass plans.
static std::vector<LaunchRec> record_series(Context& c, const HostField& T, const SeriesReq& q, const ScanOut& s,
                                            std::vector<void*>& bufs) {
  Recorder R;
  c.rec = &R;
  c.rec_merged = false;
  c.arena->set_deferral(&bufs);
  try {
    series_emit(c, T, q, s);
  } catch (...) {
    c.rec = nullptr;
    c.rec_merged = true;
    c.arena->set_deferral(nullptr);
    c.arena->flush_deferred(bufs);
    throw;
  }
  c.rec = nullptr;
  c.rec_merged = true;
  c.arena->set_deferral(nullptr);
  return std::move(R.recs);
}

// poly::inv / log / exp / sqrt on device pointers, through the graph cache.
static void series_call(Context& c, const HostField& T, int op, const uint32_t* a, size_t la, size_t n,
                        uint32_t* out) {
  if (op < FPXG_INV || op > FPXG_SQRT) throw Error(E_ARG, "unknown series op");
  if (!c.graphs_enabled || (op == FPXG_SQRT && T.p == 2)) {
    series_dev(c, T, op, a, la, n, out);
    return;
  }
  const SeriesReq q{op, a, la, n, out};
  const ScanOut s = scan_inputs(c, {scan_rec_of(q)})[0];
  // What series_emit reads from the scan: nonzero positions, whether a0 is 0 or
  // 1, and (sqrt only) the value whose square root is taken.
  const uint64_t a0_class = (uint64_t)(s.a0 == 0) | ((uint64_t)(s.a0 == T.from_int(1)) << 1);
  const std::array<uint64_t, 11> key{(uint64_t)op, T.p, n, la, (uint64_t)(uintptr_t)a, (uint64_t)(uintptr_t)out,
                                     s.first, s.second, s.last1, a0_class,
                                     op == FPXG_SQRT ? (uint64_t)s.afirst : 0u};
  auto it = c.graphs.find(key);
  if (it == c.graphs.end()) {
    // A key seen for the first time runs eagerly; capture pays off from the
    // second call on (a loop over the same buffers).
    uint64_t h = 1469598103934665603ull;
    for (uint64_t v : key) h = (h ^ v) * 1099511628211ull;
    if (c.graph_seen.insert(h).second) {
      if (c.graph_seen.size() > 4096) c.graph_seen.clear();
      series_emit(c, T, q, s);
      return;
    }
    Context::GraphEntry e;
    const size_t before = c.arena->in_use();
    std::vector<LaunchRec> recs = record_series(c, T, q, s, e.bufs);
    e.bytes = c.arena->in_use() - before;
    hipGraph_t g = nullptr;
    FPXG_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
    try {
      for (size_t i = 0; i < recs.size();) {
        const LaunchRec& r = recs[i];
        const size_t cap = r.kind == LK_COLS   ? (size_t)kColPack
                           : r.kind == LK_COLG ? (size_t)kCrtPack
                           : r.kind == LK_ROWS ? (size_t)kRowPack
                                               : 1;
        size_t e = i + 1;
        while (e < recs.size() && e - i < cap && same_launch(r, recs[e])) ++e;
        if (e - i > 1) issue_group(c, recs.data() + i, e - i);
        else issue_one(c, r);
        i = e;
      }
    } catch (...) {
      (void)hipStreamEndCapture(c.stream, &g);
      if (g) (void)hipGraphDestroy(g);
      c.arena->flush_deferred(e.bufs);
      throw;
    }
    FPXG_HIP(hipStreamEndCapture(c.stream, &g));
    const hipError_t ie = hipGraphInstantiate(&e.exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ie != hipSuccess) {
      c.arena->flush_deferred(e.bufs);
      throw Error(E_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
    }
    c.graph_bytes += e.bytes;
    it = c.graphs.emplace(key, std::move(e)).first;
  }
  it->second.stamp = ++c.graph_clock;
  FPXG_HIP(hipGraphLaunch(it->second.exec, c.stream));
  evict_graphs(c);
}

}  // namespace fpxg
