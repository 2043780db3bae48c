#!/usr/bin/env python3
"""Benchmark of the fastpoly NTT convolution path on one MI355X.

Workload (one "step"), mod 998244353 with synthetic uniform [0, p) limbs:
  * conv of two 2^23-coefficient polynomials: product of 2^24 - 1 limbs, one
    transform size past 998244353's largest (2^23): four 2^23-point spectra of
    the operands' halves, three block sums inverted and accumulated;
  * Newton series at n = 2^20: inv (random a0), log (a0 = 1), exp (b0 = 0),
    sqrt (c = d^2 mod x^n, d0 = 1).
Inputs are resident in HBM before the timed region; `value` counts output
coefficients per second of the whole step.

Also reported on the same line:
  roofline        2^24-point NTT passes (column pass / row pass) timed with HIP
                  events on the library stream; algorithmic bytes 8N per pass
                  (read + write N uint32), 16N per transform; traffic from the
                  FETCH_SIZE / WRITE_SIZE counter runs summarised in profiles/;
  pcie_inclusive  the 2^23 conv through the host-buffer entry point;
  config1_conv_2p16  conv 2^16 x 2^16 mod 998244353 (device and host buffers);
  config3_mod_1e9p7  conv 2^22 x 2^22 and inv 2^20 mod 1000000007 (auxiliary primes);
  batched         the uneven batch (8192 convs + 1024 inv/exp), merged vs per call;
  serving         a Poisson request stream through the continuous-batching
                  server: throughput, p50 / p99 latency from arrival;
  cpu_baseline    the reference library (compiled from the reference sources into
                  oracle/_ref) timed on this host, one thread, on a bounded
                  sample of the step; its outputs are compared limb for limb.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-batch] [--no-serve]
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

P = 998244353
AUX = 167772161
N_CONV = 1 << 23
N_SERIES = 1 << 20
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--ntt-reps", type=int, default=100)
    ap.add_argument("--ntt-warmup", type=int, default=200,
                    help="untimed launches of each 2^24 pass before the timed ones (clock ramp)")
    ap.add_argument("--no-batch", action="store_true", help="skip the uneven-batch leg")
    ap.add_argument("--batch-convs", type=int, default=8192)
    ap.add_argument("--batch-series", type=int, default=1024)
    ap.add_argument("--batch-reps", type=int, default=5)
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg")
    ap.add_argument("--ntt-only", action="store_true",
                    help="only the 2^24 transform passes and a 64 MiB device copy (for counter runs)")
    ap.add_argument("--series-only", default=None, metavar="OPS",
                    help="profiling: only the Newton series chains at n = 2^20 (comma-separated ops), "
                         "each called --steps times after --warmup calls (graph replay)")
    ap.add_argument("--batch-only", action="store_true",
                    help="profiling: only the merged uneven batch, --batch-reps times after one warm-up")
    ap.add_argument("--serve-requests", type=int, default=1000)
    ap.add_argument("--serve-rate", type=float, default=1000.0, help="offered load, requests/s (Poisson)")
    ap.add_argument("--serve-levels", type=int, default=16)
    return ap.parse_args()


def montgomery_one(p):
    return (1 << 32) % p


def make_inputs(rng):
    a = rng.integers(0, P, N_CONV, dtype=np.uint32)
    b = rng.integers(0, P, N_CONV, dtype=np.uint32)
    s = rng.integers(0, P, N_SERIES, dtype=np.uint32)
    inv_in = s.copy()
    inv_in[0] = int(rng.integers(1, P))
    log_in = s.copy()
    log_in[0] = montgomery_one(P)
    exp_in = s.copy()
    exp_in[0] = 0
    d = rng.integers(0, P, N_SERIES, dtype=np.uint32)
    d[0] = montgomery_one(P)
    return a, b, inv_in, log_in, exp_in, d


def output_coeffs():
    return (2 * N_CONV - 1) + 4 * N_SERIES


def traffic_from_profiles():
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not offered "
              "(batches and serving streams would shard across GPUs by request, with no exchange)",
              file=sys.stderr)
        sys.exit(2)
    import fastpoly_amd as fa

    fa.init(0, 0)
    if args.ntt_only:
        ntt_only(fa, args)
        return
    if args.series_only:
        series_only(fa, args)
        return
    if args.batch_only:
        batch_only(fa, args)
        return
    rng = np.random.default_rng(20261203)
    roofline = roofline_leg(fa, args, np.random.default_rng(24))
    t_tune = time.perf_counter()
    tuned = fa.autotune(13, 24, 3)  # pass splits for single calls, untimed setup
    t_tune = time.perf_counter() - t_tune
    a, b, inv_in, log_in, exp_in, d = make_inputs(rng)
    D = fa.DeviceArray
    da, db = D.from_host(a), D.from_host(b)
    dinv, dlog, dexp = D.from_host(inv_in), D.from_host(log_in), D.from_host(exp_in)
    dd = D.from_host(d)
    dsq = D(N_SERIES)
    fa.conv_dev(P, dd, N_SERIES, dd, N_SERIES, dsq, N_SERIES)  # sqrt input, untimed setup
    out_conv = D(2 * N_CONV - 1)
    out_s = [D(N_SERIES) for _ in range(4)]
    fa.synchronize()

    ops = (("conv", lambda: fa.conv_dev(P, da, N_CONV, db, N_CONV, out_conv)),
           ("inv", lambda: fa.series_dev("inv", P, dinv, N_SERIES, N_SERIES, out_s[0])),
           ("log", lambda: fa.series_dev("log", P, dlog, N_SERIES, N_SERIES, out_s[1])),
           ("exp", lambda: fa.series_dev("exp", P, dexp, N_SERIES, N_SERIES, out_s[2])),
           ("sqrt", lambda: fa.series_dev("sqrt", P, dsq, N_SERIES, N_SERIES, out_s[3])))

    def step():
        for _, fn in ops:
            fn()

    for _ in range(args.warmup):
        step()
    fa.synchronize()
    fa.event_record(0)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    fa.synchronize()
    t1 = time.perf_counter()
    fa.event_record(1)
    wall = t1 - t0
    ms_per_step = 1e3 * wall / args.steps
    value = output_coeffs() * args.steps / wall
    # Per-operation split of the step (after the timed region; HIP events on
    # the library stream, 3 back-to-back calls of each operation).
    by_op = {}
    for name, fn in ops:
        fa.synchronize()
        fa.event_record(2)
        for _ in range(3):
            fn()
        fa.event_record(3)
        by_op[name] = round(fa.event_elapsed_ms(2, 3) / 3, 4)

    # Host-buffer form of the same conv (fpxg_conv_u32): includes the PCIe copies
    # of both operands and the product; reported beside `value`, never as it.
    fa.conv(P, a, b)
    t0 = time.perf_counter()
    host_prod = fa.conv(P, a, b)
    t_host = time.perf_counter() - t0
    pcie = {"conv_2p23_host_buffers_ms": round(1e3 * t_host, 3),
            "coeffs_per_s": (2 * N_CONV - 1) / t_host,
            "bytes_over_pcie": int(4 * (2 * N_CONV + 2 * N_CONV - 1))}

    gpu_out = {"conv": out_conv.to_host(), "inv": out_s[0].to_host(), "log": out_s[1].to_host(),
               "exp": out_s[2].to_host(), "sqrt": out_s[3].to_host()}
    sq_host = dsq.to_host()

    result = {
        "metric": "bit-exact coeffs/s: batched vs per-call GPU; 2^24 NTT HBM GB/s frac; p99 latency",
        "value": value, "unit": "coeffs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u32", "data": "synthetic (uniform [0,p) Montgomery limbs, numpy PCG64 seed 20261203)",
        "config": {"workload": "single large transforms mod 998244353: conv 2^23 x 2^23 (coset product of 2^23-point native NTTs) "
                               "+ inv/log/exp/sqrt at n = 2^20",
                   "modulus": P, "conv_len": N_CONV, "series_n": N_SERIES},
        "step_ms_by_op": by_op,
        "roofline": roofline,
        "cpu_baseline": None,
        "pcie_inclusive": pcie | {"same_as_device_path": bool(np.array_equal(host_prod, gpu_out["conv"]))},
        "autotune": {"seconds": round(t_tune, 3),
                     "col_levels_by_log2n": {str(k): v[0] for k, v in tuned.items()},
                     "product_col_levels_by_log2n": {str(k): fa.product_plan_info(1 << k)["col_levels"]
                                                     for k in tuned if k >= 13},
                     "forward_us_by_log2n": {str(k): round(v[1], 2) for k, v in tuned.items()}},
    }

    result["pow_divmod"] = pow_divmod_leg(fa, args)
    result["config1_conv_2p16"] = config1_leg(fa, args)
    result["config3_mod_1e9p7"] = config3_leg(fa, args)
    if not args.no_batch:
        result["batched"] = batched_leg(fa, args)
    if not args.no_serve:
        result["serving"] = serving_leg(fa, args)
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(args, a, b, inv_in, log_in, exp_in, sq_host, gpu_out)
    print(json.dumps(result), flush=True)


def roofline_leg(fa, args, rng):
    """The 2^24-point transform's two passes (column pass, row pass; HIP events
    on the library stream).  Run before the autotuner, so the split is the
    fixed default one (2^10 column levels in 2^10 x 16 tiles, rows of 2^14
    words) that the committed rocprofv3 summaries profile with --ntt-only."""
    D = fa.DeviceArray
    n = 1 << 24
    plan = fa.ntt_plan_info(n)
    x = D.from_host(rng.integers(0, AUX, n, dtype=np.uint32))
    # The clocks ramp up over the first tens of milliseconds of sustained
    # work: both passes run args.ntt_warmup times (untimed) first.
    for _ in range(args.ntt_warmup):
        fa.ntt_pass_dev(AUX, x, 0)
        fa.ntt_pass_dev(AUX, x, 1)
    fa.synchronize()
    passes = {}
    for pidx, name in ((0, "ntt_col_pass"), (1, "ntt_row_pass")):
        for _ in range(3):
            fa.ntt_pass_dev(AUX, x, pidx)
        fa.synchronize()
        fa.event_record(2)
        for _ in range(args.ntt_reps):
            fa.ntt_pass_dev(AUX, x, pidx)
        fa.event_record(3)
        ms = fa.event_elapsed_ms(2, 3) / args.ntt_reps
        passes[name] = ms
    alg_pass = 8 * n
    dominant = max(passes, key=passes.get)
    achieved = alg_pass / (passes[dominant] * 1e-3) / 1e9
    traffic = traffic_from_profiles().get(dominant)
    t_transform = sum(passes.values())
    roofline = {
        "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
        "kernel": dominant, "alg_bytes_per_launch": alg_pass, "avg_launch_ms": round(passes[dominant], 5),
        "ntt_2p24": {
            "plan": plan, "pass_ms": {k: round(v, 5) for k, v in passes.items()},
            "transform_ms": round(t_transform, 5),
            "achieved_GBps": round(16 * n / (t_transform * 1e-3) / 1e9, 1),
            "frac": round(16 * n / (t_transform * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        },
    }

    x.free()
    return roofline


def ntt_only(fa, args):
    """The two passes of the 2^24-point transform (args.ntt_reps each) and
    args.ntt_reps device-to-device copies of the same 64 MiB: the copy is the
    calibration kernel for the FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes
    n = 1 << 24
    rng = np.random.default_rng(1)
    x = fa.DeviceArray.from_host(rng.integers(0, AUX, n, dtype=np.uint32))
    y = fa.DeviceArray(n)
    for _ in range(args.ntt_warmup):
        fa.ntt_pass_dev(AUX, x, 0)
        fa.ntt_pass_dev(AUX, x, 1)
    fa.synchronize()
    for pidx in (0, 1):
        for _ in range(args.ntt_reps):
            fa.ntt_pass_dev(AUX, x, pidx)
        fa.synchronize()
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpyDtoD.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
    for _ in range(args.ntt_reps):
        if hip.hipMemcpyDtoD(y.ptr, x.ptr, 4 * n) != 0:
            raise RuntimeError("hipMemcpyDtoD failed")
    hip.hipDeviceSynchronize()
    print(json.dumps({"ntt_only": True, "n": n, "reps": args.ntt_reps, "copy_bytes": 4 * n}), flush=True)


def series_only(fa, args):
    """The step's Newton series chains alone (after the autotuner, inputs as
    in the step), each op called args.warmup + args.steps times on the same
    buffers: the second call captures its hipGraph, later calls replay it.
    The JSON line names the fused products' pass splits the autotuner chose
    (a run under a profiler may time, and so choose, differently)."""
    fa.autotune(13, 24, 3)
    plans = {lg: fa.product_plan_info(1 << lg)["col_levels"] for lg in range(13, 22)}
    _, _, inv_in, log_in, exp_in, d = make_inputs(np.random.default_rng(20261203))
    D = fa.DeviceArray
    ins = {"inv": D.from_host(inv_in), "log": D.from_host(log_in), "exp": D.from_host(exp_in)}
    dd = D.from_host(d)
    ins["sqrt"] = D(N_SERIES)
    fa.conv_dev(P, dd, N_SERIES, dd, N_SERIES, ins["sqrt"], N_SERIES)
    out = D(N_SERIES)
    res = {}
    for op in args.series_only.split(","):
        for _ in range(max(args.warmup, 2)):
            fa.series_dev(op, P, ins[op], N_SERIES, N_SERIES, out)
        fa.synchronize()
        fa.event_record(2)
        for _ in range(args.steps):
            fa.series_dev(op, P, ins[op], N_SERIES, N_SERIES, out)
        fa.event_record(3)
        fa.synchronize()
        res[op + "_ms"] = round(fa.event_elapsed_ms(2, 3) / args.steps, 4)
    print(json.dumps({"series_only": True, "n": N_SERIES, "steps": args.steps,
                      "product_col_levels_by_log2n": plans} | res), flush=True)


def batch_only(fa, args):
    """The merged uneven batch alone (one warm-up submit, then
    args.batch_reps timed submits)."""
    from fastpoly_amd.workloads import uneven_batch
    reqs = uneven_batch(args.batch_convs, args.batch_series)
    sin = [(r[2].size, r[3].size if r[0] == "conv" else 0) for r in reqs]
    sout = [r[2].size + r[3].size - 1 if r[0] == "conv" else r[3] for r in reqs]
    oi = np.concatenate([[0], np.cumsum([x + y for x, y in sin])])
    oo = np.concatenate([[0], np.cumsum(sout)])
    host = np.empty(int(oi[-1]), np.uint32)
    for i, r in enumerate(reqs):
        o = int(oi[i])
        host[o:o + sin[i][0]] = r[2]
        if r[0] == "conv":
            host[o + sin[i][0]:o + sin[i][0] + sin[i][1]] = r[3]
    din, dout = fa.DeviceArray.from_host(host), fa.DeviceArray(int(oo[-1]))
    arr = (fa.Req * len(reqs))()
    for i, r in enumerate(reqs):
        q = arr[i]
        q.op, q.p = fa.OPS[r[0]], r[1]
        q.a, q.la = din.ptr + 4 * int(oi[i]), sin[i][0]
        if r[0] == "conv":
            q.b, q.lb, q.n = din.ptr + 4 * (int(oi[i]) + sin[i][0]), sin[i][1], fa.SIZE_MAX
        else:
            q.n = r[3]
        q.out, q.out_cap = dout.ptr + 4 * int(oo[i]), sout[i]
    fa.wait(fa.batch_submit(arr))
    times = []
    for _ in range(args.batch_reps):
        t0 = time.perf_counter()
        fa.wait(fa.batch_submit(arr))
        times.append(round(time.perf_counter() - t0, 5))
    print(json.dumps({"batch_only": True, "requests": len(reqs), "batched_s": times}), flush=True)


def _dev_time(fa, fn, reps):
    """Mean seconds per call of fn (device-resident inputs), HIP events on the
    library stream, after one untimed call."""
    fn()
    fa.synchronize()
    fa.event_record(2)
    for _ in range(reps):
        fn()
    fa.event_record(3)
    fa.synchronize()
    return fa.event_elapsed_ms(2, 3) / 1e3 / reps


def _uniform(rng, p, n):
    return rng.integers(0, p, n, dtype=np.uint32)


def pow_divmod_leg(fa, args):
    """poly::pow(a, 1000003, 2^20) (a0 = 1) and poly::divmod of 2^20 by 2^19
    limbs mod 998244353 on uniform inputs in HBM."""
    rng = np.random.default_rng(31)
    n = N_SERIES
    a = _uniform(rng, P, n)
    a[0] = montgomery_one(P)
    x, y = _uniform(rng, P, n), _uniform(rng, P, n // 2)
    D = fa.DeviceArray
    da, dx, dy = D.from_host(a), D.from_host(x), D.from_host(y)
    out, q, r = D(n), D(n), D(n)
    t_pow = _dev_time(fa, lambda: fa.pow_dev(P, da, n, 1000003, n, out), 5)
    lens = []
    t_div = _dev_time(fa, lambda: lens.append(fa.divmod_dev(P, dx, n, dy, n // 2, q, r)), 5)
    lq, lr = lens[-1]
    res = {"workload": "pow(a, 1000003) at n = 2^20 and divmod of 2^20 by 2^19 limbs, mod 998244353",
           "pow_ms": round(1e3 * t_pow, 3), "divmod_ms": round(1e3 * t_div, 3),
           "pow_coeffs_per_s": n / t_pow, "divmod_coeffs_per_s": (lq + lr) / t_div}
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline_pow_divmod(a, x, y, out.to_host(), q.to_host(lq), r.to_host(lr))
    for d in (da, dx, dy, out, q, r):
        d.free()
    return res


def config1_leg(fa, args):
    """Plain truncated convolution mod 998244353 of two 2^16-coefficient
    polynomials (full product 2^17 - 1 limbs): inputs in HBM (fpxg_conv_dev),
    and through the host-buffer entry point (PCIe-inclusive)."""
    rng = np.random.default_rng(41)
    n = 1 << 16
    a, b = _uniform(rng, P, n), _uniform(rng, P, n)
    D = fa.DeviceArray
    da, db, do = D.from_host(a), D.from_host(b), D(2 * n - 1)
    t_dev = _dev_time(fa, lambda: fa.conv_dev(P, da, n, db, n, do), 200)
    fa.conv(P, a, b)
    t0 = time.perf_counter()
    for _ in range(20):
        host = fa.conv(P, a, b)
    t_host = (time.perf_counter() - t0) / 20
    gpu = do.to_host()
    res = {"workload": "conv 2^16 x 2^16 mod 998244353 (full product, 2^17 - 1 limbs)",
           "ms": round(1e3 * t_dev, 4), "coeffs_per_s": (2 * n - 1) / t_dev,
           "host_buffers_ms": round(1e3 * t_host, 4), "host_same_as_device": bool(np.array_equal(host, gpu)),
           "published_ms": {"neon": 0.7005, "avx2_rosetta": 5.2948}}
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline_ops([("conv", P, a, b)], [gpu], 20)
    for d in (da, db, do):
        d.free()
    return res


def config3_leg(fa, args):
    """Non-NTT-friendly prime 1000000007: conv of two 2^22-coefficient
    polynomials through the auxiliary primes 167772161, 469762049, 754974721
    with Garner back to 1e9+7, and inv at n = 2^20 with a0 uniform in [1, p);
    inputs in HBM."""
    Q = 1000000007
    rng = np.random.default_rng(43)
    n, m = 1 << 22, 1 << 20
    a, b = _uniform(rng, Q, n), _uniform(rng, Q, n)
    s = _uniform(rng, Q, m)
    s[0] = int(fa.from_int(Q, int(rng.integers(1, Q))))
    D = fa.DeviceArray
    da, db, dc, ds, di = D.from_host(a), D.from_host(b), D(2 * n - 1), D.from_host(s), D(m)
    t_conv = _dev_time(fa, lambda: fa.conv_dev(Q, da, n, db, n, dc), 10)
    t_inv = _dev_time(fa, lambda: fa.series_dev("inv", Q, ds, m, m, di), 10)
    gc, gi = dc.to_host(), di.to_host()
    res = {"workload": "mod 1000000007: conv 2^22 x 2^22 (three auxiliary primes, Garner) and inv at n = 2^20 "
                       "(a0 uniform in [1, p))",
           "conv_ms": round(1e3 * t_conv, 3), "inv_ms": round(1e3 * t_inv, 3),
           "conv_coeffs_per_s": (2 * n - 1) / t_conv, "inv_coeffs_per_s": m / t_inv,
           "published_ms_2p16": {"conv_neon": 2.4671, "inv_neon": 4.7863}}
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline_ops([("conv", Q, a, b), ("inv", Q, s, m)], [gc, gi], 1)
    for d in (da, db, dc, ds, di):
        d.free()
    return res


def serving_leg(fa, args):
    """Poisson stream of conv 50% / inv 20% / log 10% / exp 10% / sqrt 10%
    requests, lengths log-uniform in [2^10, 2^20], moduli 998244353 (70%) /
    1000000007 (30%), host inputs, submitted at their arrival times to the
    continuous-batching server by four client threads (request i by thread
    i mod 4, so a large input being copied in does not hold back the next
    arrival).  Latency = completion - arrival on the server's clock (client
    queueing, the input copy, upload, compute and download); the same from
    the submit call is reported beside it."""
    import threading
    from fastpoly_amd.workloads import serving_stream

    stream = serving_stream(args.serve_requests, args.serve_rate)
    reqs = []
    for _, t in stream:
        reqs.append(fa.Request("conv", t[1], t[2], b=t[3]) if t[0] == "conv" else fa.Request(t[0], t[1], t[2], n=t[3]))
    out_coeffs = sum(r.out_capacity() for r in reqs)
    feeders = 4
    with fa.Server(levels_per_iteration=args.serve_levels) as srv:
        # Warm-up: one request of each kind (fields, plans, pinned blocks).
        for r in reqs[:16]:
            srv.wait(srv.submit(r))
        t0 = srv.now() + 0.01
        ids, lag, due = [None] * len(reqs), [0.0] * len(reqs), [t0 + arr for arr, _ in stream]

        def feed(k):
            for i in range(k, len(reqs), feeders):
                while True:
                    now = srv.now()
                    if now >= due[i]:
                        break
                    if due[i] - now > 0.0005:
                        time.sleep(due[i] - now - 0.0003)
                ids[i] = srv.submit(reqs[i])
                lag[i] = srv.now() - due[i]

        threads = [threading.Thread(target=feed, args=(k,)) for k in range(feeders)]
        for th in threads:
            th.start()
        for th in threads:
            th.join()
        res = [srv.wait(i) for i in ids]
        st = srv.stats()
    lat_submit = np.array([t["latency_s"] for _, t in res]) * 1e3
    lat = np.array([t["done_s"] - d for (_, t), d in zip(res, due)]) * 1e3
    first = min(t["submit_s"] for _, t in res)
    last = max(t["done_s"] for _, t in res)
    errors = sum(isinstance(g, Exception) for g, _ in res)
    # Each result against the same request run alone (single call, GPU).
    exact = True
    for r, (g, _) in zip(reqs, res):
        want = fa.conv(r.p, r.a, r.b) if r.op == "conv" else fa.series(r.op, r.p, r.a, r.n)
        if isinstance(g, Exception) or not np.array_equal(g, want):
            exact = False
            break
    span = last - first
    cpu = None
    if not args.no_cpu:
        tuples = [("conv", r.p, r.a, r.b) if r.op == "conv" else (r.op, r.p, r.a, r.n) for r in reqs]
        cpu = cpu_parallel_baseline(tuples, [g for g, _ in res], args.cpu_seconds)
    return {
        "cpu_baseline": cpu,
        "workload": "Poisson stream: conv 50% / inv 20% / log 10% / exp 10% / sqrt 10%, lengths log-uniform "
                    "[2^10,2^20], moduli 998244353 70% / 1000000007 30%, host inputs",
        "requests": len(reqs), "offered_rate_per_s": args.serve_rate,
        "achieved_rate_per_s": round(len(reqs) / span, 1), "output_coeffs": int(out_coeffs),
        "coeffs_per_s": out_coeffs / span, "p50_ms": round(float(np.percentile(lat, 50)), 3),
        "p99_ms": round(float(np.percentile(lat, 99)), 3), "mean_ms": round(float(lat.mean()), 3),
        "max_ms": round(float(lat.max()), 3), "latency_from": "arrival (Poisson due time)",
        "from_submit_p50_ms": round(float(np.percentile(lat_submit, 50)), 3),
        "from_submit_p99_ms": round(float(np.percentile(lat_submit, 99)), 3), "client_threads": feeders,
        "submit_lag_p99_ms": round(float(np.percentile(lag, 99)) * 1e3, 3),
        "errors": errors, "identical_to_single_calls": exact, "levels_per_iteration": args.serve_levels,
        "server": st,
    }


def batched_leg(fa, args):
    """Uneven batch: 8192 convs (half mod 998244353, half mod 1e9+7; operand
    lengths log-uniform in [2^6, 2^18]) + 1024 inv/exp at n log-uniform in
    [2^8, 2^16].  Inputs resident in HBM; the same requests run once per call
    (fpxg_conv_dev / fpxg_series_dev in a loop) and as one merged batch
    (fpxg_batch_submit + fpxg_wait); outputs of the two are compared limb for limb."""
    import ctypes
    from fastpoly_amd.workloads import uneven_batch

    t_gen = time.perf_counter()
    reqs = uneven_batch(args.batch_convs, args.batch_series)
    sizes_in, sizes_out = [], []
    for r in reqs:
        if r[0] == "conv":
            sizes_in.append((r[2].size, r[3].size))
            sizes_out.append(r[2].size + r[3].size - 1)
        else:
            sizes_in.append((r[2].size, 0))
            sizes_out.append(r[3])
    tot_in = sum(x + y for x, y in sizes_in)
    tot_out = sum(sizes_out)
    host_in = np.empty(tot_in, np.uint32)
    offs_a, offs_b, offs_o = [], [], []
    o = 0
    for r, (x, y) in zip(reqs, sizes_in):
        offs_a.append(o); host_in[o:o + x] = r[2]; o += x
        offs_b.append(o)
        if y:
            host_in[o:o + y] = r[3]; o += y
    oo = 0
    for n in sizes_out:
        offs_o.append(oo); oo += n
    D = fa.DeviceArray
    din = D.from_host(host_in)
    dout_call, dout_batch = D(tot_out), D(tot_out)
    gen_s = time.perf_counter() - t_gen
    L = fa.lib()
    k = len(reqs)

    def make_reqs(dout):
        arr = (fa.Req * k)()
        for i, r in enumerate(reqs):
            q = arr[i]
            q.op, q.p = fa.OPS[r[0]], r[1]
            q.a, q.la = din.ptr + 4 * offs_a[i], sizes_in[i][0]
            if r[0] == "conv":
                q.b, q.lb, q.n = din.ptr + 4 * offs_b[i], sizes_in[i][1], fa.SIZE_MAX
            else:
                q.n = r[3]
            q.out, q.out_cap = dout.ptr + 4 * offs_o[i], sizes_out[i]
        return arr

    arr_b = make_reqs(dout_batch)
    arr_c = make_reqs(dout_call)
    got = ctypes.c_size_t(0)

    def per_call():
        for i in range(k):
            q = arr_c[i]
            if q.op == 4:
                rc = L.fpxg_conv_dev(q.p, q.a, q.la, q.b, q.lb, q.n, q.out, q.out_cap, ctypes.byref(got))
            else:
                rc = L.fpxg_series_dev(q.op, q.p, q.a, q.la, q.n, q.out)
            if rc:
                raise RuntimeError(f"request {i} failed: {rc}")
        fa.synchronize()

    def batched():
        fa.wait(fa.batch_submit(arr_b))

    # Per call = one launch sequence per request, as a one-off call runs (no
    # graph replay: 9216 distinct requests would only churn the graph cache).
    fa.set_graphs(False)
    per_call()  # warm-up: fields, plans, arena
    batched()
    t0 = time.perf_counter()
    per_call()
    t_call = time.perf_counter() - t0
    fa.set_graphs(True)
    times = []
    for _ in range(max(1, args.batch_reps)):
        t0 = time.perf_counter()
        batched()
        times.append(time.perf_counter() - t0)
    t_batch = min(times)
    st = [arr_b[i].status for i in range(k)]
    host_batch = dout_batch.to_host()
    exact = bool(np.array_equal(dout_call.to_host(), host_batch)) and not any(st)
    for d in (din, dout_call, dout_batch):
        d.free()
    cpu = None
    if not args.no_cpu:
        outs = [host_batch[offs_o[i]:offs_o[i] + sizes_out[i]] for i in range(k)]
        cpu = cpu_parallel_baseline(reqs, outs, args.cpu_seconds)
    return {
        "workload": f"{args.batch_convs} convs (half mod 998244353, half mod 1000000007; lengths log-uniform "
                    f"[2^6,2^18]) + {args.batch_series} inv/exp (n log-uniform [2^8,2^16]), inputs in HBM",
        "requests": k, "input_coeffs": int(tot_in), "output_coeffs": int(tot_out),
        "batched_s": round(t_batch, 5), "per_call_s": round(t_call, 5),
        "batched_coeffs_per_s": tot_out / t_batch, "per_call_coeffs_per_s": tot_out / t_call,
        "speedup_vs_per_call": round(t_call / t_batch, 3), "bit_exact_vs_per_call": exact,
        "batched_reps_s": [round(x, 5) for x in times], "setup_s": round(gen_s, 2),
        "cpu_baseline": cpu,
    }


def cpu_parallel_baseline(reqs, gpu_outs, seconds, max_threads=16):
    """The reference library on this host's cores (one request per thread at a
    time, up to 16 threads: the CPU share of one GPU), over a shuffled sample of
    the same requests for about `seconds`; every CPU result is compared with the
    GPU's.  Reports output limbs/s and the single-request CPU latency."""
    import threading
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po

    kind = "reference" if po.ref() is not None else "port"
    conv = po.ref_conv if kind == "reference" else po.conv
    series = po.ref_series if kind == "reference" else po.series
    threads = max(1, min(max_threads, os.cpu_count() or 1))
    order = list(np.random.default_rng(7).permutation(len(reqs)))
    lock = threading.Lock()
    it = iter(order)
    done = []
    t_start = time.perf_counter()
    deadline = t_start + seconds

    def worker():
        while time.perf_counter() < deadline:
            with lock:
                i = next(it, None)
            if i is None:
                return
            r = reqs[i]
            t0 = time.perf_counter()
            out = conv(r[1], r[2], r[3]) if r[0] == "conv" else series(r[0], r[1], r[2], r[3])
            dt = time.perf_counter() - t0
            g = gpu_outs[i]
            ok = not isinstance(g, Exception) and np.array_equal(out, g)
            with lock:
                done.append((out.size, dt, ok))

    ts = [threading.Thread(target=worker) for _ in range(threads)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    el = time.perf_counter() - t_start
    lat = np.array([d[1] for d in done]) * 1e3 if done else np.zeros(1)
    return {"value": sum(d[0] for d in done) / el, "unit": "coeffs/s", "cores": threads, "kind": kind,
            "sample": f"{len(done)} of {len(reqs)} requests (shuffled), {el:.1f} s on {threads} threads",
            "bit_exact_vs_gpu": bool(done) and all(d[2] for d in done),
            "request_ms_p50": round(float(np.percentile(lat, 50)), 3),
            "request_ms_p99": round(float(np.percentile(lat, 99)), 3)}


def _ref_lib():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    return po, po.ref() is not None


def cpu_baseline_ops(ops, gpu_outs, reps):
    """The reference library (oracle/_ref, one thread) on the same requests:
    best of `reps` runs of each, outputs compared limb for limb with the GPU's."""
    po, have_ref = _ref_lib()
    if not have_ref:
        return None
    times, exact = [], True
    for op, g in zip(ops, gpu_outs):
        best = 1e30
        for _ in range(reps):
            t0 = time.perf_counter()
            out = po.ref_conv(op[1], op[2], op[3]) if op[0] == "conv" else po.ref_series(op[0], op[1], op[2], op[3])
            best = min(best, time.perf_counter() - t0)
        times.append(round(1e3 * best, 3))
        exact = exact and bool(np.array_equal(out, g))
    return {"ms": times, "cores": 1, "kind": "reference", "backend": po.ref().fpref_backend().decode(),
            "sample": f"best of {reps} run(s) of each operation", "bit_exact_vs_gpu": exact}


def cpu_baseline_pow_divmod(a, x, y, g_pow, g_q, g_r):
    po, have_ref = _ref_lib()
    if not have_ref or not hasattr(po.ref(), "fpref_pow"):
        return None
    n = a.size
    t0 = time.perf_counter()
    cp = po.ref_pow(P, a, 1000003, n)
    t1 = time.perf_counter()
    cq, cr = po.ref_divmod(P, x, y)
    t2 = time.perf_counter()
    return {"pow_ms": round(1e3 * (t1 - t0), 1), "divmod_ms": round(1e3 * (t2 - t1), 1), "cores": 1,
            "kind": "reference", "bit_exact_vs_gpu": bool(np.array_equal(cp, g_pow) and np.array_equal(cq, g_q)
                                                            and np.array_equal(cr, g_r))}


def cpu_baseline(args, a, b, inv_in, log_in, exp_in, sq_host, gpu_out):
    """The reference library on this host (1 thread), same step, same inputs."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po

    kind = "reference" if po.ref() is not None else "port"
    conv = po.ref_conv if kind == "reference" else po.conv
    series = po.ref_series if kind == "reference" else po.series
    outs = {}
    t0 = time.perf_counter()
    reps = 0
    while True:
        outs["conv"] = conv(P, a, b)
        outs["inv"] = series("inv", P, inv_in, N_SERIES)
        outs["log"] = series("log", P, log_in, N_SERIES)
        outs["exp"] = series("exp", P, exp_in, N_SERIES)
        outs["sqrt"] = series("sqrt", P, sq_host, N_SERIES)
        reps += 1
        el = time.perf_counter() - t0
        if el >= args.cpu_seconds or reps >= 20:
            break
    exact = all(np.array_equal(outs[k], gpu_out[k]) for k in outs)
    cpu = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    cpu = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    backend = po.ref().fpref_backend().decode() if kind == "reference" else "scalar C"
    return {"value": output_coeffs() * reps / el, "unit": "coeffs/s", "cores": 1, "kind": kind,
            "sample": f"{reps} full steps (conv 2^23x2^23 + inv/log/exp/sqrt n=2^20), {el:.1f} s, "
                      f"backend {backend}, {cpu}",
            "bit_exact_vs_gpu": bool(exact), "ms_per_step": 1e3 * el / reps}


if __name__ == "__main__":
    main()
