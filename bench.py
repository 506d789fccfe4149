#!/usr/bin/env python3
"""Batch decode benchmark on one MI355X.

One step = one pass of the batch decoder over the whole synthetic batch,
which is device-resident (compressed arena, pointer / size tables, output
arena) before the timed region starts.

Default workload: the uneven LZ4 batch -- 20000 chunks, sizes log-uniform
over [1 KiB, 4 MiB), 80% token-process / 10% uniform-random / 10%
repetitive chunks, shuffled, compressed with LZ4_compress_default. The
tuned form (cudec_lz4_decompress_batch) is the measured value; the plain
one-wave-per-chunk form (cudec_lz4_decompress_batch_plain) is timed on the
same batch for the "vs plain" ratio. Every chunk of every form is verified
byte-exact against its original (and status OK, bytes_written == size)
before timing.

Output: one JSON line (see the keys at the bottom of main()).
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0  # MI355X published HBM3E bandwidth, GB/s


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse_args():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--workload", choices=["uneven", "token64k"], default="uneven")
    p.add_argument("--format", choices=["lz4", "snappy"], default="lz4")
    p.add_argument("--chunks", type=int, default=None)
    p.add_argument("--seed", type=int, default=0x5EED)
    p.add_argument("--forms", default="tuned,plain")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--cpu-threads", type=int, default=None)
    p.add_argument("--traffic-file", default=None,
                   help="JSON with measured HBM bytes per launch (rocprofv3 PMC)")
    p.add_argument("--malformed-pct", type=float, default=None,
                   help="percent of chunks made malformed by a seeded truncation or a "
                        "single-byte flip after the preamble (default: 2 for snappy, 0 for lz4)")
    p.add_argument("--copy-probe-mib", type=int, default=0,
                   help="before timing, run a plain elementwise copy (torch.neg over int32) "
                        "of this many MiB three times: the known-size kernel that HBM "
                        "counters are scaled by")
    p.add_argument("--no-graph", action="store_true",
                   help="skip the HIP-graph-captured replay of the tuned launch")
    p.add_argument("--no-autotune", action="store_true",
                   help="keep the default LZ4 pipelining threshold instead of tuning it "
                        "on the batch before timing")
    p.add_argument("--threshold", type=int, default=None,
                   help="set the LZ4 pipelining threshold (bytes of cost; -1: none) instead of "
                        "tuning it, so a profiled run repeats one launch shape")
    p.add_argument("--serve-rates", default="2000,8000,32000",
                   help="request rates (per second) of the serving leg; empty to skip it")
    p.add_argument("--serve-seconds", type=float, default=0.5,
                   help="scheduled arrival span of each serving run")
    p.add_argument("--serve-pool", type=int, default=2048,
                   help="distinct chunks the serving requests draw from")
    p.add_argument("--serve-budget-gib", type=float, default=64.0)
    return p.parse_args()


def build_batch(args):
    from cudec_amd import workload
    codec = workload.CODEC_LZ4 if args.format == "lz4" else workload.CODEC_SNAPPY
    t0 = time.perf_counter()
    if args.workload == "uneven":
        n = args.chunks or 20000
        b = workload.uneven(n, seed=args.seed, codec=codec)
        name = f"uneven-{args.format}: {n} chunks log-uniform [1 KiB, 4 MiB), 80/10/10 token/random/repetitive"
    else:
        n = args.chunks or 65536
        b = workload.token64k(n, seed=args.seed, codec=codec)
        name = f"token64k-{args.format}: {n} x 64 KiB token-process chunks"
    log(f"generated + compressed {n} chunks in {time.perf_counter() - t0:.1f}s: "
        f"{b.total_uncompressed / 2**30:.2f} GiB -> {b.total_compressed / 2**30:.2f} GiB "
        f"({workload.codec_version(codec)})")
    pct = args.malformed_pct if args.malformed_pct is not None else \
        (2.0 if args.format == "snappy" else 0.0)
    malformed = malform(b, pct, args.seed)
    if len(malformed):
        name += f", {pct:g}% malformed ({len(malformed)} chunks: truncation or byte flip)"
    return b, name, malformed


def malform(b, pct, seed):
    """Makes pct percent of the chunks malformed in place, seeded: half lose
    1-16 trailing bytes, half get one byte after the first 5 XOR-ed with a
    nonzero value. Returns the sorted indices. A flip may leave a chunk
    valid; the decoders are then held to the CPU reference's verdict."""
    import numpy as np
    k = int(round(b.n * pct / 100.0))
    if k == 0:
        return np.zeros(0, dtype=np.int64)
    rng = np.random.default_rng(seed ^ 0xBAD)
    idx = np.sort(rng.choice(b.n, size=k, replace=False))
    for i in idx:
        size = int(b.comp_size[i])
        if size < 2:
            continue
        if rng.integers(0, 2) == 0:
            b.comp_size[i] = size - int(rng.integers(1, min(16, size - 1) + 1))
        else:
            at = int(b.comp_off[i]) + int(rng.integers(min(5, size - 1), size))
            b.comp[at] ^= np.uint8(rng.integers(1, 256))
    return idx


def verify(db, form, orig_dev, batch, valid):
    """Every intact chunk: status OK, bytes_written = size, output = original.
    Returns the results read back."""
    import numpy as np
    import torch
    res = db.results().copy()
    ok = bool(np.all(res["status"][valid] == 0)) and \
        bool(np.array_equal(res["bytes_written"][valid], batch.sizes[valid])) and \
        bool(np.all(res["reserved"] == 0))
    if bool(np.all(valid)):
        same = bool(torch.equal(db.d_out[: orig_dev.numel()], orig_dev))
    else:
        # malformed chunks may leave anything in their slots: take those
        # slots from the decoder's output and compare everything else
        expect = orig_dev.clone()
        for i in np.nonzero(~valid)[0]:
            o, n = int(batch.orig_off[i]), int(batch.sizes[i])
            expect[o:o + n] = db.d_out[o:o + n]
        same = bool(torch.equal(db.d_out[: orig_dev.numel()], expect))
        del expect
    if not (ok and same):
        bad = int(np.sum(res["status"][valid] != 0))
        raise SystemExit(f"[bench] {form}: verification FAILED (intact non-OK chunks: {bad}, "
                         f"arena equal: {same})")
    return res


def time_form(db, fmt, form, steps, warmup, stream, orig_dev, batch, valid):
    """Warmup, a byte-exact check of every intact chunk, then `steps`
    event-timed launches. Returns (wall seconds, per-launch ms, the checked
    results)."""
    import torch
    for _ in range(warmup):
        st = db.run(fmt, form, stream=stream)
        assert st == 0, f"{form} submission failed: {st}"
    torch.cuda.synchronize()
    res = verify(db, form, orig_dev, batch, valid)
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        starts[k].record(stream)
        st = db.run(fmt, form, stream=stream)
        ends[k].record(stream)
        assert st == 0
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = [s.elapsed_time(e) for s, e in zip(starts, ends)]
    return wall, kernel_ms, res


def graph_leg(db, fmt, steps, stream, orig_dev, batch, valid, eager):
    """The tuned launch captured once into a HIP graph (cudec_amd/graph.py)
    and replayed `steps` times, event-timed; the output is cleared before
    the replays and checked after them, and the results must equal the
    eager launch's."""
    import numpy as np
    import torch
    from cudec_amd.graph import CapturedLaunch
    with CapturedLaunch(lambda: db.run(fmt, "tuned", stream=stream), stream) as g:
        with torch.cuda.stream(stream):
            db.d_out.zero_()
            db.reset_results()
        starts = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
        ends = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
        for k in range(steps):
            starts[k].record(stream)
            g.replay()
            ends[k].record(stream)
        torch.cuda.synchronize()
    res = verify(db, "graph", orig_dev, batch, valid)
    if not np.array_equal(res, eager):
        raise SystemExit("[bench] graph replay results differ from the eager launch")
    ms = [s.elapsed_time(e) for s, e in zip(starts, ends)]
    return {"replays": steps, "avg_kernel_ms": round(sum(ms) / len(ms), 4), "verified": True}


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(batch, fmt, threads, valid, runs=30):
    """The reference ladder compiled from its own source (oracle/_ref) when
    that build is present, else the C oracle: `threads` host threads, one
    pass over the whole batch into pre-sized buffers, decode calls only.
    Returns the baseline record and every chunk's (status, bytes_written),
    which the GPU results are held to."""
    import numpy as np
    ref_path = os.path.join(ROOT, "oracle", "_ref", "libcudec_ref.so")
    port_path = os.path.join(ROOT, "oracle", "liboracle.so")
    u64 = ctypes.c_uint64
    if os.path.exists(ref_path):
        lib = ctypes.CDLL(ref_path)
        fn = lib.ref_lz4_decode if fmt == "lz4" else lib.ref_snappy_decode
        kind = "reference"
    else:
        lib = ctypes.CDLL(port_path)
        fn = lib.oracle_lz4_decode if fmt == "lz4" else lib.oracle_snappy_decode
        kind = "port"
    extra = [] if fmt == "lz4" else [ctypes.c_void_p]
    fn.argtypes = [ctypes.c_void_p, u64, ctypes.c_void_p, u64, ctypes.POINTER(u64)] + extra
    fn.restype = ctypes.c_int32
    out = np.empty(batch.total_uncompressed, dtype=np.uint8)
    status = np.zeros(batch.n, dtype=np.int32)
    written = np.zeros(batch.n, dtype=np.uint64)
    cbase = batch.comp.ctypes.data
    obase = out.ctypes.data
    n = batch.n
    bounds = [n * t // threads for t in range(threads + 1)]

    def work(t):
        w = u64(0)
        for i in range(bounds[t], bounds[t + 1]):
            args = [cbase + int(batch.comp_off[i]), int(batch.comp_size[i]),
                    obase + int(batch.orig_off[i]), int(batch.sizes[i]), ctypes.byref(w)]
            if extra:
                args.append(None)
            status[i] = fn(*args)
            written[i] = w.value if status[i] == 0 else 0
        return 0

    with cf.ThreadPoolExecutor(max_workers=threads) as ex:
        t0 = time.perf_counter()
        list(ex.map(work, range(threads)))
        first = time.perf_counter() - t0
        intact_ok = bool(np.all(status[valid] == 0)) and \
            bool(np.array_equal(written[valid], batch.sizes[valid]))
        if intact_ok and bool(np.all(valid)):
            intact_ok = bool(np.array_equal(out, batch.orig))
        elif intact_ok:
            for i in np.nonzero(valid)[0]:
                o, m = int(batch.orig_off[i]), int(batch.sizes[i])
                if not np.array_equal(out[o:o + m], batch.orig[o:o + m]):
                    intact_ok = False
                    break
        if not intact_ok:
            raise SystemExit("[bench] CPU baseline produced wrong output")
        # the checked pass is the warmup; then timed passes, bounded to about
        # 10 s of wall time
        runs = max(1, min(runs, int(10.0 / max(first, 1e-3))))
        secs = []
        for _ in range(runs):
            t0 = time.perf_counter()
            list(ex.map(work, range(threads)))
            secs.append(time.perf_counter() - t0)
    from cudec_amd.serve import percentile

    def rate(t, nbytes):
        return round(nbytes / t / 1e9, 3)
    # one thread over a sample: the batch's first chunks (the batch order is
    # shuffled) up to 256 MiB of output, 3 warmup + 30 timed passes
    m = int(np.searchsorted(np.cumsum(batch.sizes.astype(np.int64)), 256 << 20, side="right"))
    m = max(1, min(m, n))
    sample_bytes = int(batch.sizes[:m].sum())
    single_bounds = (0, m)

    def single():
        w = u64(0)
        for i in range(*single_bounds):
            args = [cbase + int(batch.comp_off[i]), int(batch.comp_size[i]),
                    obase + int(batch.orig_off[i]), int(batch.sizes[i]), ctypes.byref(w)]
            if extra:
                args.append(None)
            fn(*args)

    one = []
    for k in range(33):
        t0 = time.perf_counter()
        single()
        if k >= 3:
            one.append(time.perf_counter() - t0)
    total = batch.total_uncompressed
    # rates at the nearest-rank p50 / p90 / p99 pass times
    rec = {"value": rate(percentile(secs, 50), total), "unit": "GB/s (uncompressed)",
           "cores": threads, "kind": kind,
           "sample": f"the full batch ({n} chunks, {batch.total_uncompressed / 2**30:.2f} GiB "
                     f"out), {threads} threads, static contiguous partition; 1 checked warmup "
                     f"pass + {len(secs)} timed passes, nearest-rank p50 "
                     f"({min(secs):.2f}-{max(secs):.2f} s wall per pass)",
           "at_p90_time": rate(percentile(secs, 90), total),
           "at_p99_time": rate(percentile(secs, 99), total),
           "single_thread": {"value": rate(percentile(one, 50), sample_bytes),
                             "at_p90_time": rate(percentile(one, 90), sample_bytes),
                             "sample": f"first {m} chunks ({sample_bytes / 2**20:.0f} MiB out), "
                                       "1 thread, 3 warmup + 30 timed passes, nearest-rank p50"},
           "cpu_model": cpu_model()}
    return rec, status, written


TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "lz4_uneven_hbm_traffic.md")
KERNEL_SOURCE = os.path.join(ROOT, "cudec_amd", "csrc", "tuned.hip")


def committed_traffic():
    """The HBM bytes per launch recorded in the committed PMC summary for the
    default uneven LZ4 batch, used only while the kernel source is the one
    that summary measured (its sha256 is recorded beside the figure)."""
    import hashlib
    import re
    if not os.path.exists(TRAFFIC_SUMMARY):
        return None
    text = open(TRAFFIC_SUMMARY).read()
    src = re.search(r"tuned\.hip sha256: ([0-9a-f]{64})", text)
    val = re.search(r"bytes per launch: (\d+)", text)
    if not (src and val):
        return None
    with open(KERNEL_SOURCE, "rb") as f:
        if hashlib.sha256(f.read()).hexdigest() != src.group(1):
            return None
    return int(val.group(1))


def serving_leg(args):
    """Poisson request streams through the native serving loop. Per rate:
    one run with device digests, checked chunk by chunk against the
    originals, then the timed run (no digests) whose latencies are
    reported: nearest-rank p50/p99 from scheduled arrival to completion."""
    import numpy as np
    import torch

    from cudec_amd import serve
    rates = [float(r) for r in args.serve_rates.split(",") if r]
    t0 = time.perf_counter()
    pool = serve.chunk_pool(args.serve_pool, seed=args.seed ^ 0x5E5E)
    want = np.array([serve.host_digest(pool.orig[int(o):int(o) + int(n)])
                     for o, n in zip(pool.orig_off, pool.sizes)], dtype=np.uint64)
    d_comp = torch.from_numpy(pool.comp.copy()).to("cuda")
    torch.cuda.synchronize()
    log(f"serving pool: {pool.n} chunks, {pool.total_uncompressed / 2**20:.0f} MiB out, "
        f"built in {time.perf_counter() - t0:.1f}s")
    budget = int(args.serve_budget_gib * 2**30)
    out = []
    for k, rate in enumerate(rates):
        nreq = max(16, int(rate * args.serve_seconds))
        req, ids = serve.make_trace(rate, nreq, pool.n, seed=args.seed + k)
        check = serve.serve(d_comp.data_ptr(), pool, req, ids, budget=budget, inflight=4,
                            digests=True)
        ok = check.status == 0 and bool(np.all(check.results["status"] == 0)) and \
            bool(np.array_equal(check.results["bytes_written"], pool.sizes[ids])) and \
            bool(np.array_equal(check.digests, want[ids]))
        if not ok:
            raise SystemExit(f"[bench] serving at {rate:.0f}/s: verification FAILED")
        run = serve.serve(d_comp.data_ptr(), pool, req, ids, budget=budget, inflight=4,
                          digests=False)
        if run.status != 0 or not np.array_equal(run.results, check.results):
            raise SystemExit(f"[bench] serving at {rate:.0f}/s: timed run differs")
        st = run.stats
        algo = st.output_bytes + st.compressed_bytes
        row = {"rate_per_s": rate, "requests": nreq, "chunks": int(len(ids)),
               "p50_ms": round(1e3 * serve.percentile(run.latency_s, 50), 3),
               "p99_ms": round(1e3 * serve.percentile(run.latency_s, 99), 3),
               "gbps_uncompressed": round(st.output_bytes / st.elapsed_s / 1e9, 2),
               "gbps_algorithmic": round(algo / st.elapsed_s / 1e9, 2),
               "offered_gbps_uncompressed": round(st.output_bytes / max(
                   float(req["arrival_s"][-1]), 1e-9) / 1e9, 2),
               "launches": int(st.launches), "max_batch_requests": int(st.max_batch_requests),
               "max_inflight": int(st.max_inflight),
               "peak_pool_mib": round(st.peak_pool_bytes / 2**20, 1)}
        log(f"serving {rate:.0f}/s: {row}")
        out.append(row)
    return {"budget_bytes": budget, "inflight": 4, "verified": True,
            "workload": f"{pool.n} distinct token chunks, log-uniform [4 KiB, 1 MiB); "
                        "requests of 1-256 chunks (log-uniform), Poisson arrivals",
            "rates": out}


def main():
    args = parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                          "(chunks shard across GPUs with no exchange; see DESIGN.md)"}))
        sys.exit(2)
    import numpy as np
    import torch

    from cudec_amd.device import DeviceBatch

    torch.cuda.init()
    batch, workload_name, malformed = build_batch(args)
    valid = np.ones(batch.n, dtype=bool)
    valid[malformed] = False
    t0 = time.perf_counter()
    db = DeviceBatch(batch.comp, batch.comp_off, batch.comp_size, batch.sizes,
                     out_off=batch.orig_off)
    orig_dev = torch.from_numpy(batch.orig).to("cuda")
    torch.cuda.synchronize()
    log(f"uploaded in {time.perf_counter() - t0:.1f}s")
    stream = torch.cuda.Stream()  # default priority
    if args.copy_probe_mib:
        n = args.copy_probe_mib << 18
        a = torch.arange(n, dtype=torch.int32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            torch.neg(a, out=b)
        torch.cuda.synchronize()
        log(f"copy probe: 3 x torch.neg over {args.copy_probe_mib} MiB of int32 "
            f"({n * 4} bytes read and written per dispatch)")
        del a, b
    uncomp = batch.total_uncompressed
    algo = uncomp + batch.total_compressed
    forms = [f for f in args.forms.split(",") if f]
    default_config = (args.workload == "uneven" and args.format == "lz4" and args.chunks is None
                      and args.seed == 0x5EED and forms[0] == "tuned")
    measured = {}
    tuning = None
    if args.threshold is not None:
        import cudec_amd
        t = (1 << 64) - 1 if args.threshold < 0 else args.threshold
        if cudec_amd.set_lz4_pipeline_threshold(t) != 0:
            raise SystemExit("[bench] could not set the pipelining threshold")
        log(f"pipelining threshold set to {t}")
    elif args.format == "lz4" and "tuned" in forms and not args.no_autotune:
        import cudec_amd
        t0 = time.perf_counter()
        st, best, sweep = cudec_amd.autotune_lz4(
            db.src_ptrs.data_ptr(), db.src_sizes.data_ptr(), db.dst_ptrs.data_ptr(),
            db.dst_caps.data_ptr(), db.n, db.d_results.data_ptr(), stream.cuda_stream)
        if st != 0:
            raise SystemExit(f"[bench] autotune failed: {st}")
        tuning = {"knob": "LZ4 chunks with cost (2 x capacity - compressed) >= threshold decode "
                          "on two pipelined waves, cheaper ones on one",
                  "best_threshold_bytes": best,
                  "sweep_ms": {("all" if c == 0 else "none" if c >= 2**63 else str(c)):
                               round(ms, 3) for c, ms in sweep}}
        log(f"autotune ({time.perf_counter() - t0:.1f}s): best threshold {best}; {tuning['sweep_ms']}")
    for form in forms:
        wall, kms, res = time_form(db, args.format, form, args.steps, args.warmup, stream,
                                   orig_dev, batch, valid)
        avg_ms = sum(kms) / len(kms)
        measured[form] = {"wall_s": wall, "kernel_ms": kms, "avg_kernel_ms": avg_ms,
                          "results": res,
                          "gbps": uncomp * args.steps / wall / 1e9,
                          "kernel_gbps": uncomp / (avg_ms / 1e3) / 1e9,
                          "algo_gbps": algo / (avg_ms / 1e3) / 1e9}
        log(f"{form}: {avg_ms:.3f} ms/launch (event), {measured[form]['gbps']:.1f} GB/s "
            f"uncompressed (wall), {measured[form]['algo_gbps']:.1f} GB/s algorithmic "
            f"= {100 * measured[form]['algo_gbps'] / HBM_PEAK_GBPS:.1f}% of HBM peak")
    head = measured[forms[0]]
    graph = None
    if "tuned" in measured and not args.no_graph:
        graph = graph_leg(db, args.format, args.steps, stream, orig_dev, batch, valid,
                          measured["tuned"]["results"])
        log(f"graph: {graph['avg_kernel_ms']:.3f} ms/replay over {graph['replays']} replays, verified")
    cpu = None
    malformed_checked = False
    if not args.no_cpu:
        threads = args.cpu_threads or min(os.cpu_count() or 1, 16)
        cpu, cpu_status, cpu_written = cpu_baseline(batch, args.format, threads, valid)
        log(f"cpu baseline: {cpu}")
        # every chunk, malformed ones included: the GPU's status and
        # bytes_written equal the CPU reference's
        for form in forms:
            r = measured[form]["results"]
            if not (np.array_equal(r["status"], cpu_status) and
                    np.array_equal(r["bytes_written"], cpu_written)):
                raise SystemExit(f"[bench] {form}: results differ from the CPU reference")
        malformed_checked = True
    serving = serving_leg(args) if args.serve_rates else None
    traffic = None
    if args.traffic_file and os.path.exists(args.traffic_file):
        # HBM bytes per launch of the tuned kernel from rocprofv3 FETCH_SIZE and
        # WRITE_SIZE runs, scaled by a known-size copy (cudec_amd/profiling.py)
        with open(args.traffic_file) as f:
            traffic = json.load(f).get("bytes_per_launch")
    elif default_config:
        traffic = committed_traffic()
    line = {
        "metric": "LZ4 batch decode GB/s (uncompressed) + % HBM peak, vs plain; p99 request latency"
        if args.format == "lz4" else "Snappy batch decode GB/s (uncompressed) + % HBM peak, vs plain",
        "value": round(head["gbps"], 3),
        "unit": "GB/s (uncompressed)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * head["wall_s"] / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic (seeded generator, liblz4/snappy-compressed on the host)",
        "config": {"workload": workload_name, "chunks": batch.n, "form": forms[0],
                   "uncompressed_bytes": uncomp, "compressed_bytes": batch.total_compressed,
                   "seed": args.seed},
        "roofline": {"bound": "hbm", "achieved": round(head["algo_gbps"], 2),
                     "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(head["algo_gbps"] / HBM_PEAK_GBPS, 5),
                     "traffic": traffic},
        "cpu_baseline": cpu,
    }
    if "plain" in measured and forms[0] != "plain":
        line["plain"] = {"value": round(measured["plain"]["gbps"], 3),
                         "avg_kernel_ms": round(measured["plain"]["avg_kernel_ms"], 4),
                         "algo_gbps": round(measured["plain"]["algo_gbps"], 2)}
        line["speedup_vs_plain"] = round(measured["plain"]["avg_kernel_ms"] / head["avg_kernel_ms"], 3)
    line["avg_kernel_ms"] = round(head["avg_kernel_ms"], 4)
    if graph is not None:
        line["graph"] = graph
    if tuning is not None:
        line["autotune"] = tuning
    if len(malformed):
        r = head["results"]
        line["malformed"] = {"chunks": int(len(malformed)),
                             "status_counts": {str(k): int(v) for k, v in zip(
                                 *np.unique(r["status"][malformed], return_counts=True))},
                             "checked_against_cpu_reference": malformed_checked}
    if serving is not None:
        line["serving"] = serving
        line["p99_request_latency_ms"] = serving["rates"][-1]["p99_ms"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
