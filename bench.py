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
before timing. --format snappy and --format zstd run the same batch as raw
Snappy blocks and as Zstandard frames (libzstd level 3, one frame a chunk,
cudec_zstd_decompress_batch).

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
    p.add_argument("--steps", type=int, default=30)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--workload", choices=["uneven", "token64k"], default="uneven")
    p.add_argument("--format", choices=["lz4", "snappy", "zstd"], default="lz4")
    p.add_argument("--chunks", type=int, default=None)
    p.add_argument("--seed", type=int, default=0x5EED)
    p.add_argument("--forms", default="tuned,plain")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--no-parse", action="store_true", help="skip the parse-only ceiling leg")
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
                   help="keep the default decode plan instead of tuning it "
                        "on the batch before timing")
    p.add_argument("--threshold", type=int, default=None,
                   help="set the format's pipelining threshold (bytes of cost; -1: none) instead of "
                        "tuning it, so a profiled run repeats one launch shape")
    p.add_argument("--serve-rates", default="2000,8000,32000",
                   help="request rates (per second) of the serving leg; empty to skip it")
    p.add_argument("--serve-seconds", type=float, default=0.5,
                   help="scheduled arrival span of each serving run")
    p.add_argument("--serve-pool", type=int, default=2048,
                   help="distinct chunks the serving requests draw from")
    p.add_argument("--serve-budget-gib", type=float, default=64.0)
    p.add_argument("--serve-inflight", default="1,2,4",
                   help="launches in flight (streams) of the serving runs, comma-separated; "
                        "the first is the headline configuration (one in flight has the "
                        "lowest p99 at every rate measured, the metric's latency figure)")
    return p.parse_args()


def build_batch(args):
    from cudec_amd import workload
    codec = {"lz4": workload.CODEC_LZ4, "snappy": workload.CODEC_SNAPPY,
             "zstd": workload.CODEC_ZSTD}[args.format]
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


def workload_codec(fmt: str) -> str:
    from cudec_amd import workload
    return workload.codec_version({"lz4": workload.CODEC_LZ4, "snappy": workload.CODEC_SNAPPY,
                                   "zstd": workload.CODEC_ZSTD}[fmt])


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


def parse_leg(db, fmt, steps, warmup, stream, eager):
    """The parse ceiling: the tuned launch's parse-only build
    (cudec_mi355x_parse_batch) on the same batch and plan, event-timed over
    `steps` launches after `warmup`; its statuses and sizes must equal the
    decode's."""
    import numpy as np
    import torch
    import cudec_amd
    from cudec_amd.serve import percentile

    def launch():
        return cudec_amd.lib().cudec_mi355x_parse_batch(
            cudec_amd.FORMAT_CODE[fmt], db.src_ptrs.data_ptr(), db.src_sizes.data_ptr(),
            db.dst_ptrs.data_ptr(), db.dst_caps.data_ptr(), db.n, db.d_results.data_ptr(),
            stream.cuda_stream)
    for _ in range(warmup):
        assert launch() == 0
    torch.cuda.synchronize()
    if not np.array_equal(db.results(), eager):
        raise SystemExit("[bench] parse-only results differ from the decode's")
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    for k in range(steps):
        starts[k].record(stream)
        assert launch() == 0
        ends[k].record(stream)
    torch.cuda.synchronize()
    ms = [a.elapsed_time(b) for a, b in zip(starts, ends)]
    return {"avg_kernel_ms": round(sum(ms) / len(ms), 4),
            "p50_kernel_ms": round(percentile(ms, 50), 4),
            "launches": steps, "results_equal_decode": True}


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


# The box's CPU share for one GPU is 16 threads; the harness never runs more
# worker threads than that, whatever the host shows.
CPU_SHARE = 16

DEC_LIBLZ4, DEC_LIBSNAPPY, DEC_TWIN_LZ4, DEC_TWIN_SNAPPY, DEC_LIBZSTD, DEC_TWIN_ZSTD = 0, 1, 2, 3, 4, 5


def cpu_threads(requested=None) -> tuple[int, int]:
    """(threads to use, cores this process may run on)."""
    visible = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else \
        (os.cpu_count() or 1)
    return (requested or max(1, min(visible, CPU_SHARE))), visible


def twin_decoders(fmt):
    """The reference's host ladders compiled from its source (oracle/_ref,
    kind "reference") where that build is present, else the oracle's C
    restatement (kind "port"): (function address, kind, name)."""
    ref_path = os.path.join(ROOT, "oracle", "_ref",
                            "libcudec_ref_zstd.so" if fmt == "zstd" else "libcudec_ref.so")
    port_path = os.path.join(ROOT, "oracle", "liboracle.so")
    if fmt == "zstd" and not os.path.exists(ref_path):
        raise SystemExit("[bench] the Zstd CPU baseline needs oracle/_ref/libcudec_ref_zstd.so "
                         "(the reference's frame twin, built from the reference tree)")
    if os.path.exists(ref_path):
        lib, kind, prefix = ctypes.CDLL(ref_path), "reference", "ref"
    else:
        lib, kind, prefix = ctypes.CDLL(port_path), "port", "oracle"
    fn = getattr(lib, f"{prefix}_{fmt}_decode")
    return ctypes.cast(fn, ctypes.c_void_p).value, kind, f"{prefix}_{fmt}_decode", lib


class CpuBatch:
    """Host pointer tables of a batch, and output buffers sized to it."""

    def __init__(self, batch, m=None):
        import numpy as np
        m = batch.n if m is None else m
        self.n = m
        self.out = np.empty(max(int(batch.sizes[:m].sum()), 1), dtype=np.uint8)
        self.srcs = (batch.comp_off[:m] + np.uint64(batch.comp.ctypes.data)).astype(np.uint64)
        self.sizes = batch.comp_size[:m].astype(np.uint64)
        off = batch.orig_off[:m] - batch.orig_off[0] if m else batch.orig_off[:0]
        self.dsts = (off + np.uint64(self.out.ctypes.data)).astype(np.uint64)
        self.caps = batch.sizes[:m].astype(np.uint64)
        self.bytes = int(batch.sizes[:m].sum())
        self.algo = self.bytes + int(batch.comp_size[:m].sum())

    def time(self, kind, fn, threads, passes):
        """Per-pass seconds, and (status, written) of the last pass."""
        import numpy as np
        from cudec_amd import workload
        W = workload.wl()
        W.wl_time_decode.argtypes = [ctypes.c_int32, ctypes.c_void_p] + [ctypes.c_void_p] * 4 + \
            [ctypes.c_uint64, ctypes.c_int32, ctypes.c_int32] + [ctypes.c_void_p] * 3
        W.wl_time_decode.restype = ctypes.c_int32
        secs = np.zeros(passes, dtype=np.float64)
        status = np.zeros(self.n, dtype=np.int32)
        written = np.zeros(self.n, dtype=np.uint64)
        ran = W.wl_time_decode(kind, fn, self.srcs.ctypes.data, self.sizes.ctypes.data,
                               self.dsts.ctypes.data, self.caps.ctypes.data, self.n, threads,
                               passes, secs.ctypes.data, status.ctypes.data, written.ctypes.data)
        if ran < 0:
            return None, None, None
        return secs, status, written


def cpu_row(cb, kind, fn, threads, name, warmup=3, runs=30, budget_s=10.0):
    """warmup + up to `runs` timed passes (fewer when a pass is long, so the
    row takes about budget_s), nearest-rank p50/p90/p99 of the pass times."""
    from cudec_amd.serve import percentile
    first, status, written = cb.time(kind, fn, threads, 1)
    if first is None:
        return None, None, None
    runs = max(3, min(runs, int(budget_s / max(float(first[0]), 1e-4))))
    secs, _, _ = cb.time(kind, fn, threads, warmup + runs)
    secs = secs[warmup:]
    rate = lambda t: round(cb.bytes / t / 1e9, 3)
    row = {"decoder": name, "threads": threads, "chunks": cb.n,
           "uncompressed_bytes": cb.bytes,
           "p50": rate(percentile(secs, 50)), "at_p90_time": rate(percentile(secs, 90)),
           "at_p99_time": rate(percentile(secs, 99)),
           "algo_gbps_p50": round(cb.algo / percentile(secs, 50) / 1e9, 3),
           "passes": f"{warmup} warmup + {len(secs)} timed"}
    return row, status, written


def cpu_baseline(batch, fmt, threads, valid, visible):
    """The CPU decoders timed on the host's cores in this run: the
    reference's host twin (oracle/_ref) and the compression library's own
    decoder (liblz4 LZ4_decompress_safe, snappy_uncompress), on `threads`
    threads (static contiguous partition) and on one thread, decode calls
    only into pre-sized buffers (a C harness, cudec_amd/csrc/workload.c).
    The twin's statuses and sizes over the whole batch are returned for the
    GPU results to be held to; the twin's bytes are checked against the
    originals."""
    import numpy as np
    from cudec_amd import workload
    fn, kind, twin_name, _keep = twin_decoders(fmt)
    twin = {"lz4": DEC_TWIN_LZ4, "snappy": DEC_TWIN_SNAPPY, "zstd": DEC_TWIN_ZSTD}[fmt]
    lib_kind = {"lz4": DEC_LIBLZ4, "snappy": DEC_LIBSNAPPY, "zstd": DEC_LIBZSTD}[fmt]
    codec = {"lz4": workload.CODEC_LZ4, "snappy": workload.CODEC_SNAPPY,
             "zstd": workload.CODEC_ZSTD}[fmt]
    lib_name = {"lz4": "LZ4_decompress_safe", "snappy": "snappy_uncompress",
                "zstd": "ZSTD_decompress"}[fmt] + f" ({workload.codec_version(codec)})"
    full = CpuBatch(batch)
    rows = []
    row, status, written = cpu_row(full, twin, fn, threads, f"twin {twin_name} [{kind}]")
    if row is None:
        raise SystemExit("[bench] CPU baseline: the twin decoder could not run")
    ok = bool(np.all(status[valid] == 0)) and \
        bool(np.array_equal(written[valid], batch.sizes[valid]))
    if ok:
        for i in np.nonzero(valid)[0]:
            o = int(batch.orig_off[i] - batch.orig_off[0])
            n = int(batch.sizes[i])
            if not np.array_equal(full.out[o:o + n], batch.orig[int(batch.orig_off[i]):
                                                               int(batch.orig_off[i]) + n]):
                ok = False
                break
    if not ok:
        raise SystemExit("[bench] CPU baseline produced wrong output")
    rows.append(row)
    lib_row, lib_status, lib_written = cpu_row(full, lib_kind, None, threads, lib_name)
    if lib_row is not None:
        if not (np.all(lib_status[valid] == 0) and
                np.array_equal(lib_written[valid], batch.sizes[valid])):
            raise SystemExit("[bench] CPU baseline: the library decoder disagrees")
        rows.append(lib_row)
    # one thread over a sample: the batch's first chunks (the order is
    # shuffled) up to 256 MiB of output
    m = int(np.searchsorted(np.cumsum(batch.sizes.astype(np.int64)), 256 << 20, side="right"))
    sample = CpuBatch(batch, max(1, min(m, batch.n)))
    for k, f, name in ((twin, fn, f"twin {twin_name} [{kind}]"), (lib_kind, None, lib_name)):
        r, _, _ = cpu_row(sample, k, f, 1, name)
        if r is not None:
            rows.append(r)
    rec = {"value": rows[0]["p50"], "unit": "GB/s (uncompressed)", "cores": threads,
           "kind": kind,
           "sample": f"the full batch ({batch.n} chunks, {batch.total_uncompressed / 2**30:.2f} "
                     f"GiB out), {threads} threads (static contiguous partition) of the "
                     f"{visible} this process may run on (one GPU's share is {CPU_SHARE}), "
                     f"decode calls only into pre-sized buffers; {rows[0]['passes']} passes, "
                     "nearest-rank p50 of the pass times; single-thread rows decode the "
                     f"batch's first {sample.n} chunks ({sample.bytes / 2**20:.0f} MiB out)",
           "at_p90_time": rows[0]["at_p90_time"], "at_p99_time": rows[0]["at_p99_time"],
           "cores_visible": visible, "cpu_model": cpu_model(), "rows": rows}
    return rec, status, written


def cpu_config_token64k(threads, seed):
    """The CPU configuration: 1024 chunks of exactly 64 KiB from the token
    process, LZ4_compress_default; the twin and liblz4, one thread and
    `threads` threads, 3 warmup + 30 timed passes each, every chunk checked
    byte-exact first."""
    import numpy as np
    from cudec_amd import workload
    b = workload.token64k(1024, seed=seed)
    fn, kind, twin_name, _keep = twin_decoders("lz4")
    cb = CpuBatch(b)
    rows = []
    for k, f, name in ((DEC_TWIN_LZ4, fn, f"twin {twin_name} [{kind}]"),
                       (DEC_LIBLZ4, None,
                        f"LZ4_decompress_safe ({workload.codec_version(workload.CODEC_LZ4)})")):
        for t in (1, threads):
            cb.out.fill(0)
            r, status, written = cpu_row(cb, k, f, t, name, budget_s=30.0)
            if r is None:
                continue
            if not (np.all(status == 0) and np.array_equal(written, b.sizes) and
                    np.array_equal(cb.out, b.orig)):
                raise SystemExit(f"[bench] CPU configuration: {name} output differs")
            rows.append(r)
    return {"workload": "1024 x 64 KiB token-process chunks, LZ4_compress_default "
                        f"(ratio {b.total_uncompressed / b.total_compressed:.2f})",
            "rows": rows}


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


def chunk_latency_leg(seed):
    """Idle latency of one chunk per size bin: a batch of one token-process
    chunk, decoded by the tuned launch on two pipelined waves and on one
    wave, median of 10 event-timed launches after 2 warmups. A request's
    latency cannot fall below its largest chunk's."""
    import numpy as np
    import torch
    import cudec_amd
    from cudec_amd import workload
    from cudec_amd.device import DeviceBatch
    st, old = cudec_amd.pipeline_threshold("lz4")
    out = []
    try:
        for size in (4 << 10, 16 << 10, 64 << 10, 256 << 10, 1 << 20, 4 << 20):
            b = workload.build(workload.CODEC_LZ4, np.array([size], np.uint64),
                               np.zeros(1, np.int32), np.array([seed + size], np.uint64))
            db = DeviceBatch(b.comp, b.comp_off, b.comp_size, b.sizes)
            row = {"bytes": size}
            for name, t in (("two_wave_ms", 0), ("one_wave_ms", (1 << 64) - 1)):
                cudec_amd.set_pipeline_threshold("lz4", t)
                ms = []
                for r in range(12):
                    e0 = torch.cuda.Event(enable_timing=True)
                    e1 = torch.cuda.Event(enable_timing=True)
                    e0.record()
                    assert db.run("lz4", "tuned") == 0
                    e1.record()
                    torch.cuda.synchronize()
                    if r >= 2:
                        ms.append(e0.elapsed_time(e1))
                row[name] = round(float(np.median(ms)), 4)
            res = db.results()
            assert res["status"][0] == 0 and res["bytes_written"][0] == size
            out.append(row)
    finally:
        cudec_amd.set_pipeline_threshold("lz4", old)
    return out


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
    inflights = [int(x) for x in args.serve_inflight.split(",") if x]
    out = []
    sweep = []
    for inflight in inflights:
        for k, rate in enumerate(rates):
            nreq = max(16, int(rate * args.serve_seconds))
            req, ids = serve.make_trace(rate, nreq, pool.n, seed=args.seed + k)
            check = serve.serve(d_comp.data_ptr(), pool, req, ids, budget=budget,
                                inflight=inflight, digests=True)
            ok = check.status == 0 and bool(np.all(check.results["status"] == 0)) and \
                bool(np.array_equal(check.results["bytes_written"], pool.sizes[ids])) and \
                bool(np.array_equal(check.digests, want[ids]))
            if not ok:
                raise SystemExit(f"[bench] serving at {rate:.0f}/s: verification FAILED")
            run = serve.serve(d_comp.data_ptr(), pool, req, ids, budget=budget,
                              inflight=inflight, digests=False)
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
                   "launches": int(st.launches),
                   "max_batch_requests": int(st.max_batch_requests),
                   "max_inflight": int(st.max_inflight),
                   "peak_pool_mib": round(st.peak_pool_bytes / 2**20, 1)}
            log(f"serving {rate:.0f}/s, {inflight} in flight: {row}")
            if inflight == inflights[0]:
                out.append(row)
            sweep.append({"inflight": inflight, "rate_per_s": rate, "p50_ms": row["p50_ms"],
                          "p99_ms": row["p99_ms"], "gbps_uncompressed": row["gbps_uncompressed"]})
    return {"budget_bytes": budget, "inflight": inflights[0], "verified": True,
            "workload": f"{pool.n} distinct token chunks, log-uniform [4 KiB, 1 MiB); "
                        "requests of 1-256 chunks (log-uniform), Poisson arrivals",
            "rates": out, "inflight_sweep": sweep}


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
    if args.format == "zstd":
        # one launch form, and no decode plan, parse-only build or graph leg
        forms = ["tuned"]
        args.no_autotune = args.no_parse = args.no_graph = True
    default_config = (args.workload == "uneven" and args.format == "lz4" and args.chunks is None
                      and args.seed == 0x5EED and forms[0] == "tuned")
    measured = {}
    tuning = None
    if args.threshold is not None:
        import cudec_amd
        t = (1 << 64) - 1 if args.threshold < 0 else args.threshold
        if cudec_amd.set_pipeline_threshold(args.format, t) != 0:
            raise SystemExit("[bench] could not set the pipelining threshold")
        log(f"pipelining threshold set to {t}")
    elif "tuned" in forms and not args.no_autotune and args.format != "zstd":
        import cudec_amd
        t0 = time.perf_counter()
        st, lanes, ms_threshold, ms_plan = cudec_amd.autotune_plan(
            args.format, db.src_ptrs.data_ptr(), db.src_sizes.data_ptr(), db.dst_ptrs.data_ptr(),
            db.dst_caps.data_ptr(), db.n, db.d_results.data_ptr(), stream.cuda_stream)
        if st != 0:
            raise SystemExit(f"[bench] autotune failed: {st}")
        bins = sorted({cudec_amd.cost_bin(int(c), int(u))
                       for c, u in zip(batch.comp_size, batch.sizes)})
        tuning = {"knob": "lanes per chunk for each cost bin (cost = 2 x capacity - "
                          "compressed): 128 = two pipelined waves, 64 = one wave; a threshold "
                          "sweep, then each occupied octave moved to the other lane count when "
                          "that gains 1%",
                  "plan": {str(b): lanes[b] for b in bins},
                  "threshold_plan_ms": round(ms_threshold, 3), "plan_ms": round(ms_plan, 3),
                  "tuning_s": round(time.perf_counter() - t0, 2)}
        log(f"autotune: {tuning}")
    for form in forms:
        wall, kms, res = time_form(db, args.format, form, args.steps, args.warmup, stream,
                                   orig_dev, batch, valid)
        avg_ms = sum(kms) / len(kms)
        from cudec_amd.serve import percentile
        measured[form] = {"wall_s": wall, "kernel_ms": kms, "avg_kernel_ms": avg_ms,
                          "p50_kernel_ms": percentile(kms, 50),
                          "p90_kernel_ms": percentile(kms, 90),
                          "p99_kernel_ms": percentile(kms, 99),
                          "results": res,
                          "gbps": uncomp * args.steps / wall / 1e9,
                          "kernel_gbps": uncomp / (avg_ms / 1e3) / 1e9,
                          "algo_gbps": algo / (avg_ms / 1e3) / 1e9}
        log(f"{form}: {avg_ms:.3f} ms/launch (event), {measured[form]['gbps']:.1f} GB/s "
            f"uncompressed (wall), {measured[form]['algo_gbps']:.1f} GB/s algorithmic "
            f"= {100 * measured[form]['algo_gbps'] / HBM_PEAK_GBPS:.1f}% of HBM peak")
    head = measured[forms[0]]
    parse = None
    if "tuned" in measured and not args.no_parse:
        parse = parse_leg(db, args.format, args.steps, args.warmup, stream,
                          measured["tuned"]["results"])
        parse["gbps_uncompressed"] = round(uncomp / parse["avg_kernel_ms"] / 1e6, 2)
        parse["algo_gbps"] = round(algo / parse["avg_kernel_ms"] / 1e6, 2)
        parse["decode_over_parse"] = round(measured["tuned"]["avg_kernel_ms"] /
                                           parse["avg_kernel_ms"], 3)
        log(f"parse only: {parse}")
    graph = None
    if "tuned" in measured and not args.no_graph:
        graph = graph_leg(db, args.format, args.steps, stream, orig_dev, batch, valid,
                          measured["tuned"]["results"])
        log(f"graph: {graph['avg_kernel_ms']:.3f} ms/replay over {graph['replays']} replays, verified")
    cpu = None
    cpu_config = None
    malformed_checked = False
    if not args.no_cpu:
        threads, visible = cpu_threads(args.cpu_threads)
        cpu, cpu_status, cpu_written = cpu_baseline(batch, args.format, threads, valid, visible)
        log(f"cpu baseline: {cpu}")
        if args.format == "lz4":
            cpu_config = cpu_config_token64k(threads, args.seed)
            log(f"cpu configuration: {cpu_config}")
        # every chunk, malformed ones included: the GPU's status and
        # bytes_written equal the CPU reference's
        for form in forms:
            r = measured[form]["results"]
            if not (np.array_equal(r["status"], cpu_status) and
                    np.array_equal(r["bytes_written"], cpu_written)):
                raise SystemExit(f"[bench] {form}: results differ from the CPU reference")
        malformed_checked = True
    serving = serving_leg(args) if args.serve_rates else None
    latency = None
    if args.serve_rates and args.format == "lz4":
        latency = chunk_latency_leg(args.seed)
        log(f"single-chunk latency at idle: {latency}")
    traffic = None
    if args.traffic_file and os.path.exists(args.traffic_file):
        # HBM bytes per launch of the tuned kernel from rocprofv3 FETCH_SIZE and
        # WRITE_SIZE runs, scaled by a known-size copy (cudec_amd/profiling.py)
        with open(args.traffic_file) as f:
            traffic = json.load(f).get("bytes_per_launch")
    elif default_config:
        traffic = committed_traffic()
    line = {
        "metric": {"lz4": "LZ4 batch decode GB/s (uncompressed) + % HBM peak, vs plain; "
                          "p99 request latency",
                   "snappy": "Snappy batch decode GB/s (uncompressed) + % HBM peak, vs plain",
                   "zstd": "Zstd frame batch decode GB/s (uncompressed) + % HBM peak"}[args.format],
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
        "data": "synthetic (seeded generator, compressed on the host with " +
                workload_codec(args.format) + ")",
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
    line["kernel_ms_percentiles"] = {
        "p50": round(head["p50_kernel_ms"], 4), "p90": round(head["p90_kernel_ms"], 4),
        "p99": round(head["p99_kernel_ms"], 4), "launches": args.steps,
        "rule": "nearest rank over the event-timed launches"}
    if parse is not None:
        line["parse_only"] = parse
    if cpu_config is not None:
        line["cpu_config_token64k"] = cpu_config
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
    if latency is not None:
        line["single_chunk_latency_idle"] = latency
    if serving is not None:
        line["serving"] = serving
        line["p99_request_latency_ms"] = serving["rates"][-1]["p99_ms"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
