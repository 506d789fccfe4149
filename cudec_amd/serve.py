"""Serving runs over the native request loop (cudec_mi355x_serve_lz4).

A serving run replays a request trace in real time: Poisson arrivals at a
given rate, each request 1-256 LZ4 chunks (count log-uniform) drawn from a
pool of distinct chunks whose compressed bytes are device-resident, chunk
sizes log-uniform over [4 KiB, 1 MiB) from the seeded token generator. The
loop admits requests against an HBM output budget, batches whatever has
been admitted into the next launch while earlier launches run, and reports
per-request latency from scheduled arrival to completion.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import CUDEC_OK, lib
from . import workload

REQUEST_DTYPE = np.dtype([("arrival_s", "<f8"), ("first_chunk", "<u8"), ("chunk_count", "<u8")])
RESULT_DTYPE = np.dtype([("status", "<i4"), ("reserved", "<u4"), ("bytes_written", "<u8")])

_LEN_MUL = 0x9E3779B97F4A7C15
_POS_MUL = 0xD6E8FEB86659FD93


class ServeStats(ctypes.Structure):
    _fields_ = [("elapsed_s", ctypes.c_double), ("launches", ctypes.c_uint64),
                ("max_inflight", ctypes.c_uint64), ("max_batch_requests", ctypes.c_uint64),
                ("peak_pool_bytes", ctypes.c_uint64), ("output_bytes", ctypes.c_uint64),
                ("compressed_bytes", ctypes.c_uint64)]


def host_digest(data) -> int:
    """The digest cudec_mi355x_digest_batch computes, on the host (numpy):
    len * 0x9E3779B97F4A7C15 + sum_j splitmix64_finalise(w_j + j * 0xD6E8FEB86659FD93)
    over the zero-padded little-endian 8-byte words w_j, mod 2^64."""
    b = np.frombuffer(bytes(data), dtype=np.uint8) if not isinstance(data, np.ndarray) else data
    n = int(b.size)
    pad = (-n) % 8
    words = np.concatenate([b, np.zeros(pad, np.uint8)]).view("<u8") if n else \
        np.zeros(0, np.uint64)
    with np.errstate(over="ignore"):
        x = words + np.arange(words.size, dtype=np.uint64) * np.uint64(_POS_MUL)
        x ^= x >> np.uint64(30)
        x *= np.uint64(0xBF58476D1CE4E5B9)
        x ^= x >> np.uint64(27)
        x *= np.uint64(0x94D049BB133111EB)
        x ^= x >> np.uint64(31)
        total = int(x.sum(dtype=np.uint64)) if x.size else 0
    return (n * _LEN_MUL + total) & 0xFFFFFFFFFFFFFFFF


def chunk_pool(n: int, seed: int, min_size: int = 4096, octaves: int = 8) -> workload.Batch:
    """n distinct token-process chunks, sizes log-uniform over
    [min_size, min_size << octaves), LZ4-compressed."""
    sizes, kinds, seeds = workload.plan_uneven(n, seed, min_size, octaves, 0, 0)
    return workload.build(workload.CODEC_LZ4, sizes, kinds, seeds,
                          workload=f"serving pool x{n} [{min_size}, {min_size << octaves})",
                          seed=seed)


def make_trace(rate: float, n_requests: int, n_pool: int, seed: int, max_chunks: int = 256):
    """(requests, chunk_ids): Poisson arrivals at `rate` per second starting
    at 0, chunk counts log-uniform over [1, max_chunks], chunks drawn
    uniformly from the pool. Request r covers chunk_ids[first, first + count)."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, n_requests)
    arrivals = np.cumsum(gaps) - gaps[0]
    counts = np.floor(np.exp(rng.uniform(0.0, np.log(max_chunks + 1), n_requests)))
    counts = np.clip(counts, 1, max_chunks).astype(np.uint64)
    first = np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.uint64)
    chunk_ids = rng.integers(0, n_pool, int(counts.sum()), dtype=np.int64)
    req = np.zeros(n_requests, dtype=REQUEST_DTYPE)
    req["arrival_s"] = arrivals
    req["first_chunk"] = first
    req["chunk_count"] = counts
    return req, chunk_ids


@dataclass
class ServeRun:
    status: int
    results: np.ndarray
    digests: np.ndarray | None
    latency_s: np.ndarray
    stats: ServeStats


def serve(d_comp_base: int, pool: workload.Batch, requests: np.ndarray, chunk_ids: np.ndarray,
          budget: int, inflight: int = 4, digests: bool = True,
          caps: np.ndarray | None = None) -> ServeRun:
    """Runs the trace through the native loop. d_comp_base is the device
    address of the pool's compressed arena (pool.comp uploaded as is);
    capacities default to each chunk's original size."""
    ids = np.asarray(chunk_ids, dtype=np.int64)
    src = (pool.comp_off[ids] + np.uint64(d_comp_base)).astype(np.uint64)
    sizes = pool.comp_size[ids].astype(np.uint64)
    cap = (pool.sizes[ids] if caps is None else np.asarray(caps, np.uint64)).astype(np.uint64)
    requests = np.ascontiguousarray(requests, dtype=REQUEST_DTYPE)
    n = len(ids)
    res = np.zeros(n, dtype=RESULT_DTYPE)
    dig = np.zeros(n, dtype=np.uint64) if digests else None
    lat = np.zeros(len(requests), dtype=np.float64)
    stats = ServeStats()
    st = lib().cudec_mi355x_serve_lz4(
        src.ctypes.data, sizes.ctypes.data, cap.ctypes.data, n, requests.ctypes.data,
        len(requests), budget, inflight, res.ctypes.data,
        dig.ctypes.data if dig is not None else None, lat.ctypes.data, ctypes.byref(stats))
    return ServeRun(st, res, dig, lat, stats)


def percentile(values: np.ndarray, q: float) -> float:
    """Nearest-rank percentile (the reference bench's rule,
    bench/bench_stats.h:33-43)."""
    v = np.sort(np.asarray(values, dtype=np.float64))
    if v.size == 0:
        return float("nan")
    k = int(np.ceil(q / 100.0 * v.size)) - 1
    return float(v[min(max(k, 0), v.size - 1)])


__all__ = ["REQUEST_DTYPE", "ServeStats", "host_digest", "chunk_pool", "make_trace", "serve",
           "percentile", "ServeRun", "CUDEC_OK"]
