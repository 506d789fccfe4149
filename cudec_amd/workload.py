"""Seeded synthetic batches, compressed (host side).

Wraps cudec_amd/libcudec_workload.so. A batch is three packed byte arenas
(original, compressed) plus uint64 offset / size arrays, which is the layout
the device side uploads: one arena per direction and per-chunk pointers
built as base + offset.

Workloads (sizes in bytes):
* token64k(n)   -- n chunks of exactly 65536 bytes from the seeded token
                   process (p = 0.45 literal runs of 1 + Geometric(mean 12)
                   Zipf(1.1) symbols over 64; else copies of
                   4 + Geometric(mean 20) at offsets uniform in
                   [1, min(pos, 65535)], 30% forced into [1, 16]).
* uneven(n)     -- n chunks log-uniform over [1 KiB, 4 MiB): 80% token
                   process, 10% uniform random, 10% repetitive (p = 0.05
                   literals, offsets in [1, 8]); order shuffled.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
_WL = None

KIND_TOKENS, KIND_RANDOM, KIND_REPETITIVE = 0, 1, 2
CODEC_LZ4, CODEC_SNAPPY, CODEC_ZSTD = 0, 1, 2


def wl() -> ctypes.CDLL:
    global _WL
    if _WL is None:
        L = ctypes.CDLL(os.path.join(_PKG, "libcudec_workload.so"))
        vp, u64, i32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int32
        L.wl_chunk_seed.argtypes = [u64, u64]
        L.wl_chunk_seed.restype = u64
        L.wl_gen_chunk.argtypes = [vp, u64, u64, i32]
        L.wl_gen_batch.argtypes = [vp, vp, vp, vp, vp, u64, i32]
        L.wl_plan_uneven.argtypes = [u64, u64, u64, ctypes.c_uint32, ctypes.c_uint32,
                                     ctypes.c_uint32, vp, vp, vp]
        L.wl_codec_available.argtypes = [i32]
        L.wl_codec_available.restype = i32
        L.wl_codec_version.argtypes = [i32]
        L.wl_codec_version.restype = ctypes.c_char_p
        L.wl_compress_bound.argtypes = [i32, u64]
        L.wl_compress_bound.restype = u64
        L.wl_compress.argtypes = [i32, vp, u64, vp, u64]
        L.wl_compress.restype = u64
        L.wl_decompress.argtypes = [i32, vp, u64, vp, u64, ctypes.POINTER(u64)]
        L.wl_decompress.restype = i32
        L.wl_compress_batch.argtypes = [i32, vp, vp, vp, vp, vp, vp, vp, u64, i32]
        L.wl_compress_batch.restype = u64
        L.wl_pack.argtypes = [vp, vp, vp, vp, vp, u64]
        L.wl_fnv1a64.argtypes = [u64, vp, u64]
        L.wl_fnv1a64.restype = u64
        _WL = L
    return _WL


def _p(a: np.ndarray) -> int:
    return a.ctypes.data


def default_threads() -> int:
    n = os.cpu_count() or 1
    return max(1, min(n, 16))


def codec_version(codec: int) -> str:
    return wl().wl_codec_version(codec).decode()


def codec_available(codec: int) -> bool:
    return bool(wl().wl_codec_available(codec))


def fnv1a64(data: np.ndarray, h: int = 14695981039346656037) -> int:
    return wl().wl_fnv1a64(h, _p(data), data.nbytes)


def offsets_of(sizes: np.ndarray, align: int = 1) -> np.ndarray:
    s = sizes.astype(np.uint64)
    if align > 1:
        s = (s + np.uint64(align - 1)) // np.uint64(align) * np.uint64(align)
    off = np.zeros(len(sizes), dtype=np.uint64)
    if len(sizes) > 1:
        np.cumsum(s[:-1], out=off[1:])
    return off


@dataclass
class Batch:
    """A compressed batch in packed arenas."""
    codec: int
    sizes: np.ndarray            # uint64 original sizes
    kinds: np.ndarray            # int32 chunk kinds
    seeds: np.ndarray            # uint64 per-chunk seeds
    orig: np.ndarray             # uint8 arena of originals
    orig_off: np.ndarray         # uint64
    comp: np.ndarray             # uint8 arena of compressed chunks
    comp_off: np.ndarray         # uint64
    comp_size: np.ndarray        # uint64
    meta: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return len(self.sizes)

    def original(self, i: int) -> bytes:
        o = int(self.orig_off[i])
        return self.orig[o:o + int(self.sizes[i])].tobytes()

    def compressed(self, i: int) -> bytes:
        o = int(self.comp_off[i])
        return self.comp[o:o + int(self.comp_size[i])].tobytes()

    @property
    def total_uncompressed(self) -> int:
        return int(self.sizes.sum())

    @property
    def total_compressed(self) -> int:
        return int(self.comp_size.sum())


def generate(sizes, kinds, seeds, threads: int | None = None):
    sizes = np.ascontiguousarray(sizes, dtype=np.uint64)
    kinds = np.ascontiguousarray(kinds, dtype=np.int32)
    seeds = np.ascontiguousarray(seeds, dtype=np.uint64)
    off = offsets_of(sizes)
    arena = np.empty(int(sizes.sum()) if len(sizes) else 0, dtype=np.uint8)
    wl().wl_gen_batch(_p(arena), _p(off), _p(sizes), _p(seeds), _p(kinds), len(sizes),
                      threads or default_threads())
    return arena, off


def compress(codec: int, orig: np.ndarray, orig_off: np.ndarray, sizes: np.ndarray,
             threads: int | None = None):
    if not codec_available(codec):
        raise RuntimeError(f"codec {codec} library is not available: {codec_version(codec)}")
    L = wl()
    bounds = np.array([L.wl_compress_bound(codec, int(s)) for s in sizes], dtype=np.uint64)
    slot_off = offsets_of(bounds)
    staging = np.empty(int(bounds.sum()), dtype=np.uint8)
    csize = np.zeros(len(sizes), dtype=np.uint64)
    fails = L.wl_compress_batch(codec, _p(orig), _p(orig_off), _p(sizes), _p(staging),
                                _p(slot_off), _p(bounds), _p(csize), len(sizes),
                                threads or default_threads())
    if fails:
        raise RuntimeError(f"{fails} chunks failed to compress")
    comp_off = offsets_of(csize)
    L.wl_pack(_p(staging), _p(comp_off), _p(staging), _p(slot_off), _p(csize), len(sizes))
    comp = staging[:int(csize.sum())].copy()
    del staging
    return comp, comp_off, csize


def build(codec: int, sizes, kinds, seeds, threads: int | None = None, **meta) -> Batch:
    orig, orig_off = generate(sizes, kinds, seeds, threads)
    comp, comp_off, csize = compress(codec, orig, orig_off, np.asarray(sizes, dtype=np.uint64),
                                     threads)
    return Batch(codec, np.asarray(sizes, dtype=np.uint64), np.asarray(kinds, dtype=np.int32),
                 np.asarray(seeds, dtype=np.uint64), orig, orig_off, comp, comp_off, csize,
                 dict(meta))


def token64k(n: int, seed: int = 0x1A2B, codec: int = CODEC_LZ4,
             threads: int | None = None) -> Batch:
    sizes = np.full(n, 65536, dtype=np.uint64)
    kinds = np.full(n, KIND_TOKENS, dtype=np.int32)
    seeds = np.array([wl().wl_chunk_seed(seed, i) for i in range(n)], dtype=np.uint64)
    return build(codec, sizes, kinds, seeds, threads, workload=f"token64k x{n}", seed=seed)


def plan_uneven(n: int, seed: int, min_size: int = 1024, octaves: int = 12,
                pct_random: int = 10, pct_repetitive: int = 10):
    sizes = np.zeros(n, dtype=np.uint64)
    kinds = np.zeros(n, dtype=np.int32)
    seeds = np.zeros(n, dtype=np.uint64)
    wl().wl_plan_uneven(n, seed, min_size, octaves, pct_random, pct_repetitive,
                        _p(sizes), _p(kinds), _p(seeds))
    return sizes, kinds, seeds


def uneven(n: int, seed: int = 0x5EED, codec: int = CODEC_LZ4, min_size: int = 1024,
           octaves: int = 12, threads: int | None = None) -> Batch:
    sizes, kinds, seeds = plan_uneven(n, seed, min_size, octaves)
    return build(codec, sizes, kinds, seeds, threads,
                 workload=f"uneven x{n} [{min_size}, {min_size << octaves})", seed=seed)
