"""cudec_amd -- MI355X (gfx950) batch LZ4 / Snappy block and Zstandard frame
decoding behind the cudec C ABI.

The product is the native library cudec_amd/libcudec.so (HIP kernels and
the C entry points declared in include/cudec.h, include/cudec_snappy.h,
include/cudec_zstd.h and include/cudec_mi355x.h). This module binds those entry points with ctypes
under their C names, so Python callers drive exactly the ABI a C caller
does. Importing it loads the library and fails loudly if it is absent:
there is no fallback decode path.
"""
from __future__ import annotations

import ctypes
import os

# torch is imported first on purpose: its bundled HIP runtime carries the
# soname libamdhip64.so.7, so loading it before libcudec.so makes the
# library bind to that same runtime instead of a second copy from
# /opt/rocm (one process must not hold two HIP runtimes).
import torch  # noqa: F401

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libcudec.so")

# cudec_status (include/cudec.h)
CUDEC_OK = 0
CUDEC_ERR_INVALID_ARGUMENT = 1
CUDEC_ERR_CORRUPT_INPUT = 2
CUDEC_ERR_OUTPUT_TOO_SMALL = 3
CUDEC_ERR_CUDA = 4
CUDEC_ERR_NOT_IMPLEMENTED = 5
CUDEC_ERR_UNSUPPORTED = 6

CUDEC_MEM_HOST = 0
CUDEC_MEM_DEVICE = 1

# The symbols every header in include/ declares, in declaration order.
ABI_SYMBOLS = {
    "cudec.h": [
        "cudec_version", "cudec_lz4_decompress_batch", "cudec_lz4f_decompress",
        "cudec_stream_ctx_create", "cudec_lz4_decompress_stream_ctx",
        "cudec_stream_ctx_destroy", "cudec_lz4_decompress_stream",
    ],
    "cudec_snappy.h": ["cudec_snappy_decompress_batch"],
    "cudec_zstd.h": ["cudec_zstd_decompress_batch"],
    "cudec_mi355x.h": [
        "cudec_lz4_decompress_batch_plain", "cudec_snappy_decompress_batch_plain",
        "cudec_mi355x_max_batch_chunks", "cudec_mi355x_cost_bin",
        "cudec_mi355x_set_decode_plan", "cudec_mi355x_decode_plan",
        "cudec_mi355x_set_lz4_pipeline_threshold", "cudec_mi355x_lz4_pipeline_threshold",
        "cudec_mi355x_set_snappy_pipeline_threshold", "cudec_mi355x_snappy_pipeline_threshold",
        "cudec_mi355x_autotune_lz4", "cudec_mi355x_autotune_plan",
        "cudec_mi355x_lz4_decompress_batch_noalloc", "cudec_mi355x_parse_batch",
        "cudec_mi355x_xxh32_batch", "cudec_mi355x_xxh64_batch", "cudec_mi355x_digest_batch",
        "cudec_mi355x_serve_lz4",
    ],
}


class ChunkResult(ctypes.Structure):
    """cudec_chunk_result: {int32 status, uint32 reserved, uint64 bytes_written}."""
    _fields_ = [("status", ctypes.c_int32), ("reserved", ctypes.c_uint32),
                ("bytes_written", ctypes.c_uint64)]


assert ctypes.sizeof(ChunkResult) == 16 and ChunkResult.bytes_written.offset == 8

_vp = ctypes.c_void_p
_sz = ctypes.c_size_t
_BATCH_ARGS = [_vp, _vp, _vp, _vp, _sz, _vp, _vp]

_lib = None


def lib() -> ctypes.CDLL:
    """The loaded product library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                "(the decoders have no non-native fallback)")
        L = ctypes.CDLL(LIB_PATH)
        L.cudec_version.restype = ctypes.c_int
        for name in ("cudec_lz4_decompress_batch", "cudec_snappy_decompress_batch",
                     "cudec_zstd_decompress_batch", "cudec_lz4_decompress_batch_plain",
                     "cudec_snappy_decompress_batch_plain",
                     "cudec_mi355x_lz4_decompress_batch_noalloc"):
            fn = getattr(L, name)
            fn.argtypes = _BATCH_ARGS
            fn.restype = ctypes.c_int
        L.cudec_mi355x_max_batch_chunks.restype = _sz
        L.cudec_stream_ctx_create.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        L.cudec_stream_ctx_create.restype = ctypes.c_int
        L.cudec_stream_ctx_destroy.argtypes = [_vp]
        L.cudec_stream_ctx_destroy.restype = None
        L.cudec_lz4_decompress_stream_ctx.argtypes = [_vp, _vp, _vp, _vp, _vp, _sz, ctypes.c_int, _vp]
        L.cudec_lz4_decompress_stream_ctx.restype = ctypes.c_int
        L.cudec_lz4_decompress_stream.argtypes = [_vp, _vp, _vp, _vp, _sz, ctypes.c_int, _vp]
        L.cudec_lz4_decompress_stream.restype = ctypes.c_int
        L.cudec_lz4f_decompress.argtypes = [_vp, _sz, _vp, _sz, ctypes.POINTER(_sz)]
        L.cudec_lz4f_decompress.restype = ctypes.c_int
        L.cudec_mi355x_parse_batch.argtypes = [ctypes.c_int] + _BATCH_ARGS
        L.cudec_mi355x_parse_batch.restype = ctypes.c_int
        L.cudec_mi355x_xxh32_batch.argtypes = [_vp, _vp, _sz, ctypes.c_uint32, _vp, _vp]
        L.cudec_mi355x_xxh32_batch.restype = ctypes.c_int
        L.cudec_mi355x_xxh64_batch.argtypes = [_vp, _vp, _sz, ctypes.c_uint64, _vp, _vp]
        L.cudec_mi355x_xxh64_batch.restype = ctypes.c_int
        L.cudec_mi355x_digest_batch.argtypes = [_vp, _vp, _vp, _sz, _vp, _vp]
        L.cudec_mi355x_digest_batch.restype = ctypes.c_int
        L.cudec_mi355x_serve_lz4.argtypes = [_vp, _vp, _vp, _sz, _vp, _sz, _sz, ctypes.c_int,
                                             _vp, _vp, _vp, _vp]
        L.cudec_mi355x_serve_lz4.restype = ctypes.c_int
        for fmt in ("lz4", "snappy"):
            setter = getattr(L, f"cudec_mi355x_set_{fmt}_pipeline_threshold")
            setter.argtypes = [_sz]
            setter.restype = ctypes.c_int
            getter = getattr(L, f"cudec_mi355x_{fmt}_pipeline_threshold")
            getter.argtypes = [ctypes.POINTER(_sz)]
            getter.restype = ctypes.c_int
        L.cudec_mi355x_cost_bin.argtypes = [_sz, _sz]
        L.cudec_mi355x_cost_bin.restype = ctypes.c_int
        L.cudec_mi355x_set_decode_plan.argtypes = [ctypes.c_int, _vp]
        L.cudec_mi355x_set_decode_plan.restype = ctypes.c_int
        L.cudec_mi355x_decode_plan.argtypes = [ctypes.c_int, _vp]
        L.cudec_mi355x_decode_plan.restype = ctypes.c_int
        L.cudec_mi355x_autotune_plan.argtypes = [ctypes.c_int, _vp, _vp, _vp, _vp, _sz, _vp, _vp,
                                                 _vp, ctypes.POINTER(ctypes.c_float),
                                                 ctypes.POINTER(ctypes.c_float)]
        L.cudec_mi355x_autotune_plan.restype = ctypes.c_int
        L.cudec_mi355x_autotune_lz4.argtypes = [_vp, _vp, _vp, _vp, _sz, _vp, _vp,
                                                ctypes.POINTER(_sz), _vp, _vp]
        L.cudec_mi355x_autotune_lz4.restype = ctypes.c_int
        _lib = L
    return _lib


def _batch(name, src_ptrs, src_sizes, dst_ptrs, dst_caps, chunk_count, results, stream):
    return getattr(lib(), name)(src_ptrs, src_sizes, dst_ptrs, dst_caps,
                                chunk_count, results, stream)


def cudec_version() -> int:
    return lib().cudec_version()


def cudec_lz4_decompress_batch(d_src_ptrs, d_src_sizes, d_dst_ptrs, d_dst_capacities,
                               chunk_count, d_results, stream=None) -> int:
    """include/cudec.h: all arguments are device addresses (ints) except
    chunk_count; stream is a hipStream_t handle (int) or None."""
    return _batch("cudec_lz4_decompress_batch", d_src_ptrs, d_src_sizes, d_dst_ptrs,
                  d_dst_capacities, chunk_count, d_results, stream)


def cudec_snappy_decompress_batch(d_src_ptrs, d_src_sizes, d_dst_ptrs, d_dst_capacities,
                                  chunk_count, d_results, stream=None) -> int:
    return _batch("cudec_snappy_decompress_batch", d_src_ptrs, d_src_sizes, d_dst_ptrs,
                  d_dst_capacities, chunk_count, d_results, stream)


def cudec_zstd_decompress_batch(d_src_ptrs, d_src_sizes, d_dst_ptrs, d_dst_capacities,
                                chunk_count, d_results, stream=None) -> int:
    """One Zstandard frame per chunk (include/cudec_zstd.h)."""
    return _batch("cudec_zstd_decompress_batch", d_src_ptrs, d_src_sizes, d_dst_ptrs,
                  d_dst_capacities, chunk_count, d_results, stream)


def cudec_lz4_decompress_batch_plain(d_src_ptrs, d_src_sizes, d_dst_ptrs, d_dst_capacities,
                                     chunk_count, d_results, stream=None) -> int:
    return _batch("cudec_lz4_decompress_batch_plain", d_src_ptrs, d_src_sizes, d_dst_ptrs,
                  d_dst_capacities, chunk_count, d_results, stream)


def cudec_snappy_decompress_batch_plain(d_src_ptrs, d_src_sizes, d_dst_ptrs,
                                        d_dst_capacities, chunk_count, d_results,
                                        stream=None) -> int:
    return _batch("cudec_snappy_decompress_batch_plain", d_src_ptrs, d_src_sizes, d_dst_ptrs,
                  d_dst_capacities, chunk_count, d_results, stream)


def max_batch_chunks() -> int:
    return lib().cudec_mi355x_max_batch_chunks()


ENTRY = {
    ("lz4", "tuned"): cudec_lz4_decompress_batch,
    ("lz4", "plain"): cudec_lz4_decompress_batch_plain,
    ("snappy", "tuned"): cudec_snappy_decompress_batch,
    ("snappy", "plain"): cudec_snappy_decompress_batch_plain,
    ("zstd", "tuned"): cudec_zstd_decompress_batch,
}


# ----------------------------------------------------------- streaming context

def cudec_stream_ctx_create():
    """Returns (status, handle)."""
    h = ctypes.c_void_p(0)
    st = lib().cudec_stream_ctx_create(ctypes.byref(h))
    return st, h.value


def cudec_stream_ctx_destroy(ctx) -> None:
    lib().cudec_stream_ctx_destroy(ctx)


def cudec_lz4_decompress_stream_ctx(ctx, h_src_ptrs, h_src_sizes, dst_ptrs, dst_caps,
                                    chunk_count, dst_space, h_results) -> int:
    """include/cudec.h:193-197; array arguments are host addresses (ints)."""
    return lib().cudec_lz4_decompress_stream_ctx(ctx, h_src_ptrs, h_src_sizes, dst_ptrs,
                                                 dst_caps, chunk_count, dst_space, h_results)


def cudec_lz4_decompress_stream(h_src_ptrs, h_src_sizes, dst_ptrs, dst_caps, chunk_count,
                                dst_space, h_results) -> int:
    return lib().cudec_lz4_decompress_stream(h_src_ptrs, h_src_sizes, dst_ptrs, dst_caps,
                                             chunk_count, dst_space, h_results)


# ------------------------------------------------------------------ LZ4 frames

def cudec_lz4f_decompress(frame: bytes, capacity: int):
    """include/cudec.h (cudec_lz4f_decompress): one LZ4 frame in host memory
    decoded into a host buffer of `capacity` bytes. Returns (status, bytes):
    the decoded content on CUDEC_OK, b"" otherwise."""
    out = ctypes.create_string_buffer(max(capacity, 1))
    written = _sz(0)
    src = ctypes.c_char_p(frame)
    st = lib().cudec_lz4f_decompress(ctypes.cast(src, _vp), len(frame), ctypes.cast(out, _vp),
                                     capacity, ctypes.byref(written))
    return st, (out.raw[:written.value] if st == CUDEC_OK else b"")


# ------------------------------------------------------------------ tuning

TUNE_CANDIDATES = 9
COST_BINS = 256
FORMAT_CODE = {"lz4": 0, "snappy": 1}


def set_lz4_pipeline_threshold(min_cost: int) -> int:
    return lib().cudec_mi355x_set_lz4_pipeline_threshold(min_cost)


def lz4_pipeline_threshold():
    """(status, min_cost) of the current device."""
    v = _sz(0)
    st = lib().cudec_mi355x_lz4_pipeline_threshold(ctypes.byref(v))
    return st, v.value


def set_pipeline_threshold(fmt: str, min_cost: int) -> int:
    return getattr(lib(), f"cudec_mi355x_set_{fmt}_pipeline_threshold")(min_cost)


def pipeline_threshold(fmt: str):
    """(status, min_cost) of the current device for format fmt."""
    v = _sz(0)
    st = getattr(lib(), f"cudec_mi355x_{fmt}_pipeline_threshold")(ctypes.byref(v))
    return st, v.value


def cost_bin(compressed_size: int, capacity: int) -> int:
    return lib().cudec_mi355x_cost_bin(compressed_size, capacity)


def set_decode_plan(fmt: str, lanes) -> int:
    """lanes: COST_BINS lane counts (64 or 128), bin 0 the costliest."""
    arr = (ctypes.c_uint16 * COST_BINS)(*[int(x) for x in lanes])
    return lib().cudec_mi355x_set_decode_plan(FORMAT_CODE[fmt], arr)


def decode_plan(fmt: str):
    """(status, [lanes per bin]) of the current device for format fmt."""
    arr = (ctypes.c_uint16 * COST_BINS)()
    st = lib().cudec_mi355x_decode_plan(FORMAT_CODE[fmt], arr)
    return st, list(arr)


def autotune_plan(fmt: str, d_src_ptrs, d_src_sizes, d_dst_ptrs, d_dst_capacities,
                  chunk_count, d_results, stream=None):
    """Runs cudec_mi355x_autotune_plan: (status, [lanes per bin],
    best threshold plan ms, final plan ms)."""
    arr = (ctypes.c_uint16 * COST_BINS)()
    ms_t, ms_p = ctypes.c_float(0), ctypes.c_float(0)
    st = lib().cudec_mi355x_autotune_plan(FORMAT_CODE[fmt], d_src_ptrs, d_src_sizes, d_dst_ptrs,
                                          d_dst_capacities, chunk_count, d_results, stream, arr,
                                          ctypes.byref(ms_t), ctypes.byref(ms_p))
    return st, list(arr), ms_t.value, ms_p.value


def autotune_lz4(d_src_ptrs, d_src_sizes, d_dst_ptrs, d_dst_capacities, chunk_count,
                 d_results, stream=None):
    """Runs cudec_mi355x_autotune_lz4: (status, best_min_cost, [(candidate, ms)])."""
    best = _sz(0)
    cands = (_sz * TUNE_CANDIDATES)()
    ms = (ctypes.c_float * TUNE_CANDIDATES)()
    st = lib().cudec_mi355x_autotune_lz4(d_src_ptrs, d_src_sizes, d_dst_ptrs, d_dst_capacities,
                                         chunk_count, d_results, stream, ctypes.byref(best),
                                         cands, ms)
    return st, best.value, list(zip(list(cands), list(ms)))
