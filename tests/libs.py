"""ctypes access to the test-side libraries: the CPU oracle (oracle/), the
optional reference-ladder checker (oracle/_ref/, present only where the
reference tree was available at build time), and the codec libraries
(liblz4, libsnappy) used to compress corpora and as parity oracles.

Test infrastructure only; the product never imports this module.
"""
from __future__ import annotations

import ctypes
import os
from functools import lru_cache

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

u8p = ctypes.POINTER(ctypes.c_uint8)
u64 = ctypes.c_uint64
u64p = ctypes.POINTER(ctypes.c_uint64)
i32 = ctypes.c_int32
i32p = ctypes.POINTER(ctypes.c_int32)

OK, INVALID, CORRUPT, TOO_SMALL, DEVICE = 0, 1, 2, 3, 4


def _first(names):
    for n in names:
        try:
            return ctypes.CDLL(n)
        except OSError:
            continue
    return None


@lru_cache(None)
def lz4lib():
    lib = _first(["liblz4.so.1", "/opt/conda/lib/liblz4.so.1",
                  "/usr/lib/x86_64-linux-gnu/liblz4.so.1"])
    if lib is None:
        return None
    lib.LZ4_compress_default.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    lib.LZ4_compress_default.restype = ctypes.c_int
    lib.LZ4_decompress_safe.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    lib.LZ4_decompress_safe.restype = ctypes.c_int
    lib.LZ4_compressBound.argtypes = [ctypes.c_int]
    lib.LZ4_compressBound.restype = ctypes.c_int
    lib.LZ4_versionString.restype = ctypes.c_char_p
    return lib


@lru_cache(None)
def snappylib():
    lib = _first(["libsnappy.so.1", "/opt/conda/lib/libsnappy.so.1",
                  "/usr/lib/x86_64-linux-gnu/libsnappy.so.1"])
    if lib is None:
        return None
    lib.snappy_compress.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.POINTER(ctypes.c_size_t)]
    lib.snappy_compress.restype = ctypes.c_int
    lib.snappy_uncompress.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.POINTER(ctypes.c_size_t)]
    lib.snappy_uncompress.restype = ctypes.c_int
    lib.snappy_max_compressed_length.argtypes = [ctypes.c_size_t]
    lib.snappy_max_compressed_length.restype = ctypes.c_size_t
    lib.snappy_uncompressed_length.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
    lib.snappy_uncompressed_length.restype = ctypes.c_int
    lib.snappy_validate_compressed_buffer.argtypes = [ctypes.c_char_p, ctypes.c_size_t]
    lib.snappy_validate_compressed_buffer.restype = ctypes.c_int
    return lib


def lz4_compress(data: bytes) -> bytes:
    """tests/fixtures.cpp:66-82 (LZ4_compress_default, acceleration 1)."""
    lib = lz4lib()
    bound = lib.LZ4_compressBound(len(data))
    out = ctypes.create_string_buffer(bound)
    w = lib.LZ4_compress_default(data, out, len(data), bound)
    assert w > 0, "liblz4 compression failed"
    return out.raw[:w]


def lz4_decompress(stream: bytes, capacity: int):
    """tests/fixtures.cpp:147-162 (OracleDecodes): (accepted, bytes). An
    empty stream is never handed to liblz4."""
    lib = lz4lib()
    if len(stream) == 0:
        return False, b""
    buf = ctypes.create_string_buffer(max(capacity, 1))
    w = lib.LZ4_decompress_safe(stream, buf, len(stream), capacity)
    if w < 0:
        return False, b""
    return True, buf.raw[:w]


def snappy_compress(data: bytes) -> bytes:
    """tests/fixtures.cpp:443-461 (RawCompress via the C API)."""
    lib = snappylib()
    cap = lib.snappy_max_compressed_length(len(data))
    out = ctypes.create_string_buffer(cap)
    n = ctypes.c_size_t(cap)
    assert lib.snappy_compress(data, len(data), out, ctypes.byref(n)) == 0
    return out.raw[:n.value]


SNAPPY_ORACLE_MAX_OUTPUT = 256 * 1024 * 1024


def snappy_decompress(stream: bytes):
    """tests/fixtures.cpp:173-198 (SnappyOracleDecodes): exact-consume,
    exact-produce decode sized from the stream's own declaration, refused
    above the 256 MiB wrapper bound."""
    lib = snappylib()
    if len(stream) == 0:
        return False, b""
    declared = ctypes.c_size_t(0)
    if lib.snappy_uncompressed_length(stream, len(stream), ctypes.byref(declared)) != 0:
        return False, b""
    if declared.value > SNAPPY_ORACLE_MAX_OUTPUT:
        return False, b""
    buf = ctypes.create_string_buffer(max(declared.value, 1))
    n = ctypes.c_size_t(declared.value)
    if lib.snappy_uncompress(stream, len(stream), buf, ctypes.byref(n)) != 0:
        return False, b""
    return True, buf.raw[:n.value]


# ------------------------------------------------------------------ oracle

@lru_cache(None)
def oracle():
    path = os.path.join(ROOT, "oracle", "liboracle.so")
    lib = ctypes.CDLL(path)
    lib.oracle_lz4_decode.argtypes = [ctypes.c_char_p, u64, ctypes.c_void_p, u64, u64p]
    lib.oracle_lz4_decode.restype = i32
    lib.oracle_lz4_parse.argtypes = [ctypes.c_char_p, u64, u64, u64p, u64p, u64p]
    lib.oracle_lz4_parse.restype = i32
    lib.oracle_snappy_declared.argtypes = [ctypes.c_char_p, u64, u64p, u64p, i32p]
    lib.oracle_snappy_declared.restype = i32
    lib.oracle_snappy_decode.argtypes = [ctypes.c_char_p, u64, ctypes.c_void_p, u64, u64p, i32p]
    lib.oracle_snappy_decode.restype = i32
    lib.oracle_snappy_parse.argtypes = [ctypes.c_char_p, u64, u64, u64p, u64p, u64p, i32p]
    lib.oracle_snappy_parse.restype = i32
    lib.oracle_decode_batch_timed.argtypes = [i32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                              ctypes.c_void_p, u64, i32, ctypes.c_void_p, ctypes.c_void_p]
    lib.oracle_decode_batch_timed.restype = ctypes.c_double
    return lib


@lru_cache(None)
def ref():
    """The reference ladders compiled from the reference tree, or None where
    that build is absent (e.g. on a host without the reference)."""
    path = os.path.join(ROOT, "oracle", "_ref", "libcudec_ref.so")
    if not os.path.exists(path):
        return None
    lib = ctypes.CDLL(path)
    lib.ref_lz4_decode.argtypes = [ctypes.c_char_p, u64, ctypes.c_void_p, u64, u64p]
    lib.ref_lz4_decode.restype = i32
    lib.ref_lz4_parse.argtypes = [ctypes.c_char_p, u64, u64, u64p]
    lib.ref_lz4_parse.restype = i32
    lib.ref_snappy_decode.argtypes = [ctypes.c_char_p, u64, ctypes.c_void_p, u64, u64p, i32p]
    lib.ref_snappy_decode.restype = i32
    lib.ref_snappy_declared.argtypes = [ctypes.c_char_p, u64, u64p, u64p, i32p]
    lib.ref_snappy_declared.restype = i32
    return lib


def _decode(fn, stream: bytes, capacity: int, *extra):
    buf = ctypes.create_string_buffer(max(capacity, 1))
    w = u64(0)
    st = fn(stream, len(stream), buf, capacity, ctypes.byref(w), *extra)
    return st, (buf.raw[:w.value] if st == OK else b"")


def oracle_lz4(stream: bytes, capacity: int):
    """(status, output) from the C oracle."""
    return _decode(oracle().oracle_lz4_decode, stream, capacity)


def oracle_snappy(stream: bytes, capacity: int):
    """(status, output, reject rung) from the C oracle."""
    r = i32(0)
    st, out = _decode(oracle().oracle_snappy_decode, stream, capacity, ctypes.byref(r))
    return st, out, r.value


def ref_lz4(stream: bytes, capacity: int):
    return _decode(ref().ref_lz4_decode, stream, capacity)


def ref_snappy(stream: bytes, capacity: int):
    r = i32(0)
    st, out = _decode(ref().ref_snappy_decode, stream, capacity, ctypes.byref(r))
    return st, out, r.value


# ------------------------------------------------------------- LZ4 frames

class LZ4FFrameInfo(ctypes.Structure):
    _fields_ = [("blockSizeID", ctypes.c_int), ("blockMode", ctypes.c_int),
                ("contentChecksumFlag", ctypes.c_int), ("frameType", ctypes.c_int),
                ("contentSize", ctypes.c_ulonglong), ("dictID", ctypes.c_uint),
                ("blockChecksumFlag", ctypes.c_int)]


class LZ4FPreferences(ctypes.Structure):
    _fields_ = [("frameInfo", LZ4FFrameInfo), ("compressionLevel", ctypes.c_int),
                ("autoFlush", ctypes.c_uint), ("favorDecSpeed", ctypes.c_uint),
                ("reserved", ctypes.c_uint * 3)]


def lz4_frame(data: bytes, block_size_id: int = 4, independent: bool = True,
              content_checksum: bool = False, block_checksum: bool = False,
              content_size: bool = False) -> bytes:
    """liblz4's LZ4F_compressFrame (block size id 4..7 = 64 KiB..4 MiB)."""
    lib = lz4lib()
    lib.LZ4F_compressFrameBound.argtypes = [ctypes.c_size_t, ctypes.c_void_p]
    lib.LZ4F_compressFrameBound.restype = ctypes.c_size_t
    lib.LZ4F_compressFrame.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_char_p,
                                       ctypes.c_size_t, ctypes.c_void_p]
    lib.LZ4F_compressFrame.restype = ctypes.c_size_t
    prefs = LZ4FPreferences()
    prefs.frameInfo.blockSizeID = block_size_id
    prefs.frameInfo.blockMode = 1 if independent else 0
    prefs.frameInfo.contentChecksumFlag = int(content_checksum)
    prefs.frameInfo.blockChecksumFlag = int(block_checksum)
    prefs.frameInfo.contentSize = len(data) if content_size else 0
    bound = lib.LZ4F_compressFrameBound(len(data), ctypes.byref(prefs))
    out = ctypes.create_string_buffer(bound)
    n = lib.LZ4F_compressFrame(out, bound, data, len(data), ctypes.byref(prefs))
    assert n < (1 << 63), "LZ4F_compressFrame failed"
    return out.raw[:n]


def lz4_frame_decode(frame: bytes, capacity: int):
    """liblz4's streaming frame API (LZ4F_decompress) as the frame oracle:
    (accepted, bytes). A frame that ends before its end mark, or that needs
    more than `capacity` bytes of output, is refused."""
    lib = lz4lib()
    vp, sz = ctypes.c_void_p, ctypes.c_size_t
    lib.LZ4F_createDecompressionContext.argtypes = [ctypes.POINTER(vp), ctypes.c_uint]
    lib.LZ4F_createDecompressionContext.restype = sz
    lib.LZ4F_freeDecompressionContext.argtypes = [vp]
    lib.LZ4F_freeDecompressionContext.restype = sz
    lib.LZ4F_decompress.argtypes = [vp, vp, ctypes.POINTER(sz), vp, ctypes.POINTER(sz), vp]
    lib.LZ4F_decompress.restype = sz
    lib.LZ4F_isError.argtypes = [sz]
    lib.LZ4F_isError.restype = ctypes.c_uint
    dctx = vp()
    if lib.LZ4F_isError(lib.LZ4F_createDecompressionContext(ctypes.byref(dctx), 100)):
        return False, b""
    src = ctypes.create_string_buffer(frame, max(len(frame), 1))
    out = ctypes.create_string_buffer(max(capacity, 1))
    so = do = 0
    hint = 1
    ok = True
    try:
        while so < len(frame):
            sl, dl = sz(len(frame) - so), sz(capacity - do)
            hint = lib.LZ4F_decompress(dctx, ctypes.addressof(out) + do, ctypes.byref(dl),
                                       ctypes.addressof(src) + so, ctypes.byref(sl), None)
            if lib.LZ4F_isError(hint):
                ok = False
                break
            so += sl.value
            do += dl.value
            if hint == 0:
                break
            if sl.value == 0 and dl.value == 0:
                ok = False
                break
    finally:
        lib.LZ4F_freeDecompressionContext(dctx)
    if ok and hint != 0:
        ok = False
    return ok, (out.raw[:do] if ok else b"")



_P1, _P2, _P3, _P4, _P5 = 0x9E3779B1, 0x85EBCA77, 0xC2B2AE3D, 0x27D4EB2F, 0x165667B1
_M32 = 0xFFFFFFFF


def _rotl(x: int, r: int) -> int:
    return ((x << r) | (x >> (32 - r))) & _M32


def xxh32(data: bytes, seed: int = 0) -> int:
    """XXH32 from the published xxHash specification (four 32-bit lanes
    over 16-byte stripes, then 4-byte and 1-byte tails and the avalanche).
    Used to repair header checksums of hand-mutated frames; pinned against
    the header-checksum byte liblz4 writes and the specification's
    empty-input value."""
    n = len(data)
    i = 0
    if n >= 16:
        v = [(seed + _P1 + _P2) & _M32, (seed + _P2) & _M32, seed & _M32, (seed - _P1) & _M32]
        while i + 16 <= n:
            for k in range(4):
                lane = int.from_bytes(data[i + 4 * k:i + 4 * k + 4], "little")
                v[k] = (_rotl((v[k] + lane * _P2) & _M32, 13) * _P1) & _M32
            i += 16
        h = (_rotl(v[0], 1) + _rotl(v[1], 7) + _rotl(v[2], 12) + _rotl(v[3], 18)) & _M32
    else:
        h = (seed + _P5) & _M32
    h = (h + n) & _M32
    while i + 4 <= n:
        h = (_rotl((h + int.from_bytes(data[i:i + 4], "little") * _P3) & _M32, 17) * _P4) & _M32
        i += 4
    while i < n:
        h = (_rotl((h + data[i] * _P5) & _M32, 11) * _P1) & _M32
        i += 1
    h ^= h >> 15
    h = (h * _P2) & _M32
    h ^= h >> 13
    h = (h * _P3) & _M32
    h ^= h >> 16
    return h


_Q1, _Q2, _Q3, _Q4, _Q5 = (0x9E3779B185EBCA87, 0xC2B2AE3D27D4EB4F, 0x165667B19E3779F9,
                           0x85EBCA77C2B2AE63, 0x27D4EB2F165667C5)
_M64 = (1 << 64) - 1


def _rotl64(x: int, r: int) -> int:
    return ((x << r) | (x >> (64 - r))) & _M64


def _round64(acc: int, lane: int) -> int:
    return (_rotl64((acc + lane * _Q2) & _M64, 31) * _Q1) & _M64


def xxh64(data: bytes, seed: int = 0) -> int:
    """XXH64 from the published xxHash specification (four 64-bit lanes over
    32-byte stripes, the lane merge, 8-, 4- and 1-byte tails, the
    avalanche). Pinned against the specification's empty-input value and,
    where the system's libxxhash is present, against it."""
    n = len(data)
    i = 0
    if n >= 32:
        v = [(seed + _Q1 + _Q2) & _M64, (seed + _Q2) & _M64, seed & _M64, (seed - _Q1) & _M64]
        while i + 32 <= n:
            for k in range(4):
                v[k] = _round64(v[k], int.from_bytes(data[i + 8 * k:i + 8 * k + 8], "little"))
            i += 32
        h = (_rotl64(v[0], 1) + _rotl64(v[1], 7) + _rotl64(v[2], 12) + _rotl64(v[3], 18)) & _M64
        for k in range(4):
            h = ((h ^ _round64(0, v[k])) * _Q1 + _Q4) & _M64
    else:
        h = (seed + _Q5) & _M64
    h = (h + n) & _M64
    while i + 8 <= n:
        h ^= _round64(0, int.from_bytes(data[i:i + 8], "little"))
        h = (_rotl64(h, 27) * _Q1 + _Q4) & _M64
        i += 8
    if i + 4 <= n:
        h ^= (int.from_bytes(data[i:i + 4], "little") * _Q1) & _M64
        h = (_rotl64(h, 23) * _Q2 + _Q3) & _M64
        i += 4
    while i < n:
        h ^= (data[i] * _Q5) & _M64
        h = (_rotl64(h, 11) * _Q1) & _M64
        i += 1
    h ^= h >> 33
    h = (h * _Q2) & _M64
    h ^= h >> 29
    h = (h * _Q3) & _M64
    h ^= h >> 32
    return h


@lru_cache(None)
def xxhash_lib():
    """The system's libxxhash (XXH32, XXH64), or None where it is absent."""
    for name in ("libxxhash.so.0", "/usr/lib/x86_64-linux-gnu/libxxhash.so.0"):
        try:
            lib = ctypes.CDLL(name)
        except OSError:
            continue
        lib.XXH32.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint32]
        lib.XXH32.restype = ctypes.c_uint32
        lib.XXH64.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint64]
        lib.XXH64.restype = ctypes.c_uint64
        return lib
    return None


# ------------------------------------------------------------------ zstd

ZSTD_C_LEVEL, ZSTD_C_WINDOW_LOG = 100, 101
ZSTD_C_CONTENT_SIZE, ZSTD_C_CHECKSUM = 200, 201
ZSTD_C_LITERAL_MODE, ZSTD_C_TARGET_BLOCK = 1002, 1003  # experimental in 1.4.9
UNSUPPORTED = 6


@lru_cache(None)
def zstdlib():
    """libzstd (1.4.9 in this image), the compressor of the Zstd corpora and
    the byte oracle of valid frames; None where it is absent."""
    lib = _first(["/opt/conda/lib/libzstd.so.1", "libzstd.so.1"])
    if lib is None:
        return None
    lib.ZSTD_versionString.restype = ctypes.c_char_p
    lib.ZSTD_compressBound.argtypes = [ctypes.c_size_t]
    lib.ZSTD_compressBound.restype = ctypes.c_size_t
    lib.ZSTD_isError.argtypes = [ctypes.c_size_t]
    lib.ZSTD_isError.restype = ctypes.c_uint
    lib.ZSTD_createCCtx.restype = ctypes.c_void_p
    lib.ZSTD_freeCCtx.argtypes = [ctypes.c_void_p]
    lib.ZSTD_CCtx_setParameter.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    lib.ZSTD_CCtx_setParameter.restype = ctypes.c_size_t
    lib.ZSTD_compress2.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                   ctypes.c_char_p, ctypes.c_size_t]
    lib.ZSTD_compress2.restype = ctypes.c_size_t
    lib.ZSTD_decompress.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_char_p,
                                    ctypes.c_size_t]
    lib.ZSTD_decompress.restype = ctypes.c_size_t
    return lib


def zstd_compress(data: bytes, level: int = 3, **params) -> bytes:
    """One frame of `data` through ZSTD_compress2 with the given parameters
    (ids above, by keyword: window_log, content_size, checksum,
    literal_mode, target_block)."""
    z = zstdlib()
    ids = {"window_log": ZSTD_C_WINDOW_LOG, "content_size": ZSTD_C_CONTENT_SIZE,
           "checksum": ZSTD_C_CHECKSUM, "literal_mode": ZSTD_C_LITERAL_MODE,
           "target_block": ZSTD_C_TARGET_BLOCK}
    cctx = z.ZSTD_createCCtx()
    try:
        for pid, val in [(ZSTD_C_LEVEL, level)] + [(ids[k], v) for k, v in params.items()]:
            r = z.ZSTD_CCtx_setParameter(cctx, pid, int(val))
            assert not z.ZSTD_isError(r), f"libzstd refused parameter {pid}={val}"
        cap = z.ZSTD_compressBound(len(data)) + 64
        buf = ctypes.create_string_buffer(cap)
        n = z.ZSTD_compress2(cctx, buf, cap, data, len(data))
        assert not z.ZSTD_isError(n)
        return buf.raw[:n]
    finally:
        z.ZSTD_freeCCtx(cctx)


def zstd_decompress(frame: bytes, capacity: int):
    """(ok, bytes) from libzstd's ZSTD_decompress."""
    z = zstdlib()
    buf = ctypes.create_string_buffer(max(capacity, 1))
    n = z.ZSTD_decompress(buf, capacity, frame, len(frame))
    if z.ZSTD_isError(n):
        return False, b""
    return True, buf.raw[:n]


@lru_cache(None)
def ref_zstd():
    """The reference's Zstd frame twin compiled from the reference tree
    (oracle/_ref/libcudec_ref_zstd.so), or None where that build is absent."""
    path = os.path.join(ROOT, "oracle", "_ref", "libcudec_ref_zstd.so")
    if not os.path.exists(path):
        return None
    lib = ctypes.CDLL(path)
    lib.ref_zstd_decode.argtypes = [ctypes.c_char_p, u64, ctypes.c_void_p, u64, u64p, i32p]
    lib.ref_zstd_decode.restype = i32
    return lib


def ref_zstd_decode(frame: bytes, capacity: int):
    """(status, output) of one frame through the reference's twin."""
    buf = ctypes.create_string_buffer(max(capacity, 1))
    w, stage = u64(0), i32(0)
    st = ref_zstd().ref_zstd_decode(frame, len(frame), buf, capacity, ctypes.byref(w),
                                    ctypes.byref(stage))
    return st, (buf.raw[:w.value] if st == OK else b"")
