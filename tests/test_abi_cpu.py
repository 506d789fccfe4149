"""The C ABI without a GPU: every symbol include/*.h declares is exported,
the headers compile as C99 / C11 / C++ with the fixed result layout
(tests/conformance_c.c), every synchronous argument reject happens without
a device call, a validation-passing call with no visible device reports
CUDEC_ERR_CUDA (tests/launch_fail.cpp), and the frame parser's host-side
verdicts (tests/frame_host_negative.cpp) -- including the XXH32 header,
block and content checksums on stored-block frames, which need no device."""
import ctypes
import os
import re
import subprocess
import sys
import textwrap

import numpy as np
import pytest

import cudec_amd
from tests import libs

ROOT = libs.ROOT
INCLUDE = os.path.join(ROOT, "include")


def declared_functions():
    out = {}
    for h in sorted(os.listdir(INCLUDE)):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(INCLUDE, h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        out[h] = re.findall(r"\b(cudec_\w+)\s*\(", text)
    return out


def test_every_declared_symbol_is_exported():
    lib = cudec_amd.lib()
    decl = declared_functions()
    assert decl == cudec_amd.ABI_SYMBOLS
    for h, names in decl.items():
        for n in names:
            assert hasattr(lib, n), f"{h}: {n} not exported"
    nm = subprocess.run(["nm", "-D", "--defined-only", cudec_amd.LIB_PATH],
                        capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (cudec_\w+)", nm))
    assert exported == {n for names in decl.values() for n in names}
    assert os.path.exists(os.path.join(os.path.dirname(cudec_amd.LIB_PATH), "libcudec.a"))


def test_version():
    assert cudec_amd.cudec_version() == 0 * 10000 + 1 * 100 + 0


@pytest.mark.parametrize("std,cc", [("c99", "gcc"), ("c11", "gcc"), ("c++11", "g++"),
                                    ("c++17", "g++")])
def test_headers_compile_and_link(tmp_path, std, cc):
    """tests/conformance_c.c: the headers stand alone in C and C++ and a
    consumer links against the library; run, it exercises synchronous
    rejects that touch no device."""
    src = tmp_path / ("t.c" if cc == "gcc" else "t.cpp")
    src.write_text(textwrap.dedent("""
        #include "cudec.h"
        #include "cudec_snappy.h"
        #include "cudec_zstd.h"
        #include "cudec_mi355x.h"
        #include <stdio.h>
        int main(void) {
            cudec_chunk_result r;
            size_t w = 7;
            if (sizeof(r) != 16 || cudec_version() != 100) return 1;
            if (cudec_lz4_decompress_batch(0, 0, 0, 0, 1, 0, 0) != CUDEC_ERR_INVALID_ARGUMENT) return 2;
            if (cudec_snappy_decompress_batch(0, 0, 0, 0, 1, 0, 0) != CUDEC_ERR_INVALID_ARGUMENT) return 3;
            if (cudec_lz4f_decompress(0, 0, 0, 0, &w) != CUDEC_ERR_INVALID_ARGUMENT || w != 0) return 4;
            if (cudec_stream_ctx_create(0) != CUDEC_ERR_INVALID_ARGUMENT) return 5;
            if (cudec_lz4_decompress_stream_ctx(0, 0, 0, 0, 0, 1, CUDEC_MEM_HOST, 0) != CUDEC_ERR_INVALID_ARGUMENT) return 6;
            if (cudec_lz4_decompress_stream(0, 0, 0, 0, 1, CUDEC_MEM_HOST, 0) != CUDEC_ERR_INVALID_ARGUMENT) return 7;
            cudec_stream_ctx_destroy(0);
            if (cudec_mi355x_max_batch_chunks() != (size_t)0x7fffffff * 64) return 8;
            if (cudec_zstd_decompress_batch(0, 0, 0, 0, 1, 0, 0) != CUDEC_ERR_INVALID_ARGUMENT) return 9;
            puts("ok");
            return 0;
        }
    """))
    exe = tmp_path / "t"
    libdir = os.path.dirname(cudec_amd.LIB_PATH)
    flags = ["-std=" + std, "-Wall", "-Wextra", "-Werror", "-pedantic"]
    subprocess.run([cc, *flags, f"-I{INCLUDE}", str(src), "-o", str(exe), f"-L{libdir}",
                    "-lcudec", f"-Wl,-rpath,{libdir}"], check=True)
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0 and run.stdout.strip() == "ok", (run.returncode, run.stderr)


def test_batch_rejects_without_a_device_call():
    """src/lz4_decode.cuh:187-201 (and tests/batch_limit_parity.cpp): NULL
    arrays, misaligned results, 0, SIZE_MAX and limit+1 chunks, for every
    batch entry."""
    limit = cudec_amd.max_batch_chunks()
    assert limit == 0x7FFFFFFF * 64
    for entry in cudec_amd.ENTRY.values():
        good = [16, 16, 16, 16]
        for k in range(4):
            bad = list(good)
            bad[k] = None
            assert entry(*bad, 1, 16, None) == cudec_amd.CUDEC_ERR_INVALID_ARGUMENT
        assert entry(*good, 1, None, None) == cudec_amd.CUDEC_ERR_INVALID_ARGUMENT
        assert entry(*good, 1, 24, None) == cudec_amd.CUDEC_ERR_INVALID_ARGUMENT
        for n in (0, (1 << 64) - 1, limit + 1):
            assert entry(*good, n, 16, None) == cudec_amd.CUDEC_ERR_INVALID_ARGUMENT


def test_stream_rejects_without_a_device_call():
    src = np.zeros(4, np.uint8)
    ptrs = np.array([src.ctypes.data], np.uint64)
    null_ptrs = np.array([0], np.uint64)
    sizes = np.array([1], np.uint64)
    caps = np.array([4], np.uint64)
    res = np.zeros(2, np.uint64)
    P = lambda a: a.ctypes.data  # noqa: E731
    f = cudec_amd.cudec_lz4_decompress_stream
    I = cudec_amd.CUDEC_ERR_INVALID_ARGUMENT
    assert f(None, P(sizes), P(ptrs), P(caps), 1, 0, P(res)) == I
    assert f(P(ptrs), P(sizes), P(ptrs), P(caps), 0, 0, P(res)) == I
    assert f(P(ptrs), P(sizes), P(ptrs), P(caps), 1, 2, P(res)) == I          # unknown space
    assert f(P(null_ptrs), P(sizes), P(ptrs), P(caps), 1, 0, P(res)) == I     # NULL src, size 1
    assert f(P(ptrs), P(sizes), P(null_ptrs), P(caps), 1, 0, P(res)) == I     # NULL dst, cap 4
    assert f(P(ptrs), P(sizes), P(ptrs), P(caps), cudec_amd.max_batch_chunks() + 1, 0, P(res)) == I
    assert cudec_amd.cudec_lz4_decompress_stream_ctx(None, P(ptrs), P(sizes), P(ptrs), P(caps),
                                                     1, 0, P(res)) == I


def test_no_visible_device_reports_device_error():
    """tests/launch_fail.cpp: with no visible device, a call that passes
    validation returns CUDEC_ERR_CUDA (twice: nothing sticks), and so do
    context creation and the one-shot streaming entry."""
    code = textwrap.dedent("""
        import numpy as np, cudec_amd
        E = cudec_amd.CUDEC_ERR_CUDA
        for _ in range(2):
            assert cudec_amd.cudec_lz4_decompress_batch(16, 16, 16, 16, 1, 16, None) == E
            assert cudec_amd.cudec_snappy_decompress_batch_plain(16, 16, 16, 16, 1, 16, None) == E
        st, h = cudec_amd.cudec_stream_ctx_create()
        assert st == E and not h
        src = np.zeros(4, np.uint8); dst = np.zeros(4, np.uint8)
        p = np.array([src.ctypes.data], np.uint64); d = np.array([dst.ctypes.data], np.uint64)
        one = np.array([1], np.uint64); cap = np.array([4], np.uint64); res = np.zeros(2, np.uint64)
        assert cudec_amd.cudec_lz4_decompress_stream(p.ctypes.data, one.ctypes.data, d.ctypes.data,
                                                     cap.ctypes.data, 1, 0, res.ctypes.data) == E
        print("ok")
    """)
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1", ROCR_VISIBLE_DEVICES="-1",
               CUDA_VISIBLE_DEVICES="-1")
    run = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                         timeout=300, cwd=ROOT, env=env)
    assert run.returncode == 0 and "ok" in run.stdout, run.stderr[-2000:]


def _lz4f(frame: bytes, cap: int):
    out = ctypes.create_string_buffer(max(cap, 1))
    w = ctypes.c_size_t(99)
    lib = cudec_amd.lib()
    lib.cudec_lz4f_decompress.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_void_p,
                                          ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
    st = lib.cudec_lz4f_decompress(frame, len(frame), out, cap, ctypes.byref(w))
    return st, w.value, out.raw[: w.value]


needs_lz4 = pytest.mark.skipif(libs.lz4lib() is None, reason="liblz4 absent")


@needs_lz4
def test_frame_host_verdicts():
    """tests/frame_host_negative.cpp: the verdicts the parser reaches before
    any block is decoded."""
    from tests import corpus
    data = corpus.random_bytes(5000, 7)
    good = libs.lz4_frame(data, content_checksum=True, block_checksum=True, content_size=True)
    C, U = cudec_amd.CUDEC_ERR_CORRUPT_INPUT, cudec_amd.CUDEC_ERR_UNSUPPORTED
    assert _lz4f(b"\x04\x22\x4d\x19", 100)[0] == C                 # short
    assert _lz4f(b"\x05" + good[1:], 9000)[0] == C                  # magic
    bad = bytearray(good); bad[4] ^= 0x80                            # version bits
    assert _lz4f(bytes(bad), 9000)[0] == C
    bad = bytearray(good); bad[4] |= 0x02                            # FLG reserved bit
    assert _lz4f(bytes(bad), 9000)[0] == C
    bad = bytearray(good); bad[5] |= 0x01                            # BD reserved bit
    assert _lz4f(bytes(bad), 9000)[0] == C
    bad = bytearray(good); bad[5] = (bad[5] & 0x8F) | (3 << 4)       # block max id 3
    assert _lz4f(bytes(bad), 9000)[0] == C
    bad = bytearray(good); bad[14] ^= 1                              # header checksum byte
    assert _lz4f(bytes(bad), 9000)[0] == C
    # liblz4 marks a frame that fits one block independent by itself, so
    # the linked frame spans several blocks
    big = corpus.random_bytes(200000, 8)
    linked = libs.lz4_frame(big, independent=False)
    assert linked[4] & 0x20 == 0
    assert _lz4f(linked, len(big))[0] == U
    assert _lz4f(good[:-9], 9000)[0] == C                            # no end mark
    st, w, out = _lz4f(b"", 0)
    assert (st, w) == (C, 0)


@needs_lz4
@pytest.mark.parametrize("flags", [(False, False, False), (True, False, True), (True, False, False),
                                   (False, False, True)])
def test_stored_block_frames_decode_on_host(flags):
    """Incompressible data makes liblz4 store every block verbatim, so a
    frame without block checksums decodes with no device work at all: this
    pins XXH32 (header and content checksums) and the block walk against
    liblz4's frames. Block checksums are checked on the device
    (tests/test_gpu_frame.py)."""
    from tests import corpus
    cs, bs, sz = flags
    data = corpus.random_bytes(300000, 11)
    frame = libs.lz4_frame(data, block_size_id=4, content_checksum=cs, block_checksum=bs,
                           content_size=sz)
    st, w, out = _lz4f(frame, len(data))
    assert (st, w) == (0, len(data)) and out == data
    assert _lz4f(frame, len(data) - 1)[:2] == (cudec_amd.CUDEC_ERR_OUTPUT_TOO_SMALL, 0)
    if cs:
        bad = bytearray(frame); bad[-1] ^= 0x40                       # content checksum
        assert _lz4f(bytes(bad), len(data))[:2] == (cudec_amd.CUDEC_ERR_CORRUPT_INPUT, 0)
    if bs:
        bad = bytearray(frame); bad[len(frame) // 2] ^= 0x01          # inside a block
        assert _lz4f(bytes(bad), len(data))[:2] == (cudec_amd.CUDEC_ERR_CORRUPT_INPUT, 0)


@needs_lz4
def test_test_side_xxh32_is_pinned():
    """The test-side XXH32 (used to repair header checksums of mutated
    frames) against the specification's empty-input value and against
    liblz4's header and content checksums over every tail length."""
    assert libs.xxh32(b"") == 0x02CC5D05
    from tests import corpus
    for n in list(range(0, 40)) + [1000, 65537]:
        d = corpus.random_bytes(n, n)
        f = libs.lz4_frame(d, content_checksum=True, content_size=True)
        hc = 14 if f[4] & 0x08 else 6     # a zero size is written as "unknown"
        assert (libs.xxh32(f[4:hc]) >> 8) & 0xFF == f[hc]
        assert int.from_bytes(f[-4:], "little") == libs.xxh32(d), n


def test_serve_rejects_without_a_device_call():
    """cudec_mi355x_serve_lz4's synchronous argument checks."""
    from cudec_amd import serve
    L = cudec_amd.lib()
    src = np.array([4096, 8192], np.uint64)
    sizes = np.array([10, 10], np.uint64)
    caps = np.array([1000, 1000], np.uint64)
    res = np.zeros(2, dtype=serve.RESULT_DTYPE)
    lat = np.zeros(2)

    def call(req, n=2, budget=1 << 20, inflight=2, s=src, z=sizes, c=caps):
        req = np.ascontiguousarray(req, dtype=serve.REQUEST_DTYPE)
        return L.cudec_mi355x_serve_lz4(s.ctypes.data if s is not None else None, z.ctypes.data,
                                        c.ctypes.data, n, req.ctypes.data, len(req), budget,
                                        inflight, res.ctypes.data, None, lat.ctypes.data, None)

    def reqs(*rows):
        r = np.zeros(len(rows), dtype=serve.REQUEST_DTYPE)
        for i, (a, f, k) in enumerate(rows):
            r[i] = (a, f, k)
        return r

    I = cudec_amd.CUDEC_ERR_INVALID_ARGUMENT
    good = reqs((0.0, 0, 1), (0.1, 1, 1))
    assert call(good, s=None) == I
    assert call(good, n=0) == I
    assert call(good, inflight=0) == I and call(good, inflight=17) == I
    assert call(reqs((0.0, 0, 3))) == I                    # range past the tables
    assert call(reqs((0.0, 2, 1))) == I
    assert call(reqs((0.0, 0, 0))) == I                    # empty request
    assert call(reqs((0.2, 0, 1), (0.1, 1, 1))) == I       # arrivals out of order
    assert call(reqs((float("nan"), 0, 1))) == I
    assert call(good, budget=1000) == I                    # 1000 rounds up to 1024 > budget
    assert call(reqs((0.0, 0, 2)), budget=1024 + 1023) == I  # two chunks need 2048


def test_host_digest_definition():
    """The host digest against values worked out by hand from its definition
    (splitmix64 finaliser over position-mixed words)."""
    from cudec_amd import serve

    def mix(x):
        M = (1 << 64) - 1
        x ^= x >> 30
        x = (x * 0xBF58476D1CE4E5B9) & M
        x ^= x >> 27
        x = (x * 0x94D049BB133111EB) & M
        x ^= x >> 31
        return x

    assert serve.host_digest(b"") == 0
    assert serve.host_digest(b"\x01") == (0x9E3779B97F4A7C15 + mix(1)) & ((1 << 64) - 1)
    nine = bytes(range(1, 10))
    w0 = int.from_bytes(nine[:8], "little")
    w1 = 9 + 0xD6E8FEB86659FD93
    want = (9 * 0x9E3779B97F4A7C15 + mix(w0) + mix(w1)) & ((1 << 64) - 1)
    assert serve.host_digest(nine) == want
    assert serve.host_digest(b"ab") != serve.host_digest(b"ba")


def test_tuning_entries_reject_without_a_device_call():
    I = cudec_amd.CUDEC_ERR_INVALID_ARGUMENT
    assert cudec_amd.autotune_lz4(16, 16, 16, 16, 1, 24, None)[0] == I     # misaligned results
    assert cudec_amd.autotune_lz4(None, 16, 16, 16, 1, 16, None)[0] == I
    assert cudec_amd.autotune_lz4(16, 16, 16, 16, 0, 16, None)[0] == I
    assert cudec_amd.lib().cudec_mi355x_lz4_pipeline_threshold(None) == I
    assert cudec_amd.lib().cudec_mi355x_snappy_pipeline_threshold(None) == I
    assert cudec_amd.autotune_plan("lz4", 16, 16, 16, 16, 1, 24, None)[0] == I
    assert cudec_amd.autotune_plan("snappy", None, 16, 16, 16, 1, 16, None)[0] == I
    assert cudec_amd.lib().cudec_mi355x_autotune_plan(7, 16, 16, 16, 16, 1, 16, None,
                                                      16, None, None) == I
    L = cudec_amd.lib()
    ok = (ctypes.c_uint16 * cudec_amd.COST_BINS)(*([64] * cudec_amd.COST_BINS))
    bad = (ctypes.c_uint16 * cudec_amd.COST_BINS)(*([64] * 255 + [96]))
    assert L.cudec_mi355x_set_decode_plan(2, ok) == I          # unknown format
    assert L.cudec_mi355x_set_decode_plan(0, None) == I
    assert L.cudec_mi355x_set_decode_plan(0, bad) == I          # 96 lanes is no plan entry
    assert L.cudec_mi355x_decode_plan(1, None) == I


def test_cost_bins_follow_the_cost_order():
    """cost = capacity + (capacity - compressed): four bins per octave,
    bin 0 the costliest, and a costlier chunk never gets a larger bin."""
    cb = cudec_amd.cost_bin
    assert cb(0, 0) == 255 and cb(1, 0) == 255
    assert cb(0, 1) == 253                       # cost 2
    costs = sorted({c for k in range(2, 40) for c in (1 << k, (1 << k) + (1 << (k - 2)),
                                                       (1 << k) + 3 * (1 << (k - 2)))})
    bins = [cb(c, c) for c in costs]            # compressed == capacity: cost == capacity
    assert bins == sorted(bins, reverse=True)
    assert cb(1 << 20, 1 << 20) - cb(1 << 21, 1 << 21) == 4    # one octave = four bins
    # matches cost time: a 4 MiB chunk that compresses to 256 KiB is costlier
    # than a 4 MiB literal-only one
    assert cb(256 << 10, 4 << 20) < cb(4 << 20, 4 << 20)


def test_noalloc_rejects_without_a_device_call():
    """The allocation-free entry takes the batch entry's argument checks and
    refuses a batch above its fixed scratch, all before any device call."""
    fn = cudec_amd.lib().cudec_mi355x_lz4_decompress_batch_noalloc
    I = cudec_amd.CUDEC_ERR_INVALID_ARGUMENT
    assert fn(16, 16, 16, 16, (1 << 20) + 1, 16, None) == I
    assert fn(16, 16, 16, 16, 0, 16, None) == I
    assert fn(None, 16, 16, 16, 1, 16, None) == I
    assert fn(16, 16, 16, 16, 1, 24, None) == I


def test_parse_batch_rejects_without_a_device_call():
    fn = cudec_amd.lib().cudec_mi355x_parse_batch
    I = cudec_amd.CUDEC_ERR_INVALID_ARGUMENT
    assert fn(2, 16, 16, 16, 16, 1, 16, None) == I             # unknown format
    assert fn(0, 16, 16, 16, 16, 0, 16, None) == I
    assert fn(1, None, 16, 16, 16, 1, 16, None) == I
    assert fn(0, 16, 16, 16, 16, 1, 24, None) == I
    assert fn(0, 16, 16, 16, 16, 1 << 32, 16, None) == I        # the queue indexes in 32 bits


def test_test_side_xxh64_is_pinned():
    """The specification's empty-input value, and the system's libxxhash
    (an implementation independent of this build) on lengths that reach
    every tail path, at two seeds; the same for the test-side XXH32."""
    assert libs.xxh64(b"") == 0xEF46DB3751D8E999
    x = libs.xxhash_lib()
    if x is None:
        pytest.skip("libxxhash absent")
    rng = np.random.default_rng(5)
    for n in list(range(0, 80)) + [255, 256, 1000, 4099]:
        d = rng.integers(0, 256, n, dtype=np.uint8).tobytes()
        for seed in (0, 0x9E3779B1):
            assert libs.xxh64(d, seed) == x.XXH64(d, n, seed), (n, seed)
            assert libs.xxh32(d, seed) == x.XXH32(d, n, seed), (n, seed)


def test_checksum_entries_reject_without_a_device_call():
    L = cudec_amd.lib()
    I = cudec_amd.CUDEC_ERR_INVALID_ARGUMENT
    for fn in (L.cudec_mi355x_xxh32_batch, L.cudec_mi355x_xxh64_batch):
        assert fn(None, 16, 1, 0, 16, None) == I
        assert fn(16, None, 1, 0, 16, None) == I
        assert fn(16, 16, 0, 0, 16, None) == I
        assert fn(16, 16, 1, 0, None, None) == I
