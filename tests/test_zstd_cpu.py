"""The Zstandard decoder's frame logic on the CPU: cudec_amd/csrc/zstd_decode.h
built for one lane (tests/zstd_host.cpp, with the undefined-behaviour
sanitizer), held to libzstd 1.4.9 for bytes and to the reference's own
frame twin (tests/zstd_twin_driver.h compiled from the reference tree into
oracle/_ref) for every verdict, over a corpus that reaches every block,
literals and table mode, hand-made header frames, and seeded mutants.

The content checksum is compared here, as the device's second pass does:
the frame logic returns the checksum the frame carries, and the XXH64 of
the output decides."""
import ctypes
import os
import subprocess

import pytest

from tests import libs
from tests import zstd_corpus as zc

ROOT = libs.ROOT
BUILD = os.path.join(ROOT, "tests", "_build")

needs_zstd = pytest.mark.skipif(libs.zstdlib() is None, reason="libzstd absent")
needs_twin = pytest.mark.skipif(libs.ref_zstd() is None,
                                reason="reference twin not built (oracle/_ref)")


@pytest.fixture(scope="module")
def host():
    os.makedirs(BUILD, exist_ok=True)
    out = os.path.join(BUILD, "libzstd_host.so")
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-Wall", "-Wextra", "-Werror",
                    "-fsanitize=undefined", "-fno-sanitize-recover=all", "-fPIC", "-shared",
                    "-I", os.path.join(ROOT, "include"),
                    "-I", os.path.join(ROOT, "cudec_amd", "csrc"),
                    "-o", out, os.path.join(ROOT, "tests", "zstd_host.cpp")],
                   check=True)
    lib = ctypes.CDLL(out)
    lib.host_zstd_decode.argtypes = [ctypes.c_char_p, ctypes.c_uint64, ctypes.c_void_p,
                                     ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64),
                                     ctypes.POINTER(ctypes.c_uint32),
                                     ctypes.POINTER(ctypes.c_int32)]
    lib.host_zstd_decode.restype = ctypes.c_int32
    return lib


def xxh64(data: bytes) -> int:
    x = libs.xxhash_lib()
    if x is not None:
        return x.XXH64(data, len(data), 0)
    return libs.xxh64(data)


def host_decode(lib, frame: bytes, cap: int):
    """(status, output) with the content checksum compared."""
    buf = ctypes.create_string_buffer(max(cap, 1))
    w, ck, has = ctypes.c_uint64(0), ctypes.c_uint32(0), ctypes.c_int32(0)
    st = lib.host_zstd_decode(frame, len(frame), buf, cap, ctypes.byref(w), ctypes.byref(ck),
                              ctypes.byref(has))
    if st != libs.OK:
        assert w.value == 0
        return st, b""
    out = buf.raw[:w.value]
    if has.value and (xxh64(out) & 0xFFFFFFFF) != ck.value:
        return libs.CORRUPT, b""
    return st, out


def test_code_tables_closed_form(host):
    """Each literal- and match-length baseline is the previous baseline plus
    two to the previous extra-bit width; the direct codes are the value
    itself (literals) and the value plus three (matches)."""
    ll_base = (ctypes.c_uint32 * 36)()
    ll_bits = (ctypes.c_uint8 * 36)()
    ml_base = (ctypes.c_uint32 * 53)()
    ml_bits = (ctypes.c_uint8 * 53)()
    host.host_zstd_tables(ll_base, ll_bits, ml_base, ml_bits)
    assert [ll_base[c] for c in range(16)] == list(range(16))
    assert [ml_base[c] for c in range(32)] == [c + 3 for c in range(32)]
    for base, bits, n in ((ll_base, ll_bits, 36), (ml_base, ml_bits, 53)):
        for c in range(n - 1):
            assert base[c + 1] == base[c] + (1 << bits[c]), c
    assert ll_bits[35] == 16 and ml_bits[52] == 16


@needs_zstd
def test_corpus_reaches_every_surface():
    """Raw, run-length and compressed blocks; raw, run-length, Huffman and
    treeless literals; one and four streams; and all four table modes of
    each sequence field."""
    frames = [f for _, f, _ in zc.corpus()] + [f for _, f, _, _ in zc.crafted()]
    missing = zc.ALL_SURFACES - zc.surfaces(frames)
    assert not missing, sorted(missing)


@needs_zstd
@needs_twin
def test_corpus_decodes_like_libzstd_and_the_twin(host):
    for name, frame, orig in zc.corpus():
        ok, z = libs.zstd_decompress(frame, len(orig))
        assert ok and z == orig, name
        st, out = libs.ref_zstd_decode(frame, len(orig))
        assert st == libs.OK and out == orig, f"twin: {name}"
        st, out = host_decode(host, frame, len(orig))
        assert st == libs.OK, f"{name}: status {st}"
        assert out == orig, name
        # one byte short of the content size
        if orig:
            assert host_decode(host, frame, len(orig) - 1)[0] == libs.TOO_SMALL, name


@needs_zstd
@needs_twin
def test_compressor_rejects_are_refused(host):
    rejects = zc.compressor_rejects()
    for name, frame, orig in rejects:
        want = libs.ref_zstd_decode(frame, len(orig))[0]
        assert want != libs.OK, name
        assert host_decode(host, frame, len(orig))[0] == want, name


@needs_zstd
@needs_twin
def test_crafted_frames(host):
    for name, frame, cap, want in zc.crafted():
        assert libs.ref_zstd_decode(frame, cap)[0] == want, f"twin: {name}"
        st, out = host_decode(host, frame, cap)
        assert st == want, f"{name}: {st} != {want}"
        if want == libs.OK:
            assert out == libs.ref_zstd_decode(frame, cap)[1], name


@needs_zstd
@needs_twin
@pytest.mark.parametrize("seed", [1, 2])
def test_mutants_match_the_twin(host, seed):
    frames = [c for c in zc.corpus(max_size=65536) if len(c[1]) > 8]
    counts = {}
    for name, frame, cap in zc.mutants(frames, per_frame=12, seed=seed):
        want, want_out = libs.ref_zstd_decode(frame, cap)
        st, out = host_decode(host, frame, cap)
        assert st == want, f"{name}: {st} != twin {want}"
        if st == libs.OK:
            assert out == want_out, name
        counts[st] = counts.get(st, 0) + 1
    # the mutants reach accepts, corruption and short capacities
    assert counts.get(libs.OK, 0) and counts.get(libs.CORRUPT, 0) and counts.get(libs.TOO_SMALL, 0)
