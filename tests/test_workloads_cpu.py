"""The benchmark workloads are pinned: every configuration's generator output
(FNV-1a 64 of the originals, at a small chunk count) and its compressed
bytes (with the compressor version they were taken with), and the token
process's draw statistics and compression ratio. A change in sampling or in
the compressor moves a digest here before it moves a benchmark number.

Configurations (chunk sizes and mixes as the benchmark uses them):
* CPU and plain-GPU batches: token-process chunks of exactly 64 KiB;
* uneven batch: log-uniform sizes, 80/10/10 token/random/repetitive, LZ4;
* Snappy batch: the same plan through snappy_compress;
* serving pool: token-process chunks log-uniform over [4 KiB, 64 KiB) here.
"""
import ctypes

import numpy as np
import pytest

from cudec_amd import serve, workload

SEED = 0x5EED


def digests(b):
    return (hex(workload.fnv1a64(b.orig)), hex(workload.fnv1a64(b.comp)), b.total_compressed)


CASES = {
    "token64k x8": (lambda c: workload.token64k(8, seed=SEED, codec=c), workload.CODEC_LZ4,
                    ("0x75a66be2d29e7cac", "0x944858e906269f3f", 238593)),
    "uneven x24": (lambda c: workload.uneven(24, seed=SEED, octaves=8, codec=c),
                   workload.CODEC_LZ4, ("0xc734d212abbe3aa4", "0x1582d4a8aca9878e", 522740)),
    "uneven snappy x24": (lambda c: workload.uneven(24, seed=SEED, octaves=8, codec=c),
                          workload.CODEC_SNAPPY,
                          ("0xc734d212abbe3aa4", "0x87f5217d5ff52d3a", 514901)),
}
# the compressed digests were taken with these library builds
PINNED_VERSIONS = {workload.CODEC_LZ4: "liblz4 1.9.3",
                   workload.CODEC_SNAPPY: "libsnappy (/opt/conda/lib/libsnappy.so.1.1.8)"}


@pytest.mark.parametrize("name", sorted(CASES))
def test_configuration_digests(name):
    make, codec, (want_orig, want_comp, want_size) = CASES[name]
    if not workload.codec_available(codec):
        pytest.skip(f"compressor absent: {workload.codec_version(codec)}")
    got_orig, got_comp, got_size = digests(make(codec))
    assert got_orig == want_orig, "the generator's output moved"
    if workload.codec_version(codec) == PINNED_VERSIONS[codec]:
        assert (got_comp, got_size) == (want_comp, want_size), "the compressed bytes moved"


def test_serving_pool_digest():
    p = serve.chunk_pool(16, seed=SEED ^ 0x5E5E, octaves=4)
    got = digests(p)
    assert got[0] == "0xbdc03c567a69370c"
    if workload.codec_version(workload.CODEC_LZ4) == PINNED_VERSIONS[workload.CODEC_LZ4]:
        assert got[1:] == ("0x395591e636d8abd1", 127850)


def test_token_process_draws():
    """Over 64 chunks of 64 KiB: literal runs at 45% of the steps, copies at
    offsets in [1, 16] at 30% of the copies, runs of 1 + Geometric(mean 12)
    bytes and copies of 4 + Geometric(mean 20), within sampling error."""
    W = workload.wl()
    W.wl_gen_chunk_stats.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64,
                                     ctypes.c_int32, ctypes.c_void_p]
    tot = np.zeros(6, dtype=np.uint64)
    for i in range(64):
        out = np.empty(65536, dtype=np.uint8)
        st = np.zeros(6, dtype=np.uint64)
        W.wl_gen_chunk_stats(out.ctypes.data, 65536, W.wl_chunk_seed(SEED, i),
                             workload.KIND_TOKENS, st.ctypes.data)
        plain = np.empty(65536, dtype=np.uint8)
        W.wl_gen_chunk(plain.ctypes.data, 65536, W.wl_chunk_seed(SEED, i), workload.KIND_TOKENS)
        assert np.array_equal(out, plain)
        tot += st
    steps, runs, lit_bytes, copies, copy_bytes, short = (int(x) for x in tot)
    assert runs + copies == steps
    assert abs(runs / steps - 0.45) < 0.005
    assert abs(short / copies - 0.30) < 0.005
    assert abs(lit_bytes / runs - 13.0) < 0.1
    assert abs(copy_bytes / copies - 24.0) < 0.2


def test_token64k_compression_ratio():
    """The CPU configuration (1024 x 64 KiB token chunks) compresses about
    2.1x with LZ4_compress_default; measured 2.18 with liblz4 1.9.3."""
    if not workload.codec_available(workload.CODEC_LZ4):
        pytest.skip("liblz4 absent")
    b = workload.token64k(1024, seed=SEED)
    ratio = b.total_uncompressed / b.total_compressed
    assert 2.0 < ratio < 2.35, ratio
