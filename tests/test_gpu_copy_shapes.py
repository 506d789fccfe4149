"""GPU parity for match and literal shapes that pick different copy paths in
the tuned LZ4 kernel, against the CPU oracle, on every launch shape
(all chunks on two pipelined waves, all on one wave each).

The streams are written sequence by sequence, so each shape appears many
times at many output alignments (the 4 KiB LDS ring wraps at a different
point of each): offsets 1..20 and around 16, 32, 64 and the ring size;
match lengths around 16, 19 (first extension byte), 64 (one lane's limit)
and beyond; literal runs of 0..20 bytes, so a match reads its own
literals, the previous member's match (chains of short offsets), earlier
output in the ring, or flushed history beyond it. Every stream is valid
LZ4 (the oracle accepts it) and ends with the format's literal tail.
"""
import numpy as np
import pytest

import cudec_amd
from tests import libs
from tests.gpu_util import check_against_oracle, gpu, oracle_decode

pytestmark = gpu

LIT_LENS = (0, 0, 0, 1, 3, 7, 15, 16, 20)
OFFSETS = tuple(range(1, 21)) + (31, 32, 33, 63, 64, 65, 100, 1000, 4095, 4096, 4097, 5000,
                                 20000, 65535)
MATCH_LENS = (4, 5, 8, 15, 16, 17, 18, 19, 20, 31, 32, 33, 48, 63, 64, 65, 80, 200)


def _ext(n: int) -> bytes:
    out = bytearray()
    while n >= 255:
        out.append(255)
        n -= 255
    out.append(n)
    return bytes(out)


def sequence(lits: bytes, off: int, mlen: int) -> bytes:
    L, M = len(lits), mlen - 4
    out = bytearray([(min(L, 15) << 4) | min(M, 15)])
    if L >= 15:
        out += _ext(L - 15)
    out += lits
    out += off.to_bytes(2, "little")
    if M >= 15:
        out += _ext(M - 15)
    return bytes(out)


def last_literals(lits: bytes) -> bytes:
    L = len(lits)
    out = bytearray([min(L, 15) << 4])
    if L >= 15:
        out += _ext(L - 15)
    return bytes(out + lits)


def shaped_stream(seed: int, prefix: int, target: int) -> bytes:
    """A literal prefix of `prefix` random bytes, then sequences of random
    shapes until about `target` output bytes, then a 16-byte literal tail."""
    rng = np.random.default_rng(seed)
    out = bytearray(sequence(rng.integers(0, 256, prefix, dtype=np.uint8).tobytes()[:-4],
                             1, 4)) if prefix >= 8 else bytearray()
    pos = prefix if prefix >= 8 else 0
    while pos < target:
        L = int(rng.choice(LIT_LENS))
        off = int(rng.choice(OFFSETS))
        m = int(rng.choice(MATCH_LENS))
        if pos + L == 0 or off > pos + L:
            off = max(1, min(off, pos + L))
        if pos + L == 0:
            L = 1
        lits = rng.integers(0, 256, L, dtype=np.uint8).tobytes()
        out += sequence(lits, off, m)
        pos += L + m
    out += last_literals(rng.integers(0, 256, 16, dtype=np.uint8).tobytes())
    return bytes(out)


@pytest.fixture(scope="module")
def shaped():
    streams, caps, names = [], [], []
    for k, prefix in enumerate((8, 9, 13, 16, 23, 31, 4100, 5003, 6007, 70001)):
        for target in (3000, 9000, 40000, 150000):
            s = shaped_stream(0x5A0 + 17 * k + target, prefix, target)
            st, out = oracle_decode("lz4", s, 1 << 20)
            assert st == libs.OK, f"generator wrote an invalid stream (prefix {prefix})"
            streams.append(s)
            caps.append(len(out))
            names.append(f"prefix{prefix}/target{target}")
    return streams, caps, names


@pytest.fixture()
def restore_threshold():
    st, old = cudec_amd.lz4_pipeline_threshold()
    assert st == 0
    yield
    assert cudec_amd.set_lz4_pipeline_threshold(old) == 0


@pytest.mark.parametrize("threshold", [0, (1 << 64) - 1])
def test_copy_shapes_exact(shaped, restore_threshold, threshold):
    streams, caps, names = shaped
    assert cudec_amd.set_lz4_pipeline_threshold(threshold) == 0
    print(f"[copy shapes] {len(streams)} chunks, threshold {threshold}", flush=True)
    check_against_oracle("lz4", "tuned", streams, caps, names)
