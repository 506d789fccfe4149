"""GPU parity for the edges of a window's member chain in the tuned LZ4
kernel, against the CPU oracle, on both launch shapes (two pipelined waves
walk the chain one member at a time; one wave per chunk walks it by pointer
doubling).

A window's chain ends after 64 members, at the window's end, or at a
sequence whose length needs two or more extension bytes (a stop: that
sequence is not a member and the next window starts with it). The streams
here put runs of n minimal sequences (3 to 4 bytes each, so a 512- or
1024-byte window holds far more than 64) in front of such a sequence, for
every n from 0 to 70, after literal prefixes that move the window
boundaries: the stop lands on member 0..70, including exactly on the 64th
and 65th position of a chain. Every stream is valid LZ4 (the oracle accepts
it) and ends with the format's literal tail.

The Snappy streams do the same with 2-byte copy elements (32 to a 64-byte
region of a window, where an LZ4 sequence takes at least 3 bytes) in front
of a literal too long for a window member, so a chain passes 32 and 64
members inside one region.
"""
import numpy as np
import pytest

import cudec_amd
from tests import libs
from tests.gpu_util import check_against_oracle, gpu, oracle_decode
from tests.test_gpu_copy_shapes import last_literals, sequence

pytestmark = gpu


def walk_stream(seed: int, prefix: int, long_literal: bool) -> bytes:
    """`prefix` literal bytes, then for n = 0..70: n minimal sequences
    followed by one whose literal (long_literal) or match length takes two
    extension bytes."""
    rng = np.random.default_rng(seed)
    out = bytearray(sequence(rng.integers(0, 256, prefix - 4, dtype=np.uint8).tobytes(), 1, 4))
    pos = prefix
    for n in range(71):
        for _ in range(n):
            L = int(rng.integers(0, 2))
            m = int(rng.integers(4, 19))
            off = int(rng.integers(1, min(pos + L, 300) + 1))
            out += sequence(rng.integers(0, 256, L, dtype=np.uint8).tobytes(), off, m)
            pos += L + m
        if long_literal:
            L, m = 300, 4
        else:
            L, m = 1, 600
        off = int(rng.integers(1, min(pos + L, 5000) + 1))
        out += sequence(rng.integers(0, 256, L, dtype=np.uint8).tobytes(), off, m)
        pos += L + m
    out += last_literals(rng.integers(0, 256, 16, dtype=np.uint8).tobytes())
    return bytes(out)


@pytest.fixture(scope="module")
def chains():
    streams, caps, names = [], [], []
    for k, prefix in enumerate((8, 9, 13, 64, 100, 257, 1000, 4100)):
        for long_literal in (False, True):
            s = walk_stream(0x3A1C + 31 * k + long_literal, prefix, long_literal)
            st, out = oracle_decode("lz4", s, 1 << 20)
            assert st == libs.OK, f"generator wrote an invalid stream (prefix {prefix})"
            streams.append(s)
            caps.append(len(out))
            names.append(f"prefix{prefix}/{'literal' if long_literal else 'match'}-stop")
    return streams, caps, names


def snappy_walk_stream(seed: int, prefix: int) -> bytes:
    """A `prefix`-byte literal, then for n = 0..70: n copy-1 elements of
    4..11 bytes (2 bytes each) and one 300-byte literal."""
    rng = np.random.default_rng(seed)
    body = bytearray()
    pos = 0

    def literal(data: bytes):
        nonlocal pos
        n = len(data) - 1
        if n < 60:
            body.append(n << 2)
        elif n < 256:
            body.extend((60 << 2, n))
        else:
            body.extend((61 << 2, n & 0xFF, n >> 8))
        body.extend(data)
        pos += len(data)

    literal(rng.integers(0, 256, prefix, dtype=np.uint8).tobytes())
    for n in range(71):
        for _ in range(n):
            length = int(rng.integers(4, 12))
            off = int(rng.integers(1, min(pos, 2047) + 1))
            body += bytes([((length - 4) << 2) | ((off >> 8) << 5) | 1, off & 0xFF])
            pos += length
        literal(rng.integers(0, 256, 300, dtype=np.uint8).tobytes())
    head = bytearray()
    v = pos
    while v >= 0x80:
        head.append((v & 0x7F) | 0x80)
        v >>= 7
    head.append(v)
    return bytes(head + body)


@pytest.fixture(scope="module")
def snappy_chains():
    streams, caps, names = [], [], []
    for k, prefix in enumerate((1, 2, 7, 64, 100, 1000)):
        s = snappy_walk_stream(0x5A17 + 13 * k, prefix)
        st, out = oracle_decode("snappy", s, 1 << 20)
        assert st == libs.OK, f"generator wrote an invalid stream (prefix {prefix})"
        streams.append(s)
        caps.append(len(out))
        names.append(f"snappy-prefix{prefix}")
    return streams, caps, names


@pytest.fixture()
def restore_threshold():
    st, old = cudec_amd.lz4_pipeline_threshold()
    assert st == 0
    yield
    assert cudec_amd.set_lz4_pipeline_threshold(old) == 0


@pytest.mark.parametrize("threshold", [0, (1 << 64) - 1])
def test_chain_edges_exact(chains, restore_threshold, threshold):
    streams, caps, names = chains
    assert cudec_amd.set_lz4_pipeline_threshold(threshold) == 0
    print(f"[walk edges] {len(streams)} chunks, threshold {threshold}", flush=True)
    check_against_oracle("lz4", "tuned", streams, caps, names)


@pytest.mark.parametrize("threshold", [0, (1 << 64) - 1])
def test_snappy_chain_edges_exact(snappy_chains, restore_threshold, threshold):
    streams, caps, names = snappy_chains
    assert cudec_amd.set_lz4_pipeline_threshold(threshold) == 0
    print(f"[walk edges] snappy {len(streams)} chunks, threshold {threshold}", flush=True)
    check_against_oracle("snappy", "tuned", streams, caps, names)
