"""GPU parity for LZ4 sequences a window cannot hold, against the CPU oracle,
on every launch shape (all chunks on two pipelined waves, all on one wave
each).

A sequence whose literal or match length takes two or more extension bytes
(literal >= 270 bytes, match >= 275) is decided by the scalar ladder and
executed on its own: a literal of up to 64 bytes goes into the output ring
and a longer one runs global to global; a match whose period (offset up to
1024) is in the ring is expanded there in pieces (3040 bytes on the 4 KiB
ring of the one-wave form, 7136 on the 8 KiB ring of the two-wave form),
flushing finished blocks between pieces; any other match runs global to
global. The streams put such sequences at a chunk's very first output
bytes (where an aligned block address precedes the destination), after
literal prefixes of many lengths, back to back, and between ordinary
sequences, with match lengths on both sides of the piece sizes, offsets on
both sides of 1024 and of the ring sizes, and literal lengths on both sides
of 64, and often with a burst of short sequences right after a slow match,
so that the next window's group (up to its batch limit) follows a piece
still in the ring. Destinations start at every alignment. Capacities are
exact, and one byte short for a share of chunks.
"""
import numpy as np
import pytest

import cudec_amd
from tests import libs
from tests.gpu_util import check_against_oracle, gpu, oracle_decode
from tests.test_gpu_copy_shapes import last_literals, sequence

pytestmark = gpu

SLOW_MATCH = (275, 300, 1000, 3039, 3040, 3041, 3060, 7136, 7137, 7200, 20000)
OFFSETS = (1, 2, 3, 5, 8, 15, 16, 17, 31, 64, 100, 1023, 1024, 1025, 4000, 4096, 8192, 20000)
LITS = (0, 0, 1, 5, 63, 64, 65, 300)


def slow_stream(seed: int) -> bytes:
    rng = np.random.default_rng(seed)
    out = bytearray()
    pos = 0
    # first sequence: a short literal prefix and, often, a slow match at once
    first = int(rng.choice((1, 2, 3, 5, 16, 17, 70, 300, 3000)))
    lits = rng.integers(0, 256, first, dtype=np.uint8).tobytes()
    off = int(rng.integers(1, first + 1))
    m = int(rng.choice(SLOW_MATCH)) if rng.integers(0, 3) else 4
    out += sequence(lits, off, m)
    pos = first + m
    for _ in range(int(rng.integers(2, 9))):
        if rng.integers(0, 3) == 0:
            # an ordinary sequence between the slow ones
            L = int(rng.integers(0, 12))
            m = int(rng.integers(4, 40))
        else:
            L = int(rng.choice(LITS))
            m = int(rng.choice(SLOW_MATCH))
        off = int(rng.choice(OFFSETS))
        off = max(1, min(off, pos + L))
        out += sequence(rng.integers(0, 256, L, dtype=np.uint8).tobytes(), off, m)
        pos += L + m
        if rng.integers(0, 2) == 0:
            # a burst of short sequences right after: one window's group
            # fills its batch limit while the slow match's last piece is
            # still in the ring
            for _ in range(int(rng.integers(40, 200))):
                lits = rng.integers(0, 256, 4, dtype=np.uint8).tobytes()
                off = int(rng.integers(1, min(pos + 4, 3000) + 1))
                out += sequence(lits, off, 36)
                pos += 40
    out += last_literals(rng.integers(0, 256, int(rng.choice((5, 16, 300))),
                                      dtype=np.uint8).tobytes())
    return bytes(out)


@pytest.fixture(scope="module")
def slow_chunks():
    streams, caps, names = [], [], []
    for k in range(160):
        s = slow_stream(0x510 + k)
        st, out = oracle_decode("lz4", s, 1 << 22)
        assert st == libs.OK, f"generator wrote an invalid stream ({k})"
        streams.append(s)
        # exact capacity, and one byte short for every fifth chunk
        caps.append(len(out) - (1 if k % 5 == 4 else 0))
        names.append(f"slow{k}")
    return streams, caps, names


@pytest.fixture()
def restore_threshold():
    st, old = cudec_amd.lz4_pipeline_threshold()
    assert st == 0
    yield
    assert cudec_amd.set_lz4_pipeline_threshold(old) == 0


@pytest.mark.parametrize("threshold", [0, (1 << 64) - 1])
def test_slow_sequences_exact(slow_chunks, restore_threshold, threshold):
    streams, caps, names = slow_chunks
    assert cudec_amd.set_lz4_pipeline_threshold(threshold) == 0
    print(f"[slow sequences] {len(streams)} chunks, threshold {threshold}", flush=True)
    res = check_against_oracle("lz4", "tuned", streams, caps, names)
    assert np.sum(res["status"] == 0) >= 100 and np.sum(res["status"] != 0) >= 20
