"""GPU parity for the Snappy batch decoders, through the C ABI, against the
CPU oracle: the reference's Snappy corpus and mutants
(tests/fixtures.cpp:289-441), its crafted negatives, capacity negatives and
accept set (tests/snappy_parser_twin.cpp:226-530), short streams, and
seeded generated batches with malformed chunks. Exact status, byte and
poison parity on every chunk."""
import numpy as np
import pytest

from tests import corpus, libs, snappy_vectors as sv
from tests.gpu_util import FORMS, LAUNCHES, check_against_oracle, gpu, run_device

pytestmark = gpu


@pytest.mark.parametrize("form", LAUNCHES)
def test_fixtures_and_mutants(form, snappy_fixtures):
    streams, caps, names = [], [], []
    for f in snappy_fixtures:
        streams.append(f.compressed)
        caps.append(len(f.original) + 8)
        names.append(f.name)
        for m in corpus.mutate_snappy_stream(f.compressed, f.seed):
            streams.append(m.stream)
            caps.append(len(f.original) + 8)
            names.append(f"{f.name}/{m.description}")
    res = check_against_oracle("snappy", form, streams, caps, names)
    if libs.snappylib() is not None:
        # where the device accepts, snappy accepts with the same bytes
        for i, s in enumerate(streams):
            if res["status"][i] == 0:
                ok, out = libs.snappy_decompress(s)
                assert ok, names[i]


@pytest.mark.parametrize("form", LAUNCHES)
def test_crafted_capacity_and_accept_sets(form):
    streams, caps, names = [], [], []
    for c in sv.crafted_negatives():
        streams.append(c.stream)
        caps.append(1 << 17)
        names.append(c.name)
    for name, s, cap, _ in sv.capacity_negatives():
        streams.append(s)
        caps.append(cap)
        names.append(name)
    for name, s, exp in sv.accept_cases():
        streams.append(s)
        caps.append(len(exp) + 5)
        names.append(name)
    for off in (65535, 65536, 65537, 131072):
        s = sv.wide_offset_stream(off, 4)
        streams.append(s)
        caps.append(off + 4)
        names.append(f"wide-{off}")
    # the 32-bit literal-length wrap the reference's snappy accepts
    streams += [b"\x00\xFC\xFF\xFF\xFF\xFF", b"\x04\xFC\xFF\xFF\xFF\xFF\x0Cabcd"]
    caps += [16, 16]
    names += ["wrap", "wrap-then-literal"]
    res = check_against_oracle("snappy", form, streams, caps, names)
    n_neg = len(sv.crafted_negatives())
    assert np.all(res["status"][:n_neg] == libs.CORRUPT)


@pytest.mark.parametrize("form", LAUNCHES)
def test_short_streams_exhaustive(form):
    streams, caps = [b""], [0]
    for cap in (0, 1, 64):
        for a in range(256):
            streams.append(bytes([a]))
            caps.append(cap)
            for b in range(0, 256, 3):
                streams.append(bytes([a, b]))
                caps.append(cap)
    check_against_oracle("snappy", form, streams, caps)


@pytest.mark.parametrize("form", LAUNCHES)
def test_generated_batches_with_malformed_chunks(form):
    """Seeded uneven chunks through snappy_compress; a seeded 2% are
    truncated or have one tag-stream byte flipped; the malformed ones get
    the oracle's verdict, the rest decode to their originals."""
    from cudec_amd import workload
    b = workload.uneven(150, seed=0x91, codec=workload.CODEC_SNAPPY, min_size=1024, octaves=8)
    streams = [b.compressed(i) for i in range(b.n)]
    g = corpus.SplitMix64(0x92)
    for i in range(b.n):
        if g.next() % 100 < 8:
            s = bytearray(streams[i])
            if g.next() % 2 == 0:
                s = s[: int(g.next() % len(s))]
            else:
                s[len(s) // 2 + int(g.next() % (len(s) - len(s) // 2))] ^= 1 << (g.next() % 8)
            streams[i] = bytes(s)
    caps = [int(x) for x in b.sizes]
    check_against_oracle("snappy", form, streams, caps, poison=None)
