"""Pins the CPU LZ4 oracle (oracle/cudec_oracle.c) before anything is held to
it. The checks follow the reference's own twin test (tests/parser_twin.cpp),
its corpus pin (tests/oracle_lz4.cpp:20), its fuzz seeds
(fuzz/corpus/fuzz_lz4_block/*) and its termination test
(tests/termination.cpp); where the reference ladders were compiled from the
reference tree (oracle/_ref), the oracle must agree with them exactly.

Each family is chosen so a plausible restatement mistake fails it: a dropped
LASTLITERALS rule fails the last-5 sweep, a wrong read margin the
extension-read sweep, a missing zero-capacity rule the 1-of-256 sweep, a
non-chasing match copy the RLE fixtures, a wrong status order the crafted
negatives.
"""
import pytest

from tests import corpus, libs

OK, CORRUPT, TOO_SMALL = libs.OK, libs.CORRUPT, libs.TOO_SMALL

# tests/oracle_lz4.cpp:20 -- FNV-1a over (name, compressed) of the fixture
# corpus. Reproduced here with liblz4 1.9.3 because LZ4_compress_default
# emits the same bytes for this corpus as the pinned 1.10.0.
REFERENCE_CORPUS_DIGEST = 0x04E2FB5078DB9E2B

needs_lz4 = pytest.mark.skipif(libs.lz4lib() is None, reason="liblz4 absent")


@needs_lz4
def test_corpus_digest_matches_reference_pin(lz4_fixtures):
    d = corpus.fnv1a64(x for f in lz4_fixtures for x in (f.name.encode(), f.compressed))
    assert d == REFERENCE_CORPUS_DIGEST, hex(d)


@needs_lz4
def test_fixtures_round_trip(lz4_fixtures):
    for f in lz4_fixtures:
        st, out = libs.oracle_lz4(f.compressed, len(f.original))
        assert st == OK, f.name
        assert out == f.original, f.name


@needs_lz4
def test_mutant_parity_both_directions(lz4_fixtures):
    """tests/parser_twin.cpp:295-342: where the oracle accepts, liblz4
    accepts with identical bytes; where liblz4 rejects, the oracle rejects;
    the only stricter case is the one offset==0 mutant, pinned by identity
    (tests/gpu_fixture.cu:220-221)."""
    total = rejected = 0
    stricter = []
    for f in lz4_fixtures:
        for m in corpus.mutate_stream(f.compressed, f.seed):
            ok, ref_out = libs.lz4_decompress(m.stream, len(f.original))
            st, out = libs.oracle_lz4(m.stream, len(f.original))
            total += 1
            if st == OK:
                assert ok, f"{f.name}/{m.description}"
                assert out == ref_out, f"{f.name}/{m.description}"
            else:
                assert st in (CORRUPT, TOO_SMALL)
                if ok:
                    stricter.append(f"{f.name}/{m.description}")
                else:
                    rejected += 1
    assert total > 0 and rejected > 0
    assert stricter == ["text-256/flip-bit-at-88"]


def test_crafted_negatives():
    """tests/parser_twin.cpp:49-77, each with its exact status."""
    cases = [
        ("empty-src", b"", 16, CORRUPT),
        ("literal-extension-truncated", b"\xF0", 1024, CORRUPT),
        ("literal-length-past-capacity", b"\xF0\xFF" + b"\x00" * 16, 40, TOO_SMALL),
        ("terminal-inexact-consumption", b"\x10\x41\x01\x00", 16, CORRUPT),
        ("offset-beyond-written", b"\x00\x05" + b"\x00" * 7, 16, CORRUPT),
        ("match-length-past-capacity", b"\x1F\x41\x01\x00\xFF\xFF\x00\x00\x00\x00", 100, TOO_SMALL),
    ]
    for name, s, cap, want in cases:
        st, _ = libs.oracle_lz4(s, cap)
        assert st == want, name
        if libs.lz4lib() is not None:
            assert not libs.lz4_decompress(s, cap)[0], name


@needs_lz4
def test_last_literals_sweep():
    """tests/parser_twin.cpp:88-130: a match ending T bytes before the end
    is accepted iff T >= 5, on both sides, with identical bytes."""
    acc = rej = 0
    for L in (4, 8, 12):
        for ext in (0, 10, 100):
            for T in (4, 5, 6, 10):
                s = bytes([(L << 4) | 15]) + bytes(0x41 + i % 26 for i in range(L)) + \
                    b"\x01\x00" + bytes([ext, T << 4]) + bytes(0x61 + i % 26 for i in range(T))
                cap = L + 19 + ext + T
                st, out = libs.oracle_lz4(s, cap)
                ok, ref_out = libs.lz4_decompress(s, cap)
                assert (st == OK) == ok == (T >= 5), (L, ext, T)
                if ok:
                    assert out == ref_out
                    acc += 1
                else:
                    rej += 1
    assert acc > 0 and rej > 0


@needs_lz4
def test_extension_read_bound_sweep():
    """tests/parser_twin.cpp:138-209: liblz4's read margins on both length
    extensions, swept across the input end."""
    acc = rej = 0
    for ff in (0, 1, 2):
        for last in (0, 30, 200):
            for T in (0, 1, 2, 3, 4, 6):
                s = b"\x1F\x41\x01\x00" + b"\xFF" * ff + bytes([last, T << 4]) + \
                    bytes(0x61 + i for i in range(T))
                cap = 1 + 4 + 15 + ff * 255 + last + T
                st, out = libs.oracle_lz4(s, cap)
                ok, ref_out = libs.lz4_decompress(s, cap)
                assert (st == OK) == ok, (ff, last, T)
                if ok:
                    assert out == ref_out
                    acc += 1
                else:
                    rej += 1
    for ff in (0, 1):
        for last in (0, 30, 200):
            lit = 15 + ff * 255 + last
            base = b"\xF0" + b"\xFF" * ff + bytes([last]) + bytes(0x41 + i % 26 for i in range(lit))
            for cap in range(lit, lit + 19):
                st, _ = libs.oracle_lz4(base, cap)
                ok, _ = libs.lz4_decompress(base, cap)
                assert (st == OK) == ok, (ff, last, cap)
                acc += ok
                rej += not ok
    assert acc > 0 and rej > 0


@needs_lz4
def test_zero_capacity_sweep():
    """tests/parser_twin.cpp:224-265: at capacity 0 exactly the one-byte
    stream 0x00 is accepted; 1 accept, 511 rejects."""
    lib = libs.lz4lib()
    import ctypes
    scratch = ctypes.create_string_buffer(1)
    acc = rej = 0
    for tok in range(256):
        s = bytes([tok])
        st, out = libs.oracle_lz4(s, 0)
        ok = lib.LZ4_decompress_safe(s, scratch, 1, 0) >= 0
        assert (st == OK) == ok, hex(tok)
        if st == OK:
            assert tok == 0 and out == b""
            acc += 1
        else:
            assert st in (CORRUPT, TOO_SMALL)
            rej += 1
    for tok in range(256):
        s = bytes([tok, 0])
        assert libs.oracle_lz4(s, 0)[0] != OK
        assert lib.LZ4_decompress_safe(s, scratch, 2, 0) < 0
        rej += 1
    assert (acc, rej) == (1, 511)


def test_offset_zero_divergence_and_empty_block():
    """tests/parser_twin.cpp:385-402."""
    assert libs.oracle_lz4(b"\x00" * 9, 16)[0] == CORRUPT
    assert libs.oracle_lz4(b"\x00", 16) == (OK, b"")


@needs_lz4
def test_capacity_beyond_convention_and_size_max(lz4_fixtures):
    """tests/parser_twin.cpp:408-440."""
    import ctypes
    f = lz4_fixtures[0]
    st, out = libs.oracle_lz4(f.compressed, 1 << 20)
    ok, ref_out = libs.lz4_decompress(f.compressed, 1 << 20)
    assert st == OK and ok and out == ref_out
    f = lz4_fixtures[-1]
    dp, seqs, stalls = libs.u64(), libs.u64(), libs.u64()
    st = libs.oracle().oracle_lz4_parse(f.compressed, len(f.compressed), (1 << 64) - 1,
                                        ctypes.byref(dp), ctypes.byref(seqs), ctypes.byref(stalls))
    assert st == OK and dp.value == len(f.original) and stalls.value == 0


def test_fuzz_seed_vectors():
    """fuzz/corpus/fuzz_lz4_block/*: a 2-byte big-endian capacity, then the
    stream (fuzz/fuzz_lz4_block.cpp:63-73)."""
    seeds = {
        "literals-only-accept": (b"\x00\x05\x50hello", (OK, b"hello")),
        "match-then-literal-tail-accept": (b"\x00\x11\x44abcd\x04\x00\x50wxyz!", (OK, b"abcdabcdabcdwxyz!")),
        "offset-zero-reject": (b"\x00\x11\x44abcd\x00\x00\x50wxyz!", (CORRUPT, b"")),
        "truncated-literal-run-reject": (b"\x00\x20\x50he", (CORRUPT, b"")),
        "issue-315-zero-capacity-nonzero-token": (b"\x00\x00\x06", (CORRUPT, b"")),
    }
    for name, (data, want) in seeds.items():
        cap = (data[0] << 8) | data[1]
        assert libs.oracle_lz4(data[2:], cap) == want, name


def test_adversarial_termination():
    """tests/termination.cpp:53-114: every OK call advances src_pos, over the
    adversarial corpus at capacities {0, 1, 12, 1 MiB, SIZE_MAX}."""
    import ctypes
    for b in corpus.make_adversarial_blocks():
        for cap in (0, 1, 12, 1 << 20, (1 << 64) - 1, b.capacity):
            dp, seqs, stalls = libs.u64(), libs.u64(), libs.u64()
            st = libs.oracle().oracle_lz4_parse(b.stream, len(b.stream), cap, ctypes.byref(dp),
                                                ctypes.byref(seqs), ctypes.byref(stalls))
            assert st in (OK, CORRUPT, TOO_SMALL)
            assert stalls.value == 0, b.name
            assert seqs.value <= len(b.stream) + 1


@needs_lz4
def test_exhaustive_short_streams_against_liblz4():
    """Brute force on tiny inputs: every 1- and 2-byte stream at capacities
    0, 1, 16, 64. Oracle accept implies liblz4 accept with equal bytes, and
    liblz4 reject implies oracle reject."""
    for cap in (0, 1, 16, 64):
        for a in range(256):
            for s in [bytes([a])] + [bytes([a, b]) for b in range(256)]:
                st, out = libs.oracle_lz4(s, cap)
                ok, ref_out = libs.lz4_decompress(s, cap)
                if st == OK:
                    assert ok and out == ref_out, (s.hex(), cap)
                elif not ok:
                    pass
                else:
                    # the one deliberate divergence is offset == 0, which
                    # needs at least 3 bytes; none may appear here
                    raise AssertionError(("stricter", s.hex(), cap))


@pytest.mark.skipif(libs.ref() is None, reason="reference ladders not built here")
def test_agrees_with_reference_ladders(lz4_fixtures):
    """Identical status and bytes against the reference's own ladder,
    compiled from its source, on every stream the suite has."""
    streams = []
    for f in lz4_fixtures:
        streams += [(f.compressed, len(f.original)), (f.compressed, len(f.original) + 8)]
        streams += [(m.stream, len(f.original)) for m in corpus.mutate_stream(f.compressed, f.seed)]
    for b in corpus.make_adversarial_blocks():
        streams += [(b.stream, c) for c in (0, 1, 12, 4096, b.capacity)]
    for a in range(256):
        for cap in (0, 1, 16):
            streams.append((bytes([a]), cap))
            streams += [(bytes([a, b]), cap) for b in range(0, 256, 7)]
    for s, cap in streams:
        assert libs.oracle_lz4(s, cap) == libs.ref_lz4(s, cap), (s[:16].hex(), cap)
