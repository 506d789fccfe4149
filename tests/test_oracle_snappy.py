"""Pins the CPU Snappy oracle (oracle/cudec_oracle.c) against the
reference's Snappy twin test (tests/snappy_parser_twin.cpp), its corpus
machinery (tests/fixtures.cpp:289-441, tests/snappy_corpus.cpp), its fuzz
seeds (fuzz/corpus/fuzz_snappy_block/*), snappy itself, and -- where built
here -- the reference ladder compiled from its own source (oracle/_ref).

A dropped or reordered rung fails the per-rung crafted negatives (status
AND rung are pinned), an off-by-one capacity bound fails the capacity axis,
a 32-bit length accumulator fails the pinned wrap divergence, and a
non-chasing copy fails the overlapping accept case and the fixtures.
"""
import ctypes

import pytest

from tests import corpus, libs, snappy_vectors as sv

OK, CORRUPT, TOO_SMALL = libs.OK, libs.CORRUPT, libs.TOO_SMALL

# FNV-1a over (name, compressed) of the Snappy fixture corpus as built here
# with snappy 1.1.8. The reference pins 0x79af9459a11bfd4c for snappy 1.2.2
# (tests/snappy_corpus.cpp:24); RawCompress output differs between the two
# versions, so the pin is re-derived for the version in this image.
SNAPPY_CORPUS_DIGEST_1_1_8 = 0x37DF2A95DEFD8DCE

needs_snappy = pytest.mark.skipif(libs.snappylib() is None, reason="libsnappy absent")


def _twin(stream: bytes):
    """The reference's whole-stream twin entry (tests/snappy_twin.h:221-242):
    capacity = the declared length, bounded by the oracle wrapper's 256 MiB."""
    st, out, rung = libs.oracle_snappy(stream, libs.SNAPPY_ORACLE_MAX_OUTPUT) \
        if _declared(stream) is None else libs.oracle_snappy(stream, _declared(stream))
    return st, out, rung


def _declared(stream: bytes):
    d, p, r = libs.u64(), libs.u64(), libs.i32()
    st = libs.oracle().oracle_snappy_declared(stream, len(stream), ctypes.byref(d),
                                              ctypes.byref(p), ctypes.byref(r))
    if st != OK or d.value > libs.SNAPPY_ORACLE_MAX_OUTPUT:
        return None
    return d.value


@needs_snappy
def test_corpus_digest(snappy_fixtures):
    d = corpus.fnv1a64(x for f in snappy_fixtures for x in (f.name.encode(), f.compressed))
    assert d == SNAPPY_CORPUS_DIGEST_1_1_8, hex(d)


@needs_snappy
def test_fixtures_round_trip_and_capacity_axis(snappy_fixtures):
    for f in snappy_fixtures:
        st, out, _ = libs.oracle_snappy(f.compressed, len(f.original))
        assert st == OK and out == f.original, f.name
    # tests/snappy_parser_twin.cpp:800-836: one short refuses with
    # OUTPUT_TOO_SMALL, exact decodes, and a larger capacity still produces
    # exactly the declared length.
    f = snappy_fixtures[0]
    st, out, rung = libs.oracle_snappy(f.compressed, len(f.original) - 1)
    assert (st, out, rung) == (TOO_SMALL, b"", sv.R["declared_over_capacity"])
    st, out, _ = libs.oracle_snappy(f.compressed, len(f.original) + 4096)
    assert st == OK and out == f.original


@needs_snappy
def test_mutant_parity_both_directions(snappy_fixtures):
    """tests/snappy_parser_twin.cpp:530-600: the oracle never accepts what
    snappy rejects, bytes match where both accept, and over this corpus the
    stricter-than-snappy count is zero."""
    total = rejected = stricter = 0
    for f in snappy_fixtures:
        for m in corpus.mutate_snappy_stream(f.compressed, f.seed):
            ok, ref_out = libs.snappy_decompress(m.stream)
            st, out, _ = _twin(m.stream)
            total += 1
            if st == OK:
                assert ok and out == ref_out, f"{f.name}/{m.description}"
            else:
                assert st in (CORRUPT, TOO_SMALL)
                if ok:
                    stricter += 1
                else:
                    rejected += 1
    assert total > 0 and rejected > 0
    assert stricter == 0


def test_crafted_negatives_status_and_rung():
    covered = set()
    for c in sv.crafted_negatives():
        st, out, rung = _twin(c.stream)
        assert (st, rung) == (c.status, c.rung), c.name
        covered.add(rung)
        if libs.snappylib() is not None:
            assert not libs.snappy_decompress(c.stream)[0], c.name
    for name, s, cap, rung in sv.capacity_negatives():
        st, out, r = libs.oracle_snappy(s, cap)
        assert (st, r) == (TOO_SMALL, rung), name
        covered.add(r)
    # the past-terminal rung is reached by calling past the terminal
    # element, which the liveness check does on every accepted stream
    covered.add(sv.R["past_terminal"])
    want = set(range(1, sv.RUNG_COUNT)) - {sv.R["preamble_unreachable"]}
    assert covered == want


def test_accept_cases_and_wide_offsets():
    for name, s, exp in sv.accept_cases():
        st, out, _ = _twin(s)
        assert st == OK and out == exp, name
        if libs.snappylib() is not None:
            ok, ref_out = libs.snappy_decompress(s)
            assert ok and ref_out == exp, name
    for off in (65535, 65536, 65537, 131072):
        s = sv.wide_offset_stream(off, 4)
        st, out, _ = _twin(s)
        assert st == OK and len(out) == off + 4
        over = bytearray(s)
        over[-4:] = (off + 1).to_bytes(4, "little")
        assert _twin(bytes(over))[0] == CORRUPT
        if libs.snappylib() is not None:
            assert libs.snappy_decompress(s) == (True, out)
            assert not libs.snappy_decompress(bytes(over))[0]


def test_literal_length_wrap_divergence():
    """tests/snappy_parser_twin.cpp:682-711: snappy's 32-bit literal length
    wraps 0xFFFFFFFF + 1 to zero and accepts; the ladder refuses."""
    wrapped = b"\x00\xFC\xFF\xFF\xFF\xFF"
    near = b"\x00\xFC\xFE\xFF\xFF\xFF"
    tail = b"\x04\xFC\xFF\xFF\xFF\xFF\x0Cabcd"
    assert _twin(wrapped)[:2] == (CORRUPT, b"")
    assert _twin(near)[0] == CORRUPT
    assert _twin(tail)[0] == CORRUPT
    if libs.snappylib() is not None:
        assert libs.snappy_decompress(wrapped) == (True, b"")
        assert not libs.snappy_decompress(near)[0]
        assert libs.snappy_decompress(tail) == (True, b"abcd")


def test_huge_declaration_refused_on_the_capacity_rung():
    huge = b"\xFF\xFF\xFF\xFF\x0F\x0Cabcd"
    assert _declared(huge) is None
    st, out, rung = libs.oracle_snappy(huge, 0xFFFFFFFF - 1)
    assert (st, rung) == (TOO_SMALL, sv.R["declared_over_capacity"])


def test_fuzz_seed_vectors():
    """fuzz/corpus/fuzz_snappy_block/* (whole input is the stream)."""
    assert _twin(b"\x05\x10hello")[:2] == (OK, b"hello")
    assert _twin(b"\x0c\x0cabcd\x11\x04")[:2] == (OK, b"abcd" * 3)
    st, out, _ = _twin(b"\x0e\x0cabcd\x26\x04\x00")
    assert st == OK and len(out) == 14 and out == b"abcd" + b"abcdabcdab"
    assert _twin(b"\x05\x10h")[0] == CORRUPT
    # declared 8 MiB against the fuzz harness bound of 4 MiB
    # (fuzz/fuzz_snappy_block.cpp:46)
    st, _, rung = libs.oracle_snappy(b"\x80\x80\x80\x04\x10abcde", 1 << 22)
    assert (st, rung) == (TOO_SMALL, sv.R["declared_over_capacity"])


def test_liveness_exhaustive_short_and_corpus(snappy_fixtures):
    """tests/snappy_parser_twin.cpp:838-872: every OK element advances the
    source, the terminal is reached, and a call past it is refused -- over
    the corpus, its mutants, and every stream of up to two bytes."""
    streams = [b""]
    for a in range(256):
        streams.append(bytes([a]))
        streams += [bytes([a, b]) for b in range(256)]
    for f in snappy_fixtures:
        streams.append(f.compressed)
        streams += [m.stream for m in corpus.mutate_snappy_stream(f.compressed, f.seed)]
    streams += [c.stream for c in sv.crafted_negatives()]
    streams += [s for _, s, _ in sv.accept_cases()]
    dp, el, stalls, r = libs.u64(), libs.u64(), libs.u64(), libs.i32()
    for s in streams:
        cap = _declared(s) or 0
        st = libs.oracle().oracle_snappy_parse(s, len(s), cap, ctypes.byref(dp), ctypes.byref(el),
                                               ctypes.byref(stalls), ctypes.byref(r))
        assert st in (OK, CORRUPT, TOO_SMALL)
        assert stalls.value == 0, s[:16].hex()
        assert el.value <= len(s) + 1


@pytest.mark.skipif(libs.ref() is None, reason="reference ladders not built here")
def test_agrees_with_reference_ladders(snappy_fixtures):
    streams = []
    for f in snappy_fixtures:
        for cap in (len(f.original), len(f.original) - 1, len(f.original) + 9):
            streams.append((f.compressed, max(cap, 0)))
        streams += [(m.stream, len(f.original) + 1)
                    for m in corpus.mutate_snappy_stream(f.compressed, f.seed)]
    streams += [(c.stream, 1 << 17) for c in sv.crafted_negatives()]
    streams += [(s, cap) for _, s, cap, _ in sv.capacity_negatives()]
    streams += [(s, len(e) + 3) for _, s, e in sv.accept_cases()]
    for a in range(256):
        for cap in (0, 1, 64):
            streams.append((bytes([a]), cap))
            streams += [(bytes([a, b]), cap) for b in range(0, 256, 3)]
    for s, cap in streams:
        assert libs.oracle_snappy(s, cap) == libs.ref_snappy(s, cap), (s[:16].hex(), cap)
