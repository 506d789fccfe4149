"""Hand-built Snappy streams: the element builders and the crafted negative,
capacity-negative and accept sets of the reference's Snappy twin test
(tests/snappy_parser_twin.cpp:143-203 builders, 226-404 negatives, 419-429
capacity negatives, 441-509 accepts, 516-530 wide offsets).
"""
from __future__ import annotations

from dataclasses import dataclass

from . import libs


def preamble(length: int) -> bytes:
    out = bytearray()
    v = length
    while v >= 0x80:
        out.append((v & 0x7F) | 0x80)
        v >>= 7
    out.append(v)
    return bytes(out)


def literal(data: bytes) -> bytes:
    return bytes([((len(data) - 1) << 2) & 0xFF]) + data


def long_literal(data: bytes, extra: int) -> bytes:
    enc = len(data) - 1
    return bytes([((59 + extra) << 2) & 0xFF]) + bytes(
        (enc >> (8 * i)) & 0xFF for i in range(extra)) + data


def copy1(length: int, offset: int) -> bytes:
    return bytes([(((offset >> 8) << 5) | ((length - 4) << 2) | 1) & 0xFF, offset & 0xFF])


def copy2(length: int, offset: int) -> bytes:
    return bytes([(((length - 1) << 2) | 2) & 0xFF, offset & 0xFF, (offset >> 8) & 0xFF])


def copy4(length: int, offset: int) -> bytes:
    return bytes([(((length - 1) << 2) | 3) & 0xFF]) + bytes(
        (offset >> (8 * i)) & 0xFF for i in range(4))


R = {
    "preamble_byte_missing": 1, "preamble_final_group": 2, "preamble_unreachable": 3,
    "declared_over_capacity": 4, "past_terminal": 5, "under_production": 6,
    "literal_header_truncated": 7, "literal_payload_truncated": 8,
    "literal_over_production": 9, "copy_header_truncated": 10,
    "copy_offset_zero": 11, "copy_offset_before_start": 12,
    "copy_over_production": 13,
}
RUNG_COUNT = 14


@dataclass
class Crafted:
    name: str
    stream: bytes
    status: int
    rung: int


def crafted_negatives() -> list[Crafted]:
    abcd = b"abcd"
    C = libs.CORRUPT
    out = [
        Crafted("preamble-absent", b"", C, R["preamble_byte_missing"]),
        Crafted("preamble-truncated", b"\x80", C, R["preamble_byte_missing"]),
        Crafted("preamble-six-groups", b"\x80\x80\x80\x80\x80\x00", C, R["preamble_final_group"]),
        Crafted("preamble-overflows-32-bit", b"\x80\x80\x80\x80\x10", C, R["preamble_final_group"]),
        Crafted("under-production-empty", preamble(4), C, R["under_production"]),
        Crafted("under-production-short", preamble(8) + literal(abcd), C, R["under_production"]),
        Crafted("literal-header-truncated", preamble(300) + b"\xF0", C, R["literal_header_truncated"]),
        Crafted("literal-header-truncated-wide", preamble(70000) + b"\xF4", C, R["literal_header_truncated"]),
        Crafted("literal-payload-truncated", (preamble(4) + literal(abcd))[:-2], C, R["literal_payload_truncated"]),
        Crafted("literal-over-production", preamble(2) + literal(abcd), C, R["literal_over_production"]),
        Crafted("copy-header-truncated", preamble(8) + literal(abcd) + b"\x0E", C, R["copy_header_truncated"]),
        Crafted("copy1-header-truncated", preamble(8) + literal(abcd) + b"\x01", C, R["copy_header_truncated"]),
        Crafted("copy4-header-truncated", (preamble(8) + literal(abcd) + copy4(4, 4))[:-3], C, R["copy_header_truncated"]),
        Crafted("copy-offset-zero-tag-10", preamble(8) + literal(abcd) + copy2(4, 0), C, R["copy_offset_zero"]),
        Crafted("copy-offset-zero-tag-11", preamble(8) + literal(abcd) + copy4(4, 0), C, R["copy_offset_zero"]),
        Crafted("copy-offset-zero-tag-01", preamble(8) + literal(abcd) + copy1(4, 0), C, R["copy_offset_zero"]),
        Crafted("copy-offset-zero-at-start", preamble(4) + copy2(4, 0), C, R["copy_offset_zero"]),
        Crafted("copy-offset-before-start", preamble(8) + literal(abcd) + copy2(4, 5), C, R["copy_offset_before_start"]),
        Crafted("copy-offset-huge", preamble(8) + literal(abcd) + copy4(4, 0xFFFFFFFF), C, R["copy_offset_before_start"]),
        Crafted("copy-over-production", preamble(6) + literal(abcd) + copy2(4, 1), C, R["copy_over_production"]),
        Crafted("trailing-element", preamble(4) + literal(abcd) + literal(b"x"), C, R["literal_over_production"]),
        Crafted("trailing-tag", preamble(4) + literal(abcd) + b"\x00", C, R["literal_payload_truncated"]),
    ]
    return out


def capacity_negatives():
    """(name, stream, capacity, rung): snappy accepts the bytes; the caller's
    capacity is what refuses them."""
    s = preamble(4) + literal(b"abcd")
    return [("capacity-one-short", s, 3, R["declared_over_capacity"]),
            ("capacity-zero", s, 0, R["declared_over_capacity"])]


def accept_cases():
    abcd = b"abcd"
    out = [("empty-output", preamble(0), b""),
           ("preamble-non-minimal", b"\x84\x00" + literal(abcd), b"abcd"),
           ("preamble-five-groups", b"\x84\x80\x80\x80\x00" + literal(abcd), b"abcd"),
           ("copy-tag-11", preamble(8) + literal(abcd) + copy4(4, 4), b"abcdabcd")]
    for length in (1, 2, 3):
        exp = b"abcd" + b"d" * length
        out.append((f"copy-length-{length}-tag-10", preamble(4 + length) + literal(abcd) + copy2(length, 1), exp))
        out.append((f"copy-length-{length}-tag-11", preamble(4 + length) + literal(abcd) + copy4(length, 1), exp))
    out.append(("copy-tag-01", preamble(9) + literal(b"abcde") + copy1(4, 5), b"abcdeabcd"))
    out.append(("copy-overlapping", preamble(8) + literal(b"ab") + copy2(6, 1), b"abbbbbbb"))
    for extra in (1, 2, 3, 4):
        out.append((f"literal-long-form-{extra}-bytes", preamble(64) + long_literal(b"z" * 64, extra), b"z" * 64))
    return out


def wide_offset_stream(offset: int, length: int) -> bytes:
    payload = bytes(0x41 + (i % 26) for i in range(offset))
    return preamble(offset + length) + long_literal(payload, 4) + copy4(length, offset)
