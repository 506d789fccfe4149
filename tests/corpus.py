"""Deterministic test corpora: LZ4 and Snappy fixture pairs, mutants, and the
hand-crafted adversarial blocks.

Restates the reference's corpus machinery so the same seeds give the same
bytes:

* SplitMix64, RandomBytes, TextLike         -- tests/fixtures.cpp:15-61
* MakeLz4BlockFixtures                      -- tests/fixtures.cpp:84-104
* MutateStream                              -- tests/fixtures.cpp:106-145
* MakeSnappyFixtures                        -- tests/fixtures.cpp:289-320
* MutateSnappyStream (+ the element walk)   -- tests/fixtures.cpp:206-441
* MakeAdversarialBlocks                     -- tests/adversarial_blocks.h:31-99

Compression goes through the codec libraries (liblz4 LZ4_compress_default,
snappy_compress) via tests/libs.py. The reference pins a digest of its LZ4
fixture corpus (tests/oracle_lz4.cpp:20); the corpus built here reproduces
that digest when the installed liblz4 emits the same bytes, which the oracle
tests check.
"""
from __future__ import annotations

from dataclasses import dataclass

from . import libs

MASK64 = (1 << 64) - 1


class SplitMix64:
    """tests/fixtures.cpp:15-21."""

    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)


def random_bytes(size: int, seed: int) -> bytes:
    """tests/fixtures.cpp:23-30: one SplitMix64 draw per byte, low byte."""
    g = SplitMix64(seed)
    return bytes(g.next() & 0xFF for _ in range(size))


_WORDS = [b"the", b"quick", b"brown", b"fox", b"jumps", b"over", b"lazy",
          b"dog", b"gpu", b"decode", b"batch", b"chunk", b"warp", b"stream",
          b"block", b"lane"]


def text_like(size: int, seed: int) -> bytes:
    """tests/fixtures.cpp:35-61: words with splices copied from distances
    that double up to 65535 and wrap to 1."""
    out = bytearray()
    g = SplitMix64(seed)
    distance = 1
    while len(out) < size:
        word = _WORDS[g.next() % 16]
        out += word
        out.append(0x20)
        if len(out) > distance and g.next() % 4 == 0:
            start = len(out) - distance
            length = distance if distance < 16 else 16
            for i in range(length):
                out.append(out[start + i])
            if distance >= 65535:
                distance = 1
            else:
                distance = distance * 2 if distance * 2 < 65535 else 65535
    return bytes(out[:size])


@dataclass
class Fixture:
    name: str
    seed: int
    original: bytes
    compressed: bytes


@dataclass
class Mutant:
    description: str
    stream: bytes


def make_lz4_fixtures() -> list[Fixture]:
    """tests/fixtures.cpp:84-104."""
    out = []

    def add(name, seed, original):
        out.append(Fixture(name, seed, original, libs.lz4_compress(original)))

    add("zeros-65536", 0x01, bytes(65536))
    add("random-65536", 0x11, random_bytes(65536, 0x11))
    for n in (1, 31, 32, 33, 255, 256, 257, 4095, 4096, 65536):
        add(f"text-{n}", 0x22 + n, text_like(n, 0x22 + n))
    return out


def mutate_stream(stream: bytes, seed: int) -> list[Mutant]:
    """tests/fixtures.cpp:106-145: truncations at 1/4, 2/4, 3/4; 1..8 bytes
    off either end; eight seeded single-bit flips."""
    out: list[Mutant] = []
    n = len(stream)
    if n == 0:
        return out
    for quarters in (1, 2, 3):
        keep = n * quarters // 4
        if keep < n:
            out.append(Mutant(f"truncate-to-{keep}", stream[:keep]))
    k = 1
    while k <= 8 and k < n:
        out.append(Mutant(f"drop-{k}-tail", stream[:n - k]))
        out.append(Mutant(f"drop-{k}-head", stream[k:]))
        k += 1
    g = SplitMix64(seed)
    for _ in range(8):
        flipped = bytearray(stream)
        offset = g.next() % n
        flipped[offset] ^= (1 << (g.next() % 8)) & 0xFF
        out.append(Mutant(f"flip-bit-at-{offset}", bytes(flipped)))
    return out


# ------------------------------------------------------------------ Snappy

def make_snappy_fixtures() -> list[Fixture]:
    """tests/fixtures.cpp:289-320."""
    out = []

    def add(name, seed, original):
        out.append(Fixture(name, seed, original,
                           libs.snappy_compress(original)))

    add("zeros-65536", 0x41, bytes(65536))
    add("random-65536", 0x51, random_bytes(65536, 0x51))
    add("random-70000", 0x52, random_bytes(70000, 0x52))
    add("zeros-70000", 0x42, bytes(70000))
    for n in (1, 2, 59, 60, 61, 62, 63, 64, 65, 127, 128, 255, 256, 257,
              4095, 4096, 65535):
        add(f"text-{n}", 0x62 + n, text_like(n, 0x62 + n))
    return out


def _snappy_preamble_size(stream: bytes) -> int:
    """tests/fixtures.cpp:220-227."""
    for i in range(min(5, len(stream))):
        if stream[i] & 0x80 == 0:
            return i + 1
    return 0


def _snappy_elements(stream: bytes):
    """tests/fixtures.cpp:232-274: element boundaries (offset, size, tag),
    stopping at the first bytes the walk cannot account for."""
    out = []
    at = _snappy_preamble_size(stream)
    if at == 0:
        return out
    while at < len(stream):
        tag = stream[at]
        if tag & 3 == 0:
            length_class = tag >> 2
            if length_class < 60:
                size = 1 + (length_class + 1)
            else:
                extra = length_class - 59
                if at + 1 + extra > len(stream):
                    break
                length = 0
                for k in range(extra):
                    length |= stream[at + 1 + k] << (8 * k)
                size = 1 + extra + (length + 1)
        elif tag & 3 == 1:
            size = 2
        elif tag & 3 == 2:
            size = 3
        else:
            size = 5
        if size == 0 or at + size > len(stream):
            break
        out.append((at, size, tag))
        at += size
    return out


def _snappy_with_declared(stream: bytes, declared: int) -> bytes:
    """tests/fixtures.cpp:279-292: the declaration as a 5-byte varint."""
    pre = _snappy_preamble_size(stream)
    out = bytearray()
    v = declared & 0xFFFFFFFF
    for _ in range(4):
        out.append((v & 0x7F) | 0x80)
        v >>= 7
    out.append(v & 0x0F)
    return bytes(out) + stream[pre:]


def mutate_snappy_stream(stream: bytes, seed: int) -> list[Mutant]:
    """tests/fixtures.cpp:322-441."""
    out = mutate_stream(stream, seed)
    n = len(stream)
    if n == 0:
        return out
    pre = _snappy_preamble_size(stream)
    if pre > 0:
        cont = bytearray(stream)
        cont[pre - 1] |= 0x80
        out.append(Mutant("preamble-continuation-bit-set", bytes(cont)))
        declared = 0
        for k in range(pre):
            declared |= (stream[k] & 0x7F) << (7 * k)
        out.append(Mutant("preamble-declares-0", _snappy_with_declared(stream, 0)))
        out.append(Mutant("preamble-declares-1", _snappy_with_declared(stream, 1)))
        out.append(Mutant("preamble-declares-one-less",
                          _snappy_with_declared(stream, declared - 1 if declared > 0 else 0)))
        out.append(Mutant("preamble-declares-one-more",
                          _snappy_with_declared(stream, declared + 1)))
        out.append(Mutant("preamble-declares-max",
                          _snappy_with_declared(stream, 0xFFFFFFFF)))
    elements = _snappy_elements(stream)
    targets = list(range(min(3, len(elements))))
    if len(elements) > 3:
        targets.append(len(elements) - 1)
    for index in targets:
        e_off, _e_size, e_tag = elements[index]
        at = f"element-{index}"
        for kind in range(4):
            if e_tag & 3 == kind:
                continue
            retag = bytearray(stream)
            retag[e_off] = (e_tag & ~3 & 0xFF) | kind
            out.append(Mutant(f"{at}-tag-class-{kind}", bytes(retag)))
        for delta in (-1, 1):
            high = e_tag >> 2
            if delta < 0 and high == 0:
                continue
            if delta > 0 and high == 63:
                continue
            moved = high - 1 if delta < 0 else high + 1
            rel = bytearray(stream)
            rel[e_off] = ((moved << 2) | (e_tag & 3)) & 0xFF
            out.append(Mutant(f"{at}-length-bits-{'minus-1' if delta < 0 else 'plus-1'}",
                              bytes(rel)))
        kind = e_tag & 3
        if kind != 0:
            offset_bytes = 1 if kind == 1 else (2 if kind == 2 else 4)
            for value in (0, 0xFFFFFFFF):
                reo = bytearray(stream)
                for k in range(offset_bytes):
                    reo[e_off + 1 + k] = (value >> (8 * k)) & 0xFF
                if kind == 1:
                    reo[e_off] = (e_tag & 0x1F) | (0 if value == 0 else 0xE0)
                out.append(Mutant(f"{at}-copy-offset-{'0' if value == 0 else 'max'}",
                                  bytes(reo)))
    out.append(Mutant("trailing-byte", stream + b"\x00"))
    return out


# ------------------------------------------------------------- adversarial

@dataclass
class AdversarialBlock:
    name: str
    stream: bytes
    capacity: int


def lz4_literal_extension_run(run: int) -> bytes:
    """tests/adversarial_blocks.h:34-41."""
    return b"\xF0" + b"\xFF" * run + b"\x00" + b"\x41" * 64


def lz4_match_extension_run(run: int) -> bytes:
    """tests/adversarial_blocks.h:45-51."""
    return bytes([0x4F, 0x41, 0x42, 0x43, 0x44, 0x01, 0x00]) + b"\xFF" * run + \
        b"\x00" + b"\x41" * 64


def make_adversarial_blocks() -> list[AdversarialBlock]:
    """tests/adversarial_blocks.h:53-99."""
    out = [
        AdversarialBlock("empty", b"", 64),
        AdversarialBlock("lone-zero-token", b"\x00", 64),
        AdversarialBlock("lone-max-token", b"\xFF", 64),
        AdversarialBlock("all-ff-4096", b"\xFF" * 4096, 1 << 20),
        AdversarialBlock("all-zero-4096", b"\x00" * 4096, 1 << 20),
        AdversarialBlock("literal-extension-run-1", lz4_literal_extension_run(1), 1 << 20),
        AdversarialBlock("literal-extension-run-4096", lz4_literal_extension_run(4096), 1 << 20),
        AdversarialBlock("match-extension-run-1", lz4_match_extension_run(1), 1 << 20),
        AdversarialBlock("match-extension-run-4096", lz4_match_extension_run(4096), 1 << 20),
        AdversarialBlock("zero-offset", bytes([0x40, 0x41, 0x42, 0x43, 0x44, 0x00, 0x00]) + b"\x41" * 64, 1 << 20),
        AdversarialBlock("offset-past-output", bytes([0x40, 0x41, 0x42, 0x43, 0x44, 0xFF, 0xFF]) + b"\x41" * 64, 1 << 20),
        AdversarialBlock("self-referential-rle", libs.lz4_compress(b"\x5A" * 65536), 65536),
    ]
    period2 = bytes(0xFF if i % 2 else 0x00 for i in range(65536))
    out.append(AdversarialBlock("self-referential-period-2", libs.lz4_compress(period2), 65536))
    return out


def fnv1a64(chunks, h: int = 14695981039346656037) -> int:
    """FNV-1a 64, chained over byte strings (tests/oracle_lz4.cpp:22-29)."""
    for data in chunks:
        for b in data:
            h ^= b
            h = (h * 1099511628211) & MASK64
    return h
