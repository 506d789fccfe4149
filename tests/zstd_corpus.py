"""Zstandard test corpora: frames from libzstd with parameters chosen to
reach every decode surface, a header walker that proves which surfaces a
frame reaches, seeded mutants, and hand-made frames for the header rungs.

The walker reads headers only (block headers, literals-section headers,
the sequence count and the mode byte are all in the clear), in the manner
of the reference's tests/zstd_corpus.h:1-17, so the coverage claim needs
no second entropy decoder. Test infrastructure only.
"""
from __future__ import annotations

import random
import struct

import numpy as np

from cudec_amd import workload

from . import libs

SEED = 0x2575


# ------------------------------------------------------------ the walker

def walk(frame: bytes):
    """The blocks of a frame as dicts: type (0 raw, 1 RLE, 2 compressed),
    and for compressed blocks the literals type (0 raw, 1 RLE, 2 Huffman,
    3 treeless), its stream count, the sequence count and the three table
    modes (0 predefined, 1 RLE, 2 FSE, 3 repeat). None for a frame the
    walker cannot follow."""
    if len(frame) < 6 or struct.unpack_from("<I", frame)[0] != 0xFD2FB528:
        return None
    d = frame[4]
    single = (d >> 5) & 1
    fcs = [single, 2, 4, 8][d >> 6]
    pos = 5 + (0 if single else 1) + [0, 1, 2, 4][d & 3] + fcs
    blocks = []
    while pos + 3 <= len(frame):
        bh = int.from_bytes(frame[pos:pos + 3], "little")
        last, btype, bsize = bh & 1, (bh >> 1) & 3, bh >> 3
        body = frame[pos + 3:pos + 3 + (1 if btype == 1 else bsize)]
        b = {"type": btype}
        if btype == 2 and body:
            lt, sf = body[0] & 3, (body[0] >> 2) & 3
            b["literals"] = lt
            if lt <= 1:
                lh = {0: 1, 1: 2, 2: 1, 3: 3}[sf]
                regen = int.from_bytes(body[:lh], "little") >> (3 if lh == 1 else 4)
                used = lh + (regen if lt == 0 else 1)
                b["streams"] = 0
            else:
                lh, fb = {0: (3, 10), 1: (3, 10), 2: (4, 14), 3: (5, 18)}[sf]
                raw = int.from_bytes(body[:lh], "little")
                used = lh + ((raw >> (4 + fb)) & ((1 << fb) - 1))
                b["streams"] = 1 if sf == 0 else 4
            sec = body[used:]
            if sec:
                c = sec[0]
                nb = 1
                if c >= 128:
                    c, nb = (((c - 128) << 8) + sec[1], 2) if c < 255 else (
                        sec[1] + (sec[2] << 8) + 0x7F00, 3)
                b["sequences"] = c
                if c and len(sec) > nb:
                    m = sec[nb]
                    b["modes"] = (m >> 6, (m >> 4) & 3, (m >> 2) & 3)
        blocks.append(b)
        pos += 3 + len(body)
        if last:
            break
    return blocks


def surfaces(frames):
    """Every (kind, value) surface the frames reach."""
    out = set()
    for f in frames:
        for b in walk(f) or []:
            out.add(("block", b["type"]))
            if "literals" in b:
                out.add(("literals", b["literals"]))
                if b["literals"] >= 2:
                    out.add(("streams", b["streams"]))
            for field, mode in zip(("ll", "of", "ml"), b.get("modes", ())):
                out.add((field, mode))
    return out


ALL_SURFACES = ({("block", t) for t in range(3)} | {("literals", t) for t in range(4)}
                | {("streams", 1), ("streams", 4)}
                | {(f, m) for f in ("ll", "of", "ml") for m in range(4)})


# ------------------------------------------------------------ the corpus

def _data(kind: int, size: int, seed: int) -> bytes:
    out = np.empty(size, dtype=np.uint8)
    workload.wl().wl_gen_chunk(out.ctypes.data, size, seed, kind)
    return out.tobytes()


def sources():
    """(name, bytes): token, random and repetitive chunks from the benchmark
    generator at several sizes, plus shapes that force particular block and
    literal types."""
    rng = random.Random(SEED)
    out = []
    for kind, kname in ((workload.KIND_TOKENS, "token"), (workload.KIND_RANDOM, "random"),
                        (workload.KIND_REPETITIVE, "repetitive")):
        for size in (1, 17, 300, 4096, 65536, 200_000, 700_000):
            out.append((f"{kname}-{size}", _data(kind, size, rng.getrandbits(63))))
    out.append(("zeros-300000", bytes(300_000)))                      # RLE blocks
    out.append(("one-byte-literals", b"\x07" * 5000 + _data(workload.KIND_TOKENS, 5000, 7)))
    text = b"".join(b"line %d of a small text corpus, %s\n" % (i, b"abc" * (i % 7))
                    for i in range(3000))
    out.append(("text", text))
    out.append(("empty", b""))
    return out


# (name, level, parameters)
SETTINGS = [
    ("l1", 1, {}),
    ("l3-checksum", 3, {"checksum": 1}),
    ("l9", 9, {}),
    ("l19-checksum", 19, {"checksum": 1}),
    ("fast", -5, {}),
    ("raw-literals", 3, {"literal_mode": 2}),
    ("small-blocks", 6, {"target_block": 340}),
    ("mid-blocks", 12, {"target_block": 1300}),
    ("window-10", 3, {"window_log": 10}),
    ("window-17-checksum", 5, {"window_log": 17, "checksum": 1}),
]


def _frames(max_size):
    for sname, data in sources():
        if max_size is not None and len(data) > max_size:
            continue
        for tname, level, params in SETTINGS:
            yield f"{sname}/{tname}", libs.zstd_compress(data, level, **params), data


def corpus(max_size: int | None = None):
    """(name, frame, original) over sources() x SETTINGS, for every frame
    that libzstd decodes back to its input."""
    return [(n, f, d) for n, f, d in _frames(max_size)
            if libs.zstd_decompress(f, len(d)) == (True, d)]


def compressor_rejects(max_size: int | None = None):
    """(name, frame, original) for the frames libzstd 1.4.9 writes but does
    not decode back (its experimental target block size on long zero runs
    emits frames whose blocks overrun the declared content size): a
    decoder must refuse them."""
    return [(n, f, d) for n, f, d in _frames(max_size)
            if libs.zstd_decompress(f, len(d)) != (True, d)]


# ------------------------------------------------------------ hand-made

def _frame(descriptor: int, rest: bytes) -> bytes:
    return struct.pack("<I", 0xFD2FB528) + bytes([descriptor]) + rest


def crafted():
    """(name, frame, capacity, status): header and block rungs, with the
    status RFC 8878 and the v1 subset give them (checked against the
    reference twin by the tests)."""
    U, C, T, OK = libs.UNSUPPORTED, libs.CORRUPT, libs.TOO_SMALL, libs.OK
    raw_block_last = lambda body: (((len(body) << 3) | 1)).to_bytes(3, "little") + body
    hello = b"hello, zstd"
    single = _frame(0x20, bytes([len(hello)]) + raw_block_last(hello))  # single segment, 1-byte FCS
    out = [
        ("single-segment raw", single, len(hello), OK),
        ("capacity one short", single, len(hello) - 1, T),
        ("skippable", struct.pack("<II", 0x184D2A53, 4) + b"abcd", 16, U),
        ("magic truncated", b"\x28\xB5\x2F", 16, C),
        ("magic wrong", b"\x28\xB5\x2F\xFE\x20\x00\x01\x00\x00", 16, C),
        ("descriptor truncated", struct.pack("<I", 0xFD2FB528), 16, C),
        ("reserved bit", _frame(0x28, b"\x00\x01\x00\x00"), 16, C),
        ("dictionary id", _frame(0x21, b"\x05\x00" + raw_block_last(b"")), 16, U),
        ("no content size", _frame(0x00, b"\x00" + raw_block_last(b"")), 16, U),
        ("window 16 MiB", _frame(0x40, b"\x70\x00\x00" + raw_block_last(b"")), 1 << 20, U),
        ("window 8 MiB", _frame(0x40, b"\x68\x05\x00" + raw_block_last(b"\x00" * 261)), 1 << 20,
         OK),
        ("header truncated", _frame(0x80, b"\x01\x00"), 16, C),
        ("block header truncated", _frame(0x20, b"\x00\x01"), 16, C),
        ("block type 3", _frame(0x20, b"\x00" + bytes([0x07, 0, 0])), 16, C),
        ("block past window", _frame(0x20, b"\x02" + (((3 << 3) | 1)).to_bytes(3, "little")
                                     + b"abc"), 16, C),
        ("block body truncated", _frame(0x20, b"\x05" + (((5 << 3) | 1)).to_bytes(3, "little")
                                        + b"abc"), 16, C),
        ("content size mismatch", _frame(0x20, b"\x06" + raw_block_last(b"abc")), 16, C),
        ("trailing byte", single + b"\x00", len(hello), C),
        ("rle block", _frame(0x20, b"\x40" + (((64 << 3) | 3)).to_bytes(3, "little") + b"z"),
         64, OK),
        ("empty compressed block", _frame(0x20, b"\x00" + (((0 << 3) | 5)).to_bytes(3, "little")),
         16, C),
    ]
    # compressed blocks with run-length literals (libzstd emits them only
    # for long single-byte literal runs): five 'Q', then one sequence with
    # run-length tables for all three fields (literal length code 5, offset
    # value 1 = repeat offset 1, match length code 0 = 3 bytes), whose
    # bitstream is the start marker alone; and twenty 'x' with no sequences
    seq_body = bytes([0x29, ord("Q"), 0x01, 0x54, 5, 0, 0, 0x01])
    out += [("rle literals, rle tables", _frame(0x20, bytes([8]) + ((len(seq_body) << 3) | 5)
                                                .to_bytes(3, "little") + seq_body), 8, OK),
            ("rle literals only", _frame(0x20, bytes([20]) + ((3 << 3) | 5).to_bytes(3, "little")
                                         + bytes([0xA1, ord("x"), 0x00])), 20, OK)]
    with_sum = libs.zstd_compress(b"checksummed content " * 40, 3, checksum=1)
    bad_sum = with_sum[:-1] + bytes([with_sum[-1] ^ 0x40])
    out += [("checksum", with_sum, 800, OK), ("checksum flipped", bad_sum, 800, C),
            ("checksum truncated", with_sum[:-2], 800, C)]
    return out


# ------------------------------------------------------------ mutants

def mutants(frames, per_frame: int, seed: int):
    """(name, frame, capacity) with one of: a flipped bit, a byte replaced, a
    truncation, or a capacity below the original; frames are (name, frame,
    original) triples."""
    rng = random.Random(seed)
    out = []
    for name, frame, orig in frames:
        for k in range(per_frame):
            m = bytearray(frame)
            kind = rng.randrange(4)
            if kind == 0:
                p = rng.randrange(len(m))
                m[p] ^= 1 << rng.randrange(8)
            elif kind == 1:
                p = rng.randrange(len(m))
                m[p] = rng.randrange(256)
            elif kind == 2:
                m = m[:rng.randrange(len(m))]
            cap = len(orig) if kind != 3 else rng.randrange(len(orig) + 1)
            out.append((f"{name}#{k}", bytes(m), cap))
    return out
