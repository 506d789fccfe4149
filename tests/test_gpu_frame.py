"""cudec_lz4f_decompress on the device, held to liblz4's frame API the way
the reference's frame test is (tests/frame_twin.cu:137-456): independent
frames of every flag combination and size class decode byte-exact,
declared content sizes are checked, payload mutations reach the kernel's
status seam with the oracle's direction kept, blocks that sit apart in the
device slots and together in the output (and the reverse) are placed
right, and every reject branch fails closed with bytes_written 0."""
import numpy as np
import pytest

import cudec_amd
from tests import libs
from tests.gpu_util import gpu

pytestmark = [gpu, pytest.mark.skipif(libs.lz4lib() is None, reason="liblz4 absent")]

OK = cudec_amd.CUDEC_OK
CORRUPT = cudec_amd.CUDEC_ERR_CORRUPT_INPUT
UNSUPPORTED = cudec_amd.CUDEC_ERR_UNSUPPORTED
TOO_SMALL = cudec_amd.CUDEC_ERR_OUTPUT_TOO_SMALL
BLOCK64K = 64 << 10


def make_data(n: int, seed: int, mode: int) -> bytes:
    """mode 0: repeating letters (compresses), 1: PRNG bytes (stored
    blocks), 2: 8-byte runs alternating zeros and PRNG bytes (mixed)."""
    rng = np.random.default_rng(seed)
    i = np.arange(n)
    if mode == 0:
        return (ord("A") + (i // 37) % 26).astype(np.uint8).tobytes()
    noise = rng.integers(0, 256, n, dtype=np.uint8)
    if mode == 1:
        return noise.tobytes()
    return np.where((i // 8) % 2 == 1, noise, 0).astype(np.uint8).tobytes()


def block_table(frame: bytes):
    """(offset, length, stored) of each block of a frame with a 7-byte
    header and no block checksums."""
    pos, out = 7, []
    while True:
        bs = int.from_bytes(frame[pos:pos + 4], "little")
        pos += 4
        if bs == 0:
            return out
        out.append((pos, bs & 0x7FFFFFFF, bool(bs >> 31)))
        pos += bs & 0x7FFFFFFF


def decode(frame: bytes, cap: int):
    return cudec_amd.cudec_lz4f_decompress(frame, cap)


def expect_reject(frame: bytes, status: int, cap: int = 1 << 20):
    st, out = decode(frame, cap)
    assert st == status and out == b""


@pytest.mark.parametrize("n", [0, 1, 100, 65535, 65536, 65537, 200000])
@pytest.mark.parametrize("mode", [0, 1, 2])
def test_independent_frames_all_flags(n, mode):
    data = make_data(n, 0x100 + n + mode, mode)
    for flags in range(4):
        frame = libs.lz4_frame(data, content_checksum=bool(flags & 1),
                               block_checksum=bool(flags & 2))
        st, out = decode(frame, n + 16)
        assert st == OK and out == data, (n, mode, flags)
        ok, want = libs.lz4_frame_decode(frame, n + 16)
        assert ok and want == data


@pytest.mark.parametrize("bsid", [5, 6, 7])
def test_larger_block_maxima(bsid):
    """256 KiB, 1 MiB and 4 MiB block maxima over several blocks each."""
    n = (3 * (64 << 10) << (2 * (bsid - 4))) + 12345
    data = make_data(n, bsid, 2)
    frame = libs.lz4_frame(data, block_size_id=bsid, content_checksum=True, block_checksum=True,
                           content_size=True)
    assert (frame[5] >> 4) & 7 == bsid
    st, out = decode(frame, n)
    assert st == OK and out == data
    assert decode(frame, n - 1) == (TOO_SMALL, b"")


@pytest.mark.parametrize("n", [0, 100, 65537, 200000])
def test_content_size_frames(n):
    data = make_data(n, 0x777 + n, 0)
    frame = libs.lz4_frame(data, content_size=True)
    assert bool(frame[4] & 0x08) == (n > 0)   # liblz4 writes a zero size as "unknown"
    st, out = decode(frame, n + 4)
    assert st == OK and out == data


def test_declared_size_mismatch_rejected():
    data = make_data(100000, 0x778, 0)
    frame = bytearray(libs.lz4_frame(data, content_size=True))
    frame[6] ^= 0xFF
    frame[14] = (libs.xxh32(bytes(frame[4:14])) >> 8) & 0xFF   # repaired HC
    assert libs.lz4_frame_decode(bytes(frame), len(data) + 4096)[0] is False
    expect_reject(bytes(frame), CORRUPT)


def test_payload_mutations_reach_the_kernel_seam():
    """Only bytes inside one compressed block's payload change, so every
    reject comes from the device decode; whenever liblz4 rejects, so does
    this decoder, and where both accept the bytes agree."""
    data = make_data(40000, 0x321, 0)
    base = libs.lz4_frame(data)
    (off, blen, stored), = block_table(base)
    assert not stored and off + blen + 4 == len(base)
    cap = BLOCK64K + 4096
    step = max(blen // 96, 1)
    seam_rejects = 0
    for p in range(off, off + blen, step):
        frame = bytearray(base)
        frame[p] ^= 0x5A
        frame = bytes(frame)
        ok, want = libs.lz4_frame_decode(frame, cap)
        st, out = decode(frame, cap)
        if st != OK:
            assert out == b""
            seam_rejects += 1
        if not ok:
            assert st != OK, p
        elif st == OK:
            assert out == want, p
    assert seam_rejects > 0


def test_stored_block_between_compressed_ones():
    """Blocks 0 and 2 sit back to back in device memory but 64 KiB apart in
    the output: a merged copy that looked only at the device side would
    overwrite the stored block."""
    mixed = b"".join(make_data(BLOCK64K, 0x900 + part, 1 if part == 1 else 0)
                     for part in range(3))
    frame = libs.lz4_frame(mixed)
    assert [s for _, _, s in block_table(frame)] == [False, True, False]
    st, out = decode(frame, len(mixed))
    assert st == OK and out == mixed


def test_short_block_then_full_block():
    """A short compressed block followed by a full one: the device slots
    are strided by the block maximum, so a merged copy across them would
    copy the gap. Built by splicing two frames' block records under one
    header; liblz4 accepts it."""
    short = make_data(1000, 0xA01, 0)
    full = make_data(BLOCK64K, 0xA02, 0)
    fa, fb = libs.lz4_frame(short), libs.lz4_frame(full)
    assert fa[4:6] == fb[4:6]
    (oa, la, sa), = block_table(fa)
    (ob, lb, sb), = block_table(fb)
    assert not sa and not sb
    spliced = fa[:7] + fa[oa - 4:oa + la] + fb[ob - 4:ob + lb] + b"\0\0\0\0"
    expect = short + full
    ok, want = libs.lz4_frame_decode(spliced, len(expect) + 4096)
    assert ok and want == expect
    st, out = decode(spliced, len(expect))
    assert st == OK and out == expect


def test_reject_branches():
    data = make_data(200000, 0x999, 2)
    expect_reject(libs.lz4_frame(data, independent=False), UNSUPPORTED)

    frame = bytearray(libs.lz4_frame(data))
    frame[4] |= 0x01                                           # dictionary id
    frame[6] = (libs.xxh32(bytes(frame[4:6])) >> 8) & 0xFF
    expect_reject(bytes(frame), UNSUPPORTED)

    frame = bytearray(libs.lz4_frame(data))
    frame[0] ^= 0xFF
    expect_reject(bytes(frame), CORRUPT)

    frame = bytearray(libs.lz4_frame(data))
    frame[6] ^= 0xFF
    expect_reject(bytes(frame), CORRUPT)

    frame = bytearray(libs.lz4_frame(data, content_checksum=True))
    frame[-1] ^= 0xFF
    expect_reject(bytes(frame), CORRUPT)

    frame = bytearray(libs.lz4_frame(data, block_checksum=True))
    frame[len(frame) // 2] ^= 0xFF
    expect_reject(bytes(frame), CORRUPT)

    whole = libs.lz4_frame(data, content_checksum=True, block_checksum=True)
    for cut in (3, 6, 10, len(whole) // 2, len(whole) - 1):
        st, out = decode(whole[:cut], 1 << 20)
        assert st != OK and out == b"", cut

    frame = libs.lz4_frame(data)
    expect_reject(frame, TOO_SMALL, cap=len(data) // 2)


def test_kernel_reject_status_passes_through():
    """A compressed block whose last sequence is cut short decodes to
    CORRUPT on the device; the frame reports that status."""
    data = make_data(5000, 3, 0)
    frame = bytearray(libs.lz4_frame(data))
    (off, blen, stored), = block_table(bytes(frame))
    assert not stored
    # the final byte of a compressed block is a literal of the last
    # sequence; shorten the block by one byte and rewrite its size field
    cut = bytes(frame[:off - 4]) + (blen - 1).to_bytes(4, "little") + \
        bytes(frame[off:off + blen - 1]) + b"\0\0\0\0"
    assert libs.lz4_frame_decode(cut, 1 << 20)[0] is False
    expect_reject(cut, CORRUPT)


@pytest.mark.parametrize("content_sum", [False, True])
def test_block_checksums_checked_on_the_device(content_sum):
    """Block checksums are computed on the device for every block, stored
    and compressed alike. Stored-block frames (incompressible data) and a
    many-block frame decode when every checksum holds; one flipped data
    byte or checksum byte anywhere makes the frame CORRUPT, and a bad
    checksum wins over a capacity that is too small, as the host parse
    order has it (checksums before any decode result)."""
    from tests import corpus
    stored = corpus.random_bytes(300000, 11)
    mixed = make_data(40 * 65536 + 123, 0x4C, 2)
    for data, bsid in ((stored, 4), (mixed, 4), (mixed, 5)):
        frame = libs.lz4_frame(data, block_size_id=bsid, content_checksum=content_sum,
                               block_checksum=True)
        st, out = decode(frame, len(data))
        assert st == OK and out == data
        # blocks: walk the frame to find each block's data and checksum bytes
        pos = 7
        spans = []
        while True:
            bs = int.from_bytes(frame[pos:pos + 4], "little")
            pos += 4
            if bs == 0:
                break
            n = bs & 0x7FFFFFFF
            spans.append((pos, n))
            pos += n + 4
        for k in sorted({0, len(spans) // 2, len(spans) - 1}):
            off, n = spans[k]
            for at in (off, off + n // 2, off + n + 1):        # data bytes, checksum byte
                bad = bytearray(frame)
                bad[at] ^= 0x10
                expect_reject(bytes(bad), CORRUPT)
                expect_reject(bytes(bad), CORRUPT, cap=len(data) // 3)
