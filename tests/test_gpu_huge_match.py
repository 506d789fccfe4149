"""A match longer than 2^32 bytes (tests/huge_match_gpu.cu:44-176): the
literals "ABC", one match at offset 3 of 2^32 + 4096 bytes, then the
literal tail "wxyz.". The output is "ABC"[i mod 3] for i < 2^32 + 4099,
then "wxyz."; 2^32 mod 3 = 1, so a length or position truncated to 32 bits
shows up as a phase error. Checked on every launch shape: bytes_written, a
window astride 4 GiB, the head and the tail, and the whole output in
256 MiB slices against the closed form."""
import pytest
import torch

from cudec_amd.device import DeviceBatch
from tests.gpu_util import LAUNCHES, gpu, launch_shape

pytestmark = gpu

MATCH = (1 << 32) + 4096


def huge_stream() -> bytes:
    rest = MATCH - 4 - 15
    ext = b"\xff" * (rest // 255) + bytes([rest % 255])
    return bytes([0x3F]) + b"ABC" + b"\x03\x00" + ext + bytes([0x50]) + b"wxyz."


def expected(lo: int, hi: int) -> bytes:
    """Output bytes [lo, hi) by the closed form."""
    n = 3 + MATCH
    out = bytearray()
    for i in range(lo, hi):
        out.append(b"ABC"[i % 3] if i < n else b"wxyz."[i - n])
    return bytes(out)


@pytest.fixture(scope="module")
def stream():
    return huge_stream()


@pytest.mark.parametrize("form", LAUNCHES)
def test_match_over_4gib(stream, form):
    total = 3 + MATCH + 5
    print(f"[huge match] {form}: {total} bytes out", flush=True)
    # the arena is read back in slices, not whole
    db = DeviceBatch.from_streams([stream], [total], poison=0xA5)
    db.reset_results()
    with launch_shape("lz4", form) as f:
        assert db.run("lz4", f) == 0
        torch.cuda.synchronize()
    res = db.results()
    assert int(res["status"][0]) == 0 and int(res["bytes_written"][0]) == total
    out = db.d_out
    for lo, hi in ((0, 64), ((1 << 32) - 64, (1 << 32) + 64), (total - 70, total)):
        assert out[lo:hi].cpu().numpy().tobytes() == expected(lo, hi), (lo, hi)
    # the whole output, a slice at a time: byte i is "ABC"[i mod 3]
    slice_len = 3 << 26
    pattern = torch.tensor(list(b"ABC"), dtype=torch.uint8, device="cuda").repeat(slice_len // 3)
    n = 3 + MATCH
    for lo in range(0, n, slice_len):
        hi = min(lo + slice_len, n)
        assert torch.equal(out[lo:hi], pattern[: hi - lo]), lo   # lo is a multiple of 3
    del pattern
