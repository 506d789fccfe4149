"""Device XXH32 and XXH64 (cudec_mi355x_xxh32_batch, _xxh64_batch) against
the system's libxxhash (or the test-side restatements pinned to it): one
batch of segments of every length 0..130 plus 255..257, 4095, 64 KiB + 7
and 1 MiB + 3, at unaligned starts, at two seeds. The batch is wider than
one launch's grid stride holds at once only in the large case."""
import ctypes

import numpy as np
import pytest
import torch

import cudec_amd
from tests import libs
from tests.gpu_util import gpu

pytestmark = gpu


def host_hash(bits, data, seed):
    x = libs.xxhash_lib()
    if x is not None:
        buf = ctypes.create_string_buffer(data, len(data))
        return (x.XXH32 if bits == 32 else x.XXH64)(buf, len(data), seed)
    return (libs.xxh32 if bits == 32 else libs.xxh64)(data, seed)


@pytest.mark.parametrize("bits", [32, 64])
@pytest.mark.parametrize("seed", [0, 0x9E3779B1])
def test_device_xxhash_matches(bits, seed):
    rng = np.random.default_rng(bits + seed)
    lengths = list(range(0, 131)) + [255, 256, 257, 4095, 65536 + 7, (1 << 20) + 3]
    bufs = [rng.integers(0, 256, n, dtype=np.uint8) for n in lengths]
    offs = []
    pos = 0
    for n in lengths:
        pos += 1 + int(rng.integers(0, 15))       # unaligned starts
        offs.append(pos)
        pos += n
    arena = np.zeros(pos + 16, dtype=np.uint8)
    for o, b in zip(offs, bufs):
        arena[o:o + len(b)] = b
    d = torch.from_numpy(arena).to("cuda")
    ptrs = torch.from_numpy((np.array(offs, np.uint64) + np.uint64(d.data_ptr())).view(np.int64)).to("cuda")
    sizes = torch.from_numpy(np.array(lengths, np.uint64).view(np.int64)).to("cuda")
    out = torch.zeros(len(lengths), dtype=torch.int64, device="cuda")
    fn = cudec_amd.lib().cudec_mi355x_xxh32_batch if bits == 32 else \
        cudec_amd.lib().cudec_mi355x_xxh64_batch
    assert fn(ptrs.data_ptr(), sizes.data_ptr(), len(lengths), seed, out.data_ptr(), None) == 0
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    got = got.view(np.uint32)[: len(lengths)] if bits == 32 else got.view(np.uint64)
    for n, b, g in zip(lengths, bufs, got):
        assert int(g) == host_hash(bits, b.tobytes(), seed), (bits, seed, n)
