"""The parse-only build (cudec_mi355x_parse_batch): for every chunk the same
status and bytes_written as the format's batch entry, and no destination
byte written, on both tuned launch shapes, for LZ4 and Snappy batches with
malformed and short-capacity chunks mixed in."""
import numpy as np
import pytest
import torch

import cudec_amd
from cudec_amd import workload
from cudec_amd.device import DeviceBatch
from tests.gpu_util import gpu, launch_shape

pytestmark = gpu


@pytest.mark.parametrize("fmt", ["lz4", "snappy"])
@pytest.mark.parametrize("launch", ["tuned-pipe", "tuned-split"])
def test_parse_only_matches_the_decode(fmt, launch):
    codec = workload.CODEC_LZ4 if fmt == "lz4" else workload.CODEC_SNAPPY
    b = workload.uneven(200, seed=0x9A55, codec=codec, min_size=1024, octaves=11)
    comp = [b.compressed(i) for i in range(b.n)]
    caps = [int(s) for s in b.sizes]
    for i in range(0, b.n, 7):
        comp[i] = comp[i][: len(comp[i]) // 2]
    for i in range(2, b.n, 9):
        caps[i] = max(caps[i] - 3, 0)
    with launch_shape(fmt, launch):
        want = DeviceBatch.from_streams(comp, caps, poison=0x6B)
        want.reset_results()
        assert want.run(fmt, "tuned") == 0
        parsed = DeviceBatch.from_streams(comp, caps, poison=0x6B)
        parsed.reset_results()
        st = cudec_amd.lib().cudec_mi355x_parse_batch(
            cudec_amd.FORMAT_CODE[fmt], parsed.src_ptrs.data_ptr(), parsed.src_sizes.data_ptr(),
            parsed.dst_ptrs.data_ptr(), parsed.dst_caps.data_ptr(), parsed.n,
            parsed.d_results.data_ptr(), None)
        assert st == 0
        torch.cuda.synchronize()
    r_want, r_got = want.results(), parsed.results()
    assert np.any(r_want["status"] == 0) and np.any(r_want["status"] != 0)
    assert np.array_equal(r_got, r_want)
    assert bool(torch.all(parsed.d_out == 0x6B)), "the parse-only build wrote output bytes"
