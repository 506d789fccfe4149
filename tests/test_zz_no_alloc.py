"""The allocation-free LZ4 entry (cudec_mi355x_lz4_decompress_batch_noalloc)
against the allocating one (cudec_lz4_decompress_batch), through the C ABI:
the same results and the same output bytes, poison past bytes_written
included, on two calls in a row (the second with a smaller batch, so it
must not see the first call's queue) on both tuned launch shapes, with
malformed chunks mixed in."""
import numpy as np
import pytest
import torch

import cudec_amd
from cudec_amd import workload
from cudec_amd.device import DeviceBatch
from tests.gpu_util import gpu, launch_shape

pytestmark = gpu


def noalloc(db):
    return cudec_amd.lib().cudec_mi355x_lz4_decompress_batch_noalloc(
        db.src_ptrs.data_ptr(), db.src_sizes.data_ptr(), db.dst_ptrs.data_ptr(),
        db.dst_caps.data_ptr(), db.n, db.d_results.data_ptr(), None)


def batch(b, n):
    comp = [b.compressed(i) for i in range(n)]
    caps = [int(s) for s in b.sizes[:n]]
    for i in range(0, n, 9):
        comp[i] = comp[i][:-1]
    for i in range(4, n, 13):
        caps[i] = max(caps[i] - 1, 0)
    return comp, caps


def decode(comp, caps, entry):
    db = DeviceBatch.from_streams(comp, caps, poison=0x5A)
    db.reset_results()
    assert entry(db) == 0
    torch.cuda.synchronize()
    return db.results().copy(), db.output_arena().copy()


@pytest.mark.parametrize("launch", ["tuned-pipe", "tuned-split"])
def test_noalloc_equals_the_allocating_entry(launch):
    b = workload.uneven(240, seed=0xA110C, min_size=1024, octaves=9)
    with launch_shape("lz4", launch):
        for n in (240, 77):
            comp, caps = batch(b, n)
            want_res, want_out = decode(comp, caps, lambda db: db.run("lz4", "tuned"))
            got_res, got_out = decode(comp, caps, noalloc)
            assert np.any(want_res["status"] == 0) and np.any(want_res["status"] != 0)
            assert np.array_equal(got_res, want_res), (launch, n)
            assert np.array_equal(got_out, want_out), (launch, n)
