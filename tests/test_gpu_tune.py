"""The LZ4 launch's pipelining threshold and its autotuner
(cudec_mi355x_set_lz4_pipeline_threshold, cudec_mi355x_autotune_lz4).
Output bytes, statuses and bytes_written must not depend on the setting:
every chunk on two waves, every chunk on one, and the split in between
(the queue runs costliest first, so a block splits mid-launch) all equal
the plain form and the originals, with malformed chunks mixed in."""
import numpy as np
import pytest
import torch

import cudec_amd
from cudec_amd import workload
from cudec_amd.device import DeviceBatch
from tests.gpu_util import gpu

pytestmark = gpu

THRESHOLDS = [0, 64 << 10, 512 << 10, 2 << 20, (1 << 64) - 1]


@pytest.fixture()
def restore_threshold():
    st, old = cudec_amd.lz4_pipeline_threshold()
    assert st == 0
    yield
    assert cudec_amd.set_lz4_pipeline_threshold(old) == 0


@pytest.fixture(scope="module")
def mixed():
    """A few hundred chunks from 1 KiB to 4 MiB; every seventh truncated by
    one byte and every eleventh given a capacity one byte short."""
    b = workload.uneven(160, seed=0x7E57, min_size=1024, octaves=12)
    comp = [b.compressed(i) for i in range(b.n)]
    caps = [int(s) for s in b.sizes]
    for i in range(0, b.n, 7):
        comp[i] = comp[i][:-1]
    for i in range(3, b.n, 11):
        caps[i] = max(caps[i] - 1, 0)
    return b, comp, caps


def run(comp, caps, form):
    db = DeviceBatch.from_streams(comp, caps, poison=0xA5)
    db.reset_results()
    assert db.run("lz4", form) == 0
    torch.cuda.synchronize()
    return db.results().copy(), db.output_arena().copy(), db


def test_results_do_not_depend_on_the_threshold(mixed, restore_threshold):
    b, comp, caps = mixed
    want_res, want_out, db = run(comp, caps, "plain")
    assert np.any(want_res["status"] != 0) and np.any(want_res["status"] == 0)
    ok = want_res["status"] == 0
    for i in np.nonzero(ok)[0]:
        assert db.slot(want_out, int(i))[:want_res["bytes_written"][i]].tobytes() == b.original(int(i))
    for t in THRESHOLDS:
        assert cudec_amd.set_lz4_pipeline_threshold(t) == 0
        assert cudec_amd.lz4_pipeline_threshold() == (0, t)
        res, out, db2 = run(comp, caps, "tuned")
        assert np.array_equal(res, want_res), t
        for i in np.nonzero(ok)[0]:
            n = int(want_res["bytes_written"][i])
            assert np.array_equal(db2.slot(out, int(i))[:n], db.slot(want_out, int(i))[:n]), (t, i)
            # bytes past bytes_written are never written on OK
            assert np.all(db2.slot(out, int(i))[n:] == 0xA5), (t, i)


def test_autotune_picks_a_candidate_and_keeps_it(mixed, restore_threshold):
    b, comp, caps = mixed
    db = DeviceBatch.from_streams(comp, caps)
    st, best, sweep = cudec_amd.autotune_lz4(db.src_ptrs.data_ptr(), db.src_sizes.data_ptr(),
                                             db.dst_ptrs.data_ptr(), db.dst_caps.data_ptr(), db.n,
                                             db.d_results.data_ptr(), None)
    assert st == 0
    cands = [c for c, _ in sweep]
    assert cands == [0, 64 << 10, 128 << 10, 256 << 10, 512 << 10, 1 << 20, 2 << 20, 4 << 20,
                     (1 << 64) - 1]
    assert best in cands
    assert all(ms > 0 for _, ms in sweep)
    assert min(sweep, key=lambda x: x[1])[0] == best
    assert cudec_amd.lz4_pipeline_threshold() == (0, best)
    # the batch's results are those of a normal launch
    torch.cuda.synchronize()
    res = db.results()
    ref, _, _ = run(comp, caps, "plain")
    assert np.array_equal(res, ref)

