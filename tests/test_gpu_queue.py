"""GPU parity for batches of many small chunks, which the tuned launch
decodes from eight interleaved sub-queues (16 or more chunks per wave:
on a 256-CU MI355X, 65536 chunks with two-wave blocks and 81920 when
every chunk runs on one wave). Fewer chunks per wave keep one queue, which
the other tests cover. 120000 chunks of 16 B to 2 KiB, LZ4 and Snappy, on
both launch shapes, with malformed chunks mixed in, must equal the plain
form chunk by chunk: statuses, bytes_written, output bytes and the poison
beyond bytes_written."""
import numpy as np
import pytest
import torch

import cudec_amd
from cudec_amd import workload
from cudec_amd.device import DeviceBatch
from tests.gpu_util import gpu

pytestmark = gpu

N_CHUNKS = 120_000


@pytest.fixture()
def restore_threshold():
    st, old = cudec_amd.lz4_pipeline_threshold()
    assert st == 0
    yield
    assert cudec_amd.set_lz4_pipeline_threshold(old) == 0


def small_batch(codec: int, seed: int):
    rng = np.random.default_rng(seed)
    sizes = rng.integers(16, 2049, N_CHUNKS).astype(np.uint64)
    kinds = rng.choice([workload.KIND_TOKENS, workload.KIND_REPETITIVE, workload.KIND_RANDOM],
                       N_CHUNKS, p=[0.8, 0.1, 0.1]).astype(np.int32)
    seeds = np.arange(N_CHUNKS, dtype=np.uint64) + seed
    b = workload.build(codec, sizes, kinds, seeds)
    comp = [b.compressed(i) for i in range(b.n)]
    caps = [int(s) for s in b.sizes]
    for i in range(5, b.n, 13):
        comp[i] = comp[i][:-1]
    for i in range(8, b.n, 17):
        caps[i] = max(caps[i] - 1, 0)
    return b, comp, caps


def run(fmt, comp, caps, form):
    db = DeviceBatch.from_streams(comp, caps, poison=0xA5)
    db.reset_results()
    assert db.run(fmt, form) == 0
    torch.cuda.synchronize()
    return db.results().copy(), db.output_arena().copy(), db


@pytest.mark.parametrize("fmt,codec", [("lz4", workload.CODEC_LZ4),
                                       ("snappy", workload.CODEC_SNAPPY)])
def test_many_small_chunks_match_plain(fmt, codec, restore_threshold):
    b, comp, caps = small_batch(codec, 0x51C0 + codec)
    print(f"[queue] {fmt}: {b.n} chunks", flush=True)
    want_res, want_out, db = run(fmt, comp, caps, "plain")
    ok = np.nonzero(want_res["status"] == 0)[0]
    assert 0 < len(ok) < b.n
    for i in ok[::97]:
        n = int(want_res["bytes_written"][i])
        assert db.slot(want_out, int(i))[:n].tobytes() == b.original(int(i))
    for t in (0, (1 << 64) - 1):
        assert cudec_amd.set_lz4_pipeline_threshold(t) == 0
        print(f"[queue] {fmt}: threshold {t}", flush=True)
        res, out, db2 = run(fmt, comp, caps, "tuned")
        assert np.array_equal(res, want_res), (fmt, t)
        for i in ok:
            n = int(want_res["bytes_written"][i])
            got = db2.slot(out, int(i))
            assert np.array_equal(got[:n], db.slot(want_out, int(i))[:n]), (fmt, t, i)
            assert np.all(got[n:] == 0xA5), (fmt, t, i)
