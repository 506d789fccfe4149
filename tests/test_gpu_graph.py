"""A tuned batch launch (LZ4, Snappy, Zstandard) captured into a HIP graph and replayed
(cudec_amd/graph.py over the unchanged batch entries). Every replay must
give the eager launch's statuses, bytes_written and output bytes, with the
program's own eager work and synchronisations between replays; malformed
chunks are mixed in, so failed chunks replay their verdicts too."""
import numpy as np
import pytest
import torch

from cudec_amd import workload
from cudec_amd.device import DeviceBatch
from cudec_amd.graph import CapturedLaunch
from tests.gpu_util import gpu

pytestmark = gpu

REPLAYS = 4


def mixed_streams(codec):
    """Chunks from 1 KiB to 4 MiB, every seventh truncated by one byte and
    every eleventh given a capacity one byte short."""
    b = workload.uneven(96, seed=0x6A4F, codec=codec, min_size=1024, octaves=12)
    comp = [b.compressed(i) for i in range(b.n)]
    caps = [int(s) for s in b.sizes]
    for i in range(0, b.n, 7):
        comp[i] = comp[i][:-1]
    for i in range(3, b.n, 11):
        caps[i] = max(caps[i] - 1, 0)
    return b, comp, caps


@pytest.mark.parametrize("fmt,codec", [("lz4", workload.CODEC_LZ4),
                                       ("snappy", workload.CODEC_SNAPPY),
                                       ("zstd", workload.CODEC_ZSTD)])
def test_graph_replays_equal_the_eager_launch(fmt, codec):
    b, comp, caps = mixed_streams(codec)
    print(f"[graph] {fmt}: {b.n} chunks, {sum(caps)} output bytes", flush=True)
    db = DeviceBatch.from_streams(comp, caps, poison=0x5A)
    db.reset_results()
    assert db.run(fmt, "tuned") == 0
    torch.cuda.synchronize()
    want_res = db.results().copy()
    want_out = db.output_arena().copy()
    ok = want_res["status"] == 0
    assert np.any(ok) and np.any(~ok)
    for i in np.nonzero(ok)[0]:
        assert db.slot(want_out, int(i))[:want_res["bytes_written"][i]].tobytes() == \
            b.original(int(i))

    # other eager work between replays: a second, smaller batch
    other = DeviceBatch.from_streams(comp[:9], caps[:9])
    s = torch.cuda.Stream()
    with CapturedLaunch(lambda: db.run(fmt, "tuned", stream=s), s) as g:
        for k in range(REPLAYS):
            with torch.cuda.stream(s):
                db.d_out.fill_(0x5A)
                db.reset_results()
            g.replay()
            s.synchronize()
            assert np.array_equal(db.results(), want_res), f"replay {k}: results differ"
            assert np.array_equal(db.output_arena(), want_out), f"replay {k}: output differs"
            other.reset_results()
            assert other.run(fmt, "tuned") == 0
            torch.cuda.synchronize()
            assert np.array_equal(other.results(), want_res[:9])
