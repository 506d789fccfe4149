"""GPU parity for the Zstandard batch decoder (cudec_zstd_decompress_batch),
through the C ABI, against the reference's own frame twin
(tests/zstd_twin_driver.h, compiled from the reference tree into
oracle/_ref) for every status and byte, and against libzstd 1.4.9 for the
bytes of valid frames: a corpus reaching every block, literals and table
mode, hand-made header frames, capacities one short, seeded mutants, content
checksums flipped, and a batch of more frames than the persistent grid has
waves. Exact status, byte and poison parity on every chunk."""
import random

import numpy as np
import pytest

from tests import libs
from tests import zstd_corpus as zc
from tests.gpu_util import check_against_oracle, gpu, run_device

pytestmark = [gpu,
              pytest.mark.skipif(libs.zstdlib() is None, reason="libzstd absent"),
              pytest.mark.skipif(libs.ref_zstd() is None,
                                 reason="reference twin not built (oracle/_ref)")]


def test_corpus():
    c = zc.corpus()
    res = check_against_oracle("zstd", "tuned", [f for _, f, _ in c], [len(d) for _, _, d in c],
                               [n for n, _, _ in c])
    assert np.all(res["status"] == libs.OK)


def test_corpus_with_room_and_one_short():
    c = [x for x in zc.corpus() if x[2]]
    streams = [f for _, f, _ in c] * 2
    caps = [len(d) + 100 for _, _, d in c] + [len(d) - 1 for _, _, d in c]
    names = [f"{n}+100" for n, _, _ in c] + [f"{n}-1" for n, _, _ in c]
    res = check_against_oracle("zstd", "tuned", streams, caps, names)
    assert np.all(res["status"][len(c):] == libs.TOO_SMALL)


def test_crafted_and_compressor_rejects():
    cr = zc.crafted()
    rj = zc.compressor_rejects()
    streams = [f for _, f, _, _ in cr] + [f for _, f, _ in rj]
    caps = [c for _, _, c, _ in cr] + [len(d) for _, _, d in rj]
    names = [n for n, _, _, _ in cr] + [n for n, _, _ in rj]
    res = check_against_oracle("zstd", "tuned", streams, caps, names)
    for i, (name, _, _, want) in enumerate(cr):
        assert int(res["status"][i]) == want, name


@pytest.mark.parametrize("seed", [1, 2])
def test_mutants(seed):
    frames = [c for c in zc.corpus(max_size=65536) if len(c[1]) > 8]
    m = zc.mutants(frames, per_frame=12, seed=seed)
    res = check_against_oracle("zstd", "tuned", [f for _, f, _ in m], [c for _, _, c in m],
                               [n for n, _, _ in m])
    got = set(int(s) for s in res["status"])
    assert {libs.OK, libs.CORRUPT, libs.TOO_SMALL} <= got


def test_content_checksum_flips():
    """Frames that carry a content checksum: the original decodes, and every
    single-bit flip of the stored checksum is CORRUPT (the decode itself
    succeeds; the checksum pass refuses it)."""
    rng = random.Random(7)
    streams, caps, names = [], [], []
    for name, frame, orig in zc.corpus(max_size=200_000):
        if not name.endswith("checksum") or not orig:
            continue
        streams.append(frame)
        caps.append(len(orig))
        names.append(name)
        bit = rng.randrange(32)
        bad = bytearray(frame)
        bad[len(bad) - 4 + bit // 8] ^= 1 << (bit % 8)
        streams.append(bytes(bad))
        caps.append(len(orig))
        names.append(f"{name}/checksum bit {bit}")
    res = check_against_oracle("zstd", "tuned", streams, caps, names)
    assert np.all(res["status"][0::2] == libs.OK)
    assert np.all(res["status"][1::2] == libs.CORRUPT)


def test_many_frames():
    """5000 frames (more than the persistent grid's waves), token, random and
    repetitive data of 1 B to 256 KiB at levels 1, 3 and 9, checked against
    libzstd's output."""
    rng = random.Random(11)
    kinds = (0, 1, 2)
    streams, origs = [], []
    for i in range(5000):
        size = int(2 ** rng.uniform(0, 18))
        data = zc._data(kinds[i % 3], size, rng.getrandbits(63))
        streams.append(libs.zstd_compress(data, (1, 3, 9)[i % 3], checksum=i % 2))
        origs.append(data)
    db, res, arena = run_device("zstd", "tuned", streams, [len(d) for d in origs])
    assert np.all(res["status"] == libs.OK)
    for i, d in enumerate(origs):
        assert int(res["bytes_written"][i]) == len(d)
        assert db.slot(arena, i)[:len(d)].tobytes() == d, i


def test_two_streams_at_once():
    """Two batches decoding at the same time on two streams, each with its
    own scratch: both equal libzstd's output."""
    import torch
    from cudec_amd.device import DeviceBatch
    from cudec_amd import workload
    batches = [workload.uneven(300, seed=s, codec=workload.CODEC_ZSTD, octaves=8)
               for s in (0x51, 0x52)]
    dbs = [DeviceBatch(b.comp, b.comp_off, b.comp_size, b.sizes, out_off=b.orig_off)
           for b in batches]
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    torch.cuda.synchronize()
    for db, s in zip(dbs, streams):
        assert db.run("zstd", "tuned", stream=s) == 0
    torch.cuda.synchronize()
    for b, db in zip(batches, dbs):
        res = db.results()
        assert np.all(res["status"] == libs.OK)
        assert np.array_equal(res["bytes_written"], b.sizes)
        assert np.array_equal(db.output_arena()[:b.orig.size], b.orig)
