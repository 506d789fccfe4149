"""The serving loop (cudec_mi355x_serve_lz4) and the device digest it uses.

Every chunk decoded in the mix must equal its original (digest and
bytes_written) and equal what the same request decodes to alone; malformed
chunks get the batch decoder's own status in the mix; admission never
holds more than the budget; and the device digest equals its host
definition across lengths that exercise every tail."""
import numpy as np
import pytest
import torch

import cudec_amd
from cudec_amd import serve
from tests import corpus
from tests.gpu_util import gpu, oracle_decode

pytestmark = gpu


@pytest.fixture(scope="module")
def pool():
    p = serve.chunk_pool(48, seed=0x5E7, min_size=1024, octaves=7)   # [1 KiB, 128 KiB)
    d = torch.from_numpy(p.comp.copy()).to("cuda")
    torch.cuda.synchronize()
    return p, d


def test_device_digest_matches_host():
    rng = np.random.default_rng(3)
    lengths = list(range(0, 70)) + [255, 256, 257, 4095, 65536 + 3, 1 << 20]
    bufs = [rng.integers(0, 256, n, dtype=np.uint8) for n in lengths]
    arena = torch.from_numpy(np.concatenate(bufs + [np.zeros(16, np.uint8)])).to("cuda")
    offs = np.concatenate([[0], np.cumsum(lengths[:-1])]).astype(np.uint64) + 3   # unaligned
    arena = torch.cat([torch.zeros(3, dtype=torch.uint8, device="cuda"), arena])
    ptrs = torch.from_numpy((offs + np.uint64(arena.data_ptr())).view(np.int64)).to("cuda")
    sizes = torch.from_numpy(np.array(lengths, np.uint64).view(np.int64)).to("cuda")
    out = torch.zeros(len(lengths), dtype=torch.int64, device="cuda")
    st = cudec_amd.lib().cudec_mi355x_digest_batch(ptrs.data_ptr(), sizes.data_ptr(), None,
                                                   len(lengths), out.data_ptr(), None)
    assert st == 0
    got = out.cpu().numpy().view(np.uint64)
    for n, b, g in zip(lengths, bufs, got):
        assert int(g) == serve.host_digest(b), n


def test_mix_equals_alone_equals_original(pool):
    p, d = pool
    want = np.array([serve.host_digest(p.original(i)) for i in range(p.n)], dtype=np.uint64)
    req, ids = serve.make_trace(rate=20000.0, n_requests=120, n_pool=p.n, seed=11,
                                max_chunks=32)
    budget = 6 << 20
    run = serve.serve(d.data_ptr(), p, req, ids, budget=budget, inflight=4)
    assert run.status == 0
    assert np.all(run.results["status"] == 0)
    assert np.array_equal(run.results["bytes_written"], p.sizes[ids])
    assert np.array_equal(run.digests, want[ids])
    assert np.all(run.latency_s >= 0)
    s = run.stats
    assert s.peak_pool_bytes <= budget
    assert s.launches >= 2 and s.max_batch_requests >= 2   # requests were batched together
    assert s.output_bytes == int(p.sizes[ids].sum())
    # each of a few requests alone: the same records and digests
    for r in (0, 7, 63, 119):
        first, count = int(req["first_chunk"][r]), int(req["chunk_count"][r])
        one = np.zeros(1, dtype=serve.REQUEST_DTYPE)
        one["chunk_count"] = count
        sub = ids[first:first + count]
        alone = serve.serve(d.data_ptr(), p, one, sub, budget=budget, inflight=1)
        assert alone.status == 0
        assert np.array_equal(alone.results, run.results[first:first + count])
        assert np.array_equal(alone.digests, run.digests[first:first + count])


def test_malformed_chunks_in_the_mix(pool):
    """Chunks truncated by one byte sit among good ones: each gets the status
    the batch decoder (and the oracle) gives it alone."""
    p, d = pool
    req, ids = serve.make_trace(rate=50000.0, n_requests=40, n_pool=p.n, seed=12,
                                max_chunks=16)
    caps = p.sizes[ids].copy()
    caps[::5] = caps[::5] // 2          # too small for every fifth chunk
    run = serve.serve(d.data_ptr(), p, req, ids, budget=4 << 20, inflight=3, caps=caps)
    assert run.status == 0
    for i, c in enumerate(ids):
        st, out = oracle_decode("lz4", p.compressed(int(c)), int(caps[i]))
        assert run.results["status"][i] == st, i
        if st == 0:
            assert run.results["bytes_written"][i] == len(out)
            assert int(run.digests[i]) == serve.host_digest(out)
        else:
            assert run.results["bytes_written"][i] == 0 and run.digests[i] == 0


def test_budget_smaller_than_the_load_still_completes(pool):
    """A budget that holds about one request at a time: admission waits for
    frees, every request completes, the pool never exceeds the budget."""
    p, d = pool
    req, ids = serve.make_trace(rate=1e6, n_requests=30, n_pool=p.n, seed=13, max_chunks=8)
    per_req = [int(np.sum((p.sizes[ids[int(f):int(f) + int(c)]] + 255) // 256 * 256))
               for f, c in zip(req["first_chunk"], req["chunk_count"])]
    budget = max(per_req)
    run = serve.serve(d.data_ptr(), p, req, ids, budget=budget, inflight=4)
    assert run.status == 0 and np.all(run.results["status"] == 0)
    assert run.stats.peak_pool_bytes <= budget


def test_zero_budget_with_zero_capacity_chunks_returns():
    """A request whose chunks all have capacity 0 holds no pool bytes, so it
    is admitted even with a zero budget: the loop must finish, and each
    one-byte block {00} decodes to OK with 0 bytes (the empty LZ4 block)."""
    from types import SimpleNamespace
    n = 6
    blocks = SimpleNamespace(comp_off=np.zeros(n, np.uint64), comp_size=np.ones(n, np.uint64),
                             sizes=np.zeros(n, np.uint64))
    d = torch.zeros(16, dtype=torch.uint8, device="cuda")
    req = np.zeros(3, dtype=serve.REQUEST_DTYPE)
    req["arrival_s"] = [0.0, 0.0, 0.001]
    req["first_chunk"] = [0, 2, 4]
    req["chunk_count"] = [2, 2, 2]
    run = serve.serve(d.data_ptr(), blocks, req, np.arange(n), budget=0, inflight=2)
    assert run.status == 0
    assert np.all(run.results["status"] == 0) and np.all(run.results["bytes_written"] == 0)
    assert run.stats.peak_pool_bytes == 0
