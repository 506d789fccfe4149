"""The streaming context on the device (include/cudec.h:140-218), following
the reference's stream tests (tests/stream_twin.cu): host- and
device-output parity with the oracle, bytes beyond bytes_written untouched
in both spaces, bit-identical output from a fresh and a reused context
(including across staging growth and multi-wave calls), the aggregate
status rule, and poisoning through an impossible staging allocation."""
import ctypes

import numpy as np
import pytest

import cudec_amd
from tests import corpus, libs
from tests.gpu_util import POISON, gpu, oracle_decode

pytestmark = gpu

RES = np.dtype([("status", "<i4"), ("reserved", "<u4"), ("bytes_written", "<u8")])


def _host_inputs(streams):
    bufs = [np.frombuffer(s, dtype=np.uint8).copy() if s else np.zeros(1, np.uint8)
            for s in streams]
    ptrs = np.array([b.ctypes.data for b in bufs], dtype=np.uint64)
    sizes = np.array([len(s) for s in streams], dtype=np.uint64)
    return bufs, ptrs, sizes


def _run_host_out(ctx, streams, caps):
    bufs, ptrs, sizes = _host_inputs(streams)
    outs = [np.full(max(c, 1), POISON, dtype=np.uint8) for c in caps]
    dptr = np.array([o.ctypes.data for o in outs], dtype=np.uint64)
    capa = np.array(caps, dtype=np.uint64)
    res = np.zeros(len(streams), dtype=RES)
    st = cudec_amd.cudec_lz4_decompress_stream_ctx(
        ctx, ptrs.ctypes.data, sizes.ctypes.data, dptr.ctypes.data, capa.ctypes.data,
        len(streams), cudec_amd.CUDEC_MEM_HOST, res.ctypes.data)
    return st, res, outs


def _check(res, outs, streams, caps):
    first_bad = None
    for i, (s, cap) in enumerate(zip(streams, caps)):
        want_st, want = oracle_decode("lz4", s, cap)
        assert res["status"][i] == want_st, i
        assert res["reserved"][i] == 0
        if want_st == 0:
            assert res["bytes_written"][i] == len(want)
            assert outs[i][: len(want)].tobytes() == want, i
            assert np.all(outs[i][len(want):cap] == POISON), i
        else:
            assert res["bytes_written"][i] == 0
            if first_bad is None:
                first_bad = want_st
    return first_bad


@pytest.fixture()
def ctx():
    st, h = cudec_amd.cudec_stream_ctx_create()
    assert st == 0 and h
    yield h
    cudec_amd.cudec_stream_ctx_destroy(h)


def test_host_output_parity_and_reuse(ctx, lz4_fixtures):
    streams = [f.compressed for f in lz4_fixtures]
    caps = [len(f.original) + 8 for f in lz4_fixtures]
    st, res, outs = _run_host_out(ctx, streams, caps)
    assert st == 0
    assert _check(res, outs, streams, caps) is None
    # reuse with a larger, mixed batch (grows the staging), then the first
    # batch again: identical results and bytes
    mix = [m.stream for f in lz4_fixtures for m in corpus.mutate_stream(f.compressed, f.seed)]
    mix_caps = [70000] * len(mix)
    st2, res2, outs2 = _run_host_out(ctx, mix, mix_caps)
    first_bad = _check(res2, outs2, mix, mix_caps)
    assert st2 == first_bad  # aggregate: the first non-OK chunk in index order
    st3, res3, outs3 = _run_host_out(ctx, streams, caps)
    assert st3 == 0 and np.array_equal(res3, res)
    assert all(np.array_equal(a, b) for a, b in zip(outs3, outs))


def test_device_output_parity(ctx):
    import torch

    from cudec_amd import workload
    b = workload.uneven(64, seed=0x51, min_size=1024, octaves=8)
    streams = [b.compressed(i) for i in range(b.n)]
    caps = [int(s) + 5 for s in b.sizes]
    arena = torch.full((sum(caps),), POISON, dtype=torch.uint8, device="cuda")
    offs = np.concatenate([[0], np.cumsum(caps[:-1])]).astype(np.uint64)
    dptr = (offs + np.uint64(arena.data_ptr())).astype(np.uint64)
    bufs, ptrs, sizes = _host_inputs(streams)
    capa = np.array(caps, dtype=np.uint64)
    res = np.zeros(b.n, dtype=RES)
    torch.cuda.synchronize()
    st = cudec_amd.cudec_lz4_decompress_stream_ctx(
        ctx, ptrs.ctypes.data, sizes.ctypes.data, dptr.ctypes.data, capa.ctypes.data, b.n,
        cudec_amd.CUDEC_MEM_DEVICE, res.ctypes.data)
    assert st == 0
    host = arena.cpu().numpy()
    for i in range(b.n):
        o = int(offs[i])
        assert res["status"][i] == 0 and res["bytes_written"][i] == b.sizes[i]
        assert host[o:o + int(b.sizes[i])].tobytes() == b.original(i)
        assert np.all(host[o + int(b.sizes[i]):o + caps[i]] == POISON)


def test_many_waves_and_one_shot():
    """More chunks than one wave holds (double-buffered staging on both
    slots), through the one-shot entry."""
    n = 40000
    streams = [b"\x00", b"\x44abcd\x04\x00\x50wxyz!"] * (n // 2)
    caps = [1, 17] * (n // 2)
    bufs, ptrs, sizes = _host_inputs(streams)
    outs = np.full(sum(caps), POISON, dtype=np.uint8)
    offs = np.concatenate([[0], np.cumsum(caps[:-1])]).astype(np.uint64)
    dptr = (offs + np.uint64(outs.ctypes.data)).astype(np.uint64)
    capa = np.array(caps, dtype=np.uint64)
    res = np.zeros(n, dtype=RES)
    st = cudec_amd.cudec_lz4_decompress_stream(ptrs.ctypes.data, sizes.ctypes.data,
                                               dptr.ctypes.data, capa.ctypes.data, n,
                                               cudec_amd.CUDEC_MEM_HOST, res.ctypes.data)
    assert st == 0
    assert np.all(res["status"] == 0)
    assert np.array_equal(res["bytes_written"], np.array([0, 17] * (n // 2), dtype=np.uint64))
    assert outs[1:18].tobytes() == b"abcdabcdabcdwxyz!"
    assert outs[-17:].tobytes() == b"abcdabcdabcdwxyz!"


def test_impossible_staging_poisons_the_context(ctx):
    """tests/stream_twin.cu:392-450: a host-output call whose staging cannot
    be allocated poisons the context; every later call returns
    CUDEC_ERR_CUDA with every result stamped, and destroy stays valid."""
    streams = [b"\x00"]
    bufs, ptrs, sizes = _host_inputs(streams)
    dummy = np.zeros(16, np.uint8)
    dptr = np.array([dummy.ctypes.data], dtype=np.uint64)
    capa = np.array([1 << 60], dtype=np.uint64)
    res = np.zeros(1, dtype=RES)
    st = cudec_amd.cudec_lz4_decompress_stream_ctx(ctx, ptrs.ctypes.data, sizes.ctypes.data,
                                                   dptr.ctypes.data, capa.ctypes.data, 1,
                                                   cudec_amd.CUDEC_MEM_HOST, res.ctypes.data)
    assert st == cudec_amd.CUDEC_ERR_CUDA
    assert res["status"][0] == cudec_amd.CUDEC_ERR_CUDA and res["bytes_written"][0] == 0
    # an argument error still reports as such and does not touch the device
    assert cudec_amd.cudec_lz4_decompress_stream_ctx(ctx, None, sizes.ctypes.data,
                                                     dptr.ctypes.data, capa.ctypes.data, 1,
                                                     cudec_amd.CUDEC_MEM_HOST,
                                                     res.ctypes.data) == \
        cudec_amd.CUDEC_ERR_INVALID_ARGUMENT
    capa[0] = 1
    res[:] = 0
    st = cudec_amd.cudec_lz4_decompress_stream_ctx(ctx, ptrs.ctypes.data, sizes.ctypes.data,
                                                   dptr.ctypes.data, capa.ctypes.data, 1,
                                                   cudec_amd.CUDEC_MEM_HOST, res.ctypes.data)
    assert st == cudec_amd.CUDEC_ERR_CUDA
    assert res["status"][0] == cudec_amd.CUDEC_ERR_CUDA
