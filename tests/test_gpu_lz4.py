"""GPU parity for the LZ4 batch decoders, through the C ABI, against the CPU
oracle. Follows the reference's device tests: tests/smoke.cu (contract,
pending-error semantics, multi-block empty blocks), tests/gpu_fixture.cu
(pristine pairs with poison beyond bytes_written, two-direction mutant
parity, grid-stride wraparound), tests/termination_gpu.cu (the adversarial
corpus terminates; a hang is caught by the per-test timeout),
tests/determinism_gpu.cu (identical bytes across launch forms, repeated
runs and stream splits). Every comparison here is exact: status, reserved,
bytes_written, output bytes, and untouched poison.
"""
import ctypes

import numpy as np
import pytest

from tests import corpus, libs
from tests.gpu_util import FORMS, LAUNCHES, POISON, check_against_oracle, gpu, run_device

pytestmark = gpu


@pytest.fixture(scope="module")
def hip():
    import torch
    torch.cuda.init()
    L = ctypes.CDLL("libamdhip64.so.7")
    L.hipSetDevice.argtypes = [ctypes.c_int]
    return L


def _tables(n, device="cuda"):
    import torch
    t = [torch.zeros(n, dtype=torch.int64, device=device) for _ in range(4)]
    res = torch.empty(2 * n, dtype=torch.int64, device=device)
    return t, res


@pytest.mark.parametrize("form", FORMS)
def test_contract_and_empty_blocks(form, hip):
    """tests/smoke.cu:31-141: every argument reject, the pending-error
    contract, and 257 empty-block chunks across many blocks."""
    import torch

    import cudec_amd
    entry = cudec_amd.ENTRY[("lz4", form)]
    n = 257
    src = torch.zeros(n * 64, dtype=torch.uint8, device="cuda")
    dst = torch.empty(n * 64, dtype=torch.uint8, device="cuda")
    ptr = torch.arange(n, dtype=torch.int64, device="cuda") * 64
    src_ptrs = ptr + src.data_ptr()
    dst_ptrs = ptr + dst.data_ptr()
    sizes = torch.ones(n, dtype=torch.int64, device="cuda")
    caps = torch.full((n,), 64, dtype=torch.int64, device="cuda")
    res = torch.empty(2 * n, dtype=torch.int64, device="cuda")
    a = [src_ptrs.data_ptr(), sizes.data_ptr(), dst_ptrs.data_ptr(), caps.data_ptr()]
    R = res.data_ptr()
    for k in range(4):
        bad = list(a)
        bad[k] = None
        assert entry(*bad, n, R, None) == cudec_amd.CUDEC_ERR_INVALID_ARGUMENT
    assert entry(*a, n, None, None) == cudec_amd.CUDEC_ERR_INVALID_ARGUMENT
    assert entry(*a, n, R + 8, None) == cudec_amd.CUDEC_ERR_INVALID_ARGUMENT
    assert entry(*a, 0, R, None) == cudec_amd.CUDEC_ERR_INVALID_ARGUMENT
    assert entry(*a, (1 << 64) - 1, R, None) == cudec_amd.CUDEC_ERR_INVALID_ARGUMENT
    assert entry(*a, cudec_amd.max_batch_chunks() + 1, R, None) == \
        cudec_amd.CUDEC_ERR_INVALID_ARGUMENT
    res.fill_(-1)
    torch.cuda.synchronize()
    # A planted pending error survives a rejected call and is consumed by a
    # passing one (tests/smoke.cu:111-120).
    assert hip.hipSetDevice(12345) != 0
    assert hip.hipPeekAtLastError() != 0
    assert entry(*a, 0, R, None) == cudec_amd.CUDEC_ERR_INVALID_ARGUMENT
    assert hip.hipPeekAtLastError() != 0
    assert entry(*a, n, R, None) == cudec_amd.CUDEC_OK
    assert hip.hipPeekAtLastError() == 0
    torch.cuda.synchronize()
    r = res.cpu().numpy().view(np.dtype([("s", "<i4"), ("r", "<u4"), ("w", "<u8")]))
    assert np.all(r["s"] == 0) and np.all(r["r"] == 0) and np.all(r["w"] == 0)


@pytest.mark.parametrize("form", LAUNCHES)
def test_fixture_pairs_with_poison(form, lz4_fixtures):
    """tests/gpu_fixture.cu:172-186: capacity original + 8, poison beyond
    bytes_written must survive."""
    streams = [f.compressed for f in lz4_fixtures]
    caps = [len(f.original) + 8 for f in lz4_fixtures]
    res = check_against_oracle("lz4", form, streams, caps, [f.name for f in lz4_fixtures])
    assert np.all(res["status"] == 0)
    for i, f in enumerate(lz4_fixtures):
        assert int(res["bytes_written"][i]) == len(f.original)


@pytest.mark.parametrize("form", LAUNCHES)
def test_mutant_parity(form, lz4_fixtures):
    """tests/gpu_fixture.cu:188-230: the mutant corpus, held here to exact
    status and byte parity with the oracle, plus the reference's
    two-direction liblz4 contract with the stricter set pinned by identity."""
    streams, caps, names = [], [], []
    for f in lz4_fixtures:
        for m in corpus.mutate_stream(f.compressed, f.seed):
            streams.append(m.stream)
            caps.append(len(f.original))
            names.append(f"{f.name}/{m.description}")
    res = check_against_oracle("lz4", form, streams, caps, names)
    if libs.lz4lib() is not None:
        stricter, rejected = [], 0
        for i, s in enumerate(streams):
            ok, _ = libs.lz4_decompress(s, caps[i])
            if res["status"][i] == 0:
                assert ok, names[i]
            elif ok:
                stricter.append(names[i])
            else:
                rejected += 1
        assert rejected > 0
        assert stricter == ["text-256/flip-bit-at-88"]


@pytest.mark.parametrize("form", FORMS)
def test_grid_stride_wraparound(form):
    """tests/gpu_fixture.cu:231-260: 40000 chunks sharing one empty-block
    source and one destination all decode exactly once."""
    import torch

    import cudec_amd
    n = 40000
    src = torch.zeros(16, dtype=torch.uint8, device="cuda")
    dst = torch.zeros(16, dtype=torch.uint8, device="cuda")
    src_ptrs = torch.full((n,), src.data_ptr(), dtype=torch.int64, device="cuda")
    dst_ptrs = torch.full((n,), dst.data_ptr(), dtype=torch.int64, device="cuda")
    ones = torch.ones(n, dtype=torch.int64, device="cuda")
    res = torch.full((2 * n,), -1, dtype=torch.int64, device="cuda")
    st = cudec_amd.ENTRY[("lz4", form)](src_ptrs.data_ptr(), ones.data_ptr(), dst_ptrs.data_ptr(),
                                       ones.data_ptr(), n, res.data_ptr(), None)
    assert st == 0
    torch.cuda.synchronize()
    r = res.cpu().numpy().view(np.dtype([("s", "<i4"), ("r", "<u4"), ("w", "<u8")]))
    assert np.all(r["s"] == 0) and np.all(r["w"] == 0)


@pytest.mark.parametrize("form", LAUNCHES)
def test_adversarial_corpus_terminates(form):
    """tests/termination_gpu.cu: the hand-crafted hostile corpus at several
    capacities decodes to the oracle's verdict and terminates."""
    streams, caps, names = [], [], []
    for b in corpus.make_adversarial_blocks():
        for cap in (0, 1, 12, 4096, b.capacity):
            streams.append(b.stream)
            caps.append(cap)
            names.append(f"{b.name}@{cap}")
    check_against_oracle("lz4", form, streams, caps, names)


@pytest.mark.parametrize("form", LAUNCHES)
def test_short_streams_exhaustive(form):
    """Every 1-byte stream, and every 2-byte stream with a second byte on a
    stride, at capacities 0, 1, 16 and 64: exact oracle parity."""
    streams, caps = [], []
    for cap in (0, 1, 16, 64):
        for a in range(256):
            streams.append(bytes([a]))
            caps.append(cap)
            for b in range(0, 256, 5):
                streams.append(bytes([a, b]))
                caps.append(cap)
    check_against_oracle("lz4", form, streams, caps)


@pytest.mark.parametrize("form", LAUNCHES)
def test_last_literals_and_extension_sweeps(form):
    """The boundary families of tests/parser_twin.cpp:88-209 on the device."""
    streams, caps = [], []
    for L in (4, 8, 12):
        for ext in (0, 10, 100):
            for T in (4, 5, 6, 10):
                s = bytes([(L << 4) | 15]) + bytes(0x41 + i % 26 for i in range(L)) + \
                    b"\x01\x00" + bytes([ext, T << 4]) + bytes(0x61 + i % 26 for i in range(T))
                streams.append(s)
                caps.append(L + 19 + ext + T)
    for ff in (0, 1, 2):
        for last in (0, 30, 200):
            for T in (0, 1, 2, 3, 4, 6):
                s = b"\x1F\x41\x01\x00" + b"\xFF" * ff + bytes([last, T << 4]) + \
                    bytes(0x61 + i for i in range(T))
                streams.append(s)
                caps.append(1 + 4 + 15 + ff * 255 + last + T)
    for ff in (0, 1):
        for last in (0, 30, 200):
            lit = 15 + ff * 255 + last
            base = b"\xF0" + b"\xFF" * ff + bytes([last]) + bytes(0x41 + i % 26 for i in range(lit))
            for cap in range(lit, lit + 19):
                streams.append(base)
                caps.append(cap)
    check_against_oracle("lz4", form, streams, caps)


@pytest.mark.parametrize("form", LAUNCHES)
def test_generated_workloads_against_oracle(form):
    """The benchmark configurations' chunk kinds at reduced count, every
    chunk held to the oracle (status, bytes_written, bytes, untouched
    poison): token-process chunks of exactly 64 KiB (the shape of the CPU
    and plain-GPU configurations), and uneven chunks of all three kinds
    (token process, uniform random, repetitive) from 1 KiB to 4 MiB, with
    every seventh truncated by 1-9 bytes and every eleventh given a
    capacity one byte short."""
    from cudec_amd import workload
    b1 = workload.token64k(48, seed=0x77)
    b2 = workload.uneven(72, seed=0x78, min_size=1024, octaves=12)
    for b in (b1, b2):
        streams = [b.compressed(i) for i in range(b.n)]
        caps = [int(s) for s in b.sizes]
        for i in range(0, b.n, 7):
            streams[i] = streams[i][: max(len(streams[i]) - 1 - i % 9, 0)]
        for i in range(3, b.n, 11):
            caps[i] = max(caps[i] - 1, 0)
        res = check_against_oracle("lz4", form, streams, caps,
                                   [f"{b.meta['workload']}#{i}" for i in range(b.n)])
        assert np.sum(res["status"] == 0) > b.n // 2 and np.sum(res["status"] != 0) > 0


def test_forms_and_streams_deterministic(lz4_fixtures):
    """tests/determinism_gpu.cu:163-265 under docs/DETERMINISM.md:20-32,63-68:
    every result record, and the whole destination slot of every chunk that
    decoded OK (its bytes and the untouched poison beyond them), are
    identical across both launch forms, repeated runs, and a three-stream
    split of the batch. Rejected chunks' leftover bytes are outside the
    promise and are not compared."""
    import torch

    import cudec_amd
    from cudec_amd import workload
    from cudec_amd.device import DeviceBatch
    b = workload.uneven(90, seed=0x79, min_size=1024, octaves=7)
    streams = [b.compressed(i) for i in range(b.n)]
    streams += [m.stream for f in lz4_fixtures[:6] for m in corpus.mutate_stream(f.compressed, f.seed)]
    caps = [int(s) + 3 for s in b.sizes] + [70000] * (len(streams) - b.n)

    def ok_slots(db, res, arena):
        return [db.slot(arena, i).tobytes() if res["status"][i] == 0 else None
                for i in range(db.n)]

    ref_res = ref_slots = None
    for form in FORMS:
        for _ in range(3):
            db, res, arena = run_device("lz4", form, streams, caps, poison=0x3C)
            slots = ok_slots(db, res, arena)
            if ref_res is None:
                ref_res, ref_slots = res, slots
                assert np.sum(res["status"] == 0) > b.n // 2
                assert np.sum(res["status"] != 0) > 0
            assert np.array_equal(res, ref_res), form
            assert slots == ref_slots, form
    db = DeviceBatch.from_streams(streams, caps, poison=0x3C)
    db.reset_results()
    torch.cuda.synchronize()
    n = db.n
    cuts = [0, n // 3, 2 * n // 3, n]
    ss = [torch.cuda.Stream() for _ in range(3)]
    for k in range(3):
        lo, hi = cuts[k], cuts[k + 1]
        st = cudec_amd.cudec_lz4_decompress_batch(
            db.src_ptrs.data_ptr() + 8 * lo, db.src_sizes.data_ptr() + 8 * lo,
            db.dst_ptrs.data_ptr() + 8 * lo, db.dst_caps.data_ptr() + 8 * lo, hi - lo,
            db.d_results.data_ptr() + 16 * lo, ss[k].cuda_stream)
        assert st == 0
    torch.cuda.synchronize()
    res = db.results()
    assert np.array_equal(res, ref_res)
    assert ok_slots(db, res, db.output_arena()) == ref_slots


@pytest.mark.parametrize("form", LAUNCHES)
def test_long_periodic_matches(form):
    """Periodic data compresses to long matches (two or more extension
    bytes) at the period's offset, ending at the chunk's last byte: the
    wide copy paths for offsets below, at and above the staged-period limit
    (1024), with capacity exactly the output size so no staged read may pass
    it, plus one byte short (OUTPUT_TOO_SMALL)."""
    rng = np.random.default_rng(0x9E41)
    streams, caps, names = [], [], []
    for period in (1, 2, 3, 5, 7, 8, 15, 16, 17, 31, 64, 100, 1000, 1023, 1024, 1025, 2000):
        for n in (300, 1000, 5000, 70000):
            if period * 2 > n:
                continue
            base = rng.integers(0, 256, period, dtype=np.uint8).tobytes()
            data = (base * (n // period + 1))[:n]
            s = libs.lz4_compress(data)
            for cap in (n, n - 1):
                streams.append(s)
                caps.append(cap)
                names.append(f"period{period}/n{n}/cap{cap}")
    check_against_oracle("lz4", form, streams, caps, names)
