"""The benchmark's own workloads at full size on the GPU, through the C ABI,
with every chunk held to the CPU oracle.

* Uneven LZ4: 20000 chunks, sizes log-uniform over [1 KiB, 4 MiB), 80%
  token process / 10% uniform random / 10% repetitive, seed 0x5EED (the
  benchmark's default batch, 9.35 GiB out).
* Token-process LZ4: 65536 chunks of exactly 64 KiB (4 GiB out).

Each runs the tuned form with the default plan: every chunk must be OK with
bytes_written equal to its size, the output arena must equal the
originals byte for byte, and every status and bytes_written must equal
what the oracle decodes on the host (oracle_decode_batch_timed, on up to 16
threads)."""
import ctypes
import os

import numpy as np
import pytest
import torch

from cudec_amd import workload
from cudec_amd.device import DeviceBatch
from tests import libs
from tests.gpu_util import gpu

pytestmark = gpu


def oracle_results(b):
    """Every chunk decoded by the oracle on the host: (status, written)."""
    n = b.n
    out = np.empty(max(b.total_uncompressed, 1), dtype=np.uint8)
    srcs = (b.comp_off + np.uint64(b.comp.ctypes.data)).astype(np.uint64)
    dsts = (b.orig_off + np.uint64(out.ctypes.data)).astype(np.uint64)
    sizes = b.comp_size.astype(np.uint64)
    caps = b.sizes.astype(np.uint64)
    status = np.zeros(n, dtype=np.int32)
    written = np.zeros(n, dtype=np.uint64)
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    libs.oracle().oracle_decode_batch_timed(0, srcs.ctypes.data, sizes.ctypes.data,
                                            dsts.ctypes.data, caps.ctypes.data, n, threads,
                                            status.ctypes.data, written.ctypes.data)
    assert np.array_equal(out[: b.total_uncompressed], b.orig), "the oracle disagrees with the originals"
    return status, written


def check_full(b, name):
    print(f"[configs] {name}: {b.n} chunks, {b.total_uncompressed / 2**30:.2f} GiB out", flush=True)
    db = DeviceBatch(b.comp, b.comp_off, b.comp_size, b.sizes, out_off=b.orig_off, poison=0xA5)
    db.reset_results()
    assert db.run("lz4", "tuned") == 0
    torch.cuda.synchronize()
    res = db.results()
    assert np.all(res["status"] == 0) and np.all(res["reserved"] == 0)
    assert np.array_equal(res["bytes_written"], b.sizes)
    orig = torch.from_numpy(b.orig).to("cuda")
    assert torch.equal(db.d_out[: b.total_uncompressed], orig), f"{name}: output differs"
    del orig
    status, written = oracle_results(b)
    assert np.array_equal(res["status"], status)
    assert np.array_equal(res["bytes_written"], written)


def test_uneven_batch_full_size():
    b = workload.uneven(20000, seed=0x5EED)
    assert b.total_uncompressed == 10038222939 and b.total_compressed == 5773967479
    check_full(b, "uneven LZ4")


def test_token64k_batch_full_size():
    b = workload.token64k(65536, seed=0x5EED)
    check_full(b, "token64k LZ4")
