"""Helpers for the GPU parity tests: run streams through the device decoders
via the C ABI and compare each chunk with the CPU oracle."""
from __future__ import annotations

import numpy as np
import pytest

from tests import libs

POISON = 0xA5


def device_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def oracle_decode(fmt: str, stream: bytes, cap: int):
    if fmt == "lz4":
        return libs.oracle_lz4(stream, cap)
    st, out, _ = libs.oracle_snappy(stream, cap)
    return st, out


def run_device(fmt: str, form: str, streams, caps, poison=POISON, stream=None):
    import torch

    from cudec_amd.device import DeviceBatch
    db = DeviceBatch.from_streams(list(streams), caps, poison=poison)
    db.reset_results()
    st = db.run(fmt, form, stream=stream)
    assert st == 0, f"submission failed: {st}"
    torch.cuda.synchronize()
    res = db.results()
    arena = db.output_arena()
    return db, res, arena


def check_against_oracle(fmt: str, form: str, streams, caps, names=None, poison=POISON):
    """Every chunk: identical status to the oracle, reserved == 0,
    bytes_written == oracle's, output bytes identical, and the poison beyond
    bytes_written untouched."""
    db, res, arena = run_device(fmt, form, streams, caps, poison)
    for i, (s, cap) in enumerate(zip(streams, caps)):
        name = names[i] if names else i
        want_st, want_out = oracle_decode(fmt, s, cap)
        got_st = int(res["status"][i])
        assert got_st == want_st, f"{fmt}/{form} chunk {name}: status {got_st} != oracle {want_st}"
        assert int(res["reserved"][i]) == 0, name
        slot = db.slot(arena, i)
        if want_st == libs.OK:
            w = len(want_out)
            assert int(res["bytes_written"][i]) == w, name
            assert slot[:w].tobytes() == want_out, f"{fmt}/{form} chunk {name}: bytes differ"
            if poison is not None:
                assert np.all(slot[w:] == poison), f"{fmt}/{form} chunk {name}: wrote past bytes_written"
        else:
            assert int(res["bytes_written"][i]) == 0, name
    return res


gpu = pytest.mark.gpu
FORMS = ("tuned", "plain")
