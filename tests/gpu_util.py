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
    if fmt == "zstd":
        return libs.ref_zstd_decode(stream, cap)
    st, out, _ = libs.oracle_snappy(stream, cap)
    return st, out


# Launch shapes of the parity tests: the plain form, and the tuned form with
# every chunk on two pipelined waves (threshold 0) or every chunk on one wave
# (threshold SIZE_MAX), so both of the tuned kernel's parsers and copiers
# see every case. "tuned" alone runs the tuned form with the current plan.
LAUNCHES = ("plain", "tuned-pipe", "tuned-split")
_SHAPE_THRESHOLD = {"tuned-pipe": 0, "tuned-split": (1 << 64) - 1}


class launch_shape:
    """Sets the format's pipelining threshold for a launch shape and puts the
    previous one back afterwards."""

    def __init__(self, fmt: str, launch: str):
        self.fmt, self.launch = fmt, launch
        self.old = None

    def __enter__(self):
        import cudec_amd
        if self.launch in _SHAPE_THRESHOLD:
            st, self.old = cudec_amd.pipeline_threshold(self.fmt)
            assert st == 0
            assert cudec_amd.set_pipeline_threshold(self.fmt, _SHAPE_THRESHOLD[self.launch]) == 0
        return "plain" if self.launch == "plain" else "tuned"

    def __exit__(self, *exc):
        import cudec_amd
        if self.old is not None:
            assert cudec_amd.set_pipeline_threshold(self.fmt, self.old) == 0
        return False


def run_device(fmt: str, form: str, streams, caps, poison=POISON, stream=None):
    import torch

    from cudec_amd.device import DeviceBatch
    db = DeviceBatch.from_streams(list(streams), caps, poison=poison)
    db.reset_results()
    with launch_shape(fmt, form) as f:
        st = db.run(fmt, f, stream=stream)
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
