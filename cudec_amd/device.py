"""Device-resident batches over torch memory (torch is only the allocator and
stream provider here; the decode is the native library).

A DeviceBatch holds one compressed arena, one output arena, the four
per-chunk tables the C ABI takes (source pointers, source sizes,
destination pointers, capacities; uint64 each) and a 16-byte aligned result
array, all on the device.
"""
from __future__ import annotations

import numpy as np
import torch

from . import ENTRY, CUDEC_OK
from .workload import offsets_of

RESULT_DTYPE = np.dtype([("status", "<i4"), ("reserved", "<u4"), ("bytes_written", "<u8")])


def _u64_tensor(a: np.ndarray, device) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.uint64).view(np.int64)).to(device)


class DeviceBatch:
    def __init__(self, comp: np.ndarray, comp_off: np.ndarray, comp_size: np.ndarray,
                 caps, device: str = "cuda", poison: int | None = None,
                 out_off: np.ndarray | None = None):
        self.device = torch.device(device)
        self.n = int(len(comp_size))
        comp = np.ascontiguousarray(comp, dtype=np.uint8)
        if not comp.flags.writeable:
            comp = comp.copy()
        self.d_comp = torch.from_numpy(comp).to(self.device) if comp.size else \
            torch.zeros(16, dtype=torch.uint8, device=self.device)
        self.caps = np.ascontiguousarray(caps, dtype=np.uint64)
        self.out_off = offsets_of(self.caps) if out_off is None else np.asarray(out_off, np.uint64)
        total = int(self.caps.sum())
        self.d_out = torch.empty(max(total, 16), dtype=torch.uint8, device=self.device)
        if poison is not None:
            self.d_out.fill_(poison)
        cbase = self.d_comp.data_ptr()
        obase = self.d_out.data_ptr()
        comp_off = np.asarray(comp_off, dtype=np.uint64)
        self.src_ptrs = _u64_tensor(comp_off + np.uint64(cbase), self.device)
        self.src_sizes = _u64_tensor(comp_size, self.device)
        self.dst_ptrs = _u64_tensor(self.out_off + np.uint64(obase), self.device)
        self.dst_caps = _u64_tensor(self.caps, self.device)
        self.d_results = torch.empty(2 * max(self.n, 1), dtype=torch.int64, device=self.device)
        self.poison = poison

    @classmethod
    def from_streams(cls, streams: list[bytes], caps, **kw) -> "DeviceBatch":
        sizes = np.array([len(s) for s in streams], dtype=np.uint64)
        off = offsets_of(sizes)
        arena = np.frombuffer(b"".join(streams), dtype=np.uint8) if streams else \
            np.zeros(0, np.uint8)
        return cls(arena, off, sizes, caps, **kw)

    def reset_results(self, fill: int = -1) -> None:
        self.d_results.fill_(fill)

    def run(self, fmt: str = "lz4", form: str = "tuned", stream=None) -> int:
        handle = None
        if stream is not None:
            handle = stream.cuda_stream if hasattr(stream, "cuda_stream") else stream
        return ENTRY[(fmt, form)](self.src_ptrs.data_ptr(), self.src_sizes.data_ptr(),
                                  self.dst_ptrs.data_ptr(), self.dst_caps.data_ptr(), self.n,
                                  self.d_results.data_ptr(), handle)

    def results(self) -> np.ndarray:
        raw = self.d_results[: 2 * self.n].cpu().numpy()
        return raw.view(RESULT_DTYPE)

    def output_arena(self) -> np.ndarray:
        return self.d_out.cpu().numpy()

    def outputs(self, results: np.ndarray | None = None) -> list[bytes]:
        arena = self.output_arena()
        res = self.results() if results is None else results
        out = []
        for i in range(self.n):
            o = int(self.out_off[i])
            w = int(res["bytes_written"][i]) if res["status"][i] == CUDEC_OK else 0
            out.append(arena[o:o + w].tobytes())
        return out

    def slot(self, arena: np.ndarray, i: int) -> np.ndarray:
        o = int(self.out_off[i])
        return arena[o:o + int(self.caps[i])]
