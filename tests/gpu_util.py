"""Helpers of the GPU tests: device copies of frames, and reads of the library's internal buffers."""
import ctypes

import numpy as np
import torch

import fgr

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipMemcpy.restype = ctypes.c_int
    return _hip


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).to("cuda:0")


def read_device(ptr, count, dtype):
    """Copy `count` elements of `dtype` from a raw device pointer (after a device synchronise)."""
    torch.cuda.synchronize()
    out = np.empty(count, dtype)
    rc = hip().hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2)
    assert rc == 0, f"hipMemcpy failed: {rc}"
    return out


class Internals:
    """The latest estimate's intermediates, read through fgr_debug_internal."""

    def __init__(self, fi):
        d = fi if isinstance(fi, fgr.DebugInternal) else fi.debug_internal()
        assert d is not None
        self.d = d
        self.luma = read_device(d.luma, d.lumaWords, np.float32)
        self.field = read_device(d.field, 2 * d.fieldEntries, np.float32).reshape(-1, 2)
        self.u = read_device(d.counters, 1152, np.uint32)

    def luma_level(self, i, which):
        d = self.d
        off = d.lumaOffA[i] if which == "a" else d.lumaOffC[i]
        return self.luma[off:off + d.w[i] * d.h[i]].reshape(d.h[i], d.w[i])

    def block_field(self, off, i=0):
        d = self.d
        return self.field[off:off + d.bw[i] * d.bh[i]].reshape(d.bh[i], d.bw[i], 2)

    def fs(self, i):
        return self.block_field(self.d.fS[i], i)

    def fm(self, i):
        return self.block_field(self.d.fM[i], i)

    def sub(self):
        return self.block_field(self.d.fieldSub, 0)

    def pf(self):
        d = self.d
        return self.field[d.fieldPF:d.fieldPF + d.fw * d.fh].reshape(d.fh, d.fw, 2)

    def sadsum(self):
        return int(self.u[64:128].astype(np.uint64).sum()) & 0xFFFFFFFF

    def median(self, i):
        return self.u[4 + 2 * i:6 + 2 * i].view(np.float32)
