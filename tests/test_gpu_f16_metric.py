"""RgbaF16's metric level (fgMetric format.glsl:70-73) on every one of the 65536 half bit patterns, through the C ABI.

Each 2x2 cell of a 512x256 pair holds one pattern in R, G and B, so its luma16 is (L (4899 + 9617 + 1868) + 8192) >> 14
= L and the level-0 luma (flow scale 2) is the float32 L / 16, exact: frame a holds the patterns 0 .. 32767, frame c
32768 .. 65535 (NaNs, infinities, negatives, subnormals and both zeros among them). Bar: bit-exact against
oracle/formats.py's float64 level table, encoded and linear light at three white levels.
"""
import numpy as np
import pytest
import torch

import fgr
from gpu_util import Internals
from oracle import formats as F

pytestmark = pytest.mark.gpu

W, H = 512, 256


def frame_of_patterns(first):
    h = (first + np.arange(32768, dtype=np.uint32)).astype(np.uint16).reshape(128, 256)
    cells = np.repeat(np.repeat(h, 2, axis=0), 2, axis=1)
    f = np.empty((H, W, 4), np.uint16)
    f[..., 0] = f[..., 1] = f[..., 2] = cells
    f[..., 3] = 0x3C00
    return f


@pytest.mark.parametrize("linear,white", [(False, 80.0), (True, 80.0), (True, 203.0), (True, 10000.0)])
def test_every_half_pattern_level(linear, white):
    fi = fgr.Interpolator(W, H, fmt=F.F16, transfer=fgr.TRANSFER_LINEAR if linear else fgr.TRANSFER_ENCODED,
                          linear_white_nits=white)
    ta = torch.from_numpy(frame_of_patterns(0).view(np.int16)).cuda()
    tc = torch.from_numpy(frame_of_patterns(32768).view(np.int16)).cuda()
    assert fi.record_estimate(ta, tc)
    I = Internals(fi)
    fi.close()
    want = F.f16_level_table(linear, white).astype(np.float32) / np.float32(16.0)
    got = np.concatenate([I.luma_level(0, "a").ravel(), I.luma_level(0, "c").ravel()])
    assert got.shape == (65536,)
    bad = np.flatnonzero(got.view(np.uint32) != want.view(np.uint32))
    assert bad.size == 0, f"{bad.size} patterns differ, first 0x{bad[0]:04x}: {got[bad[0]]} vs {want[bad[0]]}"
