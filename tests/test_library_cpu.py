"""The C ABI library without a GPU: it loads, exports every entry point include/fgr.h declares, and its host-only
logic (settings validation, size layout, refusal reasons) matches the reference's rules."""
import ctypes
import math
import os
import re

import pytest

import fgr
from oracle import v1

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if h.endswith(".h"):
            text = open(os.path.join(ROOT, "include", h)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(fgr_\w+)\s*\(", text))
    return sorted(names)


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) >= 27
    lib = ctypes.CDLL(os.path.join(ROOT, "frame-gen-ritz_amd", "fgr", "libfgr.so"))
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_defaults_are_interp_constants():
    s = fgr.default_settings()
    d = v1.DEFAULTS
    for k in d:
        assert getattr(s, k) == pytest.approx(float(d[k])), k
    assert fgr.validate(s)
    assert fgr.settings_error_for_format(s) is None


@pytest.mark.parametrize("field,value", [
    ("trustLow", -1.0), ("trustHigh", 256.0), ("trustHigh", 12.5), ("globalFallback", 0.0), ("globalFallback", 1.5),
    ("sceneCutSad", 0.0), ("sceneCutSad", 300.0), ("hudBonus", -0.1), ("smoothBonus", 256.0),
    ("searchPenalty", math.nan), ("trustLow", math.inf), ("searchRadius", 0), ("searchRadius", 13),
    ("searchRadiusTop", 0), ("searchRadiusTop", 13), ("flowScale", 3), ("flowScale", 8), ("pyramidMinSide", 15),
    ("pyramidMinSide", 257), ("blurSamples", 17), ("blurWeights", 2), ("subpixel", 2), ("motionVersion", 4),
])
def test_validate_rejects_out_of_range(field, value):
    # framegen.cpp:1679-1700
    s = fgr.default_settings(**{field: value})
    assert not fgr.validate(s)
    assert fgr.settings_error_for_format(s).startswith("invalid settings")


def test_validate_accepts_range_ends():
    for kw in (dict(trustLow=0.0, trustHigh=1.0), dict(trustLow=254.0, trustHigh=255.0), dict(globalFallback=1.0),
               dict(sceneCutSad=255.0), dict(searchRadius=12, searchRadiusTop=1), dict(flowScale=4),
               dict(pyramidMinSide=16), dict(pyramidMinSide=256), dict(blurSamples=16)):
        assert fgr.validate(fgr.default_settings(**kw)), kw


def test_other_motion_versions_are_refused_with_a_reason():
    s = fgr.default_settings(motionVersion=1)
    assert fgr.validate(s)                    # a valid value of the reference's enum
    assert "V1" in fgr.settings_error_for_format(s)


@pytest.mark.parametrize("W,H,S,m", [(1920, 1080, 2, 48), (3840, 2160, 2, 48), (1280, 720, 2, 48), (640, 360, 2, 48),
                                     (2560, 1440, 2, 48), (1921, 1081, 2, 48), (203, 117, 2, 16), (1920, 1080, 1, 48),
                                     (1919, 1077, 4, 16), (16, 16, 2, 16), (16384, 16384, 2, 48)])
def test_size_layout_matches_oracle(W, H, S, m):
    fw, fh, lv = fgr.size_layout(W, H, S, m)
    ofw, ofh, olv = v1.layout(W, H, S, m)
    assert (fw, fh) == (ofw, ofh)
    assert lv == olv


def test_create_refusals_name_the_reason():
    # refused before any device call: unsupported format, invalid settings, sizes
    with pytest.raises(RuntimeError, match="format"):
        fgr.Interpolator(640, 360, fmt=7)
    with pytest.raises(RuntimeError, match="even width"):
        fgr.Interpolator(641, 360, fmt=fgr.FORMAT_NV12)
    with pytest.raises(RuntimeError, match="invalid settings"):
        fgr.Interpolator(640, 360, settings=fgr.default_settings(trustHigh=5.0))
    with pytest.raises(RuntimeError, match="kMinSize|FGR_MIN_SIZE"):
        fgr.Interpolator(15, 360)
    with pytest.raises(RuntimeError, match="2\\^30"):
        fgr.Interpolator(32768, 32769)
    with pytest.raises(RuntimeError, match="linearWhiteNits"):
        fgr.Interpolator(64, 64, fmt=fgr.FORMAT_RGBAF16, transfer=fgr.TRANSFER_LINEAR, linear_white_nits=0.0)


def _batch_pair(fmt, W, H, pitch, chroma=0x2000000, cpitch=None, n=2):
    bp = fgr.BatchPair()
    bp.width, bp.height, bp.format, bp.multiplier = W, H, fmt, n
    for f in [bp.prev, bp.curr] + [bp.out[k] for k in range(n - 1)]:
        f.ptr, f.pitch = 0x1000000, pitch     # never dereferenced: the checks run before any device call
        f.chroma, f.chromaPitch = chroma, cpitch if cpitch is not None else pitch
    return bp


@pytest.mark.parametrize("fmt,W,H,pitch,chroma,cpitch", [
    (fgr.FORMAT_P010, 64, 32, 128, 0, 128),               # P010 without its CbCr plane
    (fgr.FORMAT_NV12, 64, 32, 64, 0, 64),                 # Nv12 without its CbCr plane
    (fgr.FORMAT_RGBA8, 64, 32, 255, 0, None),             # pitch below a row (64 x 4 bytes)
    (fgr.FORMAT_RGBAF16, 64, 32, 511, 0, None),           # RgbaF16: 8 bytes a pixel
    (fgr.FORMAT_RGBA8, 64, 32, 1 << 24, 0, None),         # pitch of 2^24: rows are 24-bit products
    (fgr.FORMAT_RGBA8, 4096, 32768, 131076 + 4, 0, None), # last row ends past 2^32
    (fgr.FORMAT_P010, 64, 32, 128, 0x2000000, 127),       # CbCr pitch below 32 x 4 bytes
])
def test_batch_refuses_frames_it_cannot_address(fmt, W, H, pitch, chroma, cpitch):
    bp = _batch_pair(fmt, W, H, pitch, chroma, cpitch)
    info = fgr.BatchInfo()
    info.settings = fgr.default_settings()
    h = fgr.lib.fgr_batch_create(ctypes.byref(info), ctypes.byref(bp), 1)
    assert not h
    assert "pitch" in fgr.lib.fgr_batch_last_error().decode()
