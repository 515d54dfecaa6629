"""Python binding of libfgr.so, the MI355X (gfx950) V1 frame interpolator (C ABI: include/fgr.h).

Mirrors framegen::Interpolator (gpu/framegen.h:443-822): ``Interpolator(width, height, ...)`` = init(),
``record_estimate`` / ``record_synth`` / ``record`` enqueue HIP kernels on the current torch stream (the reference
records into a VkCommandBuffer), returning True, or False when the call is refused (nothing enqueued). Frames are
torch uint8 / uint16 / float16 CUDA tensors (device memory); torch supplies memory and streams only.

The library must be built (``python -c "import __graft_entry__ as g; g.build()"``); importing this package without
it raises ImportError: there is no CPU fallback.
"""
import ctypes
import os

__all__ = ["Settings", "Interpolator", "Batch", "DebugInternal", "lib", "default_settings", "validate", "settings_error_for_format",
           "size_layout", "FORMAT_RGBA8", "FORMAT_RGB10", "FORMAT_RGBAF16", "FORMAT_NV12", "FORMAT_P010",
           "MIN_SIZE", "SET_RING", "frame_of"]

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libfgr.so")

FORMAT_RGBA8, FORMAT_RGB10, FORMAT_RGBAF16, FORMAT_NV12, FORMAT_P010 = 0, 1, 2, 3, 4
TRANSFER_ENCODED, TRANSFER_LINEAR = 0, 1
YUV_BT601, YUV_BT709, YUV_BT2020 = 0, 1, 2
MIN_SIZE, SET_RING, BLUR_MAX_SAMPLES = 16, 32, 16
UI_OFF, UI_CROSSHAIR, UI_WHOLESCREEN, UI_CROSSHAIR_V2 = 0, 1, 2, 3
UI_FRAMES, UI_FRAMES_MAX, UI_SETS = 3, 65, 64


class Settings(ctypes.Structure):
    """framegen.h:167-409 Settings (the V1 fields) == FgrSettings."""
    _fields_ = [("trustLow", ctypes.c_float), ("trustHigh", ctypes.c_float), ("globalFallback", ctypes.c_float),
                ("sceneCutSad", ctypes.c_float), ("hudBonus", ctypes.c_float), ("smoothBonus", ctypes.c_float),
                ("searchPenalty", ctypes.c_float), ("searchRadius", ctypes.c_uint32),
                ("searchRadiusTop", ctypes.c_uint32), ("subpixel", ctypes.c_uint32),
                ("motionVersion", ctypes.c_uint32), ("blurSamples", ctypes.c_uint32),
                ("blurWeights", ctypes.c_uint32), ("flowScale", ctypes.c_uint32),
                ("pyramidMinSide", ctypes.c_uint32), ("seedNet", ctypes.c_uint32),
                # static-UI protection (framegen.h:225-305)
                ("uiProtection", ctypes.c_uint32), ("uiHistory", ctypes.c_uint32),
                ("uiBoxHeightFrac", ctypes.c_float), ("uiBoxWidth", ctypes.c_uint32),
                ("uiBoxHeight", ctypes.c_uint32), ("uiCrosshairArea", ctypes.c_float),
                ("uiBoxOffsetX", ctypes.c_int32), ("uiBoxOffsetY", ctypes.c_int32),
                ("uiStillFrames", ctypes.c_uint32), ("uiStillThreshold", ctypes.c_float),
                ("uiInpaintReach", ctypes.c_uint32), ("uiMaxStillFraction", ctypes.c_float),
                ("uiContrastThreshold", ctypes.c_float), ("uiLeaveFrames", ctypes.c_uint32),
                ("uiReentryFrames", ctypes.c_uint32), ("uiEstimateFill", ctypes.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class InitInfo(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("format", ctypes.c_uint32),
                ("bgra", ctypes.c_uint32), ("transfer", ctypes.c_uint32), ("linearWhiteNits", ctypes.c_float),
                ("yuvMatrix", ctypes.c_uint32), ("yuvFullRange", ctypes.c_uint32), ("device", ctypes.c_int32),
                ("settings", Settings)]


class Frame(ctypes.Structure):
    _fields_ = [("ptr", ctypes.c_void_p), ("pitch", ctypes.c_uint64), ("chroma", ctypes.c_void_p),
                ("chromaPitch", ctypes.c_uint64)]


class DebugMapLayout(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in ("W", "H", "fw", "fh", "bw", "bh", "fscale", "eStride")] + \
               [(n, ctypes.c_uint64) for n in ("blockOff", "pixOff", "eOff", "size")]


class SizeLayout(ctypes.Structure):
    _fields_ = [("fw", ctypes.c_int32), ("fh", ctypes.c_int32), ("levels", ctypes.c_int32),
                ("w", ctypes.c_int32 * 16), ("h", ctypes.c_int32 * 16), ("bw", ctypes.c_int32 * 16),
                ("bh", ctypes.c_int32 * 16)]


class DebugInternal(ctypes.Structure):
    _fields_ = [("luma", ctypes.c_void_p), ("field", ctypes.c_void_p), ("counters", ctypes.c_void_p),
                ("levels", ctypes.c_int32), ("fw", ctypes.c_int32), ("fh", ctypes.c_int32)] + \
               [(n, ctypes.c_int32 * 16) for n in ("w", "h", "bw", "bh", "lumaOffA", "lumaOffC", "fS", "fM")] + \
               [(n, ctypes.c_int32) for n in ("fieldSub", "fieldPF", "fieldT", "lumaWords", "fieldEntries",
                                              "lastSynthSlot", "fieldSeed")]


class BatchPair(ctypes.Structure):
    _fields_ = [("prev", Frame), ("curr", Frame), ("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
                ("format", ctypes.c_uint32), ("multiplier", ctypes.c_uint32), ("out", Frame * 15)]


class BatchInfo(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int32), ("bgra", ctypes.c_uint32), ("transfer", ctypes.c_uint32),
                ("linearWhiteNits", ctypes.c_float), ("yuvMatrix", ctypes.c_uint32),
                ("yuvFullRange", ctypes.c_uint32), ("useGraph", ctypes.c_uint32), ("settings", Settings),
                ("seedNet", ctypes.c_void_p)]


if not os.path.exists(_LIB_PATH):
    raise ImportError(f"libfgr.so is not built ({_LIB_PATH}); run __graft_entry__.build()")
# One HIP runtime per process: PyTorch-ROCm ships its own libamdhip64 / libhsa-runtime64. Loaded after libfgr, torch
# would bring a second HSA runtime into the process, and the one that opened the device second cannot use it. With
# torch loaded first, libfgr's libamdhip64.so.7 resolves to torch's copy. So torch, when installed, is imported here.
try:
    import torch  # noqa: F401
except ImportError:
    pass
lib = ctypes.CDLL(_LIB_PATH)

_H = ctypes.c_void_p
_sig = {
    "fgr_default_settings": (None, [ctypes.POINTER(Settings)]),
    "fgr_validate": (ctypes.c_int, [ctypes.POINTER(Settings)]),
    "fgr_settings_error_for_format": (ctypes.c_char_p, [ctypes.POINTER(Settings), ctypes.c_uint32]),
    "fgr_size_layout": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                       ctypes.POINTER(SizeLayout)]),
    "fgr_version": (ctypes.c_uint32, []),
    "fgr_create": (_H, [ctypes.POINTER(InitInfo)]),
    "fgr_last_init_error": (ctypes.c_char_p, []),
    "fgr_destroy": (None, [_H]),
    "fgr_format": (ctypes.c_uint32, [_H]),
    "fgr_resize": (ctypes.c_int, [_H, ctypes.c_uint32, ctypes.c_uint32]),
    "fgr_settings_error": (ctypes.c_char_p, [_H, ctypes.POINTER(Settings)]),
    "fgr_size_error": (ctypes.c_char_p, [_H, ctypes.c_uint32, ctypes.c_uint32]),
    "fgr_set_settings": (ctypes.c_int, [_H, ctypes.POINTER(Settings)]),
    "fgr_reconfigure": (ctypes.c_int, [_H, ctypes.POINTER(Settings)]),
    "fgr_get_settings": (None, [_H, ctypes.POINTER(Settings)]),
    "fgr_begin_frame": (None, [_H]),
    "fgr_record_estimate": (ctypes.c_int, [_H, ctypes.c_void_p, Frame, Frame]),
    "fgr_record_synth": (ctypes.c_int, [_H, ctypes.c_void_p, Frame, Frame, Frame, ctypes.c_float]),
    "fgr_record": (ctypes.c_int, [_H, ctypes.c_void_p, Frame, Frame, Frame, ctypes.c_float]),
    "fgr_record_synth_blur": (ctypes.c_int, [_H, ctypes.c_void_p, Frame, Frame, Frame, ctypes.c_float,
                                             ctypes.c_float]),
    "fgr_record_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(InitInfo), Frame, Frame, Frame,
                                          ctypes.c_float, ctypes.POINTER(ctypes.c_char_p)]),
    "fgr_set_profiling": (ctypes.c_int, [_H, ctypes.c_int]),
    "fgr_read_timings": (ctypes.c_int, [_H, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_uint64),
                                        ctypes.c_int]),
    "fgr_record_debug_counters": (ctypes.c_int, [_H, ctypes.c_void_p, ctypes.c_void_p]),
    "fgr_debug_sad_blocks": (ctypes.c_uint32, [_H]),
    "fgr_debug_map_layout": (ctypes.c_int, [_H, ctypes.POINTER(DebugMapLayout)]),
    "fgr_record_debug_maps": (ctypes.c_int, [_H, ctypes.c_void_p, ctypes.c_void_p]),
    "fgr_debug_internal": (ctypes.c_int, [_H, ctypes.POINTER(DebugInternal)]),
    "fgr_batch_create": (_H, [ctypes.POINTER(BatchInfo), ctypes.POINTER(BatchPair), ctypes.c_uint32]),
    "fgr_batch_last_error": (ctypes.c_char_p, []),
    "fgr_batch_destroy": (None, [_H]),
    "fgr_batch_set_frames": (ctypes.c_int, [_H, ctypes.c_void_p, ctypes.POINTER(BatchPair), ctypes.c_uint32]),
    "fgr_batch_record": (ctypes.c_int, [_H, ctypes.c_void_p]),
    "fgr_batch_launches": (ctypes.c_uint32, [_H]),
    "fgr_batch_set_profiling": (ctypes.c_int, [_H, ctypes.c_int]),
    "fgr_batch_read_timings": (ctypes.c_int, [_H, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                              ctypes.POINTER(ctypes.c_uint64), ctypes.c_int]),
    "fgr_seednet_create": (ctypes.c_void_p, [ctypes.c_int32, ctypes.c_uint64]),
    "fgr_seednet_destroy": (None, [_H]),
    "fgr_seednet_param_count": (ctypes.c_uint32, []),
    "fgr_seednet_get_weights": (ctypes.c_int, [_H, ctypes.c_void_p, ctypes.c_uint32]),
    "fgr_seednet_set_weights": (ctypes.c_int, [_H, ctypes.c_void_p, ctypes.c_uint32]),
    "fgr_seednet_forward": (ctypes.c_int, [_H, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                           ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p]),
    "fgr_seednet_gradients": (ctypes.c_int, [_H, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                             ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p,
                                             ctypes.c_void_p]),
    "fgr_seednet_train_step": (ctypes.c_int, [_H, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                              ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p, ctypes.c_float,
                                              ctypes.c_void_p]),
    "fgr_seednet_search_field": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                                ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_float,
                                                ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "fgr_set_seed_net": (ctypes.c_int, [_H, _H]),
    "fgr_fgmaps_header": (ctypes.c_int, [_H, ctypes.c_void_p]),
    "fgr_batch_autotune": (ctypes.c_int, [_H, ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_uint64)]),
    "fgr_batch_set_tuning": (ctypes.c_int, [_H, ctypes.c_int32, ctypes.c_int32]),
    "fgr_batch_get_tuning": (ctypes.c_int, [_H, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]),
    "fgr_batch_set_lanes": (ctypes.c_int, [_H, ctypes.c_int32]),
    "fgr_batch_set_form": (ctypes.c_int, [_H, ctypes.c_int32]),
    "fgr_batch_get_form": (ctypes.c_int32, [_H]),
    "fgr_batch_get_lanes": (ctypes.c_int32, [_H]),
    "fgr_tuning_save": (ctypes.c_int, [ctypes.c_char_p]),
    "fgr_tuning_load": (ctypes.c_int, [ctypes.c_char_p]),
    "fgr_seednet_calibrate_fp8": (ctypes.c_int, [_H, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                                 ctypes.c_int32, ctypes.c_int32]),
    "fgr_seednet_set_fp8": (ctypes.c_int, [_H, ctypes.c_int]),
    "fgr_seednet_fp8_scales": (ctypes.c_int, [_H, ctypes.c_void_p, ctypes.c_void_p]),
    "fgr_debug_fp8_e4m3": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32]),
    "fgr_batch_debug_counters": (ctypes.c_int, [_H, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32,
                                                ctypes.c_void_p]),
    "fgr_batch_debug_internal": (ctypes.c_int, [_H, ctypes.c_uint32, ctypes.POINTER(DebugInternal)]),
    "fgr_record_observe": (ctypes.c_int, [_H, ctypes.c_void_p, Frame]),
    "fgr_record_track": (ctypes.c_int, [_H, ctypes.c_void_p, Frame]),
    "fgr_release_unused_ui": (None, [_H]),
    "fgr_box_size": (None, [ctypes.POINTER(Settings), ctypes.c_uint32, ctypes.c_uint32,
                            ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32)]),
    "fgr_box_rect": (None, [ctypes.POINTER(Settings), ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(ctypes.c_int32)]),
    "fgr_box_margin": (ctypes.c_uint32, [ctypes.POINTER(Settings)]),
    "fgr_work_rect": (None, [ctypes.POINTER(Settings), ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(ctypes.c_int32)]),
    "fgr_debug_ui_mask": (ctypes.c_int, [_H, ctypes.c_void_p, Frame, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32)]),
}
for _name, (_res, _args) in _sig.items():
    _f = getattr(lib, _name)
    _f.restype = _res
    _f.argtypes = _args


def _s(b):
    return None if b is None else b.decode()


def default_settings(**kw):
    s = Settings()
    lib.fgr_default_settings(ctypes.byref(s))
    for k, v in kw.items():
        if not hasattr(s, k):
            raise KeyError(k)
        setattr(s, k, v)
    return s


def box_size(s, w, h):
    """framegen.h:484 boxSize(): (boxW, boxH)"""
    bw, bh = ctypes.c_uint32(), ctypes.c_uint32()
    lib.fgr_box_size(ctypes.byref(s), w, h, ctypes.byref(bw), ctypes.byref(bh))
    return bw.value, bh.value


def box_rect(s, w, h):
    """framegen.h:487 boxRect(): (x, y, w, h)"""
    r = (ctypes.c_int32 * 4)()
    lib.fgr_box_rect(ctypes.byref(s), w, h, r)
    return tuple(r)


def box_margin(s):
    return lib.fgr_box_margin(ctypes.byref(s))


def work_rect(s, w, h):
    """framegen.h:494 workRect(): (x, y, w, h)"""
    r = (ctypes.c_int32 * 4)()
    lib.fgr_work_rect(ctypes.byref(s), w, h, r)
    return tuple(r)


def validate(s):
    """framegen.h:464 Interpolator::validate()."""
    return bool(lib.fgr_validate(ctypes.byref(s)))


def settings_error_for_format(s, fmt=FORMAT_RGBA8):
    return _s(lib.fgr_settings_error_for_format(ctypes.byref(s), fmt))


def size_layout(W, H, flow_scale=2, min_side=48):
    """framegen.cpp:658-685 sizeLayout: (fw, fh, [(w, h, bw, bh), ...] finest first)."""
    L = SizeLayout()
    if not lib.fgr_size_layout(W, H, flow_scale, min_side, ctypes.byref(L)):
        raise ValueError("invalid size / flow scale")
    return L.fw, L.fh, [(L.w[i], L.h[i], L.bw[i], L.bh[i]) for i in range(L.levels)]


def frame_of(t, chroma=None):
    """A torch CUDA tensor (H, W, C) or (H, W) -> FgrFrame (pitch = bytes per row)."""
    f = Frame()
    f.ptr = t.data_ptr()
    f.pitch = t.stride(0) * t.element_size()
    if chroma is not None:
        f.chroma = chroma.data_ptr()
        f.chromaPitch = chroma.stride(0) * chroma.element_size()
    return f


def _stream(stream):
    if stream is not None:
        return ctypes.c_void_p(stream)
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _frame(x):
    if isinstance(x, Frame):
        return x
    if isinstance(x, tuple):
        return frame_of(x[0], x[1])
    return frame_of(x)


NOALLOC_SCRATCH_BYTES = 48 << 20


def record_noalloc(prev, curr, out, t, width, height, fmt=FORMAT_RGBA8, settings=None, bgra=False,
                   yuv_matrix=YUV_BT709, yuv_full_range=False, device=0, stream=None):
    """fgr_record_noalloc: one estimate + one synth at t without a handle and without allocating. Returns
    (ok, reason): reason is the refusal's string, None on success."""
    info = InitInfo()
    info.width, info.height, info.format, info.bgra = width, height, fmt, int(bgra)
    info.transfer, info.linearWhiteNits = TRANSFER_ENCODED, 80.0
    info.yuvMatrix, info.yuvFullRange, info.device = yuv_matrix, int(yuv_full_range), device
    info.settings = settings if settings is not None else default_settings()
    err = ctypes.c_char_p()
    ok = bool(lib.fgr_record_noalloc(_stream(stream), ctypes.byref(info), _frame(prev), _frame(curr), _frame(out),
                                     t, ctypes.byref(err)))
    return ok, (err.value.decode() if err.value else None)


class Interpolator:
    """framegen::Interpolator over the C ABI. Raises RuntimeError(initError) if init fails."""

    def __init__(self, width, height, fmt=FORMAT_RGBA8, settings=None, bgra=False, transfer=TRANSFER_ENCODED,
                 linear_white_nits=80.0, yuv_matrix=YUV_BT709, yuv_full_range=False, device=0):
        info = InitInfo()
        info.width, info.height, info.format, info.bgra = width, height, fmt, int(bgra)
        info.transfer, info.linearWhiteNits = transfer, linear_white_nits
        info.yuvMatrix, info.yuvFullRange, info.device = yuv_matrix, int(yuv_full_range), device
        info.settings = settings if settings is not None else default_settings()
        self._h = lib.fgr_create(ctypes.byref(info))
        if not self._h:
            raise RuntimeError(_s(lib.fgr_last_init_error()))
        self._size = (width, height)

    def close(self):
        if self._h:
            lib.fgr_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def format(self):
        return lib.fgr_format(self._h)

    def resize(self, w, h):
        ok = bool(lib.fgr_resize(self._h, w, h))
        if ok:
            self._size = (w, h)
        return ok

    def settings_error(self, s):
        return _s(lib.fgr_settings_error(self._h, ctypes.byref(s)))

    def size_error(self, w, h):
        return _s(lib.fgr_size_error(self._h, w, h))

    def set_settings(self, s):
        return bool(lib.fgr_set_settings(self._h, ctypes.byref(s)))

    def reconfigure(self, s):
        return bool(lib.fgr_reconfigure(self._h, ctypes.byref(s)))

    def settings(self):
        s = Settings()
        lib.fgr_get_settings(self._h, ctypes.byref(s))
        return s

    def begin_frame(self):
        lib.fgr_begin_frame(self._h)

    def record_estimate(self, prev, curr, stream=None):
        return bool(lib.fgr_record_estimate(self._h, _stream(stream), _frame(prev), _frame(curr)))

    def record_synth(self, prev, curr, out, t, stream=None):
        return bool(lib.fgr_record_synth(self._h, _stream(stream), _frame(prev), _frame(curr), _frame(out), t))

    def record(self, prev, curr, out, t=0.5, stream=None):
        return bool(lib.fgr_record(self._h, _stream(stream), _frame(prev), _frame(curr), _frame(out), t))

    def record_synth_blur(self, prev, curr, out, t0, t1, stream=None):
        return bool(lib.fgr_record_synth_blur(self._h, _stream(stream), _frame(prev), _frame(curr), _frame(out),
                                              t0, t1))

    def record_observe(self, frame, stream=None):
        """framegen.h:610 recordObserve(): once per new real frame (static-UI protection)"""
        return bool(lib.fgr_record_observe(self._h, _stream(stream), _frame(frame)))

    def record_track(self, frame, stream=None):
        """framegen.h:686 recordTrack(): a real frame passed through without generating"""
        return bool(lib.fgr_record_track(self._h, _stream(stream), _frame(frame)))

    def release_unused_ui(self):
        lib.fgr_release_unused_ui(self._h)

    def debug_ui_mask(self, frame=None, stream=None):
        """Test / debug only: (mask uint8 (h, w) CUDA tensor, (x, y, w, h)) of a kept frame, or the box state
        (uint32 words (h, w, 4) as int32) with frame None; None if there is none."""
        import torch
        r = (ctypes.c_int32 * 4)()
        if frame is None:
            x, y, w, h = self.work_rect_now()
            dst = torch.empty((h, w, 4), dtype=torch.int32, device="cuda")
            ok = lib.fgr_debug_ui_mask(self._h, _stream(stream), Frame(), ctypes.c_void_p(dst.data_ptr()), r)
        else:   # the entry's rectangle lies in the frame
            dst = torch.empty(self._size[0] * self._size[1], dtype=torch.uint8, device="cuda")
            ok = lib.fgr_debug_ui_mask(self._h, _stream(stream), _frame(frame), ctypes.c_void_p(dst.data_ptr()), r)
            dst = dst[:r[2] * r[3]].view(r[3], r[2])
        return (dst, tuple(r)) if ok else None

    def work_rect_now(self):
        return work_rect(self.settings(), *self._size)

    def set_profiling(self, on):
        return bool(lib.fgr_set_profiling(self._h, int(on)))

    def read_timings(self):
        """[(pass group name, ns), ...] of the last profile, or None."""
        names = (ctypes.c_char_p * 64)()
        ticks = (ctypes.c_uint64 * 64)()
        n = lib.fgr_read_timings(self._h, names, ticks, 64)
        if n < 0:
            return None
        return [(names[i].decode(), int(ticks[i])) for i in range(n)]

    def record_debug_counters(self, dst, stream=None):
        """dst: a CUDA int32/uint32 tensor of >= 2 elements: [untrusted, SAD sum x 256]."""
        return bool(lib.fgr_record_debug_counters(self._h, _stream(stream), ctypes.c_void_p(dst.data_ptr())))

    def debug_sad_blocks(self):
        return lib.fgr_debug_sad_blocks(self._h)

    def debug_map_layout(self):
        l = DebugMapLayout()
        if not lib.fgr_debug_map_layout(self._h, ctypes.byref(l)):
            return None
        return {k: getattr(l, k) for k, _ in l._fields_}

    def set_seed_net(self, net):
        """Attach (or detach with None) a SeedNet; used when settings.seedNet is 1."""
        self._net = net
        return bool(lib.fgr_set_seed_net(self._h, net._h if net is not None else None))

    def debug_internal(self):
        """Test / debug only: the internal buffers' device pointers and layout (DebugInternal), or None."""
        d = DebugInternal()
        return d if lib.fgr_debug_internal(self._h, ctypes.byref(d)) else None

    def record_debug_maps(self, dst, stream=None):
        return bool(lib.fgr_record_debug_maps(self._h, _stream(stream), ctypes.c_void_p(dst.data_ptr())))

    def dump_fgmaps(self, path):
        """Write the latest record's debug maps as an fgtest --dump file (.fgmaps: 64-byte header, block field,
        per-pixel field, packed mismatch), readable by the reference's gpu/fgmaps.py."""
        import torch
        lay = self.debug_map_layout()
        if lay is None:
            return False
        buf = torch.empty(int(lay["size"]), dtype=torch.uint8, device="cuda")
        if not self.record_debug_maps(buf):
            return False
        head = (ctypes.c_uint8 * 64)()
        if not lib.fgr_fgmaps_header(self._h, head):
            return False
        torch.cuda.synchronize()
        with open(path, "wb") as f:
            f.write(bytes(head))
            f.write(buf.cpu().numpy().tobytes())
        return True


def _batch_pairs(items):
    """items: (prev, curr, multiplier, [out, ...]) with torch tensors (or (tensor, chroma) tuples), optional 5th
    element: format."""
    arr = (BatchPair * len(items))()
    for i, it in enumerate(items):
        prev, curr, n, outs = it[:4]
        fmt = it[4] if len(it) > 4 else FORMAT_RGBA8
        bp = arr[i]
        bp.prev, bp.curr = _frame(prev), _frame(curr)
        t = prev[0] if isinstance(prev, tuple) else prev
        bp.height, bp.width = int(t.shape[0]), int(t.shape[1])
        bp.format, bp.multiplier = fmt, n
        if len(outs) != n - 1:
            raise ValueError("a pair needs multiplier - 1 outputs")
        for k, o in enumerate(outs):
            bp.out[k] = _frame(o)
    return arr


class Batch:
    """Batched V1 over the C ABI (fgr_batch_*): many pairs, one launch per pass, replayed as a hipGraph."""

    def __init__(self, items, settings=None, bgra=False, use_graph=True, device=0, transfer=TRANSFER_ENCODED,
                 linear_white_nits=80.0, yuv_matrix=YUV_BT709, yuv_full_range=False, seed_net=None):
        info = BatchInfo()
        info.seedNet = seed_net._h if seed_net is not None else None
        self._net = seed_net
        info.device, info.bgra, info.useGraph = device, int(bgra), int(use_graph)
        info.transfer, info.linearWhiteNits = transfer, linear_white_nits
        info.yuvMatrix, info.yuvFullRange = yuv_matrix, int(yuv_full_range)
        info.settings = settings if settings is not None else default_settings()
        self._keep = items
        arr = _batch_pairs(items)
        self._h = lib.fgr_batch_create(ctypes.byref(info), arr, len(items))
        if not self._h:
            raise RuntimeError(_s(lib.fgr_batch_last_error()))

    def close(self):
        if self._h:
            lib.fgr_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, stream=None):
        if not lib.fgr_batch_record(self._h, _stream(stream)):
            raise RuntimeError("fgr_batch_record failed")

    def set_frames(self, items, stream=None):
        ok = bool(lib.fgr_batch_set_frames(self._h, _stream(stream), _batch_pairs(items), len(items)))
        if ok:   # a refused call keeps the batch's old frames, so their tensors stay referenced
            self._keep = items
        return ok

    def launches(self):
        return lib.fgr_batch_launches(self._h)

    def set_profiling(self, on):
        return bool(lib.fgr_batch_set_profiling(self._h, int(on)))

    TUNE_ROWS, TUNE_CHUNKS = (16, 32, 64), (16, 64, 256)

    FORM_PIXFIELD_WINDOWS, FORM_SEARCH_NARROW = 1, 2

    def autotune(self, reps=3, stream=None):
        """Time the 9 tile configurations on this batch, keep the fastest, then each estimate form knob, then 1 to 4
        lanes; returns ({(rows, chunk): ns, ("lanes", L): ns, ("pixfield", "gather" | "windows"): ns,
        ("search", "2x8" | "4x1"): ns}, (rows, chunk, lanes))."""
        ns = (ctypes.c_uint64 * 17)()
        if not lib.fgr_batch_autotune(self._h, _stream(stream), reps, ns):
            raise RuntimeError("fgr_batch_autotune failed")
        times = {(r, c): int(ns[i * 3 + j]) for i, r in enumerate(self.TUNE_ROWS) for j, c in enumerate(self.TUNE_CHUNKS)}
        for lanes in range(1, 5):
            times[("lanes", lanes)] = int(ns[8 + lanes])
        times[("pixfield", "gather")], times[("pixfield", "windows")] = int(ns[13]), int(ns[14])
        times[("search", "2x8")], times[("search", "4x1")] = int(ns[15]), int(ns[16])
        return times, self.get_tuning()

    def set_form(self, form):
        return bool(lib.fgr_batch_set_form(self._h, form))

    def get_form(self):
        return lib.fgr_batch_get_form(self._h)

    def set_tuning(self, rows, chunk, lanes=None):
        ok = bool(lib.fgr_batch_set_tuning(self._h, rows, chunk))
        return ok and (lanes is None or bool(lib.fgr_batch_set_lanes(self._h, lanes)))

    def get_tuning(self):
        """(rows, chunk, lanes)"""
        r, c = ctypes.c_int32(), ctypes.c_int32()
        lib.fgr_batch_get_tuning(self._h, ctypes.byref(r), ctypes.byref(c))
        return r.value, c.value, lib.fgr_batch_get_lanes(self._h)

    def set_lanes(self, lanes):
        return bool(lib.fgr_batch_set_lanes(self._h, lanes))

    def read_timings(self):
        """[(stage, pass, ns), ...] per launch of the last profiled record, or None."""
        n = self.launches()
        st = (ctypes.c_int32 * max(1, n))()
        ps = (ctypes.c_int32 * max(1, n))()
        tk = (ctypes.c_uint64 * max(1, n))()
        got = lib.fgr_batch_read_timings(self._h, st, ps, tk, n)
        if got < 0:
            return None
        return [(int(st[i]), int(ps[i]), int(tk[i])) for i in range(got)]

    def debug_counters(self, pair, synth, dst, stream=None):
        return bool(lib.fgr_batch_debug_counters(self._h, _stream(stream), pair, synth,
                                                 ctypes.c_void_p(dst.data_ptr())))

    def debug_internal(self, pair):
        d = DebugInternal()
        return d if lib.fgr_batch_debug_internal(self._h, pair, ctypes.byref(d)) else None


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


class SeedNet:
    """The learned seed network (fgr_seednet_*): luma tensors are CUDA float32 (n, h, w) of one level's planes."""

    def __init__(self, seed=0, device=0):
        self._h = lib.fgr_seednet_create(device, seed)
        if not self._h:
            raise RuntimeError("fgr_seednet_create failed")

    def close(self):
        if self._h:
            lib.fgr_seednet_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def param_count():
        return lib.fgr_seednet_param_count()

    def get_weights(self):
        import numpy as np
        w = np.empty(self.param_count(), np.float32)
        assert lib.fgr_seednet_get_weights(self._h, w.ctypes.data, w.size)
        return w

    def set_weights(self, w):
        import numpy as np
        w = np.ascontiguousarray(w, np.float32)
        assert lib.fgr_seednet_set_weights(self._h, w.ctypes.data, w.size)

    def forward(self, la, lc, out, stream=None):
        n, h, w = la.shape
        return bool(lib.fgr_seednet_forward(self._h, _stream(stream), _ptr(la), _ptr(lc), w, h, n, _ptr(out)))

    def gradients(self, la, lc, target, grad=None, loss=None, stream=None):
        n, h, w = la.shape
        return bool(lib.fgr_seednet_gradients(self._h, _stream(stream), _ptr(la), _ptr(lc), w, h, n, _ptr(target),
                                              _ptr(grad), _ptr(loss)))

    def calibrate_fp8(self, la, lc, stream=None):
        n, h, w = la.shape
        return bool(lib.fgr_seednet_calibrate_fp8(self._h, _stream(stream), _ptr(la), _ptr(lc), w, h, n))

    def set_fp8(self, on):
        return bool(lib.fgr_seednet_set_fp8(self._h, int(on)))

    def fp8_scales(self):
        import numpy as np
        sa, sw = np.empty(178, np.float32), np.empty(178, np.float32)
        assert lib.fgr_seednet_fp8_scales(self._h, sa.ctypes.data, sw.ctypes.data)
        return sa, sw

    def train_step(self, la, lc, target, lr, loss=None, stream=None):
        n, h, w = la.shape
        return bool(lib.fgr_seednet_train_step(self._h, _stream(stream), _ptr(la), _ptr(lc), w, h, n, _ptr(target),
                                               lr, _ptr(loss)))


def debug_fp8_e4m3(x, out, stream=None):
    """The seed network's FP8 E4M3 conversion of the CUDA float32 tensor x into the CUDA uint8 tensor out."""
    assert x.numel() == out.numel() and x.is_contiguous() and out.is_contiguous()
    return bool(lib.fgr_debug_fp8_e4m3(_stream(stream), ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                       x.numel()))


def search_field(la, lc, R, lam, raw, out, seeds=None, stream=None):
    """The coarsest-level field of n luma pairs: search with radius R around zero (or the better of zero and the
    proposal in `seeds`), then median3; the plain field (the training target) when seeds is None."""
    n, h, w = la.shape
    return bool(lib.fgr_seednet_search_field(_stream(stream), _ptr(la), _ptr(lc), w, h, n, R, lam, _ptr(seeds),
                                             _ptr(raw), _ptr(out)))


def tuning_save(path):
    """Write the process's batch tuning cache (fgr_tuning_save)."""
    return bool(lib.fgr_tuning_save(str(path).encode()))


def tuning_load(path):
    """Read a tuning cache file; returns the number of entries."""
    return int(lib.fgr_tuning_load(str(path).encode()))
