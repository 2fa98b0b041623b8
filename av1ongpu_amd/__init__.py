"""av1ongpu_amd - MI355X (gfx950) AV1 reconstruction backend.

Python access to the three native libraries built in this directory:
  libav1syn.so - seeded workload generator and the wire format of frame syntax
  libav1hip.so - the HIP reconstruction backend behind the C ABI (include/av1hip.h)
  libav1gpu.so - the C++ av1gpu::Decoder API (loaded by C++ callers, not here)

The GPU path has no fallback: HipDecoder raises when libav1hip.so is missing
or no gfx950 device is usable.
"""
import ctypes
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)

AV1H_OK = 0
AV1H_WIRE_MAGIC = b"AV1WSYN1"


class Planes(ctypes.Structure):
    _fields_ = [("plane", ctypes.POINTER(ctypes.c_uint16) * 3), ("stride", ctypes.c_int32 * 3)]


class FrameInfo(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int32), ("height", ctypes.c_int32), ("bit_depth", ctypes.c_int32),
                ("ss_x", ctypes.c_int32), ("ss_y", ctypes.c_int32)]


class Stats(ctypes.Structure):
    _fields_ = [("blocks", ctypes.c_int32), ("waves", ctypes.c_int32), ("kernel_launches", ctypes.c_int32),
                ("reconstruct_ms", ctypes.c_double), ("filter_ms", ctypes.c_double),
                ("readback_ms", ctypes.c_double), ("total_ms", ctypes.c_double),
                ("device_recon_ms", ctypes.c_double), ("device_filter_ms", ctypes.c_double)]


class DeviceInfo(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 128), ("arch", ctypes.c_char * 64),
                ("compute_units", ctypes.c_int32), ("l2_bytes", ctypes.c_int32),
                ("global_mem", ctypes.c_int64), ("clock_khz", ctypes.c_int32),
                ("mem_clock_khz", ctypes.c_int32), ("mem_bus_width", ctypes.c_int32)]


def _load(name):
    path = os.path.join(PKG_DIR, name)
    if not os.path.exists(path):
        raise RuntimeError(f"{name} is not built: run __graft_entry__.build() (make -C av1ongpu_amd)")
    return ctypes.CDLL(path)


_syn = None
_hip = None


def syn_lib():
    global _syn
    if _syn is None:
        lib = _load("libav1syn.so")
        lib.av1s_config_new.restype = ctypes.c_void_p
        lib.av1s_config_new.argtypes = [ctypes.c_char_p, ctypes.c_uint64]
        lib.av1s_config_free.argtypes = [ctypes.c_void_p]
        lib.av1s_config_set.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_double]
        lib.av1s_clip_new.restype = ctypes.c_void_p
        lib.av1s_clip_new.argtypes = [ctypes.c_void_p]
        lib.av1s_clip_free.argtypes = [ctypes.c_void_p]
        lib.av1s_clip_frames.argtypes = [ctypes.c_void_p]
        lib.av1s_clip_refs.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int32), ctypes.c_int]
        lib.av1s_clip_last_use.argtypes = [ctypes.c_void_p, ctypes.c_int]
        lib.av1s_clip_key.argtypes = [ctypes.c_void_p, ctypes.c_int]
        lib.av1s_clip_wire.restype = ctypes.c_int64
        lib.av1s_clip_wire.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.POINTER(ctypes.c_uint8))]
        lib.av1s_clip_pixels.restype = ctypes.c_int64
        lib.av1s_clip_pixels.argtypes = [ctypes.c_void_p, ctypes.c_int]
        lib.av1s_wire_roundtrip.argtypes = [ctypes.c_char_p, ctypes.c_int64]
        lib.av1s_fnv1a64.restype = ctypes.c_uint64
        lib.av1s_fnv1a64.argtypes = [ctypes.c_char_p, ctypes.c_int64]
        _syn = lib
    return _syn


class Clip:
    """A seeded synthetic clip: wire frames in decode order plus the reference
    structure (which earlier frames each frame predicts from)."""

    def __init__(self, workload="c1_mini", seed=1, **overrides):
        lib = syn_lib()
        cfg = lib.av1s_config_new(workload.encode(), seed)
        if not cfg:
            raise ValueError(f"unknown workload {workload}")
        try:
            for k, v in overrides.items():
                if lib.av1s_config_set(cfg, k.encode(), float(v)) != 0:
                    raise ValueError(f"unknown config key {k}")
            self._h = lib.av1s_clip_new(cfg)
        finally:
            lib.av1s_config_free(cfg)
        if not self._h:
            raise RuntimeError("clip generation failed")
        self.workload = workload
        self.seed = seed
        self.frames = lib.av1s_clip_frames(self._h)
        self.refs = []
        for f in range(self.frames):
            buf = (ctypes.c_int32 * 8)()
            n = lib.av1s_clip_refs(self._h, f, buf, 8)
            self.refs.append([buf[i] for i in range(n)])
        self.last_use = [lib.av1s_clip_last_use(self._h, f) for f in range(self.frames)]
        self.key = [lib.av1s_clip_key(self._h, f) for f in range(self.frames)]
        self._wire = {}

    def wire(self, f):
        """Wire bytes of frame f (reference slots hold frame indices)."""
        if f not in self._wire:
            ptr = ctypes.POINTER(ctypes.c_uint8)()
            n = syn_lib().av1s_clip_wire(self._h, f, ctypes.byref(ptr))
            self._wire[f] = ctypes.string_at(ptr, n)
        return self._wire[f]

    def pixels(self, f):
        return syn_lib().av1s_clip_pixels(self._h, f)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _syn is not None:
            _syn.av1s_clip_free(h)
            self._h = None


# ---- wire helpers ------------------------------------------------------------

_HDR_OFF = 16  # magic, version, header size


def wire_header(wire):
    """(width, coded_width, height, bit_depth, ss_x, ss_y) of a wire frame."""
    v = np.frombuffer(wire, dtype=np.int32, count=6, offset=_HDR_OFF)
    return tuple(int(x) for x in v)


def wire_set_slots(wire, slots):
    """A copy of `wire` whose reference_slots name `slots` (decoder slot ids)."""
    b = bytearray(wire)
    nref_off = _HDR_OFF + 9 * 4
    slots_off = _HDR_OFF + 10 * 4
    np.frombuffer(b, dtype=np.int32, count=1, offset=nref_off)[0] = len(slots)
    arr = np.frombuffer(b, dtype=np.int32, count=8, offset=slots_off)
    arr[:] = 0
    arr[:len(slots)] = slots
    return bytes(b)


def plane_shapes(width, height, ss_x=1, ss_y=1):
    cw = (width + ss_x) >> ss_x
    ch = (height + ss_y) >> ss_y
    return [(height, width), (ch, cw), (ch, cw)]


def alloc_planes(width, height, ss_x=1, ss_y=1):
    """Three contiguous uint16 planes and the Planes struct pointing at them."""
    arrs = [np.zeros(s, dtype=np.uint16) for s in plane_shapes(width, height, ss_x, ss_y)]
    p = Planes()
    for i, a in enumerate(arrs):
        p.plane[i] = a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16))
        p.stride[i] = a.shape[1]
    return arrs, p


def yuv16_bytes(planes):
    """frame_to_yuv16 layout: planar little-endian 16 bit rows of plane width."""
    return b"".join(np.ascontiguousarray(a).astype("<u2").tobytes() for a in planes)


def fnv1a64(data):
    """FNV-1a 64 over bytes (the reference's frame hash, tools/av1gpu_decode.cpp:133-139)."""
    if isinstance(data, np.ndarray):
        data = np.ascontiguousarray(data).tobytes()
    return syn_lib().av1s_fnv1a64(data, len(data))


def nv12_size(width, height, ss_x, ss_y, pitch):
    if width <= 0 or height <= 0 or ss_x != 1 or ss_y != 1:
        return 0
    crow = 2 * ((width + 1) // 2)
    if pitch & 1 or pitch < crow:
        return 0
    return pitch * (height + (height + 1) // 2)


# ---- the HIP backend ----------------------------------------------------------

def hip_lib():
    global _hip
    if _hip is None:
        lib = _load("libav1hip.so")
        P = ctypes.c_void_p
        lib.av1h_create.argtypes = [ctypes.c_int, ctypes.POINTER(P)]
        lib.av1h_destroy.argtypes = [P]
        lib.av1h_last_error.restype = ctypes.c_char_p
        lib.av1h_last_error.argtypes = [P]
        lib.av1h_get_device_info.argtypes = [P, ctypes.POINTER(DeviceInfo)]
        lib.av1h_decode.argtypes = [P, P, ctypes.c_int, ctypes.c_int, ctypes.POINTER(Planes),
                                    ctypes.POINTER(FrameInfo), ctypes.POINTER(ctypes.c_int)]
        lib.av1h_decode_wire.argtypes = [P, ctypes.c_char_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_int,
                                         ctypes.POINTER(Planes), ctypes.POINTER(FrameInfo),
                                         ctypes.POINTER(ctypes.c_int)]
        lib.av1h_reserve_reference.argtypes = [P]
        lib.av1h_release_reference.argtypes = [P, ctypes.c_int]
        lib.av1h_nv12_size.restype = ctypes.c_size_t
        lib.av1h_nv12_size.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_size_t]
        lib.av1h_to_nv12.argtypes = [P, ctypes.c_int, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t]
        lib.av1h_to_rgb.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t]
        lib.av1h_read_slot.argtypes = [P, ctypes.c_int, ctypes.POINTER(Planes), ctypes.POINTER(FrameInfo)]
        lib.av1h_reset.argtypes = [P]
        lib.av1h_get_stats.argtypes = [P, ctypes.POINTER(Stats)]
        lib.av1h_decode_batch.argtypes = [ctypes.POINTER(P), ctypes.POINTER(P), ctypes.c_int,
                                          ctypes.POINTER(ctypes.c_int)]
        lib.av1h_plan_frame_wire.argtypes = [P, ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(P)]
        lib.av1h_decode_planned.argtypes = [P, P, ctypes.c_int, ctypes.c_int, ctypes.POINTER(Planes),
                                            ctypes.POINTER(FrameInfo), ctypes.POINTER(ctypes.c_int)]
        lib.av1h_discard_plan.argtypes = [P, P]
        lib.av1h_set_stage_timing.argtypes = [P, ctypes.c_int]
        lib.av1h_get_stage_times.argtypes = [P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64),
                                             ctypes.POINTER(ctypes.c_int64)]
        lib.av1h_set_queue_trace.argtypes = [P, ctypes.c_int]
        lib.av1h_get_queue_trace.restype = ctypes.c_int64
        lib.av1h_get_queue_trace.argtypes = [P, ctypes.c_void_p, ctypes.c_int64]
        lib.av1h_nv12_device_buffer.restype = ctypes.c_void_p
        lib.av1h_nv12_device_buffer.argtypes = [P]
        _hip = lib
    return _hip


STAGES = ("inter", "queue", "deblock", "cdef", "superres", "lr", "store", "grain")


class HipError(RuntimeError):
    pass


class HipDecoder:
    """One decoder context on a HIP device (C ABI av1h_*)."""

    def __init__(self, device=0):
        lib = hip_lib()
        self._lib = lib
        h = ctypes.c_void_p()
        rc = lib.av1h_create(device, ctypes.byref(h))
        if rc != AV1H_OK:
            raise HipError(f"av1h_create failed ({rc}): no usable gfx950 device")
        self._h = h

    @property
    def handle(self):
        return self._h

    def error(self):
        return self._lib.av1h_last_error(self._h).decode(errors="replace")

    def device_info(self):
        info = DeviceInfo()
        self._lib.av1h_get_device_info(self._h, ctypes.byref(info))
        return info

    def decode_wire(self, wire, keep=True, reserved=-1, read_back=True):
        """Decode a wire frame; returns (slot, planes or None)."""
        w, _cw, h, _bd, sx, sy = wire_header(wire)
        arrs, planes = (alloc_planes(w, h, sx, sy) if read_back else (None, None))
        slot = ctypes.c_int(-1)
        info = FrameInfo()
        rc = self._lib.av1h_decode_wire(self._h, wire, len(wire), 1 if keep else 0, reserved,
                                        ctypes.byref(planes) if read_back else None, ctypes.byref(info),
                                        ctypes.byref(slot))
        if rc != AV1H_OK:
            raise HipError(f"decode failed ({rc}): {self.error()}")
        return slot.value, arrs

    def plan_wire(self, wire):
        """Flatten and upload a wire frame into a device-resident plan."""
        p = ctypes.c_void_p()
        rc = self._lib.av1h_plan_frame_wire(self._h, wire, len(wire), ctypes.byref(p))
        if rc != AV1H_OK:
            raise HipError(f"plan_frame failed ({rc}): {self.error()}")
        return p

    def decode_planned(self, plan, keep=True, reserved=-1):
        """Decode a plan, picture left on the device; returns the slot."""
        slot = ctypes.c_int(-1)
        rc = self._lib.av1h_decode_planned(self._h, plan, 1 if keep else 0, reserved, None, None,
                                           ctypes.byref(slot))
        if rc != AV1H_OK:
            raise HipError(f"decode_planned failed ({rc}): {self.error()}")
        return slot.value

    def discard_plan(self, plan):
        self._lib.av1h_discard_plan(self._h, plan)

    def set_stage_timing(self, enable=True):
        self._lib.av1h_set_stage_timing(self._h, 1 if enable else 0)

    def stage_times(self):
        """{stage name: (accumulated device ms, decodes counted, compulsory bytes)}"""
        ms = (ctypes.c_double * len(STAGES))()
        n = (ctypes.c_int64 * len(STAGES))()
        b = (ctypes.c_int64 * len(STAGES))()
        self._lib.av1h_get_stage_times(self._h, ms, n, b)
        return {name: (ms[i], n[i], b[i]) for i, name in enumerate(STAGES)}

    def set_queue_trace(self, enable=True):
        self._lib.av1h_set_queue_trace(self._h, 1 if enable else 0)

    def queue_trace(self):
        """(n, 4) uint64: taken, ready, published clock, workgroup per intra block."""
        n = self._lib.av1h_get_queue_trace(self._h, None, 0)
        out = np.zeros((max(n, 0), 4), dtype=np.uint64)
        if n > 0:
            self._lib.av1h_get_queue_trace(self._h, out.ctypes.data, out.size)
        return out

    def pack_nv12_device(self, slot, pitch):
        """NV12 pack left on the device (no host copy)."""
        rc = self._lib.av1h_to_nv12(self._h, slot, pitch, None, 0)
        if rc != AV1H_OK:
            raise HipError(f"to_nv12 failed ({rc}): {self.error()}")

    def reserve(self):
        return self._lib.av1h_reserve_reference(self._h)

    def release(self, slot):
        self._lib.av1h_release_reference(self._h, slot)

    def to_nv12(self, slot, pitch, width, height):
        n = nv12_size(width, height, 1, 1, pitch)
        buf = np.zeros(max(n, 1), dtype=np.uint8)
        rc = self._lib.av1h_to_nv12(self._h, slot, pitch, buf.ctypes.data, buf.nbytes)
        if rc != AV1H_OK:
            raise HipError(f"to_nv12 failed ({rc}): {self.error()}")
        return buf[:n]

    def to_rgb(self, slot, matrix, rng, width, height):
        buf = np.zeros(width * height * 3, dtype=np.uint8)
        rc = self._lib.av1h_to_rgb(self._h, slot, matrix, rng, buf.ctypes.data, buf.nbytes)
        if rc != AV1H_OK:
            raise HipError(f"to_rgb failed ({rc}): {self.error()}")
        return buf

    def read_slot(self, slot, width, height, ss_x=1, ss_y=1):
        arrs, planes = alloc_planes(width, height, ss_x, ss_y)
        rc = self._lib.av1h_read_slot(self._h, slot, ctypes.byref(planes), None)
        if rc != AV1H_OK:
            raise HipError(f"read_slot failed ({rc}): {self.error()}")
        return arrs

    def stats(self):
        s = Stats()
        self._lib.av1h_get_stats(self._h, ctypes.byref(s))
        return s

    def reset(self):
        self._lib.av1h_reset(self._h)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.av1h_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def decode_clip(decoder_factory, clip, frames=None, read_back=True):
    """Decode a clip in order through any object with decode_wire / release
    (HipDecoder, the test oracle), mapping frame references to the slots the
    decoder returned. Yields (frame index, slot, planes)."""
    dec = decoder_factory() if callable(decoder_factory) else decoder_factory
    slot_of = {}
    n = clip.frames if frames is None else min(frames, clip.frames)
    for f in range(n):
        wire = wire_set_slots(clip.wire(f), [slot_of[r] for r in clip.refs[f]])
        keep = clip.last_use[f] > f
        slot, planes = dec.decode_wire(wire, keep=keep, read_back=read_back)
        if keep:
            slot_of[f] = slot
        for r in clip.refs[f]:
            if clip.last_use[r] == f:
                dec.release(slot_of.pop(r))
        yield f, slot, planes
