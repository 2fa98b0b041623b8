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
        lib.av1s_wire_without_levels.restype = ctypes.c_int64
        lib.av1s_wire_without_levels.argtypes = [ctypes.c_char_p, ctypes.c_int64, ctypes.c_char_p, ctypes.c_int64]
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


def wire_without_levels(wire):
    """The wire frame with its coded levels and quantiser matrices dropped."""
    lib = syn_lib()
    n = lib.av1s_wire_without_levels(wire, len(wire), None, 0)
    buf = ctypes.create_string_buffer(max(n, 1))
    m = lib.av1s_wire_without_levels(wire, len(wire), buf, n)
    if m < 0:
        raise ValueError("not a wire frame")
    return buf.raw[:m]


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
        lib.av1h_decode_planned_batch.argtypes = [ctypes.POINTER(P), ctypes.POINTER(P), ctypes.POINTER(ctypes.c_int),
                                                  ctypes.POINTER(ctypes.c_int), ctypes.c_int,
                                                  ctypes.POINTER(ctypes.c_int)]
        lib.av1h_plan_frame_wire.argtypes = [P, ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(P)]
        lib.av1h_decode_planned.argtypes = [P, P, ctypes.c_int, ctypes.c_int, ctypes.POINTER(Planes),
                                            ctypes.POINTER(FrameInfo), ctypes.POINTER(ctypes.c_int)]
        lib.av1h_discard_plan.argtypes = [P, P]
        lib.av1h_set_stage_timing.argtypes = [P, ctypes.c_int]
        lib.av1h_get_stage_times.argtypes = [P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64),
                                             ctypes.POINTER(ctypes.c_int64)]
        lib.av1h_set_queue_trace.argtypes = [P, ctypes.c_int]
        lib.av1h_set_wait_limit.argtypes = [P, ctypes.c_int]
        lib.av1h_get_queue_trace.restype = ctypes.c_int64
        lib.av1h_get_queue_trace.argtypes = [P, ctypes.c_void_p, ctypes.c_int64]
        lib.av1h_set_decode_nv12.argtypes = [P, ctypes.c_size_t]
        lib.av1h_set_graphs.argtypes = [P, ctypes.c_int]
        lib.av1h_set_output_hash.argtypes = [P, ctypes.c_int]
        lib.av1h_tune_params.restype = ctypes.c_int
        lib.av1h_tune_params.argtypes = []
        lib.av1h_tune_name.restype = ctypes.c_char_p
        lib.av1h_tune_name.argtypes = [ctypes.c_int]
        lib.av1h_tune_get.argtypes = [P, ctypes.POINTER(ctypes.c_int32), ctypes.c_int]
        lib.av1h_tune_set.argtypes = [P, ctypes.POINTER(ctypes.c_int32), ctypes.c_int]
        lib.av1h_last_hash.argtypes = [P, ctypes.POINTER(ctypes.c_uint64)]
        lib.av1h_nv12_device_buffer.restype = ctypes.c_void_p
        lib.av1h_nv12_device_buffer.argtypes = [P]
        _bind_server(lib)
        _bind_costmodel(lib)
        _bind_pipe(lib)
        P = ctypes.c_void_p
        lib.av1h_decode_planned_batch_async.argtypes = [P, ctypes.POINTER(P), ctypes.POINTER(P),
                                                        ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                                        ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        lib.av1h_batch_finish.argtypes = [P]
        lib.av1h_batch_pending.argtypes = [P]
        lib.av1h_plan_needs_idle.argtypes = [P, P]
        _hip = lib
    return _hip


def _get_tuning(lib, h):
    n = lib.av1h_tune_params()
    buf = (ctypes.c_int32 * n)()
    lib.av1h_tune_get(h, buf, n)
    return {lib.av1h_tune_name(i).decode(): buf[i] for i in range(n)}


def _set_tuning(lib, h, values):
    cur = _get_tuning(lib, h)
    cur.update(values)
    names = [lib.av1h_tune_name(i).decode() for i in range(len(cur))]
    buf = (ctypes.c_int32 * len(names))(*[cur[k] for k in names])
    if lib.av1h_tune_set(h, buf, len(names)) != AV1H_OK:
        raise HipError(f"invalid launch shapes {values}")


class ServerConfig(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int32), ("max_batch", ctypes.c_int32), ("planner_threads", ctypes.c_int32),
                ("capture", ctypes.c_int32), ("hash_pictures", ctypes.c_int32), ("launchers", ctypes.c_int32),
                ("heavy_levels", ctypes.c_int32), ("heavy_launchers", ctypes.c_int32),
                ("heavy_batch", ctypes.c_int32), ("hbm_budget", ctypes.c_int64), ("nv12_pitch", ctypes.c_size_t)]


class Completion(ctypes.Structure):
    _fields_ = [("tag", ctypes.c_uint64), ("stream", ctypes.c_int32), ("slot", ctypes.c_int32),
                ("status", ctypes.c_int32), ("batch_frames", ctypes.c_int32), ("arrival_ms", ctypes.c_double),
                ("start_ms", ctypes.c_double), ("done_ms", ctypes.c_double),
                ("picture", ctypes.POINTER(ctypes.c_uint16)), ("picture_samples", ctypes.c_int64),
                ("hash", ctypes.c_uint64), ("plan_start_ms", ctypes.c_double), ("planned_ms", ctypes.c_double)]


class ServerStats(ctypes.Structure):
    _fields_ = [("frames", ctypes.c_int64), ("batches", ctypes.c_int64), ("admitted", ctypes.c_int64),
                ("waiting", ctypes.c_int64), ("active", ctypes.c_int64), ("budget_used", ctypes.c_int64),
                ("busy_ms", ctypes.c_double), ("max_batch_frames", ctypes.c_int32)]


def _bind_server(lib):
    P = ctypes.c_void_p
    lib.av1h_server_now_ms.restype = ctypes.c_double
    lib.av1h_server_now_ms.argtypes = []
    lib.av1h_server_create.argtypes = [ctypes.POINTER(ServerConfig), ctypes.POINTER(P)]
    lib.av1h_server_destroy.argtypes = [P]
    lib.av1h_server_open_stream.argtypes = [P] + [ctypes.c_int] * 6
    lib.av1h_server_submit_wire.argtypes = [P, ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t, ctypes.c_int,
                                            ctypes.POINTER(ctypes.c_int32), ctypes.c_int, ctypes.c_double,
                                            ctypes.c_uint64]
    lib.av1h_server_submit_wire_ref.argtypes = lib.av1h_server_submit_wire.argtypes
    lib.av1h_server_submit_plan.argtypes = [P, ctypes.c_int, P, ctypes.c_int, ctypes.POINTER(ctypes.c_int32),
                                            ctypes.c_int, ctypes.c_double, ctypes.c_uint64]
    lib.av1h_server_close_stream.argtypes = [P, ctypes.c_int]
    lib.av1h_server_poll.argtypes = [P, ctypes.POINTER(Completion), ctypes.c_int, ctypes.c_int]
    lib.av1h_server_free_picture.argtypes = [ctypes.POINTER(ctypes.c_uint16)]
    lib.av1h_server_get_stats.argtypes = [P, ctypes.POINTER(ServerStats)]
    lib.av1h_server_launcher.restype = P
    lib.av1h_server_launcher.argtypes = [P, ctypes.c_int]
    lib.av1h_server_last_error.restype = ctypes.c_char_p
    lib.av1h_server_last_error.argtypes = [P]
    lib.av1h_stream_bytes.restype = ctypes.c_int64
    lib.av1h_stream_bytes.argtypes = [ctypes.c_int] * 6 + [ctypes.c_size_t]
    lib.av1h_set_pool_slots.argtypes = [P, ctypes.c_int]
    lib.av1h_plan_levels.argtypes = [P]
    lib.av1h_plan_queue_counts.argtypes = [P, ctypes.POINTER(ctypes.c_int32)]


class PipeFrame(ctypes.Structure):
    _fields_ = [("slot", ctypes.c_int32), ("status", ctypes.c_int32), ("submit_ms", ctypes.c_double),
                ("plan_start_ms", ctypes.c_double), ("planned_ms", ctypes.c_double),
                ("decode_start_ms", ctypes.c_double), ("done_ms", ctypes.c_double),
                ("flatten_ms", ctypes.c_double), ("upload_ms", ctypes.c_double), ("hash", ctypes.c_uint64),
                ("launch_ms", ctypes.c_double), ("wait_ms", ctypes.c_double), ("device_ms", ctypes.c_double)]


def _bind_pipe(lib):
    P = ctypes.c_void_p
    I32P = ctypes.POINTER(ctypes.c_int32)
    lib.av1h_pipe_create.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.POINTER(P)]
    lib.av1h_pipe_submit_wire.argtypes = [P, ctypes.c_char_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_int, I32P,
                                          ctypes.c_int]
    lib.av1h_pipe_submit_wire.restype = ctypes.c_int64
    lib.av1h_pipe_wait.argtypes = [P, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)]
    lib.av1h_pipe_result.argtypes = [P, ctypes.c_int64, ctypes.POINTER(PipeFrame)]
    lib.av1h_plan_host_times.argtypes = [P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    lib.av1h_pipe_error.argtypes = [P]
    lib.av1h_pipe_error.restype = ctypes.c_char_p
    lib.av1h_pipe_destroy.argtypes = [P]


def _bind_costmodel(lib):
    P = ctypes.c_void_p
    F = ctypes.POINTER(ctypes.c_float)
    lib.av1h_plan_sb_features.argtypes = [P, F]
    lib.av1h_plan_sb_times.argtypes = [P, P, F]
    lib.av1h_plan_schedule.argtypes = [P, P, F]
    lib.av1h_plan_schedule_order.argtypes = [P, P, ctypes.c_int, F]
    lib.av1h_plan_schedule_model.argtypes = [P, P, P]
    lib.av1h_cm_create.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(P)]
    lib.av1h_cm_destroy.argtypes = [P]
    lib.av1h_cm_set_params.argtypes = [P, F]
    lib.av1h_cm_get_params.argtypes = [P, F]
    lib.av1h_cm_set_norm.argtypes = [P, F, F, ctypes.c_float, ctypes.c_float]
    lib.av1h_cm_get_norm.argtypes = [P, F, F, F, F]
    lib.av1h_cm_predict.argtypes = [P, F, ctypes.c_int64, F]
    lib.av1h_cm_predict_noalloc.argtypes = [P, F, ctypes.c_int64, F]
    lib.av1h_cm_gradients.argtypes = [P, F, F, ctypes.c_int64, F, F]
    lib.av1h_cm_train.argtypes = [P, F, F, ctypes.c_int64, ctypes.c_int, ctypes.c_float, ctypes.c_uint64,
                                  ctypes.c_int, F]


def _fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


CM_FEATURES = 24
ORDER_LEVEL, ORDER_CODING, ORDER_PATH = 0, 1, 2
CM_PARAMS = 5825


class CostModel:
    """The superblock cost model (av1h_cm_*): a 24-64-64-1 FP32 perceptron
    trained and evaluated by HIP kernels."""

    def __init__(self, device=0, seed=1):
        lib = hip_lib()
        self._lib = lib
        h = ctypes.c_void_p()
        rc = lib.av1h_cm_create(device, seed, ctypes.byref(h))
        if rc != AV1H_OK:
            raise HipError(f"av1h_cm_create failed ({rc})")
        self._h = h

    @property
    def handle(self):
        return self._h

    def params(self):
        out = np.zeros(CM_PARAMS, np.float32)
        self._lib.av1h_cm_get_params(self._h, _fptr(out))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == CM_PARAMS
        if self._lib.av1h_cm_set_params(self._h, _fptr(p)) != AV1H_OK:
            raise HipError("set_params failed")

    def norm(self):
        mean = np.zeros(CM_FEATURES, np.float32)
        dev = np.zeros(CM_FEATURES, np.float32)
        ym, yd = ctypes.c_float(), ctypes.c_float()
        self._lib.av1h_cm_get_norm(self._h, _fptr(mean), _fptr(dev), ctypes.byref(ym), ctypes.byref(yd))
        return mean, dev, ym.value, yd.value

    def set_norm(self, mean, dev, y_mean, y_dev):
        mean = np.ascontiguousarray(mean, np.float32)
        dev = np.ascontiguousarray(dev, np.float32)
        if self._lib.av1h_cm_set_norm(self._h, _fptr(mean), _fptr(dev), float(y_mean), float(y_dev)) != AV1H_OK:
            raise HipError("set_norm failed")

    def predict(self, X):
        X = np.ascontiguousarray(X, np.float32)
        out = np.zeros(X.shape[0], np.float32)
        if self._lib.av1h_cm_predict(self._h, _fptr(X), X.shape[0], _fptr(out)) != AV1H_OK:
            raise HipError("predict failed")
        return out

    def gradients(self, X, y):
        X = np.ascontiguousarray(X, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        g = np.zeros(CM_PARAMS, np.float32)
        loss = ctypes.c_float()
        if self._lib.av1h_cm_gradients(self._h, _fptr(X), _fptr(y), X.shape[0], _fptr(g),
                                       ctypes.byref(loss)) != AV1H_OK:
            raise HipError("gradients failed")
        return g, loss.value

    def train(self, X, y, epochs=10, lr=1e-3, seed=1, set_norm=True):
        X = np.ascontiguousarray(X, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        losses = np.zeros(epochs, np.float32)
        if self._lib.av1h_cm_train(self._h, _fptr(X), _fptr(y), X.shape[0], epochs, lr, seed,
                                   1 if set_norm else 0, _fptr(losses)) != AV1H_OK:
            raise HipError("train failed")
        return losses

    def close(self):
        if getattr(self, "_h", None):
            self._lib.av1h_cm_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


STAGES = ("inter", "queue", "deblock", "cdef", "superres", "lr", "store", "grain")


class HipError(RuntimeError):
    pass


class Pipe:
    """One stream's frames through planner threads and an in-order decode
    worker (av1h_pipe_*): submit wire frames, wait for them. The wire bytes of
    a frame are held here until it completes."""

    def __init__(self, decoder, planner_threads=8, depth=32):
        self._lib = decoder._lib
        self._dec = decoder
        h = ctypes.c_void_p()
        rc = self._lib.av1h_pipe_create(decoder.handle, planner_threads, depth, ctypes.byref(h))
        if rc != AV1H_OK:
            raise HipError(f"av1h_pipe_create failed ({rc})")
        self._h = h
        self._held = {}
        self._done = 0

    def submit(self, wire, keep=True, reserved=-1, release=()):
        rel = (ctypes.c_int32 * max(len(release), 1))(*release)
        seq = self._lib.av1h_pipe_submit_wire(self._h, wire, len(wire), 1 if keep else 0, reserved, rel,
                                              len(release))
        if seq < 0:
            raise HipError(f"pipe submit failed ({seq})")
        self._held[seq] = wire
        return seq

    def wait(self, seq=-1):
        """Waits for frames 0..seq (all when negative); raises on a failed frame."""
        failed = ctypes.c_int64(-1)
        rc = self._lib.av1h_pipe_wait(self._h, seq, ctypes.byref(failed))
        upto = max(self._held, default=-1) if seq < 0 else seq
        for k in [k for k in self._held if k <= upto]:
            del self._held[k]
        if rc != AV1H_OK:
            raise HipError(f"frame {failed.value} failed ({rc}): "
                           f"{self._lib.av1h_pipe_error(self._h).decode(errors='replace')}")

    def result(self, seq):
        """The completed frame's PipeFrame: slot, status, times, picture checksum or 0."""
        r = PipeFrame()
        if self._lib.av1h_pipe_result(self._h, seq, ctypes.byref(r)) != AV1H_OK:
            raise HipError(f"frame {seq} is not complete")
        return r

    def close(self):
        if self._h:
            self._lib.av1h_pipe_destroy(self._h)
            self._h = None
            self._held.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
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

    def plan_needs_idle(self, plan):
        """True when launching `plan` would wait for every batch in flight on
        this decoder (buffers that grow, stage timing, the queue trace)."""
        return self._lib.av1h_plan_needs_idle(self._h, plan) == 1

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

    def set_wait_limit(self, polls):
        """Polls of a queued block's dependency wait before the decode fails with
        a time-out (0: the default)."""
        self._lib.av1h_set_wait_limit(self._h, int(polls))

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

    def queue_counts(self, plan):
        """{blocks, packed_readers, publishers, unmarked} of a plan's queue."""
        c = (ctypes.c_int32 * 4)()
        if self._lib.av1h_plan_queue_counts(plan, c) != AV1H_OK:
            raise HipError("invalid plan")
        return dict(zip(("blocks", "packed_readers", "publishers", "unmarked"), list(c)))

    def sb_features(self, plan):
        """(superblocks x 24) features of a plan."""
        n = self._lib.av1h_plan_sb_features(plan, None)
        out = np.zeros((max(n, 0), CM_FEATURES), np.float32)
        if n > 0:
            self._lib.av1h_plan_sb_features(plan, _fptr(out))
        return out

    def sb_times(self, plan):
        """Queue work per superblock (us) of the last traced decode of `plan` alone."""
        n = self._lib.av1h_plan_sb_features(plan, None)
        out = np.zeros(max(n, 0), np.float32)
        rc = self._lib.av1h_plan_sb_times(self._h, plan, _fptr(out))
        if rc < 0:
            raise HipError(f"sb_times failed ({rc}): {self.error()}")
        return out

    def schedule(self, plan, costs=None):
        """Order the plan's queue by per-superblock costs (None: static)."""
        c = None if costs is None else np.ascontiguousarray(costs, np.float32)
        rc = self._lib.av1h_plan_schedule(self._h, plan, None if c is None else _fptr(c))
        if rc != AV1H_OK:
            raise HipError(f"schedule failed ({rc}): {self.error()}")

    def schedule_order(self, plan, order, costs=None):
        """Order the plan's queue: ORDER_LEVEL, ORDER_CODING or ORDER_PATH, by
        per-superblock costs (None: sample counts; for ORDER_LEVEL the static order)."""
        c = None if costs is None else np.ascontiguousarray(costs, np.float32)
        rc = self._lib.av1h_plan_schedule_order(self._h, plan, order, None if c is None else _fptr(c))
        if rc != AV1H_OK:
            raise HipError(f"schedule_order failed ({rc}): {self.error()}")

    def schedule_model(self, plan, model):
        rc = self._lib.av1h_plan_schedule_model(self._h, plan, model.handle)
        if rc != AV1H_OK:
            raise HipError(f"schedule_model failed ({rc}): {self.error()}")

    def set_pool_slots(self, slots):
        if self._lib.av1h_set_pool_slots(self._h, slots) != AV1H_OK:
            raise HipError(self.error())

    def tuning(self):
        """{parameter name: value} of the launch shapes of batches this context leads."""
        return _get_tuning(self._lib, self._h)

    def set_tuning(self, values):
        _set_tuning(self._lib, self._h, values)

    def set_output_hash(self, enable=True):
        """Checksum every decoded picture on the device (picture_hash for the host version)."""
        self._lib.av1h_set_output_hash(self._h, 1 if enable else 0)

    def last_hash(self):
        v = ctypes.c_uint64()
        if self._lib.av1h_last_hash(self._h, ctypes.byref(v)) != AV1H_OK:
            raise HipError(self.error())
        return v.value

    def set_graphs(self, enable=True, per_kernel_set=True):
        """Replay captured launch sequences (HIP graphs) for batches this context
        leads: one per batch size and set of kernels with work, or with
        per_kernel_set=False one per batch size with every kernel launched."""
        self._lib.av1h_set_graphs(self._h, 0 if not enable else (1 if per_kernel_set else 2))

    def set_decode_nv12(self, pitch):
        """Pack NV12 at `pitch` on the device as part of every decode (0: off)."""
        if self._lib.av1h_set_decode_nv12(self._h, pitch) != AV1H_OK:
            raise HipError(f"set_decode_nv12({pitch}) failed")

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


def decode_planned_batch(decoders, plans, keep=None, reserved=None):
    """Decode one planned frame per decoder as one batch (av1h_decode_planned_batch);
    returns the slots. The launches go to the first decoder's stream."""
    n = len(decoders)
    if n != len(plans):
        raise ValueError("one plan per decoder")
    lib = hip_lib()
    P = ctypes.c_void_p
    hs = (P * n)(*[d.handle for d in decoders])
    ps = (P * n)(*plans)
    ks = (ctypes.c_int * n)(*[1 if k else 0 for k in (keep if keep is not None else [True] * n)])
    rs = (ctypes.c_int * n)(*(reserved if reserved is not None else [-1] * n))
    slots = (ctypes.c_int * n)()
    rc = lib.av1h_decode_planned_batch(hs, ps, ks, rs, n, slots)
    if rc != AV1H_OK:
        raise HipError(f"decode_planned_batch failed ({rc}): {decoders[0].error()}")
    return list(slots)


def decode_planned_batch_async(decoders, plans, keep=None, reserved=None, launcher=None):
    """Launch one planned frame per decoder as one batch without waiting
    (av1h_decode_planned_batch_async); returns the slots. The launches go to
    `launcher`'s stream (the first decoder's by default); batch_finish takes
    the results."""
    n = len(decoders)
    if n != len(plans):
        raise ValueError("one plan per decoder")
    lib = hip_lib()
    P = ctypes.c_void_p
    lead = (launcher or decoders[0]).handle
    hs = (P * n)(*[d.handle for d in decoders])
    ps = (P * n)(*plans)
    ks = (ctypes.c_int * n)(*[1 if k else 0 for k in (keep if keep is not None else [True] * n)])
    rs = (ctypes.c_int * n)(*(reserved if reserved is not None else [-1] * n))
    slots = (ctypes.c_int * n)()
    rc = lib.av1h_decode_planned_batch_async(lead, hs, ps, ks, rs, n, slots)
    if rc != AV1H_OK:
        raise HipError(f"decode_planned_batch_async failed ({rc}): {(launcher or decoders[0]).error()}")
    return list(slots)


def batch_finish(launcher):
    """Wait for the oldest launched batch led by `launcher`; raises on its failure."""
    lib = hip_lib()
    rc = lib.av1h_batch_finish(launcher.handle)
    if rc != AV1H_OK:
        raise HipError(f"batch failed ({rc}): {launcher.error()}")


def batch_pending(launcher):
    """Batches launched on `launcher` and not yet finished."""
    return int(hip_lib().av1h_batch_pending(launcher.handle))


def decode_clips_batched(decoders, clips, frames=None, read_back=True):
    """Decode several clips, one per decoder, in lockstep: round r decodes
    frame r of every clip that has one as one batch. Yields (round, [(clip
    index, slot, planes or None)])."""
    state = [dict() for _ in clips]
    n = max(c.frames for c in clips) if frames is None else frames
    for f in range(n):
        live = [i for i, c in enumerate(clips) if f < c.frames]
        plans, keeps = [], []
        for i in live:
            c = clips[i]
            wire = wire_set_slots(c.wire(f), [state[i][r] for r in c.refs[f]])
            plans.append(decoders[i].plan_wire(wire))
            keeps.append(c.last_use[f] > f)
        try:
            slots = decode_planned_batch([decoders[i] for i in live], plans, keeps)
        finally:
            for i, p in zip(live, plans):
                decoders[i].discard_plan(p)
        out = []
        for i, slot, keep in zip(live, slots, keeps):
            c = clips[i]
            planes = None
            if read_back:
                w, _cw, h, _bd, sx, sy = wire_header(c.wire(f))
                planes = decoders[i].read_slot(-1, w, h, sx, sy)
            if keep:
                state[i][f] = slot
            for r in c.refs[f]:
                if c.last_use[r] == f:
                    decoders[i].release(state[i].pop(r))
            out.append((i, slot, planes))
        yield f, out


class Server:
    """The serving loop (av1h_server_*): streams admitted against a device
    memory budget, frames planned by host threads and decoded in continuously
    formed batches."""

    def __init__(self, device=0, max_batch=16, planner_threads=4, capture=False, hbm_budget=200 << 30,
                 nv12_pitch=0, hash_pictures=False, launchers=2, heavy_levels=0, heavy_launchers=1,
                 heavy_batch=1):
        lib = hip_lib()
        self._lib = lib
        cfg = ServerConfig(device, max_batch, planner_threads, 1 if capture else 0, 1 if hash_pictures else 0,
                           launchers, heavy_levels, heavy_launchers, heavy_batch, hbm_budget, nv12_pitch)
        h = ctypes.c_void_p()
        rc = lib.av1h_server_create(ctypes.byref(cfg), ctypes.byref(h))
        if rc != AV1H_OK:
            raise HipError(f"av1h_server_create failed ({rc})")
        self._h = h

    @staticmethod
    def now_ms():
        return hip_lib().av1h_server_now_ms()

    def open_stream(self, width, height, bit_depth, ss_x=1, ss_y=1, pool_slots=8):
        sid = self._lib.av1h_server_open_stream(self._h, width, height, bit_depth, ss_x, ss_y, pool_slots)
        if sid < 0:
            raise HipError(f"open_stream failed ({sid})")
        return sid

    def _rel(self, release):
        release = list(release or [])
        return (ctypes.c_int32 * max(1, len(release)))(*release), len(release)

    def submit_wire(self, stream, wire, reserved=-1, release=(), arrival_ms=None, tag=0):
        arr, n = self._rel(release)
        t = self.now_ms() if arrival_ms is None else arrival_ms
        rc = self._lib.av1h_server_submit_wire(self._h, stream, wire, len(wire), reserved, arr, n, t, tag)
        if rc != AV1H_OK:
            raise HipError(f"submit_wire failed ({rc})")

    def submit_wire_ref(self, stream, wire, reserved=-1, release=(), arrival_ms=None, tag=0):
        """submit_wire without the copy: `wire` (bytes) must stay alive until the
        frame's completion has been polled."""
        arr, n = self._rel(release)
        t = self.now_ms() if arrival_ms is None else arrival_ms
        rc = self._lib.av1h_server_submit_wire_ref(self._h, stream, wire, len(wire), reserved, arr, n, t, tag)
        if rc != AV1H_OK:
            raise HipError(f"submit_wire_ref failed ({rc})")

    def submit_plan(self, stream, plan, reserved=-1, release=(), arrival_ms=None, tag=0):
        arr, n = self._rel(release)
        t = self.now_ms() if arrival_ms is None else arrival_ms
        rc = self._lib.av1h_server_submit_plan(self._h, stream, plan, reserved, arr, n, t, tag)
        if rc != AV1H_OK:
            raise HipError(f"submit_plan failed ({rc})")

    def close_stream(self, stream):
        self._lib.av1h_server_close_stream(self._h, stream)

    def poll(self, max_items=64, wait_ms=10):
        """Completions as dicts; a captured picture comes as a flat uint16 array."""
        buf = (Completion * max_items)()
        n = self._lib.av1h_server_poll(self._h, buf, max_items, wait_ms)
        out = []
        for i in range(max(n, 0)):
            c = buf[i]
            d = {k: getattr(c, k) for k, _ in Completion._fields_ if k != "picture"}
            d["picture"] = None
            if c.picture:
                d["picture"] = np.ctypeslib.as_array(c.picture, shape=(c.picture_samples,)).copy()
                self._lib.av1h_server_free_picture(c.picture)
            out.append(d)
        return out

    def launcher_tuning(self, index):
        """Launch shapes of launcher `index` (heavy-frame launchers first)."""
        return _get_tuning(self._lib, self._launcher(index))

    def set_launcher_tuning(self, index, values):
        """Sets launch shapes of launcher `index`; before frames are submitted."""
        _set_tuning(self._lib, self._launcher(index), values)

    def _launcher(self, index):
        h = self._lib.av1h_server_launcher(self._h, index)
        if not h:
            raise HipError(f"no launcher {index}")
        return h

    def stats(self):
        st = ServerStats()
        self._lib.av1h_server_get_stats(self._h, ctypes.byref(st))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}

    def error(self):
        return self._lib.av1h_server_last_error(self._h).decode(errors="replace")

    def close(self):
        if getattr(self, "_h", None):
            self._lib.av1h_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def stream_bytes(width, height, bit_depth, ss_x=1, ss_y=1, pool_slots=8, nv12_pitch=0):
    return hip_lib().av1h_stream_bytes(width, height, bit_depth, ss_x, ss_y, pool_slots, nv12_pitch)


def _mix64(i):
    z = (i + 1) * np.uint64(0x9E3779B97F4A7C15)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def picture_hash(planes):
    """The device picture checksum (av1h_set_output_hash) of host planes."""
    with np.errstate(over="ignore"):
        v = np.concatenate([np.ascontiguousarray(p).reshape(-1) for p in planes]).astype(np.uint64)
        i = np.arange(v.size, dtype=np.uint64)
        return int(np.sum(v * _mix64(i), dtype=np.uint64))


def slot_schedule(clip, pool_slots=32):
    """The slots a decoder that always takes the lowest free slot uses for a
    clip: (slot per frame or -1, slots released after each frame, reference
    slots per frame)."""
    used = set()
    slot_of = {}
    keep_slot, releases, ref_slots = [], [], []
    for f in range(clip.frames):
        ref_slots.append([slot_of[r] for r in clip.refs[f]])
        if clip.last_use[f] > f:
            s = min(i for i in range(pool_slots) if i not in used)
            used.add(s)
            slot_of[f] = s
            keep_slot.append(s)
        else:
            keep_slot.append(-1)
        rel = []
        for r in clip.refs[f]:
            if clip.last_use[r] == f:
                s = slot_of.pop(r)
                used.discard(s)
                rel.append(s)
        releases.append(rel)
    return keep_slot, releases, ref_slots


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
