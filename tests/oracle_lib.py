"""Test-side access to the CPU oracle (oracle/liboracle.so) and, where it was
built, the reference's own CPU path (oracle/_ref/libav1ref.so).

Test infrastructure only: the checker, never the thing measured or shipped.
"""
import ctypes
import os

import numpy as np

import av1ongpu_amd as pkg

REPO = pkg.REPO_DIR
ORACLE_SO = os.path.join(REPO, "oracle", "liboracle.so")
REF_SO = os.path.join(REPO, "oracle", "_ref", "libav1ref.so")

_oracle = None
_ref = None


def oracle_lib():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        lib = ctypes.CDLL(ORACLE_SO)
        P = ctypes.c_void_p
        lib.av1o_create.restype = P
        lib.av1o_destroy.argtypes = [P]
        lib.av1o_dequant_wire.restype = ctypes.c_int64
        lib.av1o_dequant_wire.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_int64]
        lib.av1o_last_error.restype = ctypes.c_char_p
        lib.av1o_last_error.argtypes = [P]
        lib.av1o_decode_wire.argtypes = [P, ctypes.c_char_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_int,
                                         ctypes.POINTER(pkg.Planes), ctypes.POINTER(pkg.FrameInfo),
                                         ctypes.POINTER(ctypes.c_int)]
        lib.av1o_reserve_reference.argtypes = [P]
        lib.av1o_release_reference.argtypes = [P, ctypes.c_int]
        lib.av1o_to_nv12.argtypes = [P, ctypes.c_int, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t]
        lib.av1o_to_rgb.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t]
        lib.av1o_read_slot.argtypes = [P, ctypes.c_int, ctypes.POINTER(pkg.Planes), ctypes.POINTER(pkg.FrameInfo)]
        lib.av1o_reset.argtypes = [P]
        lib.av1o_last_waves.argtypes = [P]
        lib.av1o_set_threads.argtypes = [ctypes.c_int]
        lib.av1o_cos128.argtypes = [ctypes.c_int]
        lib.av1o_tx1d.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        lib.av1o_itx_add.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_void_p, ctypes.c_int]
        lib.av1o_tx_row_shift.argtypes = [ctypes.c_int, ctypes.c_int]
        lib.av1o_superres_params.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                             ctypes.POINTER(ctypes.c_int)]
        lib.av1o_setup_shear.argtypes = [ctypes.c_void_p] + [ctypes.POINTER(ctypes.c_int)] * 4
        lib.av1o_color_matrix.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                          ctypes.c_void_p]
        lib.av1o_spec_table_ptr.restype = ctypes.POINTER(ctypes.c_int)
        lib.av1o_spec_table_ptr.argtypes = [ctypes.POINTER(ctypes.c_int)]
        lib.av1o_grain_scaling_lut.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        _oracle = lib
    return _oracle


def ref_available():
    return os.path.exists(REF_SO)


def ref_lib():
    global _ref
    if _ref is None:
        lib = ctypes.CDLL(REF_SO)
        P = ctypes.c_void_p
        lib.av1r_create.restype = P
        lib.av1r_destroy.argtypes = [P]
        lib.av1r_decode_wire.argtypes = [P, ctypes.c_char_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int, ctypes.POINTER(pkg.Planes), ctypes.POINTER(ctypes.c_int),
                                         ctypes.c_char_p, ctypes.c_int]
        lib.av1r_reserve.argtypes = [P]
        lib.av1r_release.argtypes = [P, ctypes.c_int]
        lib.av1r_waves.argtypes = [P]
        lib.av1r_last_nv12.restype = ctypes.c_int64
        lib.av1r_last_nv12.argtypes = [P, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_int64]
        lib.av1r_last_rgb.restype = ctypes.c_int64
        lib.av1r_last_rgb.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int64]
        lib.av1r_slot_nv12.restype = ctypes.c_int64
        lib.av1r_slot_nv12.argtypes = [P, ctypes.c_int, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_int64]
        lib.av1r_slot_rgb.restype = ctypes.c_int64
        lib.av1r_slot_rgb.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int64]
        lib.av1r_dequant_check.restype = ctypes.c_int64
        lib.av1r_dequant_check.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_int64)]
        _ref = lib
    return _ref


def oracle_dequant(wire):
    """The oracle's dequantisation of a wire frame's levels: (raster int32,
    transforms dequantised), or (None, -1) without levels."""
    hsize = int(np.frombuffer(wire, dtype=np.uint32, count=1, offset=12)[0])
    pos = (16 + hsize + 7) & ~7
    ncoef = int(np.frombuffer(wire, dtype=np.int64, count=13, offset=pos)[3])
    out = np.zeros(max(ncoef, 1), np.int32)
    n = oracle_lib().av1o_dequant_wire(wire, len(wire), out.ctypes.data, out.size)
    return (out[:ncoef], int(n)) if n >= 0 else (None, int(n))


def wire_coefficients(wire):
    """The coefficient raster a wire frame carries."""
    hsize = int(np.frombuffer(wire, dtype=np.uint32, count=1, offset=12)[0])
    pos = (16 + hsize + 7) & ~7
    counts = np.frombuffer(wire, dtype=np.int64, count=13, offset=pos)
    pos += 13 * 8
    sizes = [69 * 4, 13 * 4, 10 * 4]
    for i in range(3):
        pos += (int(counts[i]) * sizes[i] + 7) & ~7
    return np.frombuffer(wire, dtype=np.int32, count=int(counts[3]), offset=pos)


def ref_dequant_check(wire):
    """(raster entries differing from the reference's dequantisation of the
    frame's levels, transforms dequantised); (-1, 0) without levels."""
    n = ctypes.c_int64(0)
    d = ref_lib().av1r_dequant_check(wire, len(wire), ctypes.byref(n))
    return int(d), int(n.value)


class OracleDecoder:
    """The CPU oracle behind the same calls HipDecoder offers."""

    def __init__(self):
        self._lib = oracle_lib()
        self._h = self._lib.av1o_create()

    def error(self):
        return self._lib.av1o_last_error(self._h).decode(errors="replace")

    def decode_wire(self, wire, keep=True, reserved=-1, read_back=True):
        w, _cw, h, _bd, sx, sy = pkg.wire_header(wire)
        arrs, planes = pkg.alloc_planes(w, h, sx, sy) if read_back else (None, None)
        slot = ctypes.c_int(-1)
        rc = self._lib.av1o_decode_wire(self._h, wire, len(wire), 1 if keep else 0, reserved,
                                        ctypes.byref(planes) if read_back else None, None, ctypes.byref(slot))
        if rc != 0:
            raise RuntimeError(f"oracle decode failed ({rc}): {self.error()}")
        return slot.value, arrs

    def reserve(self):
        return self._lib.av1o_reserve_reference(self._h)

    def release(self, slot):
        self._lib.av1o_release_reference(self._h, slot)

    def waves(self):
        return self._lib.av1o_last_waves(self._h)

    def to_nv12(self, slot, pitch, width, height):
        n = pkg.nv12_size(width, height, 1, 1, pitch)
        buf = np.zeros(max(n, 1), dtype=np.uint8)
        rc = self._lib.av1o_to_nv12(self._h, slot, pitch, buf.ctypes.data, buf.nbytes)
        if rc != 0:
            raise RuntimeError(self.error())
        return buf[:n]

    def to_rgb(self, slot, matrix, rng, width, height):
        buf = np.zeros(width * height * 3, dtype=np.uint8)
        rc = self._lib.av1o_to_rgb(self._h, slot, matrix, rng, buf.ctypes.data, buf.nbytes)
        if rc != 0:
            raise RuntimeError(self.error())
        return buf

    def read_slot(self, slot, width, height, ss_x=1, ss_y=1):
        arrs, planes = pkg.alloc_planes(width, height, ss_x, ss_y)
        rc = self._lib.av1o_read_slot(self._h, slot, ctypes.byref(planes), None)
        if rc != 0:
            raise RuntimeError(self.error())
        return arrs

    def __del__(self):
        if getattr(self, "_h", None):
            self._lib.av1o_destroy(self._h)
            self._h = None


class RefDecoder:
    """The reference's CPU backend (oracle/_ref), same calls."""

    def __init__(self, serial=False):
        self._lib = ref_lib()
        self._h = self._lib.av1r_create()
        self.serial = serial

    def decode_wire(self, wire, keep=True, reserved=-1, read_back=True):
        w, _cw, h, _bd, sx, sy = pkg.wire_header(wire)
        arrs, planes = pkg.alloc_planes(w, h, sx, sy)
        slot = ctypes.c_int(-1)
        err = ctypes.create_string_buffer(256)
        rc = self._lib.av1r_decode_wire(self._h, wire, len(wire), 1 if keep else 0, reserved,
                                        1 if self.serial else 0, ctypes.byref(planes), ctypes.byref(slot), err, 256)
        if rc != 0:
            raise RuntimeError("reference decode failed: " + err.value.decode())
        return slot.value, arrs

    def release(self, slot):
        self._lib.av1r_release(self._h, slot)

    def waves(self):
        return self._lib.av1r_waves(self._h)

    def last_nv12(self, pitch):
        n = self._lib.av1r_last_nv12(self._h, pitch, None, 0)
        if n < 0:
            return None
        buf = np.zeros(n, dtype=np.uint8)
        self._lib.av1r_last_nv12(self._h, pitch, buf.ctypes.data, n)
        return buf

    def last_rgb(self, matrix, rng):
        n = self._lib.av1r_last_rgb(self._h, matrix, rng, None, 0)
        buf = np.zeros(n, dtype=np.uint8)
        self._lib.av1r_last_rgb(self._h, matrix, rng, buf.ctypes.data, n)
        return buf

    def slot_nv12(self, slot, pitch):
        n = self._lib.av1r_slot_nv12(self._h, slot, pitch, None, 0)
        if n < 0:
            return None
        buf = np.zeros(n, dtype=np.uint8)
        self._lib.av1r_slot_nv12(self._h, slot, pitch, buf.ctypes.data, n)
        return buf

    def slot_rgb(self, slot, matrix, rng):
        n = self._lib.av1r_slot_rgb(self._h, slot, matrix, rng, None, 0)
        if n < 0:
            return None
        buf = np.zeros(n, dtype=np.uint8)
        self._lib.av1r_slot_rgb(self._h, slot, matrix, rng, buf.ctypes.data, n)
        return buf

    def __del__(self):
        if getattr(self, "_h", None):
            self._lib.av1r_destroy(self._h)
            self._h = None
