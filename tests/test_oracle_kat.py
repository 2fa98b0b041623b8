"""Known-answer and closed-form checks of the oracle's single functions and of
the specification tables both the oracle and the HIP backend carry.

What pins what (each check is chosen so that a dropped term, a wrong sign or a
swapped index fails it):
  * tables: cos128 = round(4096 cos(a pi / 128)); Mode_To_Angle,
    Dr_Intra_Derivative, Sm_Weights and filter invariants with the values of
    the reference's own table test (tests/test_kernels.cpp:98-203); the HIP
    copy equal to the oracle's entry for entry;
  * 1D inverse transforms against the real-valued transforms they round:
    DCT-II inverse (DC weighted 1/sqrt 2), ADST4 = (2 sqrt2 / 3) sin(pi (j+1)
    (2k+1) / 9), ADST8/16 = DST-IV, identity gains sqrt2, 2, 2 sqrt2, 4;
  * 2D: a lone DC coefficient gives a flat block for every size
    (test_kernels.cpp:211-223); FLIPADST is the mirror image of ADST;
  * grain scaling LUT passes through its points and is linear between them;
  * superres positions land on (x + 1/2) dw / uw - 1/2 across the row;
  * colour matrices against the floating point BT.709/2020 equations;
  * FNV-1a-64 against its published test values.
"""
import ctypes
import math
import os
import re

import numpy as np
import pytest

from oracle_lib import oracle_lib

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
L = oracle_lib()

OFF = {}
with open(os.path.join(REPO, "oracle", "av1o_spec_tables.h")) as fh:
    for m in re.finditer(r"#define AV1O_OFF_(\w+) (\d+)", fh.read()):
        OFF[m.group(1)] = int(m.group(2))


def table():
    n = ctypes.c_int()
    p = L.av1o_spec_table_ptr(ctypes.byref(n))
    return np.ctypeslib.as_array(p, shape=(n.value,)).copy()


def hip_table():
    """The table compiled into the HIP backend, parsed from its header."""
    with open(os.path.join(REPO, "av1ongpu_amd", "csrc", "spec_tables.h")) as fh:
        text = fh.read()
    body = text[text.index("{", text.index("av1t_spec_table")) + 1:]
    body = body[:body.index("};")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    return np.array([int(v) for v in body.replace("\n", " ").split(",") if v.strip()], dtype=np.int64)


def tx1d(v, log2n, kind, r=30):
    a = (ctypes.c_int32 * len(v))(*[int(x) for x in v])
    L.av1o_tx1d(a, log2n, kind, r)
    return np.array(a[:], dtype=np.int64)


# ---------------------------------------------------------------- tables ----

def test_hip_table_equals_oracle_table():
    t, h = table(), hip_table()
    assert t.shape == h.shape
    assert np.array_equal(t, h)


def test_cos128_closed_form():
    t = table()
    c = t[OFF["COS128"]:OFF["COS128"] + 65]
    assert c[0] == 4096 and c[32] == 2896
    for a in range(65):
        assert c[a] == round(4096 * math.cos(a * math.pi / 128)), a
        assert L.av1o_cos128(a) == c[a]
    # sin128(a) = cos128(64 - a) and cosine symmetry beyond 64
    for a in range(65, 128):
        assert L.av1o_cos128(a) == -c[128 - a]


def test_mode_to_angle_and_dr_derivative():
    t = table()
    assert list(t[OFF["MODE_TO_ANGLE"]:OFF["MODE_TO_ANGLE"] + 13]) == [0, 90, 180, 45, 135, 113, 157, 203, 67,
                                                                      0, 0, 0, 0]
    angles = [3, 6, 9, 14, 17, 20, 23, 26, 29, 32, 36, 39, 42, 45, 48, 51, 54, 58, 61, 64, 67, 70, 73, 76, 81,
              84, 87]
    values = [1023, 547, 372, 273, 215, 178, 151, 132, 116, 102, 90, 80, 71, 64, 57, 51, 45, 40, 35, 31, 27, 23,
              19, 15, 11, 7, 3]
    d = t[OFF["DR_DERIV"]:OFF["DR_DERIV"] + 90]
    for a, v in zip(angles, values):
        assert d[a] == v
    assert all(d[a] == 0 for a in range(90) if a not in angles)


def test_smooth_weights():
    t = table()
    base = OFF["SM_WEIGHTS"]
    for n in (4, 8, 16, 32, 64):
        w = t[base + n - 4: base + 2 * n - 4]
        assert w[0] >= 250 and w[-1] < w[0]
        assert all(w[i] <= w[i - 1] for i in range(1, n)) and all(w <= 256)
    assert list(t[base:base + 4]) == [255, 149, 85, 64]


def test_edge_and_filter_intra_kernels():
    t = table()
    k = t[OFF["EDGE_KERNEL"]:OFF["EDGE_KERNEL"] + 15].reshape(3, 5)
    assert (k.sum(axis=1) == 16).all() and (k == k[:, ::-1]).all()
    f = t[OFF["FILTER_INTRA"]:OFF["FILTER_INTRA"] + 280].reshape(5, 8, 7)
    assert (f.sum(axis=2) == 16).all()


def test_subpel_filters():
    t = table()
    f = t[OFF["SUBPEL"]:OFF["SUBPEL"] + 768].reshape(6, 16, 8)
    assert (f.sum(axis=2) == 128).all()
    copy = np.zeros(8, dtype=np.int64)
    copy[3] = 128
    for row in range(6):
        assert np.array_equal(f[row, 0], copy)
        assert np.array_equal(f[row, 8], f[row, 8][::-1])
    # bilinear: taps 128 - 8p and 8p at the centre
    for ph in range(16):
        assert f[3, ph, 3] == 128 - 8 * ph and f[3, ph, 4] == 8 * ph


def test_upscale_and_warp_filters_have_unit_gain():
    t = table()
    up = t[OFF["UPSCALE"]:OFF["UPSCALE"] + 512].reshape(64, 8)
    assert (up.sum(axis=1) == 128).all()
    wp = t[OFF["WARP"]:OFF["WARP"] + 193 * 8].reshape(193, 8)
    assert (wp.sum(axis=1) == 128).all()


def test_div_lut():
    """Div_Lut[i] = round(2^14 * 256 / (256 + i))"""
    t = table()
    d = t[OFF["DIV_LUT"]:OFF["DIV_LUT"] + 257]
    for i in range(257):
        assert abs(d[i] - (1 << 14) * 256 / (256 + i)) <= 0.5 + 1e-9, i


def test_gaussian_sequence_shape():
    t = table()
    g = t[OFF["GAUSSIAN"]:OFF["GAUSSIAN"] + 2048]
    # 12 bit Gaussian samples: zero mean, standard deviation about 512
    assert abs(g.mean()) < 16 and 450 < g.std() < 560 and np.abs(g).max() < 2048


# ---------------------------------------------------------- transforms ----

def _dct_matrix(n):
    j = np.arange(n)[:, None]
    k = np.arange(n)[None, :]
    c = np.cos(np.pi * (2 * j + 1) * k / (2 * n))
    c[:, 0] = 1 / np.sqrt(2)
    return c


# the integer butterflies round at every stage: the tolerance grows with the
# number of stages, far below what a wrong term or sign changes (hundreds)
DCT_TOL = {4: 2.5, 8: 4.5, 16: 8.5, 32: 11.0, 64: 22.0}


@pytest.mark.parametrize("log2n", [2, 3, 4, 5, 6])
def test_idct_is_the_rounded_dct(log2n):
    n = 1 << log2n
    c = _dct_matrix(n)
    for seed in range(40):
        x = np.random.default_rng(seed).integers(-4000, 4000, n)
        err = np.abs(tx1d(x, log2n, 0) - c @ x).max()
        assert err <= DCT_TOL[n], (n, seed, err)


@pytest.mark.parametrize("log2n", [2, 3, 4])
def test_iadst_is_the_rounded_dst(log2n):
    n = 1 << log2n
    j = np.arange(n)[:, None]
    k = np.arange(n)[None, :]
    if n == 4:
        s = np.sin(np.pi * (j + 1) * (2 * k + 1) / 9) * (2 * np.sqrt(2) / 3)
    else:
        s = np.sin(np.pi * (2 * j + 1) * (2 * k + 1) / (4 * n))
    for seed in range(40):
        x = np.random.default_rng(100 + seed).integers(-4000, 4000, n)
        err = np.abs(tx1d(x, log2n, 1) - s @ x).max()
        assert err <= {4: 2.0, 8: 7.5, 16: 8.5}[n], (n, seed, err)


@pytest.mark.parametrize("log2n,gain", [(2, math.sqrt(2)), (3, 2.0), (4, 2 * math.sqrt(2)), (5, 4.0)])
def test_identity_gain(log2n, gain):
    n = 1 << log2n
    x = np.random.default_rng(7).integers(-3000, 3000, n)
    assert np.abs(tx1d(x, log2n, 2) - gain * x).max() <= 1.0


def itx_add(coef, w, h, tx_type, bd=10, base=512, lossless=0):
    c = np.zeros(min(w, 32) * min(h, 32), dtype=np.int32)
    c[: len(coef)] = coef
    dst = np.full((h, w), base, dtype=np.uint16)
    L.av1o_itx_add(c.ctypes.data, w, h, tx_type, lossless, bd, dst.ctypes.data, w)
    return dst.astype(np.int64)


SIZES = [(w, h) for w in (4, 8, 16, 32, 64) for h in (4, 8, 16, 32, 64) if max(w, h) <= 4 * min(w, h)]


@pytest.mark.parametrize("w,h", SIZES)
def test_dc_gives_flat_block(w, h):
    out = itx_add([300], w, h, 0)
    assert (out == out[0, 0]).all() and out[0, 0] != 512


def test_flipadst_mirrors_adst():
    coef = np.random.default_rng(3).integers(-200, 200, 64)
    # tx types: 1 ADST_DCT, 4 FLIPADST_DCT (vertical flip); 2 DCT_ADST, 5 DCT_FLIPADST (horizontal)
    a = itx_add(coef, 8, 8, 1)
    assert np.array_equal(itx_add(coef, 8, 8, 4), a[::-1, :])
    b = itx_add(coef, 8, 8, 2)
    assert np.array_equal(itx_add(coef, 8, 8, 5), b[:, ::-1])


def test_identity_2d_is_a_scaled_copy():
    # IDTX 4x4 at 10 bit: row gain sqrt2 then row shift 0, column gain sqrt2 and >> 4
    coef = np.zeros(16, dtype=np.int64)
    coef[5] = 16 * 64
    out = itx_add(coef, 4, 4, 9) - 512
    assert out[1, 1] == 128 and (np.delete(out.ravel(), 5) == 0).all()


def test_row_shift_table():
    want = {(4, 4): 0, (8, 8): 1, (16, 16): 2, (32, 32): 2, (64, 64): 2, (4, 8): 0, (8, 4): 0, (8, 16): 1,
            (16, 8): 1, (16, 32): 1, (32, 16): 1, (32, 64): 1, (64, 32): 1, (4, 16): 1, (16, 4): 1, (8, 32): 2,
            (32, 8): 2, (16, 64): 2, (64, 16): 2}
    for (w, h), v in want.items():
        assert L.av1o_tx_row_shift(w, h) == v


# ------------------------------------------------------- film grain LUT ----

def test_scaling_lut_piecewise_linear():
    with open(os.path.join(REPO, "include", "av1hip.h")) as fh:
        text = fh.read()
    body = text[text.index("typedef struct av1h_grain"):]
    body = body[body.index("{") + 1: body.index("}")]
    names = []
    for decl in body.replace("\n", " ").split(";"):
        decl = decl.strip()
        if not decl:
            continue
        assert decl.startswith("int32_t"), decl
        for item in decl[len("int32_t"):].split(","):
            item = item.strip()
            m = re.match(r"(\w+)(?:\[(\d+)\])?", item)
            names.append((m.group(1), int(m.group(2) or 1)))
    size = sum(n for _, n in names)
    offs = {}
    pos = 0
    for name, n in names:
        offs[name] = pos
        pos += n
    g = np.zeros(size, dtype=np.int32)
    pts = [(0, 20), (64, 80), (128, 40), (255, 200)]
    g[offs["num_y_points"]] = len(pts)
    for i, (v, s) in enumerate(pts):
        g[offs["point_y_value"] + i] = v
        g[offs["point_y_scaling"] + i] = s
    lut = np.zeros(256, dtype=np.int32)
    L.av1o_grain_scaling_lut(g.ctypes.data, 0, lut.ctypes.data)
    for v, s in pts:
        assert lut[v] == s
    for (v0, s0), (v1, s1) in zip(pts, pts[1:]):
        for x in range(v0, v1):
            assert abs(lut[x] - (s0 + (s1 - s0) * (x - v0) / (v1 - v0))) <= 1.0


# --------------------------------------------------------- superres, colour ----

@pytest.mark.parametrize("dw,uw", [(1440, 1920), (192, 256), (96, 128), (100, 160), (1000, 1334)])
def test_superres_positions(dw, uw):
    """Output sample x is interpolated at input position sx / 2^14 (phase 0 of
    the upscale filter is a copy of sample sx >> 14, phase 32 the midpoint);
    the spec places output centres at (x + 1/2) dw / uw - 1/2, with the
    position rounded up by half a phase step for the >> 8 phase selection."""
    step, init = ctypes.c_int(), ctypes.c_int()
    L.av1o_superres_params(dw, uw, ctypes.byref(step), ctypes.byref(init))
    assert abs(step.value - dw * (1 << 14) / uw) <= 0.5
    for x in (0, 1, uw // 3, uw // 2, uw - 1):
        pos = (-(1 << 14) + init.value + x * step.value) / (1 << 14)
        want = (x + 0.5) * dw / uw - 0.5 + 0.5 / 64
        assert abs(pos - want) <= 1 / 128, (x, pos, want)


@pytest.mark.parametrize("matrix,kr,kb", [(1, 0.2126, 0.0722), (2, 0.2627, 0.0593)])
@pytest.mark.parametrize("rng", [0, 1])
def test_colour_matrix_float_equations(matrix, kr, kb, rng):
    """Q16 coefficients of the nominal (studio swing) equations: Y scaled by
    255/219, chroma by 255/224 (the reference uses these for both ranges and
    changes only the offsets, src/core/tables.cpp:481-533)."""
    m = np.zeros(8, dtype=np.int32)
    o = np.zeros(3, dtype=np.int32)
    L.av1o_color_matrix(matrix, rng, 8, m.ctypes.data, o.ctypes.data)
    kg = 1 - kr - kb
    ys, cs = 255 / 219, 255 / 224
    want = [ys, 2 * (1 - kr) * cs, ys, -2 * kb * (1 - kb) / kg * cs, -2 * kr * (1 - kr) / kg * cs, ys,
            2 * (1 - kb) * cs]
    got = m[:7] / 65536.0
    if matrix == 2:
        # the reference's BT.2020 Cb->B entry is 1.8556 (121634), not the
        # equation's 2.1418; kept as the reference has it
        assert m[6] == 121634
        got, want = got[:6], want[:6]
    assert np.abs(got - np.array(want)).max() < 2e-3
    # offsets: Y - 16 and C - 128 at limited range, C - 128 at full range, plus the rounding half
    yo, co = (16, 128) if rng == 0 else (0, 128)
    assert o[0] == 32768 - m[0] * yo - m[1] * co
    assert o[2] == 32768 - m[5] * yo - m[6] * co


def test_colour_matrix_601_is_the_709_table():
    """The reference has no BT.601 case: it falls to the BT.709 table."""
    a = np.zeros(8, dtype=np.int32)
    b = np.zeros(8, dtype=np.int32)
    o = np.zeros(3, dtype=np.int32)
    L.av1o_color_matrix(0, 0, 8, a.ctypes.data, o.ctypes.data)
    L.av1o_color_matrix(1, 0, 8, b.ctypes.data, o.ctypes.data)
    assert np.array_equal(a, b)


def test_fnv1a64_published_values():
    L.av1o_fnv1a64.restype = ctypes.c_uint64
    L.av1o_fnv1a64.argtypes = [ctypes.c_char_p, ctypes.c_size_t]
    assert L.av1o_fnv1a64(b"", 0) == 0xCBF29CE484222325
    assert L.av1o_fnv1a64(b"a", 1) == 0xAF63DC4C8601EC8C
    assert L.av1o_fnv1a64(b"foobar", 6) == 0x85944171F73967E8
