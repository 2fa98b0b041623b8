"""The committed golden fixtures (tests/golden/golden.json, written from the
reference's own CPU backend by tests/golden/make_golden.py) and a runner that
decodes one case through any decoder object and returns the same hashes."""
import json
import os

import av1ongpu_amd as pkg

MATRICES = [(m, r) for m in range(4) for r in range(2)]  # matrix major, as make_golden.py
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden.json")


def load_cases():
    with open(GOLDEN) as fh:
        return json.load(fh)["cases"]


def case_id(c):
    ov = ",".join(f"{k}={v}" for k, v in sorted(c["overrides"].items()))
    return f"{c['workload']}-s{c['seed']}" + (f"-{ov}" if ov else "")


def slot_view_nv12(buf, height, coded_width, pitch):
    """The bytes of an NV12 pack that a slot pack is compared on: the rows of
    the picture, the columns of the coded width. The reference CPU backend
    packs a slot through `gather` (cpu_backend.cpp:557-587), which copies the
    coded width of each row (`layout_.plane_w`) into a display sized frame, so
    on a superres picture its columns past the coded width are not the
    picture; its CUDA backend packs the display width (cuda_backend.cpp:
    2038-2046). The upscaled columns a slot holds are still pinned: every
    inter frame that predicts from the slot reads them."""
    ch, ccw = (height + 1) // 2, (coded_width + 1) // 2
    y = buf[:height * pitch].reshape(height, pitch)[:, :coded_width]
    uv = buf[height * pitch:(height + ch) * pitch].reshape(ch, pitch)[:, :2 * ccw]
    return y.tobytes() + uv.tobytes()


def slot_view_rgb(buf, width, height, coded_width):
    """RGB24 counterpart of slot_view_nv12."""
    return buf.reshape(height, width, 3)[:, :coded_width].tobytes()


def decode_hashes(dec, case):
    """Per frame {wire, yuv16, nv12, rgb, slot} of `dec` on the case's clip."""
    clip = pkg.Clip(case["workload"], case["seed"], **case["overrides"])
    out = []
    slot_of = {}
    for f in range(clip.frames):
        wire = pkg.wire_set_slots(clip.wire(f), [slot_of[r] for r in clip.refs[f]])
        w, cw, h, _bd, _sx, _sy = pkg.wire_header(wire)
        cw = cw if 0 < cw < w else w
        keep = clip.last_use[f] > f
        slot, planes = dec.decode_wire(wire, keep=keep)
        if keep:
            slot_of[f] = slot
        rec = {"wire": f"{pkg.fnv1a64(wire):016x}",
               "yuv16": f"{pkg.fnv1a64(pkg.yuv16_bytes(planes)):016x}",
               "nv12": f"{pkg.fnv1a64(dec.to_nv12(-1, case['nv12_pitch'], w, h).tobytes()):016x}",
               "rgb": f"{pkg.fnv1a64(dec.to_rgb(-1, 1, 0, w, h).tobytes()):016x}",
               "rgb_all": [f"{pkg.fnv1a64(dec.to_rgb(-1, m, r, w, h).tobytes()):016x}" for m, r in MATRICES],
               "slot": slot}
        # every held slot through the slot rule: the latest reads post-grain,
        # older ones their pre-grain reference
        pitch = case["nv12_pitch"]
        rec["slots"] = [[s, f"{pkg.fnv1a64(slot_view_nv12(dec.to_nv12(s, pitch, w, h), h, cw, pitch)):016x}",
                         f"{pkg.fnv1a64(slot_view_rgb(dec.to_rgb(s, 1, 0, w, h), w, h, cw)):016x}"]
                        for s in sorted(slot_of.values())]
        out.append(rec)
        for r in clip.refs[f]:
            if clip.last_use[r] == f:
                dec.release(slot_of.pop(r))
    return out


def compare(case, got):
    exp = case["frames"]
    assert len(got) == len(exp), f"{case_id(case)}: {len(got)} frames, golden has {len(exp)}"
    for f, (g, e) in enumerate(zip(got, exp)):
        for key in ("wire", "yuv16", "nv12", "rgb", "slot", "rgb_all", "slots"):
            assert g[key] == e[key], f"{case_id(case)} frame {f}: {key} {g[key]} != golden {e[key]}"
