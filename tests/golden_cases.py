"""The committed golden fixtures (tests/golden/golden.json, written from the
reference's own CPU backend by tests/golden/make_golden.py) and a runner that
decodes one case through any decoder object and returns the same hashes."""
import json
import os

import av1ongpu_amd as pkg

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden.json")


def load_cases():
    with open(GOLDEN) as fh:
        return json.load(fh)["cases"]


def case_id(c):
    ov = ",".join(f"{k}={v}" for k, v in sorted(c["overrides"].items()))
    return f"{c['workload']}-s{c['seed']}" + (f"-{ov}" if ov else "")


def decode_hashes(dec, case):
    """Per frame {wire, yuv16, nv12, rgb, slot} of `dec` on the case's clip."""
    clip = pkg.Clip(case["workload"], case["seed"], **case["overrides"])
    out = []
    slot_of = {}
    for f in range(clip.frames):
        wire = pkg.wire_set_slots(clip.wire(f), [slot_of[r] for r in clip.refs[f]])
        w, _cw, h, _bd, _sx, _sy = pkg.wire_header(wire)
        keep = clip.last_use[f] > f
        slot, planes = dec.decode_wire(wire, keep=keep)
        if keep:
            slot_of[f] = slot
        out.append({"wire": f"{pkg.fnv1a64(wire):016x}",
                    "yuv16": f"{pkg.fnv1a64(pkg.yuv16_bytes(planes)):016x}",
                    "nv12": f"{pkg.fnv1a64(dec.to_nv12(-1, case['nv12_pitch'], w, h).tobytes()):016x}",
                    "rgb": f"{pkg.fnv1a64(dec.to_rgb(-1, 1, 0, w, h).tobytes()):016x}",
                    "slot": slot})
        for r in clip.refs[f]:
            if clip.last_use[r] == f:
                dec.release(slot_of.pop(r))
    return out


def compare(case, got):
    exp = case["frames"]
    assert len(got) == len(exp), f"{case_id(case)}: {len(got)} frames, golden has {len(exp)}"
    for f, (g, e) in enumerate(zip(got, exp)):
        for key in ("wire", "yuv16", "nv12", "rgb", "slot"):
            assert g[key] == e[key], f"{case_id(case)} frame {f}: {key} {g[key]} != golden {e[key]}"
