#!/usr/bin/env python3
"""Writes tests/golden/golden.json: per frame FNV-1a-64 hashes of the reference
CPU backend's output on the seeded synthetic clips the parity tests use.

The reference's own CPU path (src/backends/cpu/cpu_backend.cpp with
kernels/av1gpu_kernels.h, src/core/tables.cpp, src/core/grain.cpp,
src/core/factory.cpp) is built from its sources by `make -C oracle ref` into
oracle/_ref/libav1ref.so; oracle/ref_harness.cpp feeds it the wire frames.
For every frame this records:
  wire   - hash of the wire bytes (the input, with slots as decoded)
  yuv16  - hash of frame_to_yuv16 of the decoded display picture
           (factory.cpp:194-212; hash as tools/av1gpu_decode.cpp:133-139)
  nv12   - hash of frame_to_nv12 of that picture at the case's pitch
  rgb    - hash of the reference's YUV->RGB of that picture (BT.709, limited)
  waves  - the reference scheduler's wave count for the frame
and checks that a one-block-per-wave schedule gives the same pictures (the
output does not depend on the wave grouping, only on the dependency order).

Run in the development container (needs /root/reference):
    make -C oracle ref && python tests/golden/make_golden.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "tests"))

import av1ongpu_amd as pkg  # noqa: E402
from oracle_lib import RefDecoder, ref_available  # noqa: E402

CASES = [
    ("c1_mini", 1, {}), ("c1_mini", 2, {}), ("c1_mini", 3, {}),
    ("c1_mini", 11, {"key_interval": 1, "frames": 3}),
    ("c2_mini", 1, {}), ("c2_mini", 2, {}),
    ("c2_mini", 7, {"cdef": 0, "lr": 0, "deblock_lo": 0, "deblock_hi": 0, "grain": 0}),
    ("contract_mini", 1, {}), ("contract_mini", 2, {}),
    ("tools_mini", 1, {}), ("tools_mini", 2, {}),
]


def nv12_pitch(width):
    return ((width + 63) // 64) * 64 + 64


def run_case(workload, seed, overrides, serial):
    clip = pkg.Clip(workload, seed, **overrides)
    dec = RefDecoder(serial=serial)
    slot_of = {}
    frames = []
    for f in range(clip.frames):
        wire = pkg.wire_set_slots(clip.wire(f), [slot_of[r] for r in clip.refs[f]])
        w, _cw, h, _bd, _sx, _sy = pkg.wire_header(wire)
        keep = clip.last_use[f] > f
        slot, planes = dec.decode_wire(wire, keep=keep)
        if keep:
            slot_of[f] = slot
        rec = {"wire": f"{pkg.fnv1a64(wire):016x}",
               "yuv16": f"{pkg.fnv1a64(pkg.yuv16_bytes(planes)):016x}",
               "nv12": f"{pkg.fnv1a64(dec.last_nv12(nv12_pitch(w)).tobytes()):016x}",
               "rgb": f"{pkg.fnv1a64(dec.last_rgb(1, 0).tobytes()):016x}",
               "waves": dec.waves(), "slot": slot}
        frames.append(rec)
        for r in clip.refs[f]:
            if clip.last_use[r] == f:
                dec.release(slot_of.pop(r))
    return clip, frames


def main():
    if not ref_available():
        sys.exit("oracle/_ref/libav1ref.so is not built: make -C oracle ref")
    out = {"about": "FNV-1a-64 hashes of the reference CPU backend's output, written by "
                    "tests/golden/make_golden.py", "cases": []}
    for workload, seed, ov in CASES:
        clip, frames = run_case(workload, seed, ov, serial=False)
        _, serial = run_case(workload, seed, ov, serial=True)
        for a, b in zip(frames, serial):
            if a["yuv16"] != b["yuv16"]:
                sys.exit(f"{workload} seed {seed}: wave and serial schedules differ")
        w, cw, h, bd, sx, sy = pkg.wire_header(clip.wire(0))
        out["cases"].append({"workload": workload, "seed": seed, "overrides": ov, "width": w,
                             "coded_width": cw, "height": h, "bit_depth": bd,
                             "nv12_pitch": nv12_pitch(w), "frames": frames})
        print(f"{workload} seed {seed} {ov}: {len(frames)} frames, waves {[x['waves'] for x in frames]}")
    with open(os.path.join(HERE, "golden.json"), "w") as fh:
        json.dump(out, fh, indent=1)
        fh.write("\n")


if __name__ == "__main__":
    main()
