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
  rgb_all - the same for every matrix (BT.601, BT.709, BT.2020, identity)
           and range (limited, full), in that order, matrix major
  slots  - for every slot held after the frame: [slot, NV12 hash, RGB hash]
           of the backend's own to_nv12 / to_rgb of that slot, so the latest
           slot packs the post-grain picture and older ones their pre-grain
           reference (cpu_backend.cpp:557-587), over the columns of the coded
           width (tests/golden_cases.py slot_view_nv12 says why)
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
from oracle_lib import RefDecoder, ref_available, ref_dequant_check  # noqa: E402
from golden_cases import slot_view_nv12, slot_view_rgb  # noqa: E402

CASES = [
    ("c1_mini", 1, {}), ("c1_mini", 2, {}), ("c1_mini", 3, {}),
    ("c1_mini", 11, {"key_interval": 1, "frames": 3}),
    ("c2_mini", 1, {}), ("c2_mini", 2, {}),
    ("c2_mini", 7, {"cdef": 0, "lr": 0, "deblock_lo": 0, "deblock_hi": 0, "grain": 0}),
    ("contract_mini", 1, {}), ("contract_mini", 2, {}),
    ("tools_mini", 1, {}), ("tools_mini", 2, {}),
    # availability left to the decoder (-1) and raw residual samples
    # (residual_provided): engine.cpp:779-793, av1gpu_kernels.h:933-937
    ("c1_mini", 31, {"avail_auto": 1.0}),
    ("contract_mini", 32, {"avail_auto": 0.5}),
    ("c2_mini", 33, {"avail_auto": 0.5, "residual_frac": 0.15}),
    ("c1_mini", 34, {"residual_frac": 0.3}),
    ("tools_mini", 35, {"residual_frac": 0.2, "avail_auto": 0.5}),
    # picture sizes that are not multiples of 8 (the measured 426x240 and
    # 854x480 streams among them), odd sizes and a ragged superres width
    ("c1_mini", 41, {"width": 426, "height": 240, "frames": 3}),
    ("c2_mini", 42, {"width": 854, "height": 480, "coded_width": 640, "frames": 3}),
    ("c2_mini", 43, {"width": 250, "height": 138, "coded_width": 188, "frames": 4}),
    ("c1_mini", 44, {"width": 173, "height": 99, "frames": 3}),
    # the restoration geometries of tests/test_pipeline.cpp:515-530: partial
    # stripes, a last unit taking the remainder, one 128 unit on a wider
    # picture, pictures smaller than a unit
    ("c1_mini", 51, {"width": 128, "coded_width": 112, "height": 96, "frames": 2}),
    ("c1_mini", 52, {"width": 128, "height": 100, "frames": 2, "lr": 2}),
    ("c1_mini", 53, {"width": 128, "height": 68, "frames": 2}),
    ("c1_mini", 54, {"width": 120, "height": 68, "frames": 2, "lr": 2}),
    ("c1_mini", 55, {"width": 160, "height": 96, "frames": 2, "lr_log2": 7}),
    ("c1_mini", 56, {"width": 96, "height": 72, "frames": 2, "lr": 2}),
    ("c1_mini", 57, {"width": 32, "height": 24, "frames": 2}),
    ("c1_mini", 58, {"width": 16, "height": 16, "frames": 2, "lr": 2}),
    # coded levels with their dequantisers (FrameSyntax::levels / qm_pool,
    # av1gpu.hpp:509-517): the coefficient raster the frames also carry is
    # checked against the reference's own dequantisation of the levels
    # (av1gpu_kernels.h:1024-1083) when the fixture is written
    ("c1_mini", 71, {"levels": 1}),
    ("c2_mini", 72, {"levels": 1}),
    ("contract_mini", 73, {"levels": 1}),
    ("tools_mini", 74, {"levels": 1, "residual_frac": 0.2}),
]

MATRICES = [(m, r) for m in range(4) for r in range(2)]


def nv12_pitch(width):
    return ((width + 63) // 64) * 64 + 64


def run_case(workload, seed, overrides, serial):
    clip = pkg.Clip(workload, seed, **overrides)
    dec = RefDecoder(serial=serial)
    slot_of = {}
    frames = []
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
               "nv12": f"{pkg.fnv1a64(dec.last_nv12(nv12_pitch(w)).tobytes()):016x}",
               "rgb": f"{pkg.fnv1a64(dec.last_rgb(1, 0).tobytes()):016x}",
               "rgb_all": [f"{pkg.fnv1a64(dec.last_rgb(m, r).tobytes()):016x}" for m, r in MATRICES],
               "waves": dec.waves(), "slot": slot}
        if overrides.get("levels"):
            diff, coded = ref_dequant_check(wire)
            if diff != 0:
                sys.exit(f"{workload} seed {seed} frame {f}: the raster differs from the reference's "
                         f"dequantisation of the levels in {diff} entries")
            rec["dequantised"] = coded
        held = sorted(slot_of.values())
        pitch = nv12_pitch(w)
        rec["slots"] = [[s, f"{pkg.fnv1a64(slot_view_nv12(dec.slot_nv12(s, pitch), h, cw, pitch)):016x}",
                         f"{pkg.fnv1a64(slot_view_rgb(dec.slot_rgb(s, 1, 0), w, h, cw)):016x}"] for s in held]
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
