"""Coded levels (FrameSyntax::levels / qm_pool, reference av1gpu.hpp:509-517).

The golden cases with levels were written after checking every frame's
coefficient raster against the reference's own dequantisation of its levels
(av1g_dequant_one, kernels/av1gpu_kernels.h:1024-1083, run from oracle/_ref
by tests/golden/make_golden.py). Here the oracle's restatement of that
dequantisation is held to the same rasters, transform count included; the
scans it generates (diagonal, row, column) are checked against their
definition on small rasters; and zeroed levels must change the raster, so the
check is not vacuous."""
import numpy as np
import pytest

import av1ongpu_amd as pkg
from golden_cases import load_cases
from oracle_lib import oracle_dequant, wire_coefficients

CASES = [c for c in load_cases() if c["overrides"].get("levels")]


@pytest.mark.parametrize("case", CASES, ids=[f"{c['workload']}-s{c['seed']}" for c in CASES])
def test_oracle_dequant_matches_the_reference_raster(case):
    clip = pkg.Clip(case["workload"], case["seed"], **case["overrides"])
    for f in range(clip.frames):
        wire = clip.wire(f)
        raster, n = oracle_dequant(wire)
        assert n == case["frames"][f]["dequantised"]
        assert np.array_equal(raster, wire_coefficients(wire)), f"frame {f}"


def test_levels_matter():
    case = CASES[0]
    clip = pkg.Clip(case["workload"], case["seed"], **case["overrides"])
    wire = bytearray(clip.wire(0))
    hsize = int(np.frombuffer(bytes(wire), dtype=np.uint32, count=1, offset=12)[0])
    pos = (16 + hsize + 7) & ~7
    counts = np.frombuffer(bytes(wire), dtype=np.int64, count=13, offset=pos)
    pos += 13 * 8
    for i, size in enumerate((69 * 4, 13 * 4, 10 * 4, 4, 4)):
        pos += (int(counts[i]) * size + 7) & ~7
    nlev = int(counts[5])
    assert nlev > 1000
    wire[pos:pos + 2 * nlev] = bytes(2 * nlev)
    raster, _ = oracle_dequant(bytes(wire))
    assert not raster.any() and wire_coefficients(bytes(wire)).any()


def _diag(w, h):
    """The diagonal scan by its definition: anti-diagonals in order; on a
    square raster alternating (the first runs from the top row down), on a
    wider one always from bottom-left up, on a taller one from top-right down."""
    out = []
    for d in range(w + h - 1):
        cells = [(r, d - r) for r in range(h) if 0 <= d - r < w]
        up = (d % 2 == 0) if w == h else (w > h)
        out += [r * w + c for r, c in (cells[::-1] if up else cells)]
    return out


def test_scan_definition_on_known_rasters():
    # Default_Scan_4x4 of the AV1 specification (a zig-zag)
    assert _diag(4, 4) == [0, 1, 4, 8, 5, 2, 3, 6, 9, 12, 13, 10, 7, 11, 14, 15]
    for w, h in ((4, 8), (8, 4), (16, 8), (8, 16), (32, 32), (16, 4)):
        s = _diag(w, h)
        assert sorted(s) == list(range(w * h))
        # every anti-diagonal is complete before the next starts
        diags = [(p // w) + (p % w) for p in s]
        assert diags == sorted(diags)
