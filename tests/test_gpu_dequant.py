"""Device dequantisation of coded levels: a frame whose syntax carries levels
is dequantised inside the transforms' coefficient staging (no raster pool on
the device). Its pictures must equal the golden hashes of the reference's CPU
backend (test_gpu_golden.py decodes these cases with levels) and the pictures
of the same frames decoded from their coefficient raster alone; a 1D
transform type on a transform larger than 16x16, which AV1's transform sets
never code, is refused."""
import numpy as np
import pytest

import av1ongpu_amd as pkg
from golden_cases import load_cases

pytestmark = pytest.mark.gpu
CASES = [c for c in load_cases() if c["overrides"].get("levels")]


def _hashes(clip, strip):
    keep, rel, refs = pkg.slot_schedule(clip)
    dec = pkg.HipDecoder()
    dec.set_output_hash(True)
    out = []
    for f in range(clip.frames):
        wire = pkg.wire_set_slots(clip.wire(f), refs[f])
        if strip:
            wire = pkg.wire_without_levels(wire)
        dec.decode_wire(wire, keep=keep[f] >= 0, reserved=keep[f], read_back=False)
        out.append(dec.last_hash())
        for s in rel[f]:
            dec.release(s)
    dec.close()
    return out


@pytest.mark.parametrize("case", CASES, ids=[f"{c['workload']}-s{c['seed']}" for c in CASES])
def test_levels_decode_as_the_raster(case):
    clip = pkg.Clip(case["workload"], case["seed"], **case["overrides"])
    assert _hashes(clip, strip=False) == _hashes(clip, strip=True)


def test_one_dimensional_type_beyond_16_is_refused():
    clip = pkg.Clip("c2_mini", 72, levels=1, frames=1)
    wire = bytearray(clip.wire(0))
    hsize = int(np.frombuffer(bytes(wire), dtype=np.uint32, count=1, offset=12)[0])
    pos = (16 + hsize + 7) & ~7
    counts = np.frombuffer(bytes(wire), dtype=np.int64, count=13, offset=pos)
    toff = pos + 13 * 8 + ((int(counts[0]) * 69 * 4 + 7) & ~7)
    tx = np.frombuffer(bytes(wire), dtype=np.int32, count=int(counts[1]) * 13, offset=toff).reshape(-1, 13).copy()
    big = np.nonzero((tx[:, 2] == 32) & (tx[:, 3] == 32) & (tx[:, 5] > 0))[0]
    assert len(big) > 0
    tx[big[0], 6] = 10  # V_DCT
    wire[toff:toff + tx.nbytes] = tx.tobytes()
    dec = pkg.HipDecoder()
    with pytest.raises(pkg.HipError, match="1D types only up to 16x16"):
        dec.decode_wire(bytes(wire), keep=False, read_back=False)
    # the same frame without its levels decodes from the raster
    dec.decode_wire(pkg.wire_without_levels(bytes(wire)), keep=False, read_back=False)
    dec.close()
