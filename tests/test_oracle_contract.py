"""The golden cases that pin the contract inputs really exercise them.

The cases with availability left to the decoder (have_above_right /
have_below_left = -1, engine.cpp:779-793) and with raw residual samples
(residual_provided, av1gpu_kernels.h:933-937) are held to the reference's
hashes by test_oracle_golden.py and test_gpu_golden.py. Here the same wire
frames are altered on the host - every -1 turned into 0, or the residual pool
zeroed - and the oracle's picture must then differ from the golden one: a
decoder that ignored the flag or the pool would not pass those cases by
accident.
"""
import numpy as np
import pytest

import av1ongpu_amd as pkg
from golden_cases import load_cases
from oracle_lib import OracleDecoder

BLOCK_INTS = 69          # int32 fields of av1h_block (include/av1hip.h)
AR, BL = 11, 12          # have_above_right, have_below_left
PLANE, W = 4, 2


def _arrays(wire):
    """(offset, count) of the 13 wire arrays (av1h_wire_parse's walk)."""
    hsize = int(np.frombuffer(wire, dtype=np.uint32, count=1, offset=12)[0])
    pos = (16 + hsize + 7) & ~7
    counts = np.frombuffer(wire, dtype=np.int64, count=13, offset=pos)
    pos += 13 * 8
    sizes = [BLOCK_INTS * 4, None, None, 4, 4, 2, 1, 1, None, None, None, None, None]
    out = []
    for i, n in enumerate(counts):
        out.append((pos, int(n)))
        if i >= 5:
            break
        if sizes[i] is None:
            # transforms: 13 int32; obmc: 10 int32
            sizes[i] = {1: 13 * 4, 2: 10 * 4}[i]
        pos += (int(n) * sizes[i] + 7) & ~7
    return out


def _case(workload, seed):
    for c in load_cases():
        if c["workload"] == workload and c["seed"] == seed:
            return c
    raise LookupError((workload, seed))


def _yuv16(wire):
    _slot, planes = OracleDecoder().decode_wire(wire, keep=False)
    return f"{pkg.fnv1a64(pkg.yuv16_bytes(planes)):016x}"


def test_derived_availability_matters():
    case = _case("c1_mini", 31)
    clip = pkg.Clip(case["workload"], case["seed"], **case["overrides"])
    wire = bytearray(clip.wire(0))
    (boff, nb), = _arrays(bytes(wire))[:1]
    blocks = np.frombuffer(wire, dtype=np.int32, count=nb * BLOCK_INTS, offset=boff).reshape(nb, BLOCK_INTS)
    assert set(np.unique(blocks[:, PLANE])) <= {0, 1, 2}
    assert set(np.unique(blocks[:, W])) <= {4, 8, 16, 32, 64}
    auto = (blocks[:, AR] == -1) | (blocks[:, BL] == -1)
    assert auto.sum() > nb // 2, "the case leaves most flags to the decoder"
    assert _yuv16(bytes(wire)) == case["frames"][0]["yuv16"]
    blocks = blocks.copy()
    blocks[blocks[:, AR] == -1, AR] = 0
    blocks[blocks[:, BL] == -1, BL] = 0
    wire[boff:boff + blocks.nbytes] = blocks.tobytes()
    assert _yuv16(bytes(wire)) != case["frames"][0]["yuv16"]


@pytest.mark.parametrize("workload,seed", [("c1_mini", 34), ("c2_mini", 33)])
def test_raw_residual_matters(workload, seed):
    case = _case(workload, seed)
    clip = pkg.Clip(case["workload"], case["seed"], **case["overrides"])
    wire = bytearray(clip.wire(0))
    roff, nres = _arrays(bytes(wire))[4]
    assert nres > 1000, "the key frame carries raw residual samples"
    assert _yuv16(bytes(wire)) == case["frames"][0]["yuv16"]
    wire[roff:roff + 4 * nres] = bytes(4 * nres)
    assert _yuv16(bytes(wire)) != case["frames"][0]["yuv16"]
