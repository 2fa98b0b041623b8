"""Multi-stream batches: frames of several streams (one decoder context each)
decoded together through av1h_decode_planned_batch, one launch per kernel for
the whole batch and one merged dependency queue. Each stream's output must be
bit identical to decoding it alone, and to the CPU oracle.

The streams differ in size, bit depth and tools, so one batch mixes frames
with and without superres, film grain, loop restoration, intra block copy and
key frames; streams of different lengths drop out of later batches.
"""
import numpy as np
import pytest

import av1ongpu_amd as pkg
from oracle_lib import OracleDecoder

pytestmark = pytest.mark.gpu


def _alone(clip):
    return [planes for _f, _s, planes in pkg.decode_clip(pkg.HipDecoder(), clip)]


def _check(name, f, got, want):
    for p in range(3):
        if not np.array_equal(got[p], want[p]):
            bad = np.argwhere(got[p] != want[p])
            y, x = bad[0]
            raise AssertionError(f"{name} frame {f} plane {p}: {len(bad)} mismatched samples, first at ({x},{y}) "
                                 f"batched={got[p][y, x]} alone={want[p][y, x]}")


def _run_batched(clips, per_kernel_set=True):
    decs = [pkg.HipDecoder() for _ in clips]
    if not per_kernel_set:
        decs[0].set_graphs(True, per_kernel_set=False)
    got = [[] for _ in clips]
    for _f, outs in pkg.decode_clips_batched(decs, clips):
        for i, _slot, planes in outs:
            got[i].append(planes)
    return got


MIX = [("c1_mini", 1, {}), ("c2_mini", 2, {}), ("tools_mini", 1, {}), ("contract_mini", 1, {}),
       ("c1_mini", 3, dict(width=200, height=120, bit_depth=10)),
       ("c1_mini", 4, dict(key_interval=1, frames=3)),
       ("c1_mini", 8, dict(width=426, height=240, frames=3))]


def test_batch_matches_alone():
    clips = [pkg.Clip(w, s, **o) for w, s, o in MIX]
    got = _run_batched(clips)
    for c, g in zip(clips, got):
        want = _alone(c)
        assert len(g) == len(want) == c.frames
        for f, (a, b) in enumerate(zip(g, want)):
            _check(f"{c.workload} seed {c.seed}", f, a, b)


def test_batch_matches_oracle():
    clips = [pkg.Clip("c2_mini", 5), pkg.Clip("tools_mini", 2), pkg.Clip("c1_mini", 6)]
    got = _run_batched(clips)
    for c, g in zip(clips, got):
        want = [planes for _f, _s, planes in pkg.decode_clip(OracleDecoder(), c)]
        for f, (a, b) in enumerate(zip(g, want)):
            _check(f"{c.workload} seed {c.seed} (oracle)", f, a, b)


def test_batch_of_one_is_plain():
    clip = pkg.Clip("c2_mini", 3)
    got = _run_batched([clip])[0]
    for f, (a, b) in enumerate(zip(got, _alone(clip))):
        _check("c2_mini seed 3", f, a, b)


def test_batch_rejects_repeated_context():
    clip = pkg.Clip("c1_mini", 1)
    dec = pkg.HipDecoder()
    plans = [dec.plan_wire(clip.wire(0)), dec.plan_wire(clip.wire(0))]
    try:
        with pytest.raises(pkg.HipError, match="twice"):
            pkg.decode_planned_batch([dec, dec], plans)
    finally:
        for p in plans:
            dec.discard_plan(p)
    # the context still decodes afterwards
    slot, _ = dec.decode_wire(clip.wire(0), keep=False, read_back=False)
    assert slot == -1


@pytest.mark.parametrize("extra", [0, 6])
def test_decode_time_nv12(extra):
    """NV12 packed inside the decode (set_decode_nv12) equals to_nv12 after a
    plain decode, at a pitch that is a multiple of 4 and at one that is not."""
    clip = pkg.Clip("c2_mini", 4)
    w, _cw, h, _bd, _sx, _sy = pkg.wire_header(clip.wire(0))
    pitch = 2048 if extra == 0 else 2 * ((w + 1) // 2) + extra
    a, b = pkg.HipDecoder(), pkg.HipDecoder()
    b.set_decode_nv12(pitch)
    for (f, sa, _), (_, sb, _) in zip(pkg.decode_clip(a, clip, read_back=False),
                                      pkg.decode_clip(b, clip, read_back=False)):
        na = a.to_nv12(-1, pitch, w, h)
        nb = b.to_nv12(-1, pitch, w, h)
        assert bytes(na) == bytes(nb), f"frame {f}"


@pytest.mark.parametrize("per_kernel_set", [True, False], ids=["per_kernel_set", "per_size"])
def test_graph_replays_match_eager(per_kernel_set):
    """Captured launch sequences (replayed for every batch of the same shape,
    with eager decodes and synchronisations between replays) give the same
    pictures as eager launches, for batches of 1 and of several frames, with
    graphs per batch size and set of kernels with work and per batch size
    with every kernel launched."""
    clips = [pkg.Clip("c2_mini", 8), pkg.Clip("tools_mini", 3), pkg.Clip("c1_mini", 9, bit_depth=10)]
    eager = [pkg.HipDecoder() for _ in clips]
    for d in eager:
        d.set_graphs(False)
    want = [[] for _ in clips]
    for _f, outs in pkg.decode_clips_batched(eager, clips):
        for i, _s, planes in outs:
            want[i].append(planes)
    for rep in range(2):
        got = _run_batched(clips, per_kernel_set)
        for c, g, w in zip(clips, got, want):
            for f, (a, b) in enumerate(zip(g, w)):
                _check(f"graph rep {rep} {c.workload} seed {c.seed}", f, a, b)
        # an eager decode in between
        _alone(clips[0])
    for c, w in zip(clips, want):
        for f, (a, b) in enumerate(zip(_alone(c), w)):
            _check(f"graph alone {c.workload}", f, a, b)


def test_output_hash_matches_host():
    """The device picture checksum equals the host restatement over the read
    back planes, alone and in a batch."""
    clips = [pkg.Clip("c2_mini", 10), pkg.Clip("c1_mini", 11, width=426, height=240, frames=2)]
    decs = [pkg.HipDecoder() for _ in clips]
    for d in decs:
        d.set_output_hash(True)
    for _f, outs in pkg.decode_clips_batched(decs, clips):
        for i, _slot, planes in outs:
            assert decs[i].last_hash() == pkg.picture_hash(planes)
    d = pkg.HipDecoder()
    d.set_output_hash(True)
    for f, _s, planes in pkg.decode_clip(d, clips[0]):
        assert d.last_hash() == pkg.picture_hash(planes), f


def test_launch_shapes_leave_pixels_unchanged():
    """Any launch shape (grid sizes, CDEF blocks per thread) decodes the same
    pictures; invalid shapes are refused."""
    clips = [pkg.Clip("c2_mini", 12), pkg.Clip("c1_mini", 13)]
    want = _run_batched(clips)
    decs = [pkg.HipDecoder() for _ in clips]
    for shape in ({"cdef_blocks_per_thread": 4, "queue": 1, "mc": 1, "deblock_v": 1},
                  {"cdef_blocks_per_thread": 2, "queue": 12, "lr": 32, "itx_inter": 2}):
        decs[0].set_tuning(shape)
        got = [[] for _ in clips]
        for _f, outs in pkg.decode_clips_batched(decs, clips):
            for i, _s, planes in outs:
                got[i].append(planes)
        for c, g, w in zip(clips, got, want):
            for f, (a, b) in enumerate(zip(g, w)):
                _check(f"shape {shape} {c.workload}", f, a, b)
    with pytest.raises(pkg.HipError):
        decs[0].set_tuning({"cdef_blocks_per_thread": 5})


def _async_rounds(clips, launcher=None):
    """The clips decoded in lockstep with av1h_decode_planned_batch_async:
    round r is launched before round r - 1 is finished, releases of round
    r - 1 taken before round r is launched. Returns each clip's per frame
    device checksums."""
    decs = [pkg.HipDecoder() for _ in clips]
    for d in decs:
        d.set_output_hash(True)
    scheds = [pkg.slot_schedule(c) for c in clips]
    plans = [[decs[i].plan_wire(pkg.wire_set_slots(c.wire(f), scheds[i][2][f])) for f in range(c.frames)]
             for i, c in enumerate(clips)]
    lead = launcher or decs[0]
    got = [[] for _ in clips]
    n = max(c.frames for c in clips)
    live_prev = []
    for f in range(n):
        live = [i for i, c in enumerate(clips) if f < c.frames]
        for i in live_prev:
            for s in scheds[i][1][f - 1]:
                decs[i].release(s)
        slots = pkg.decode_planned_batch_async([decs[i] for i in live], [plans[i][f] for i in live],
                                               [scheds[i][0][f] >= 0 for i in live],
                                               [scheds[i][0][f] for i in live], launcher=lead)
        assert slots == [scheds[i][0][f] for i in live]
        if f > 0:
            pkg.batch_finish(lead)
            for i in live_prev:
                got[i].append(decs[i].last_hash())
        live_prev = live
    pkg.batch_finish(lead)
    for i in live_prev:
        got[i].append(decs[i].last_hash())
    assert pkg.hip_lib().av1h_batch_pending(lead.handle) == 0
    for i, c in enumerate(clips):
        for p in plans[i]:
            decs[i].discard_plan(p)
    return got


@pytest.mark.parametrize("own_launcher", [False, True], ids=["lead", "launcher"])
def test_async_batches_match_alone(own_launcher):
    clips = [pkg.Clip(w, s, **o) for w, s, o in MIX]
    launcher = pkg.HipDecoder() if own_launcher else None
    got = _async_rounds(clips, launcher)
    for c, g in zip(clips, got):
        want = [pkg.picture_hash(p) for p in _alone(c)]
        assert len(g) == len(want) == c.frames
        for f, (a, b) in enumerate(zip(g, want)):
            assert a == b, f"{c.workload} seed {c.seed} frame {f}: asynchronous batch differs from the plain decode"


def test_async_batch_holds_its_contexts():
    """While a launched batch holds a context's frame, its pictures cannot be
    read and another launcher cannot batch it; after the finish they can."""
    clip = pkg.Clip("c1_mini", 2, frames=1)
    dec, other = pkg.HipDecoder(), pkg.HipDecoder()
    plan = dec.plan_wire(clip.wire(0))
    pkg.decode_planned_batch_async([dec], [plan], [True], [-1])
    w, _cw, h, _bd, _sx, _sy = pkg.wire_header(clip.wire(0))
    with pytest.raises(pkg.HipError, match="launched batch"):
        dec.read_slot(-1, w, h)
    with pytest.raises(pkg.HipError):
        pkg.decode_planned_batch_async([dec], [plan], [False], [-1], launcher=other)
    pkg.batch_finish(dec)
    planes = dec.read_slot(-1, w, h)
    want = _alone(clip)[0]
    for p in range(3):
        assert np.array_equal(planes[p], want[p])
    dec.discard_plan(plan)
