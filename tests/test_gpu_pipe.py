"""The ingest pipeline (av1h_pipe_*): frames planned by several host threads
while earlier frames decode, decoded in submission order. Every picture must
carry the same device checksum as the same frame decoded by the plain
synchronous path (av1h_decode_wire), which the golden tests hold to the
reference; a frame whose syntax the planner refuses fails alone, with its
number and message, and the frames after it still decode."""
import numpy as np
import pytest

import av1ongpu_amd as pkg

pytestmark = pytest.mark.gpu

CLIPS = [("c2_mini", 1, {}), ("c1_mini", 2, {"frames": 8}), ("tools_mini", 1, {}),
         ("c1_mini", 41, {"width": 426, "height": 240, "frames": 3}),
         ("contract_mini", 32, {"avail_auto": 0.5})]


def _plain_hashes(clip, keep, rel, refs):
    dec = pkg.HipDecoder()
    dec.set_output_hash(True)
    out = []
    for f in range(clip.frames):
        dec.decode_wire(pkg.wire_set_slots(clip.wire(f), refs[f]), keep=keep[f] >= 0, reserved=keep[f],
                        read_back=False)
        out.append(dec.last_hash())
        for s in rel[f]:
            dec.release(s)
    dec.close()
    return out


@pytest.mark.parametrize("workload,seed,over", CLIPS, ids=[f"{w}-{s}" for w, s, _ in CLIPS])
@pytest.mark.parametrize("threads,depth", [(1, 1), (4, 3), (8, 16)])
def test_pipe_matches_plain(workload, seed, over, threads, depth):
    clip = pkg.Clip(workload, seed, **over)
    keep, rel, refs = pkg.slot_schedule(clip)
    want = _plain_hashes(clip, keep, rel, refs)
    dec = pkg.HipDecoder()
    dec.set_output_hash(True)
    pipe = pkg.Pipe(dec, planner_threads=threads, depth=depth)
    seqs = [pipe.submit(pkg.wire_set_slots(clip.wire(f), refs[f]), keep=keep[f] >= 0, reserved=keep[f],
                        release=rel[f]) for f in range(clip.frames)]
    pipe.wait()
    assert seqs == list(range(clip.frames))
    for f in range(clip.frames):
        r = pipe.result(f)
        assert r.status == 0
        assert r.slot == keep[f]
        assert r.submit_ms <= r.plan_start_ms <= r.planned_ms <= r.decode_start_ms <= r.done_ms
        assert r.flatten_ms > 0 and r.upload_ms > 0
        assert r.hash == want[f], f"{workload} seed {seed} frame {f}: pipelined picture differs from the plain decode"
    pipe.close()
    dec.close()


def test_pipe_reports_a_refused_frame_and_goes_on():
    clip = pkg.Clip("c1_mini", 3, frames=4, key_interval=1)  # every frame a key frame: no references
    want = _plain_hashes(clip, [-1] * 4, [[]] * 4, [[]] * 4)
    bad = bytearray(clip.wire(1))
    # the first block of frame 1 moved outside the coded grid: the planner refuses the frame
    hsize = int(np.frombuffer(bytes(bad), dtype=np.uint32, count=1, offset=12)[0])
    pos = (16 + hsize + 7) & ~7
    boff = pos + 13 * 8
    blk = np.frombuffer(bytes(bad), dtype=np.int32, count=69, offset=boff).copy()
    blk[0] = 1 << 20  # x
    bad[boff:boff + blk.nbytes] = blk.tobytes()
    dec = pkg.HipDecoder()
    dec.set_output_hash(True)
    pipe = pkg.Pipe(dec, planner_threads=3, depth=4)
    for f in range(4):
        pipe.submit(bytes(bad) if f == 1 else clip.wire(f), keep=False)
    with pytest.raises(pkg.HipError, match="frame 1 failed"):
        pipe.wait()
    for f in range(4):
        r = pipe.result(f)
        if f == 1:
            assert r.status != 0
        else:
            assert r.status == 0 and r.hash == want[f]
    pipe.close()
    dec.close()


def _submit_clip(pipe, clip):
    keep, rel, refs = pkg.slot_schedule(clip)
    return [pipe.submit(pkg.wire_set_slots(clip.wire(f), refs[f]), keep=keep[f] >= 0, reserved=keep[f],
                        release=rel[f]) for f in range(clip.frames)]


def test_pipe_across_a_resolution_change():
    """A stream whose frames grow from 426x240 to 854x480 and back: the
    launches that must grow the context's buffers wait for the frames in
    flight (av1h_plan_needs_idle), and every picture still matches."""
    clips = [pkg.Clip("c1_mini", 7, width=426, height=240, frames=4),
             pkg.Clip("c1_mini", 8, width=854, height=480, frames=4),
             pkg.Clip("c1_mini", 9, width=426, height=240, frames=3)]
    want = []
    for c in clips:
        keep, rel, refs = pkg.slot_schedule(c)
        want += _plain_hashes(c, keep, rel, refs)
    dec = pkg.HipDecoder()
    dec.set_output_hash(True)
    pipe = pkg.Pipe(dec, planner_threads=4, depth=8)
    seqs = []
    for c in clips:
        seqs += _submit_clip(pipe, c)
    pipe.wait()
    for i, s in enumerate(seqs):
        r = pipe.result(s)
        assert r.status == 0 and r.hash == want[i], f"frame {i}: differs after a resolution change"
    pipe.close()
    dec.close()


def test_pipe_with_stage_timing_takes_the_waiting_path():
    """With stage timing on, every launch waits for its frame (the launch
    reports that it would wait); the worker takes each result then, and the
    pictures and per-stage times still come out."""
    clip = pkg.Clip("c2_mini", 3)
    keep, rel, refs = pkg.slot_schedule(clip)
    want = _plain_hashes(clip, keep, rel, refs)
    dec = pkg.HipDecoder()
    dec.set_output_hash(True)
    dec.set_stage_timing(True)
    plan = dec.plan_wire(pkg.wire_set_slots(clip.wire(0), refs[0]))
    assert dec.plan_needs_idle(plan)
    dec.discard_plan(plan)
    pipe = pkg.Pipe(dec, planner_threads=3, depth=6)
    seqs = _submit_clip(pipe, clip)
    pipe.wait()
    for f, s in enumerate(seqs):
        r = pipe.result(s)
        assert r.status == 0 and r.hash == want[f]
        assert r.wait_ms == 0.0  # nothing left to wait for after the launch
    pipe.close()
    assert dec.stage_times()["deblock"][1] >= clip.frames
    dec.close()


def test_older_slots_read_while_frames_are_in_flight():
    """A player's display thread packs finished frames while the pipeline
    still has later frames on the device: to_nv12 and read_slot of a
    finished frame's slot (older ones, and the latest when it has no film
    grain) are queued behind the frames in flight and give the same bytes as
    the plain decode."""
    clip = pkg.Clip("c1_mini", 12, frames=8, key_interval=1)  # independent frames, every one kept
    w, _cw, h, _bd, _sx, _sy = pkg.wire_header(clip.wire(0))
    ref = pkg.HipDecoder()
    want_nv12, want_y = [], []
    for f in range(clip.frames):
        slot, planes = ref.decode_wire(clip.wire(f), keep=True, reserved=f)
        want_nv12.append(bytes(ref.to_nv12(slot, 1024, w, h)))
        want_y.append(planes[0].copy())
    ref.close()
    dec = pkg.HipDecoder()
    pipe = pkg.Pipe(dec, planner_threads=2, depth=8)
    for f in range(clip.frames):
        pipe.submit(clip.wire(f), keep=True, reserved=f)
    for f in range(clip.frames):
        pipe.wait(f)
        r = pipe.result(f)
        assert r.status == 0 and r.slot == f
        assert bytes(dec.to_nv12(f, 1024, w, h)) == want_nv12[f], f"frame {f}: NV12 differs"
        assert np.array_equal(dec.read_slot(f, w, h)[0], want_y[f]), f"frame {f}: luma differs"
    pipe.wait()
    pipe.close()
    dec.close()
