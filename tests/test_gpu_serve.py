"""The serving loop (av1h_server_*): frames of concurrent streams admitted
against a memory budget and decoded in continuously formed batches. Every
frame must come out identical to the same stream decoded alone, whatever it
was batched with; the admission budget must hold streams back until memory
frees up.
"""
import random

import numpy as np
import pytest

import av1ongpu_amd as pkg

pytestmark = pytest.mark.gpu


def _flat(planes):
    return np.concatenate([p.reshape(-1) for p in planes])


def _alone(clip):
    return [_flat(pl) for _f, _s, pl in pkg.decode_clip(pkg.HipDecoder(), clip)]


def _geometry(clip):
    w, _cw, h, bd, sx, sy = pkg.wire_header(clip.wire(0))
    return w, h, bd, sx, sy


def _submit_all(srv, clips, sids, seed, use_plans=None):
    """Every frame of every clip in one interleaved arrival order (streams in
    a random order, each stream's frames in its own order)."""
    rng = random.Random(seed)
    sched = [pkg.slot_schedule(c, 8) for c in clips]
    nxt = [0] * len(clips)
    order = []
    while any(nxt[i] < c.frames for i, c in enumerate(clips)):
        i = rng.choice([k for k, c in enumerate(clips) if nxt[k] < c.frames])
        order.append((i, nxt[i]))
        nxt[i] += 1
    for i, f in order:
        keep, rel, refs = sched[i]
        wire = pkg.wire_set_slots(clips[i].wire(f), refs[f])
        tag = i * 1000 + f
        if use_plans is not None:
            srv.submit_plan(sids[i], use_plans[i][f], keep[f], rel[f], tag=tag)
        else:
            srv.submit_wire(sids[i], wire, keep[f], rel[f], tag=tag)
    return len(order)


def _collect(srv, n, timeout_s=120):
    import time
    got = []
    t0 = time.time()
    while len(got) < n:
        got += srv.poll(64, 50)
        assert time.time() - t0 < timeout_s, f"only {len(got)} of {n} frames completed"
    return got


def _check(clips, alone, got):
    by_stream = {}
    for c in got:
        assert c["status"] == 0, c
        by_stream.setdefault(c["stream"], []).append(c)
    for i, clip in enumerate(clips):
        seq = by_stream[i]
        assert [c["tag"] % 1000 for c in seq] == list(range(clip.frames)), "frames of a stream out of order"
        for c in seq:
            f = c["tag"] % 1000
            assert np.array_equal(c["picture"], alone[i][f]), f"{clip.workload} seed {clip.seed} frame {f}"


CLIPS = [("c1_mini", 1, {}), ("c2_mini", 2, {}), ("tools_mini", 1, {}), ("c1_mini", 3, dict(bit_depth=10)),
         ("contract_mini", 2, {}), ("c1_mini", 5, dict(width=426, height=240, frames=3))]


def test_serve_same_alone_and_in_mix():
    clips = [pkg.Clip(w, s, **o) for w, s, o in CLIPS]
    alone = [_alone(c) for c in clips]
    srv = pkg.Server(max_batch=8, planner_threads=3, capture=True)
    try:
        sids = [srv.open_stream(*_geometry(c), pool_slots=8) for c in clips]
        assert sids == list(range(len(clips)))
        n = _submit_all(srv, clips, sids, 1)
        got = _collect(srv, n)
        for s in sids:
            srv.close_stream(s)
        _check(clips, alone, got)
        st = srv.stats()
        assert st["frames"] == n
        assert st["max_batch_frames"] > 1, "frames of different streams never shared a batch"
        assert all(c["done_ms"] >= c["start_ms"] >= c["arrival_ms"] for c in got)
    finally:
        srv.close()


def test_serve_plans_shared_by_streams():
    """Two streams decode the same plans (read only, shared across contexts)."""
    clip = pkg.Clip("c2_mini", 6)
    alone = _alone(clip)
    keep, rel, refs = pkg.slot_schedule(clip, 8)
    owner = pkg.HipDecoder()
    plans = [owner.plan_wire(pkg.wire_set_slots(clip.wire(f), refs[f])) for f in range(clip.frames)]
    srv = pkg.Server(max_batch=4, planner_threads=0, capture=True)
    try:
        sids = [srv.open_stream(*_geometry(clip), pool_slots=8) for _ in range(2)]
        n = _submit_all(srv, [clip, clip], sids, 2, use_plans=[plans, plans])
        got = _collect(srv, n)
        _check([clip, clip], [alone, alone], got)
    finally:
        srv.close()
        for p in plans:
            owner.discard_plan(p)


def test_serve_admission_budget():
    """A budget that holds one stream: the next waits until the first closes
    and gives its memory back; no batch mixes them."""
    clips = [pkg.Clip("c1_mini", 7), pkg.Clip("c1_mini", 8), pkg.Clip("c1_mini", 9)]
    alone = [_alone(c) for c in clips]
    geo = _geometry(clips[0])
    one = pkg.stream_bytes(*geo, 8, 0)
    srv = pkg.Server(max_batch=8, planner_threads=2, capture=True, hbm_budget=one + one // 2)
    try:
        sids = [srv.open_stream(*geo, pool_slots=8) for _ in clips]
        st = srv.stats()
        assert st["waiting"] >= 1
        n = _submit_all(srv, clips, sids, 3)
        got = []
        # close each stream once its frames are done, which admits the next
        for k, sid in enumerate(sids):
            want = clips[k].frames
            mine = []
            while len(mine) < want:
                new = srv.poll(64, 50)
                got += new
                mine = [c for c in got if c["stream"] == sid]
            srv.close_stream(sid)
        assert len(got) == n
        _check(clips, alone, got)
        assert all(c["batch_frames"] == 1 for c in got)
        starts = {s: min(c["start_ms"] for c in got if c["stream"] == s) for s in sids}
        ends = {s: max(c["done_ms"] for c in got if c["stream"] == s) for s in sids}
        assert starts[sids[1]] >= ends[sids[0]] and starts[sids[2]] >= ends[sids[1]]
        assert srv.stats()["budget_used"] == 0
    finally:
        srv.close()


def test_serve_heavy_frames_batched_on_their_launcher():
    """Frames whose queue has many dependency levels (the key frames here)
    go to launcher 0, several per batch; each launcher keeps the launch
    shapes set on it; every frame still equals the stream decoded alone."""
    clips = [pkg.Clip(w, s, **o) for w, s, o in CLIPS[:4]]
    alone = [_alone(c) for c in clips]
    srv = pkg.Server(max_batch=8, planner_threads=2, capture=True, launchers=2, heavy_levels=8,
                     heavy_launchers=1, heavy_batch=3)
    try:
        srv.set_launcher_tuning(0, {"queue": 2, "cdef_blocks_per_thread": 2})
        srv.set_launcher_tuning(1, {"queue": 12, "mc": 4})
        assert srv.launcher_tuning(0)["queue"] == 2 and srv.launcher_tuning(0)["cdef_blocks_per_thread"] == 2
        assert srv.launcher_tuning(1)["queue"] == 12 and srv.launcher_tuning(1)["mc"] == 4
        with pytest.raises(pkg.HipError):
            srv.launcher_tuning(2)
        sids = [srv.open_stream(*_geometry(c), pool_slots=8) for c in clips]
        n = _submit_all(srv, clips, sids, 4)
        got = _collect(srv, n)
        _check(clips, alone, got)
        assert srv.stats()["frames"] == n
    finally:
        srv.close()
