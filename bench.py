#!/usr/bin/env python3
"""Decode throughput of the gfx950 AV1 reconstruction backend on one MI355X.

Workload (default): 1920x1080 10-bit 4:2:0, 60 frames in a hierarchical
16-frame group with 3 references, compound prediction (distance weighted and
average), transforms up to 64x64 with identity and flip-ADST, Gaussian motion
vectors with all three subpel filter families, deblocking + CDEF + Wiener and
self-guided restoration, superres 3/4 (coded width 1440), film grain with 24
luma AR coefficients, every frame packed to NV12 at a 2048 byte pitch. The
syntax is synthetic and seeded (av1ongpu_amd/csrc/synth.cpp).

A step decodes the whole clip in order, starting from each frame's syntax in
host memory (the FrameSyntax wire bytes): the frames go through the ingest
pipeline (av1h_pipe_*), where planner threads flatten and upload frame N+1..
while frame N decodes, the decode runs every stage on the device, keeps the
reference slots on the device, and packs NV12 on the device. `value` is
display luma Mpixel/s over the timed steps, from the first submit to the last
frame complete. Frame latency inside the pipeline (submit to NV12 packed) is
reported as p50/p99.

Also reported: `device_only`, the same clip from plans already resident in
HBM (one decode call per frame, its call latency p50/p99); `ingest_serial`,
each frame flattened, uploaded and decoded before the next on one thread; the
dominant kernel's roofline from HIP events around each stage, on the decode
stream, in one more pass over the clip after the timed steps (the timed steps
replay captured launch sequences; the event pass launches eagerly); the CPU
oracle on a bounded sample of the same clip, timed on this host, and the
bit-exactness of the GPU output on that sample; the 16-stream mix, batched vs
plain (`batched_mix`), with the learned schedule (`cost_model`); and the
serving loop (`serving`).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "decoded Mpixel/s on 1 MI355X, batched vs plain, bit-exact; p99 frame latency"
HBM_PEAK_GBS = 8000.0
NV12_PITCH = 2048
# the kernel that dominates each timed stage (the stage's events bracket it)
STAGE_KERNEL = {"inter": "k_mc", "queue": "k_queue", "deblock": "k_deblock", "cdef": "k_cdef<1>",
                "superres": "k_upscale", "lr": "k_lr", "store": "k_copy", "grain": "k_grain_apply"}
WORKLOADS = {
    "c2": "1920x1080 10-bit 4:2:0, 60 frames, hierarchical GOP 16 with 3 references, compound "
          "(distance weighted + average), transforms 4x4-64x64 incl. identity and flip-ADST, "
          "Gaussian MVs sigma 24 px with regular/smooth/sharp filters, deblock + CDEF + Wiener + "
          "self-guided restoration, superres 3/4 (coded width 1440), film grain AR lag 3, "
          "NV12 pitch 2048",
    "c1": "1280x720 8-bit 4:2:0, 30 frames (1 key + 29 inter, single reference), deblock + CDEF + "
          "Wiener restoration, NV12 pitch 2048",
}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def slot_schedule(clip):
    """The slots a decoder that always takes the lowest free slot would use:
    (slot per frame or -1, slots released after each frame, reference slots per
    frame)."""
    used = set()
    slot_of = {}
    keep_slot, releases, ref_slots = [], [], []
    for f in range(clip.frames):
        ref_slots.append([slot_of[r] for r in clip.refs[f]])
        if clip.last_use[f] > f:
            s = min(i for i in range(32) if i not in used)
            used.add(s)
            slot_of[f] = s
            keep_slot.append(s)
        else:
            keep_slot.append(-1)
        rel = []
        for r in clip.refs[f]:
            if clip.last_use[r] == f:
                s = slot_of.pop(r)
                used.discard(s)
                rel.append(s)
        releases.append(rel)
    return keep_slot, releases, ref_slots


def display_pixels(clip):
    w, _cw, h, _bd, _sx, _sy = __import__("av1ongpu_amd").wire_header(clip.wire(0))
    return w * h


def default_planners():
    """Host threads for planning: five eighths of the cores this process may
    use, at most 12. Each planner waits for its own plan's upload; on 16
    cores with 64 frames admitted the 1080p clip measured 4230-4572
    Mpixel/s with 8 planners, 4691 with 10 and 4657 with 12 (the planners
    contend for memory bandwidth, and the decode worker needs a core of its
    own)."""
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = min(cores, int(os.environ.get("OMP_NUM_THREADS", cores)))
    return max(1, min(12, cores * 5 // 8))


def pct(values, q):
    v = sorted(values)
    if not v:
        return None
    k = min(len(v) - 1, max(0, int(round(q / 100.0 * (len(v) - 1)))))
    return v[k]


def load_traffic(workload, kernel):
    """HBM bytes per launch of a kernel from the committed rocprofv3 counter
    summary (FETCH_SIZE and WRITE_SIZE, scaled by a copy kernel of known size),
    or None when the summary does not cover the kernel on this workload (said
    on stderr)."""
    path = os.path.join(REPO, "profiles", "hbm_traffic.json")
    try:
        with open(path) as fh:
            t = json.load(fh)
    except (OSError, ValueError) as e:
        log(f"WARNING: no HBM traffic summary ({e}): roofline traffic left null")
        return None
    if t.get("workload") != workload:
        log(f"WARNING: the HBM traffic summary is of workload {t.get('workload')}, not {workload}")
        return None
    k = t.get("kernels", {}).get(kernel)
    if k is None:
        log(f"WARNING: the HBM traffic summary has no kernel {kernel}: roofline traffic left null")
        return None
    return int(k["bytes_per_launch"])


def cpu_baseline(clip, seconds, gpu_hashes):
    """The CPU oracle on the first frames of the clip, in order, until
    `seconds` have passed; the GPU's per frame hashes are compared on the same
    frames."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import av1ongpu_amd as pkg
    from oracle_lib import OracleDecoder, oracle_lib

    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, int(os.environ.get("OMP_NUM_THREADS", cores))))
    oracle_lib().av1o_set_threads(cores)
    px = display_pixels(clip)
    n = 0
    mism = []
    t0 = time.perf_counter()
    for f, _slot, planes in pkg.decode_clip(OracleDecoder(), clip):
        h = pkg.fnv1a64(pkg.yuv16_bytes(planes))
        n += 1
        if f < len(gpu_hashes) and gpu_hashes[f] != h:
            mism.append(f)
        if time.perf_counter() - t0 >= seconds:
            break
    dt = time.perf_counter() - t0
    return {"value": round(n * px / dt / 1e6, 3), "unit": "Mpixel/s", "cores": cores, "kind": "port",
            "sample": f"first {n} of {clip.frames} frames of the same clip, decoded in order by the CPU "
                      f"oracle (OpenMP, {cores} threads) with syntax ingest and read back, {dt:.1f} s"}, n, mism


# The 16-stream mix: resolutions as named, 8 and 10 bit alternating inside
# each resolution group (the 3840x2160 stream is 10 bit), config-1 syntax
# statistics, 32 frames per stream.
MIX_STREAMS = [(426, 240, 8), (426, 240, 10), (426, 240, 8), (426, 240, 10),
               (854, 480, 8), (854, 480, 10), (854, 480, 8), (854, 480, 10),
               (1280, 720, 8), (1280, 720, 10), (1280, 720, 8), (1280, 720, 10),
               (1920, 1080, 8), (1920, 1080, 10), (1920, 1080, 8),
               (3840, 2160, 10)]
MIX_FRAMES = 32


def run_mix(args):
    """The 16-stream mix decoded two ways from the same device-resident plans:
    plain (one decode call per frame, streams round robin) and batched (one
    av1h_decode_planned_batch per round, the frame of every stream in it).
    Returns the report object; outputs of the two are compared frame by frame."""
    import av1ongpu_amd as pkg
    from concurrent.futures import ThreadPoolExecutor

    t = time.perf_counter()
    workers = max(1, min(8, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else 8))

    def make(i):
        w, h, bd = MIX_STREAMS[i]
        return pkg.Clip("c1", args.seed * 1000 + i, width=w, height=h, bit_depth=bd, frames=MIX_FRAMES)

    with ThreadPoolExecutor(workers) as ex:
        clips = list(ex.map(make, range(len(MIX_STREAMS))))
    ns = len(clips)
    sched = [slot_schedule(c) for c in clips]
    px = [w * h for w, h, _ in MIX_STREAMS]
    log(f"mix: {ns} streams x {MIX_FRAMES} frames, {sum(px) / 1e6:.2f} Mpixel per round, generated in "
        f"{time.perf_counter() - t:.1f} s")
    decs = [pkg.HipDecoder(0) for _ in clips]

    def plan_stream(i):
        c = clips[i]
        refs = sched[i][2]
        out = []
        for f in range(c.frames):
            out.append(decs[i].plan_wire(pkg.wire_set_slots(c.wire(f), refs[f])))
            c._wire.pop(f, None)
        return out

    t = time.perf_counter()
    with ThreadPoolExecutor(workers) as ex:
        plans = list(ex.map(plan_stream, range(ns)))
    del clips
    log(f"mix: planned {ns * MIX_FRAMES} frames in {time.perf_counter() - t:.1f} s")

    def release(i, f):
        for sl in sched[i][1][f]:
            decs[i].release(sl)

    def plain(lat, hashes=None):
        for f in range(MIX_FRAMES):
            for i in range(ns):
                t1 = time.perf_counter()
                decs[i].decode_planned(plans[i][f], keep=sched[i][0][f] >= 0, reserved=sched[i][0][f])
                lat.append(time.perf_counter() - t1)
                if hashes is not None:
                    hashes.append(decs[i].last_hash())
                release(i, f)

    def batched(lat, hashes=None):
        """Round f launched before round f - 1 is finished (the host prepares
        a round while the device runs the one before); the slots round f - 1
        last read are released before round f is launched."""
        keeps = [[sched[i][0][f] >= 0 for i in range(ns)] for f in range(MIX_FRAMES)]
        res = [[sched[i][0][f] for i in range(ns)] for f in range(MIX_FRAMES)]
        t_launch = []

        def finish(f):
            pkg.batch_finish(decs[0])
            lat.append(time.perf_counter() - t_launch[f])
            if hashes is not None:
                hashes.extend(decs[i].last_hash() for i in range(ns))

        for f in range(MIX_FRAMES):
            if f > 0:
                for i in range(ns):
                    release(i, f - 1)
            t_launch.append(time.perf_counter())
            pkg.decode_planned_batch_async(decs, [plans[i][f] for i in range(ns)], keeps[f], res[f])
            if f > 0:
                finish(f - 1)
        finish(MIX_FRAMES - 1)
        for i in range(ns):
            release(i, MIX_FRAMES - 1)

    # parity pass (untimed): every frame of every stream, plain vs batched,
    # by the device checksum of each picture
    for d in decs:
        d.set_output_hash(True)
    hp, hb = [], []
    if not args.mix_batched_only:
        plain([], hp)
    batched([], hb)
    for d in decs:
        d.set_output_hash(False)
    if args.mix_batched_only:
        hp = list(hb)
    hp = [hp[f * ns + i] for i in range(ns) for f in range(MIX_FRAMES)]
    hb = [hb[f * ns + i] for i in range(ns) for f in range(MIX_FRAMES)]
    mism = sum(1 for a, b in zip(hp, hb) if a != b)
    log(f"mix: batched vs plain on {len(hp)} frames: {mism} mismatched")
    # launch shapes of the batched path, tuned on its first rounds
    tuned = None
    if not args.no_tune:
        from av1ongpu_amd import autotune
        t = time.perf_counter()
        rounds = [([plans[i][f] for i in range(ns)], [sched[i][0][f] >= 0 for i in range(ns)],
                   [sched[i][0][f] for i in range(ns)], [sched[i][1][f] for i in range(ns)]) for f in range(8)]
        shapes, t_base, t_tuned = autotune.tune(decs, rounds, reps=1)
        tuned = {"shapes": shapes, "rounds": len(rounds), "default_ms": round(t_base * 1e3, 2),
                 "tuned_ms": round(t_tuned * 1e3, 2), "seconds": round(time.perf_counter() - t, 1)}
        log(f"mix: autotuned launch shapes in {tuned['seconds']} s: {t_base * 1e3:.1f} -> {t_tuned * 1e3:.1f} ms")
    for _ in range(args.warmup):
        if not args.mix_batched_only:
            plain([])
        batched([])
    out = {}
    for name, fn in (("plain", plain), ("batched", batched)):
        if name == "plain" and args.mix_batched_only:
            out[name] = {"value": None}
            continue
        lat = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            fn(lat)
        dt = time.perf_counter() - t0
        out[name] = {"value": round(args.steps * MIX_FRAMES * sum(px) / dt / 1e6, 2), "unit": "Mpixel/s",
                     "ms_per_step": round(dt * 1e3 / args.steps, 3),
                     "call_latency_ms": {"p50": round(pct(lat, 50) * 1e3, 3), "p99": round(pct(lat, 99) * 1e3, 3)}}
        log(f"mix {name}: {out[name]['value']} Mpixel/s")
    cost = None if args.no_costmodel else run_costmodel(args, decs, plans, sched)
    for i in range(ns):
        for p in plans[i]:
            decs[i].discard_plan(p)
        decs[i].close()
    rep = {
        "workload": "16 streams: 4x 426x240, 4x 854x480, 4x 1280x720, 3x 1920x1080, 1x 3840x2160, 8 and 10 bit "
                    "alternating (2160p 10 bit), 4:2:0, 32 frames each, 1280x720-workload syntax statistics",
        "plain": out["plain"], "batched": out["batched"],
        "speedup": round(out["batched"]["value"] / out["plain"]["value"], 3) if out["plain"]["value"] else None,
        "bit_exact": {"frames_checked": len(hp), "mismatched_frames": mism,
                      "method": "device checksum of every picture, batched decode (rounds launched ahead of "
                                "the previous round's completion) vs plain decode of the same plans"},
    }
    if cost is not None:
        rep["cost_model"] = cost
    if tuned is not None:
        rep["autotune"] = tuned
    if not args.mix_batched_only:
        rep["ingest"] = run_mix_ingest(args)
    return rep


MIX_INGEST_FRAMES = 12
MIX_INGEST_WARM = 2


def run_mix_ingest(args):
    """The 16-stream mix from syntax, ingest included: the serving loop
    (planner threads, one launcher) takes each stream's first frames and
    finishes them (its contexts, staging and plan buffers set up), then every
    later frame of every stream is submitted as wire bytes at once; wall
    time from that submit to the last completion. Once with batches of up to
    16 frames (one per stream) and once with one frame per launch; the two
    runs' picture checksums are compared frame by frame."""
    import threading
    import av1ongpu_amd as pkg

    clips = [pkg.Clip("c1", args.seed * 1000 + i, width=w, height=h, bit_depth=bd, frames=MIX_INGEST_FRAMES)
             for i, (w, h, bd) in enumerate(MIX_STREAMS)]
    scheds = [pkg.slot_schedule(c, 16) for c in clips]
    wires = [[pkg.wire_set_slots(c.wire(f), scheds[i][2][f]) for f in range(c.frames)] for i, c in enumerate(clips)]
    px = [w * h for w, h, _ in MIX_STREAMS]
    del clips
    planners = args.planners if args.planners > 0 else default_planners()
    out, hashes = {}, {}
    for name, mb in (("batched", len(MIX_STREAMS)), ("plain", 1)):
        srv = pkg.Server(0, max_batch=mb, planner_threads=planners, hbm_budget=200 << 30, hash_pictures=True,
                         launchers=1, heavy_levels=0)
        sids = [srv.open_stream(w, h, bd, 1, 1, 16) for w, h, bd in MIX_STREAMS]
        done = []

        def run(f0, f1):
            total = len(done) + len(MIX_STREAMS) * (f1 - f0)

            def poller():
                while len(done) < total:
                    done.extend(srv.poll(256, 20))

            th = threading.Thread(target=poller, daemon=True)
            th.start()
            t0 = srv.now_ms()
            for f in range(f0, f1):
                for i, sid in enumerate(sids):
                    keep, rel, _ = scheds[i]
                    srv.submit_wire_ref(sid, wires[i][f], keep[f], rel[f], arrival_ms=t0, tag=(i << 32) | f)
            th.join(timeout=300)
            if len(done) < total:
                raise RuntimeError(f"mix ingest ({name}): {len(done)} of {total} frames completed")
            return t0

        run(0, MIX_INGEST_WARM)
        b0 = srv.stats()
        t0 = run(MIX_INGEST_WARM, MIX_INGEST_FRAMES)
        st = srv.stats()
        srv.close()
        wall = max(c["done_ms"] for c in done) - t0
        hashes[name] = {c["tag"]: (c["status"], c["hash"]) for c in done}
        timed = MIX_INGEST_FRAMES - MIX_INGEST_WARM
        nb, nf = st["batches"] - b0["batches"], st["frames"] - b0["frames"]
        out[name] = {"value": round(timed * sum(px) / wall / 1e3, 2), "unit": "Mpixel/s",
                     "wall_ms": round(wall, 2), "batches": nb, "mean_batch_frames": round(nf / max(nb, 1), 2)}
        log(f"mix ingest {name}: {out[name]['value']} Mpixel/s ({out[name]['mean_batch_frames']} frames per launch)")
    mism = sum(1 for t, v in hashes["batched"].items() if v[0] != 0 or hashes["plain"].get(t) != v)
    return {"workload": f"the 16 streams, {MIX_INGEST_FRAMES} frames each: the first {MIX_INGEST_WARM} of each "
                        f"stream decoded untimed, then the other {MIX_INGEST_FRAMES - MIX_INGEST_WARM} of every "
                        f"stream submitted as syntax (wire bytes) at once; serving loop with {planners} planner "
                        f"threads and one launcher",
            "batched": out["batched"], "plain": out["plain"],
            "speedup": round(out["batched"]["value"] / out["plain"]["value"], 3),
            "bit_exact": {"frames_checked": len(hashes["batched"]), "mismatched_frames": mism,
                          "method": "device checksum of every picture, batches of 16 vs one frame per launch"}}


def run_costmodel(args, decs, plans, sched):
    """The learned schedule on the 16-stream mix: superblock timings recorded
    by the decoder on other seeds of the same stream shapes (key frames every
    other frame, so that a few clips give the samples), the perceptron
    trained on the GPU, then the mix decoded in batches under the static
    dependency-level order and under the model's order: queue device time,
    wall time and the checksums of every picture compared."""
    import numpy as np
    import av1ongpu_amd as pkg
    from av1ongpu_amd import costmodel as cmod

    t = time.perf_counter()
    Xs, ys = [], []
    total, rep = 0, 0
    while total < args.cm_samples and rep < 64:
        clips = [pkg.Clip("c1", 50000 + 100 * rep + i, width=w, height=h, bit_depth=bd, frames=6, key_interval=2)
                 for i, (w, h, bd) in enumerate(MIX_STREAMS)]
        X, y = cmod.collect_samples(clips)
        Xs.append(X)
        ys.append(y)
        total += len(y)
        rep += 1
    X, y = np.concatenate(Xs), np.concatenate(ys)
    log(f"cost model: {len(y)} superblock timings recorded in {time.perf_counter() - t:.1f} s")
    model, losses, r2, train_s = cmod.train(X, y, epochs=args.cm_epochs, lr=1e-3, seed=args.seed)
    log(f"cost model: trained {args.cm_epochs} epochs in {train_s:.1f} s, hold-out R^2 {r2:.3f}")
    for d in decs:
        d.set_output_hash(True)
    res = {}
    hashes = {}
    # (name, model, order): the static row-order schedule (coding order),
    # the static dependency-level order, critical path list scheduling with
    # sample counts as costs (no model), and the model's costs in the level
    # order and in critical path order
    schedules = [("row_order", None, pkg.ORDER_CODING), ("level", None, pkg.ORDER_LEVEL),
                 ("path_samples", None, pkg.ORDER_PATH), ("level_model", model, None),
                 ("path_model", model, pkg.ORDER_PATH)]
    for name, m, order in schedules:
        cmod.schedule_all(decs, plans, m, order)
        cmod.run_rounds(decs, plans, sched)                     # warm
        h = []
        wall = min(cmod.run_rounds(decs, plans, sched, h)[0] for _ in range(max(1, args.steps)))
        _, q = cmod.run_rounds(decs, plans, sched, timing=True)
        hashes[name] = h
        res[name] = {"wall_ms": round(wall * 1e3, 2), "queue_device_ms": round(q, 2)}
        log(f"cost model: {name} schedule: {wall * 1e3:.1f} ms wall, queue {q:.1f} ms")
    cmod.schedule_all(decs, plans, None)
    for d in decs:
        d.set_output_hash(False)
    mism = sum(1 for name in hashes for a, b in zip(hashes["level"], hashes[name]) for u, v in zip(a, b) if u != v)
    model.close()
    best_static = min(("row_order", "level", "path_samples"), key=lambda k: res[k]["wall_ms"])
    return {
        "samples": int(len(y)), "epochs": args.cm_epochs, "train_seconds": round(train_s, 2),
        "loss_first_last": [round(float(losses[0]), 4), round(float(losses[-1]), 4)],
        "holdout_r2": round(r2, 3),
        "schedules": res,
        "static": res["row_order"], "model": res["path_model"],
        "wall_speedup": round(res["row_order"]["wall_ms"] / max(res["path_model"]["wall_ms"], 1e-9), 3),
        "queue_speedup": round(res["row_order"]["queue_device_ms"] / max(res["path_model"]["queue_device_ms"], 1e-9),
                               3),
        "vs_best_static": {"schedule": best_static,
                           "wall_speedup": round(res[best_static]["wall_ms"] / max(res["path_model"]["wall_ms"], 1e-9),
                                                 3)},
        "note": "static = the static row-order schedule (coding order); model = the queue ordered by the critical "
                "path of the model's predicted superblock costs; vs_best_static compares with the fastest order "
                "that uses no model",
        "bit_exact": {"frames_checked": sum(len(a) for a in hashes["level"]) * (len(hashes) - 1),
                      "mismatched_frames": mism,
                      "method": "device checksum of every picture, every schedule vs the static level order"},
    }


# The serving workload: 64 streams, Poisson arrivals of 1200 frames/s in all,
# 1280x720-workload-style classes with the 1920x1080 full-tool statistics.
SERVE_CLASSES = [  # (name, streams, width, height, bit depth, coded width, grain)
    ("720p 8-bit", 26, 1280, 720, 8, 0, 0),
    ("1080p 10-bit + grain", 25, 1920, 1080, 10, 0, 1),
    ("2160p 10-bit + superres 3/4", 13, 3840, 2160, 10, 2880, 0),
]
SERVE_RATE = 1200.0
SERVE_GOP = 16
SERVE_PITCH = 4096


def run_serve(args, planned=False):
    """Frames of 64 streams arrive as Poisson processes and go through the
    serving loop (av1h_server_*: admission against a 200 GB budget,
    continuous batching, NV12 packed with every frame). Every completed
    frame's device checksum is compared with the same frame decoded alone.
    Latency: arrival to decoded, stored and NV12-packed. Each frame arrives
    as syntax (wire bytes), planned by the loop's planner threads inside its
    latency, or with `planned` as a plan already resident in HBM."""
    import threading
    import av1ongpu_amd as pkg
    import numpy as np

    t = time.perf_counter()
    gops, plans, scheds, want, wires = [], [], [], [], []
    owner = pkg.HipDecoder(0)
    for ci, (name, _n, w, h, bd, cw, grain) in enumerate(SERVE_CLASSES):
        clip = pkg.Clip("c2", args.seed * 100 + ci, width=w, height=h, bit_depth=bd, coded_width=cw, grain=grain,
                        frames=SERVE_GOP)
        keep, rel, refs = pkg.slot_schedule(clip, 16)
        ws = [pkg.wire_set_slots(clip.wire(f), refs[f]) for f in range(clip.frames)]
        ps = [owner.plan_wire(w) for w in ws]
        # the checksums of the group decoded alone, through the same plans
        ref = pkg.HipDecoder(0)
        ref.set_output_hash(True)
        ref.set_pool_slots(16)
        hs = []
        for f in range(clip.frames):
            ref.decode_planned(ps[f], keep=keep[f] >= 0, reserved=keep[f])
            hs.append(ref.last_hash())
            for sl in rel[f]:
                ref.release(sl)
        ref.close()
        del clip
        wires.append(ws)
        plans.append(ps)
        scheds.append((keep, rel))
        want.append(hs)
    log(f"serve: planned {len(SERVE_CLASSES)} groups of {SERVE_GOP} frames in {time.perf_counter() - t:.1f} s")

    # key frames (their queue has far more dependency levels than an inter
    # frame's) run on a launcher of their own
    levels = [max(pkg.hip_lib().av1h_plan_levels(p) for p in ps) for ps in plans]
    inter_levels = [sorted(pkg.hip_lib().av1h_plan_levels(p) for p in ps)[len(ps) // 2] for ps in plans]
    heavy = args.serve_heavy if args.serve_heavy >= 0 else 4 * max(inter_levels)
    log(f"serve: queue levels per group: max {levels}, median {inter_levels}; heavy above {heavy}")
    if not planned:
        for ps in plans:
            for p in ps:
                owner.discard_plan(p)
    srv = pkg.Server(0, max_batch=args.serve_batch, planner_threads=args.serve_planners, hbm_budget=200 << 30,
                     nv12_pitch=SERVE_PITCH, hash_pictures=True, launchers=args.serve_launchers,
                     heavy_levels=heavy, heavy_launchers=args.serve_heavy_launchers,
                     heavy_batch=args.serve_heavy_batch)
    streams = []  # (stream id, class)
    for ci, (_name, n, w, h, bd, _cw, _g) in enumerate(SERVE_CLASSES):
        for _ in range(n):
            streams.append((srv.open_stream(w, h, bd, 1, 1, 16), ci))
    rng = np.random.default_rng(args.seed)
    per = SERVE_RATE / len(streams)
    warm, dur = args.serve_warmup, args.serve_seconds
    events = []
    for si in range(len(streams)):
        tt = rng.exponential(1.0 / per)
        k = 0
        while tt < warm + dur:
            events.append((tt, si, k))
            tt += rng.exponential(1.0 / per)
            k += 1
    events.sort()
    done = []
    stop = threading.Event()

    def poller():
        deadline = None
        while len(done) < len(events):
            done.extend(srv.poll(256, 20))
            if stop.is_set():
                deadline = deadline or time.perf_counter() + 60.0
                if time.perf_counter() > deadline:
                    break

    th = threading.Thread(target=poller, daemon=True)
    th.start()
    t0 = srv.now_ms()
    for tt, si, k in events:
        sid, ci = streams[si]
        target = t0 + tt * 1e3
        while True:
            now = srv.now_ms()
            if now >= target:
                break
            time.sleep(min(0.002, (target - now) / 1e3))
        keep, rel = scheds[ci]
        g = k % SERVE_GOP
        if planned:
            srv.submit_plan(sid, plans[ci][g], keep[g], rel[g], arrival_ms=target, tag=(si << 32) | k)
        else:
            srv.submit_wire_ref(sid, wires[ci][g], keep[g], rel[g], arrival_ms=target, tag=(si << 32) | k)
    stop.set()
    th.join(timeout=120)
    st = srv.stats()
    srv.close()
    if planned:
        for ps in plans:
            for p in ps:
                owner.discard_plan(p)
    owner.close()
    px = [w * h for _n, _c, w, h, _b, _cw, _g in SERVE_CLASSES]
    lat = {ci: [] for ci in range(len(SERVE_CLASSES))}
    lat_key, lat_inter = [], []
    mism = 0
    t_lo, t_hi = t0 + warm * 1e3, t0 + (warm + dur) * 1e3
    pix = 0
    last_done = t_lo
    for c in done:
        si, k = c["tag"] >> 32, c["tag"] & 0xffffffff
        ci = streams[si][1]
        if c["status"] != 0 or c["hash"] != want[ci][k % SERVE_GOP]:
            mism += 1
        if c["arrival_ms"] >= t_lo:
            lat[ci].append(c["done_ms"] - c["arrival_ms"])
            (lat_key if k % SERVE_GOP == 0 else lat_inter).append(c["done_ms"] - c["arrival_ms"])
            pix += px[ci]
            last_done = max(last_done, c["done_ms"])
    allv = [v for ci in lat for v in lat[ci]]
    plan_ms = [c["planned_ms"] - c["plan_start_ms"] for c in done if c["arrival_ms"] >= t_lo]
    # where the slowest 1 % of frames spent their time: waiting for a planner,
    # planning, waiting for their batch to start, in their batch
    timed = sorted((c for c in done if c["arrival_ms"] >= t_lo), key=lambda c: c["done_ms"] - c["arrival_ms"])
    tail = timed[int(len(timed) * 0.99):] or timed[-1:]

    def phases(cs):
        if not cs:
            return None
        n = len(cs)
        if planned:
            return {"to_start": round(sum(c["start_ms"] - c["arrival_ms"] for c in cs) / n, 3),
                    "in_batch": round(sum(c["done_ms"] - c["start_ms"] for c in cs) / n, 3)}
        return {"to_planner": round(sum(c["plan_start_ms"] - c["arrival_ms"] for c in cs) / n, 3),
                "planning": round(sum(c["planned_ms"] - c["plan_start_ms"] for c in cs) / n, 3),
                "to_start": round(sum(c["start_ms"] - c["planned_ms"] for c in cs) / n, 3),
                "in_batch": round(sum(c["done_ms"] - c["start_ms"] for c in cs) / n, 3)}
    span = (max(last_done, t_hi) - t_lo) / 1e3
    rep = {
        "workload": "64 streams (26x 1280x720 8-bit, 25x 1920x1080 10-bit with film grain, 13x 3840x2160 10-bit "
                    "with superres 3/4), full-tool syntax statistics, Poisson arrivals of 1200 frames/s in all; "
                    f"each stream cycles through a {SERVE_GOP}-frame group (a key frame and 15 inter frames, 3 "
                    f"references); " + ("each frame submitted at its arrival as a plan resident in HBM "
                                        "(planned beforehand)" if planned else
                                        f"each frame submitted as syntax (wire bytes) at its arrival and planned "
                                        f"by {args.serve_planners} planner threads inside its latency") +
                    "; NV12 packed with every frame",
        "seconds": dur, "frames": len(allv), "completed": len(done), "submitted": len(events),
        "offered_mpixel_s": round(sum(px[streams[si][1]] for tt, si, k in events if tt >= warm) / dur / 1e6, 1),
        "value": round(pix / span / 1e6, 1), "unit": "Mpixel/s",
        "latency_ms": {"p50": round(pct(allv, 50), 3), "p99": round(pct(allv, 99), 3)},
        "latency_ms_by_class": {SERVE_CLASSES[ci][0]: {"p50": round(pct(lat[ci], 50), 3),
                                                       "p99": round(pct(lat[ci], 99), 3)}
                                for ci in lat if lat[ci]},
        "latency_ms_key_frames": {"p50": round(pct(lat_key, 50), 3), "p99": round(pct(lat_key, 99), 3),
                                  "frames": len(lat_key)},
        "latency_ms_inter_frames": {"p50": round(pct(lat_inter, 50), 3), "p99": round(pct(lat_inter, 99), 3),
                                    "frames": len(lat_inter)},
        "batches": st["batches"], "mean_batch_frames": round(st["frames"] / max(st["batches"], 1), 2),
        "device_busy_frac": round(st["busy_ms"] / (max(c["done_ms"] for c in done) - t0), 3),
        "planning_ms": {"p50": round(pct(plan_ms, 50), 3), "p99": round(pct(plan_ms, 99), 3)} if plan_ms else None,
        "latency_phases_ms": {"all_frames_mean": phases(timed), "slowest_1pct_mean": phases(tail)},
        "bit_exact": {"frames_checked": len(done), "mismatched_frames": mism,
                      "method": "device checksum of every completed picture vs the same frame decoded alone"},
    }
    log(f"serve: {rep['value']} Mpixel/s, p50 {rep['latency_ms']['p50']} ms, p99 {rep['latency_ms']['p99']} ms, "
        f"{mism} mismatched")
    return rep


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c2", choices=sorted(WORKLOADS))
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--planners", type=int, default=0, help="planner threads of the ingest pipeline (0: automatic)")
    ap.add_argument("--levels", type=int, default=0,
                    help="1: the clip's syntax carries coded levels (dequantised on the device)")
    ap.add_argument("--depth", type=int, default=64,
                    help="frames admitted to the ingest pipeline (a key frame takes about three inter frames' "
                         "planning time, so it must start planning well before its turn)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline and the parity sample")
    ap.add_argument("--no-mix", action="store_true", help="skip the 16-stream batched vs plain mix")
    ap.add_argument("--mix-only", action="store_true", help="only the 16-stream mix")
    ap.add_argument("--mix-batched-only", action="store_true",
                    help="the mix's batched path only (no plain pass, no comparison): for profiling")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving workload")
    ap.add_argument("--no-costmodel", action="store_true", help="skip the learned schedule on the mix")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch shapes of the mix")
    ap.add_argument("--cm-samples", type=int, default=200000)
    ap.add_argument("--cm-epochs", type=int, default=20)
    ap.add_argument("--serve-only", action="store_true", help="only the serving workload")
    ap.add_argument("--serve-seconds", type=float, default=3.0)
    ap.add_argument("--serve-warmup", type=float, default=1.0)
    ap.add_argument("--serve-batch", type=int, default=32)
    ap.add_argument("--serve-launchers", type=int, default=2)
    ap.add_argument("--serve-heavy-launchers", type=int, default=1)
    ap.add_argument("--serve-heavy-batch", type=int, default=1)
    ap.add_argument("--serve-planners", type=int, default=12, help="planner threads of the serving loop")
    ap.add_argument("--serve-heavy", type=int, default=-1,
                    help="dependency levels above which a frame runs on the key-frame launcher (-1: automatic, 0: off)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this backend decodes each stream on one device and has no "
              f"multi-GPU path (asked for --gpus {args.gpus})", file=sys.stderr)
        sys.exit(2)
    if int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs as a single process on one GPU", file=sys.stderr)
        sys.exit(2)

    import av1ongpu_amd as pkg

    if args.mix_only:
        print(json.dumps({"batched_mix": run_mix(args)}), flush=True)
        return
    if args.serve_only:
        print(json.dumps({"serving": run_serve(args), "serving_planned": run_serve(args, planned=True)}), flush=True)
        return
    dec = pkg.HipDecoder(0)
    log(f"device: {dec.device_info().name.decode(errors='replace')}")
    t = time.perf_counter()
    clip = pkg.Clip(args.workload, args.seed, **({"levels": 1} if args.levels else {}))
    n = clip.frames
    keep_slot, releases, ref_slots = slot_schedule(clip)
    wires = [pkg.wire_set_slots(clip.wire(f), ref_slots[f]) for f in range(n)]
    px = display_pixels(clip)
    log(f"clip {args.workload} seed {args.seed}: {n} frames, {px} display px/frame, "
        f"{sum(len(w) for w in wires) / 1e6:.0f} MB of syntax, generated in {time.perf_counter() - t:.1f} s")

    # every decode packs NV12 on the device as its last launch
    dec.set_decode_nv12(NV12_PITCH)
    # --- serial ingest: one frame at a time, flattening and upload inside
    t0 = time.perf_counter()
    for f in range(n):
        slot, _ = dec.decode_wire(wires[f], keep=keep_slot[f] >= 0, reserved=keep_slot[f], read_back=False)
        for s in releases[f]:
            dec.release(s)
    serial_s = time.perf_counter() - t0
    log(f"serial ingest pass: {serial_s * 1e3 / n:.2f} ms/frame")

    # --- the headline: every frame submitted as syntax (wire bytes) to the
    # ingest pipeline, planned by host threads while earlier frames decode
    planners = args.planners if args.planners > 0 else default_planners()
    pipe = pkg.Pipe(dec, planner_threads=planners, depth=args.depth)
    next_seq = [0]

    def submit_clip():
        base = next_seq[0]
        for f in range(n):
            pipe.submit(wires[f], keep=keep_slot[f] >= 0, reserved=keep_slot[f], release=releases[f])
        next_seq[0] = base + n
        return base

    for i in range(args.warmup):
        submit_clip()
        pipe.wait()
        log(f"warmup step {i + 1}/{args.warmup} done")
    # the timed steps: every step's frames submitted as the pipeline admits
    # them, one wait at the end (the clip restarts with a key frame each step)
    t0 = time.perf_counter()
    bases = [submit_clip() for _ in range(args.steps)]
    pipe.wait()
    elapsed = time.perf_counter() - t0
    log(f"{args.steps} steps done")
    recs = [pipe.result(b + f) for b in bases for f in range(n)]
    pipe_lat = [r.done_ms - r.submit_ms for r in recs]
    phases = {"flatten_ms": round(sum(r.flatten_ms for r in recs) / len(recs), 3),
              "upload_ms": round(sum(r.upload_ms for r in recs) / len(recs), 3),
              "launch_to_result_ms": round(sum(r.done_ms - r.decode_start_ms for r in recs) / len(recs), 3),
              "planned_to_launch_ms": round(sum(r.decode_start_ms - r.planned_ms for r in recs) / len(recs), 3),
              "worker_launch_ms": round(sum(r.launch_ms for r in recs) / len(recs), 3),
              "worker_wait_ms": round(sum(r.wait_ms for r in recs) / len(recs), 3),
              "device_ms": round(sum(r.device_ms for r in recs) / len(recs), 3)}
    log(f"pipeline per frame: {phases}")
    # the same pass with a device checksum of every picture (untimed), for
    # the comparison with the plain path below
    dec.set_output_hash(True)
    b = submit_clip()
    pipe.wait()
    pipe_hashes = [pipe.result(b + f).hash for f in range(n)]
    dec.set_output_hash(False)
    pipe.close()

    # --- device only: plans resident in HBM, one decode call per frame
    t = time.perf_counter()
    plans = [dec.plan_wire(w) for w in wires]
    log(f"planned {n} frames in {time.perf_counter() - t:.1f} s")

    def step(lat, hashes=None):
        for f in range(n):
            t1 = time.perf_counter()
            dec.decode_planned(plans[f], keep=keep_slot[f] >= 0, reserved=keep_slot[f])
            lat.append(time.perf_counter() - t1)
            if hashes is not None:
                hashes.append(dec.last_hash())
            for s in releases[f]:
                dec.release(s)

    step([])
    lat = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        step(lat)
    dev_elapsed = time.perf_counter() - t0
    # the same plans with two frames in flight (launch N + 1, then wait for
    # N): the device's own rate, which the pipeline approaches when the
    # planners keep ahead of it
    def step_queued():
        for f in range(n):
            pkg.decode_planned_batch_async([dec], [plans[f]], [keep_slot[f] >= 0], [keep_slot[f]])
            for s in releases[f]:
                dec.release(s)
            if pkg.batch_pending(dec) > 1:
                pkg.batch_finish(dec)
        while pkg.batch_pending(dec):
            pkg.batch_finish(dec)

    step_queued()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step_queued()
    queued_elapsed = time.perf_counter() - t0
    dec.set_output_hash(True)
    plain_hashes = []
    step([], plain_hashes)
    dec.set_output_hash(False)
    pipe_mism = sum(1 for a, b in zip(pipe_hashes, plain_hashes) if a != b)
    log(f"pipelined vs plain decode: {pipe_mism} of {n} pictures differ")
    # one more pass with HIP events around every stage (launched eagerly)
    # for the per stage times and the roofline
    dec.set_stage_timing(True)
    step([])
    stages = dec.stage_times()
    dec.set_stage_timing(False)
    for p in plans:
        dec.discard_plan(p)

    value = args.steps * n * px / elapsed / 1e6
    ms_step = elapsed * 1e3 / max(args.steps, 1)
    # dominant kernel: the stage with the most device time
    dom = max((s for s in stages if stages[s][1] > 0), key=lambda s: stages[s][0])
    ms, cnt, byts = stages[dom]
    avg_ms = ms / cnt
    achieved = byts / cnt / (avg_ms * 1e-3) / 1e9
    traffic = load_traffic(args.workload, STAGE_KERNEL[dom])
    roofline = {"bound": "hbm", "kernel": STAGE_KERNEL[dom], "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic, "algorithmic_bytes_per_launch": byts // cnt,
                "avg_launch_ms": round(avg_ms, 4)}
    stage_report = {s: {"avg_ms": round(v[0] / v[1], 4), "frames": v[1],
                        "gbs": round(v[2] / (v[0] * 1e-3) / 1e9, 1) if v[0] > 0 else None}
                    for s, v in stages.items() if v[1] > 0}

    result = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mpixel/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32",
        "data": "synthetic (seeded FrameSyntax of the named shape)",
        "config": {"workload": f"{args.workload}: {WORKLOADS[args.workload]}", "frames": n, "seed": args.seed,
                   "path": f"plain per-frame decode from submitted syntax (wire bytes in host memory): "
                           f"flattening and upload by {planners} planner threads overlapped with the "
                           f"decode of earlier frames (ingest pipeline, {args.depth} frames admitted, two "
                           f"on the device at a time), "
                           f"reference slots on the device, NV12 packed on the device"},
        "frame_latency_ms": {"p50": round(pct(pipe_lat, 50), 3), "p99": round(pct(pipe_lat, 99), 3),
                             "note": "submit to complete inside the pipeline, frames submitted as fast as "
                                     "the pipeline admits them"},
        "pipeline_phases": phases,
        "device_only": {"value": round(args.steps * n * px / dev_elapsed / 1e6, 2), "unit": "Mpixel/s",
                        "path": "the same clip from plans resident in HBM, one decode call per frame",
                        "call_latency_ms": {"p50": round(pct(lat, 50) * 1e3, 3),
                                            "p99": round(pct(lat, 99) * 1e3, 3)}},
        "device_queued": {"value": round(args.steps * n * px / queued_elapsed / 1e6, 2), "unit": "Mpixel/s",
                          "path": "the same plans resident in HBM, two frames in flight (frame N + 1 launched "
                                  "before frame N is waited for): the device's own rate"},
        "ingest_serial": {"value": round(n * px / serial_s / 1e6, 2), "unit": "Mpixel/s",
                          "path": "one thread: each frame flattened, uploaded and decoded before the next"},
        "pipeline_vs_plain": {"frames_checked": n, "mismatched_frames": pipe_mism,
                              "method": "device checksum of every picture, pipelined vs device-only decode"},
        "stages": stage_report,
        "roofline": roofline,
        "cpu_baseline": None,
    }

    if not args.no_cpu:
        # GPU hashes of the frames the CPU sample will cover (read back, untimed)
        import av1ongpu_amd as pkg2
        gpu_hashes = []
        t = time.perf_counter()
        for f, _slot, planes in pkg2.decode_clip(pkg2.HipDecoder(0), clip):
            gpu_hashes.append(pkg2.fnv1a64(pkg2.yuv16_bytes(planes)))
        log(f"GPU read back of {len(gpu_hashes)} frames for parity in {time.perf_counter() - t:.1f} s")
        base, checked, mism = cpu_baseline(clip, args.cpu_seconds, gpu_hashes)
        log(f"CPU oracle: {base['value']} Mpixel/s on {checked} frames, {len(mism)} mismatched")
        result["cpu_baseline"] = base
        result["bit_exact"] = {"frames_checked": checked, "mismatched_frames": len(mism),
                               "method": "FNV-1a-64 of every sample of each decoded frame, GPU vs CPU oracle"}
        if mism:
            log(f"MISMATCH on frames {mism[:10]}")
    if not args.no_mix:
        result["batched_mix"] = run_mix(args)
    if not args.no_serve:
        result["serving"] = run_serve(args)
        result["serving_planned"] = run_serve(args, planned=True)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
