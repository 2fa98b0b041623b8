"""Training data and evaluation for the superblock cost model.

The decoder records its own superblock timings: a decode with the queue
trace on stamps every queued block when its wait ends and when it is
published; av1h_plan_sb_times sums the work (publish minus ready) per
superblock. Those times and the superblocks' syntax features
(av1h_plan_sb_features) are the samples the network (CostModel, trained by
the HIP kernels) learns from. A trained model orders the queue of a plan
through av1h_plan_schedule_model; the static dependency-level order stays the
default and the baseline the model is compared with.
"""
import time

import numpy as np

from . import CostModel, HipDecoder, decode_planned_batch, slot_schedule, wire_set_slots


def collect_samples(clips, max_samples=None, min_work_us=0.0):
    """Decode each clip alone, in order, with the queue trace on; returns
    (features (n x 24), microseconds (n)) for every superblock with queued
    work."""
    feats, times = [], []
    total = 0
    for clip in clips:
        dec = HipDecoder()
        dec.set_queue_trace(True)
        keep, rel, refs = slot_schedule(clip)
        for f in range(clip.frames):
            plan = dec.plan_wire(wire_set_slots(clip.wire(f), refs[f]))
            dec.decode_planned(plan, keep=keep[f] >= 0, reserved=keep[f])
            x = dec.sb_features(plan)
            t = dec.sb_times(plan)
            dec.discard_plan(plan)
            for s in rel[f]:
                dec.release(s)
            m = (x[:, 22] > 0) & (t > min_work_us)
            feats.append(x[m])
            times.append(t[m])
            total += int(m.sum())
            if max_samples and total >= max_samples:
                break
        dec.close()
        if max_samples and total >= max_samples:
            break
    return np.concatenate(feats), np.concatenate(times)


def train(X, y, epochs=20, lr=1e-3, seed=1, holdout=0.1):
    """Trains a CostModel on (X, y) with a seeded hold-out split; returns the
    model, the loss per epoch and the hold-out R^2 of its predictions."""
    rng = np.random.default_rng(seed)
    idx = rng.permutation(len(y))
    nh = int(len(y) * holdout)
    hold, fit = idx[:nh], idx[nh:]
    cm = CostModel(seed=seed)
    t0 = time.perf_counter()
    losses = cm.train(X[fit], y[fit], epochs=epochs, lr=lr, seed=seed)
    dt = time.perf_counter() - t0
    pred = cm.predict(X[hold])
    ss = float(np.sum((y[hold] - y[hold].mean()) ** 2))
    r2 = 1.0 - float(np.sum((pred - y[hold]) ** 2)) / ss if ss > 0 else 0.0
    return cm, losses, r2, dt


def schedule_all(decoders, plans, model=None, order=None):
    """Orders every plan's queue: by dependency level with the model's costs
    breaking ties inside a level (order None; model None: the static level
    order), or in `order` (ORDER_CODING, ORDER_LEVEL, ORDER_PATH) with the
    model's predicted superblock costs (model None: sample counts)."""
    from . import ORDER_LEVEL
    for dec, ps in zip(decoders, plans):
        for p in ps:
            if order is None:
                if model is None:
                    dec.schedule(p, None)
                else:
                    dec.schedule_model(p, model)
                continue
            costs = None
            if model is not None:
                x = dec.sb_features(p)
                costs = model.predict(x) if len(x) else None
            if order == ORDER_LEVEL and costs is None:
                dec.schedule(p, None)
            else:
                dec.schedule_order(p, order, costs)


def run_rounds(decoders, plans, scheds, hashes=None, timing=False):
    """Decodes frame r of every stream as batch r; returns (wall seconds,
    queue device ms summed over the batches)."""
    n = len(decoders)
    frames = len(plans[0])
    if timing:
        decoders[0].set_stage_timing(True)
        before = decoders[0].stage_times()["queue"][0]
    t0 = time.perf_counter()
    for f in range(frames):
        decode_planned_batch(decoders, [plans[i][f] for i in range(n)],
                             [scheds[i][0][f] >= 0 for i in range(n)], [scheds[i][0][f] for i in range(n)])
        if hashes is not None:
            hashes.append([d.last_hash() for d in decoders])
        for i in range(n):
            for s in scheds[i][1][f]:
                decoders[i].release(s)
    dt = time.perf_counter() - t0
    q = 0.0
    if timing:
        q = decoders[0].stage_times()["queue"][0] - before
        decoders[0].set_stage_timing(False)
    return dt, q
