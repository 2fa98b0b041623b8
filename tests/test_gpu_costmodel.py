"""The superblock cost model on the GPU: inference and gradients of the HIP
kernels against the numpy reference, training that lowers the loss, and
queue schedules from costs that leave every decoded sample unchanged."""
import numpy as np
import pytest

import av1ongpu_amd as pkg
from mlp_ref import forward, init_params, loss_and_grad, predict_us

pytestmark = pytest.mark.gpu


def _data(rng, n):
    X = rng.normal(0, 1, (n, pkg.CM_FEATURES)).astype(np.float32) * rng.uniform(0.5, 3, pkg.CM_FEATURES)
    w = rng.normal(0, 1, pkg.CM_FEATURES)
    y = (np.maximum(X @ w, 0) + 0.3 * (X[:, 0] * X[:, 1]) + 5).astype(np.float32)
    return X, y


def test_predict_matches_reference():
    rng = np.random.default_rng(1)
    cm = pkg.CostModel(seed=3)
    p = init_params(rng).astype(np.float32)
    p[24 * 64:24 * 64 + 64] = rng.normal(0, 0.2, 64)
    cm.set_params(p)
    mean = rng.normal(0, 1, 24).astype(np.float32)
    dev = rng.uniform(0.5, 2, 24).astype(np.float32)
    cm.set_norm(mean, dev, 7.0, 3.0)
    X, _ = _data(rng, 1000)
    got = cm.predict(X)
    want = predict_us(p, X, mean, dev, 7.0, 3.0)
    np.testing.assert_allclose(got, want, rtol=1e-4, atol=1e-4)


def test_gradients_match_reference():
    rng = np.random.default_rng(2)
    cm = pkg.CostModel(seed=5)
    p = init_params(rng).astype(np.float32)
    p[24 * 64:24 * 64 + 64] = rng.normal(0, 0.2, 64)
    cm.set_params(p)
    mean = np.zeros(24, np.float32)
    dev = np.ones(24, np.float32)
    cm.set_norm(mean, dev, 0.0, 1.0)
    X, y = _data(rng, 300)  # not a multiple of the 64 sample step
    y = (y - y.mean()) / y.std()
    g, loss = cm.gradients(X, y)
    want_loss, want = loss_and_grad(p, X.astype(np.float64), y.astype(np.float64))
    assert abs(loss - want_loss) <= 1e-4 * max(1.0, want_loss)
    # FP32 accumulation against the float64 reference
    np.testing.assert_allclose(g, want, rtol=2e-3, atol=2e-5 * np.abs(want).max())


def test_training_lowers_loss():
    rng = np.random.default_rng(3)
    X, y = _data(rng, 4096)
    cm = pkg.CostModel(seed=7)
    losses = cm.train(X, y, epochs=30, lr=3e-3, seed=1)
    assert losses[-1] < 0.25 * losses[0], losses
    pred = cm.predict(X)
    r = np.corrcoef(pred, y)[0, 1]
    assert r > 0.9, r


def _decode(dec, clip, plans, keep, rel):
    out = []
    for f in range(clip.frames):
        dec.decode_planned(plans[f], keep=keep[f] >= 0, reserved=keep[f])
        w, _cw, h, _bd, sx, sy = pkg.wire_header(clip.wire(f))
        out.append(dec.read_slot(-1, w, h, sx, sy))
        for s in rel[f]:
            dec.release(s)
    return out


@pytest.mark.parametrize("workload,seed,over", [("tools_mini", 1, {}), ("c1_mini", 2, dict(key_interval=1, frames=3)),
                                                ("contract_mini", 1, {})])
def test_schedules_leave_pixels_unchanged(workload, seed, over):
    """Static, random-cost and model-cost schedules, coding order and
    critical path order decode identically (intra block copy edges
    included)."""
    clip = pkg.Clip(workload, seed, **over)
    keep, rel, refs = pkg.slot_schedule(clip)
    dec = pkg.HipDecoder()
    plans = [dec.plan_wire(pkg.wire_set_slots(clip.wire(f), refs[f])) for f in range(clip.frames)]
    base = _decode(dec, clip, plans, keep, rel)
    rng = np.random.default_rng(seed)
    cm = pkg.CostModel(seed=seed)
    for mode in ("random", "model", "coding", "path", "path_random"):
        for p in plans:
            nsb = dec.sb_features(p).shape[0]
            if mode == "random":
                dec.schedule(p, rng.uniform(0, 50, nsb))
            elif mode == "model":
                dec.schedule_model(p, cm)
            elif mode == "coding":
                dec.schedule_order(p, pkg.ORDER_CODING)
            elif mode == "path":
                dec.schedule_order(p, pkg.ORDER_PATH)
            else:
                dec.schedule_order(p, pkg.ORDER_PATH, rng.uniform(0, 50, nsb))
        got = _decode(dec, clip, plans, keep, rel)
        for f in range(clip.frames):
            for k in range(3):
                assert np.array_equal(got[f][k], base[f][k]), f"{mode} frame {f} plane {k}"
    for p in plans:
        dec.schedule(p, None)
        dec.discard_plan(p)


def test_superblock_times_recorded():
    clip = pkg.Clip("c1_mini", 9, key_interval=1, frames=2)
    dec = pkg.HipDecoder()
    plan = dec.plan_wire(clip.wire(0))
    feats = dec.sb_features(plan)
    dec.set_queue_trace(True)
    dec.decode_planned(plan, keep=False)
    t = dec.sb_times(plan)
    dec.set_queue_trace(False)
    assert t.shape[0] == feats.shape[0] > 0
    assert (t >= 0).all() and t.sum() > 0
    # superblocks with queued samples have recorded work
    assert (t[feats[:, 14] > 0] > 0).all()
    dec.discard_plan(plan)


def test_orders_in_a_batch_leave_pixels_unchanged():
    """Frames of one batch in different queue orders (merged by remaining
    path instead of by level) decode as they do alone."""
    clips = [pkg.Clip("c1_mini", 21, key_interval=1, frames=2), pkg.Clip("tools_mini", 2),
             pkg.Clip("c2_mini", 5, frames=3)]
    decs = [pkg.HipDecoder() for _ in clips]
    for d in decs:
        d.set_output_hash(True)
    scheds = [pkg.slot_schedule(c) for c in clips]
    plans = [[decs[i].plan_wire(pkg.wire_set_slots(c.wire(f), scheds[i][2][f])) for f in range(c.frames)]
             for i, c in enumerate(clips)]
    from av1ongpu_amd import costmodel as cmod
    got = {}
    for name, order in (("level", None), ("coding", pkg.ORDER_CODING), ("path", pkg.ORDER_PATH)):
        for i, d in enumerate(decs):
            for p in plans[i]:
                if order is None:
                    d.schedule(p, None)
                else:
                    d.schedule_order(p, order)
        h = []
        n = max(c.frames for c in clips)
        for f in range(n):
            live = [i for i, c in enumerate(clips) if f < c.frames]
            pkg.decode_planned_batch([decs[i] for i in live], [plans[i][f] for i in live],
                                     [scheds[i][0][f] >= 0 for i in live], [scheds[i][0][f] for i in live])
            h.append([decs[i].last_hash() for i in live])
            for i in live:
                for s in scheds[i][1][f]:
                    decs[i].release(s)
        got[name] = h
    assert got["coding"] == got["level"] and got["path"] == got["level"]
    for i, d in enumerate(decs):
        for p in plans[i]:
            d.discard_plan(p)
