"""Launch-shape autotuner for batched decodes.

Every kernel of a batch runs as a persistent grid of (workgroups per compute
unit x compute units) workgroups that walk the batch's virtual workgroups;
the dependency queue has its own grid; CDEF can stage 1 to 4 sets of blocks per
workgroup. The best values depend on the batch (frame sizes, tools) and the
device, so they are measured: coordinate descent over the parameters, each
candidate timed on the caller's own batches (replayed, captured launches
included), keeping a change only when it is faster by more than `tol`.
"""
import time

from . import decode_planned_batch

CANDIDATES = {"cdef_blocks_per_thread": (1, 2, 4), "queue": (2, 4, 6, 8, 12), "grain_template": (1, 2)}
DEFAULT_CANDIDATES = (2, 4, 8, 16, 32)


def _time_batches(decoders, batches, reps):
    """Wall time of decoding `batches` (lists of (plans, keep, reserved,
    releases)) `reps` times, after one untimed pass."""
    best = None
    for r in range(reps + 1):
        t0 = time.perf_counter()
        for plans, keep, res, rel in batches:
            decode_planned_batch(decoders, plans, keep, res)
            for d, slots in zip(decoders, rel):
                for s in slots:
                    d.release(s)
        dt = time.perf_counter() - t0
        if r > 0:
            best = dt if best is None else min(best, dt)
    return best


def tune(decoders, batches, params=None, reps=2, tol=0.01, log=None):
    """Tunes the launch shapes of decoders[0] (the batches' launcher) on
    `batches`; returns (chosen {name: value}, baseline seconds, tuned
    seconds)."""
    lead = decoders[0]
    cur = lead.tuning()
    # CDEF's chroma blocks run in the luma launch: its own shape does nothing
    names = params or [n for n in cur if n != "cdef_chroma"]
    base = _time_batches(decoders, batches, reps)
    best_t = base
    for name in names:
        cands = CANDIDATES.get(name, DEFAULT_CANDIDATES)
        best_v = cur[name]
        for v in cands:
            if v == cur[name]:
                continue
            lead.set_tuning({name: v})
            t = _time_batches(decoders, batches, reps)
            if t < best_t * (1 - tol):
                best_t, best_v = t, v
        cur[name] = best_v
        lead.set_tuning({name: best_v})
        if log:
            log(f"autotune: {name} = {best_v} ({best_t * 1e3:.2f} ms)")
    return cur, base, best_t
