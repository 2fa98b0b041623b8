"""The intra queue's packed edge protocol, as the plan describes it
(av1h_plan_queue_counts): which queued blocks take their edges as packed
words instead of waiting on dependency cells, which publish words, and which
need no cell marks. The pictures themselves are held to the reference's
hashes by test_gpu_golden.py; these check the plan's choices."""
import pytest

import av1ongpu_amd as pkg

pytestmark = pytest.mark.gpu


def _plans(clip, dec):
    keep, rel, refs = pkg.slot_schedule(clip, 8)
    return [dec.plan_wire(pkg.wire_set_slots(clip.wire(f), refs[f])) for f in range(clip.frames)]


@pytest.mark.parametrize("workload,seed", [("c1_mini", 1), ("c2_mini", 2), ("tools_mini", 1), ("contract_mini", 2)])
def test_queue_counts_are_consistent(workload, seed):
    clip = pkg.Clip(workload, seed)
    dec = pkg.HipDecoder()
    plans = _plans(clip, dec)
    try:
        for f, p in enumerate(plans):
            c = dec.queue_counts(p)
            assert 0 <= c["packed_readers"] <= c["blocks"], (f, c)
            assert 0 <= c["publishers"] <= c["blocks"], (f, c)
            assert 0 <= c["unmarked"] <= c["blocks"], (f, c)
            if c["packed_readers"] == c["blocks"]:
                # nobody waits on cells, so no queued block marks them
                assert c["unmarked"] == c["blocks"], (f, c)
            if c["blocks"]:
                assert pkg.hip_lib().av1h_plan_levels(p) >= 1
    finally:
        for p in plans:
            dec.discard_plan(p)
        dec.close()


def test_key_frames_of_plain_intra_tools_are_all_packed():
    """Intra modes without chroma-from-luma or palette (the 720p workload's
    statistics): every queued block of the key frame reads packed edges."""
    clip = pkg.Clip("c1_mini", 3)
    dec = pkg.HipDecoder()
    plans = _plans(clip, dec)
    try:
        c = dec.queue_counts(plans[0])
        assert c["blocks"] > 0
        assert c["packed_readers"] == c["blocks"], c
        assert c["publishers"] > 0
    finally:
        for p in plans:
            dec.discard_plan(p)
        dec.close()


def test_cfl_and_palette_blocks_keep_the_cell_wait():
    """Chroma from luma reads the co-located luma block and palette blocks
    read no edges through words: such blocks are not packed readers, and the
    blocks they wait on keep their cell marks."""
    clip = pkg.Clip("contract_mini", 2)
    dec = pkg.HipDecoder()
    plans = _plans(clip, dec)
    try:
        counts = [dec.queue_counts(p) for p in plans]
        assert any(c["packed_readers"] < c["blocks"] for c in counts), counts
        assert any(c["unmarked"] < c["blocks"] for c in counts), counts
    finally:
        for p in plans:
            dec.discard_plan(p)
        dec.close()


def test_a_wait_that_gives_up_fails_the_decode_and_the_next_is_clean():
    """The queue's time-out path, forced from the host with a one-poll wait
    limit on a key frame (whose blocks wait on one another): the decode fails
    with the queue's error and gives its slot back; the batch's last launch
    clears the error word, so with the limit restored the clip decodes bit
    exact against the oracle on the same context."""
    from oracle_lib import OracleDecoder
    clip = pkg.Clip("c2_mini", 1)
    dec = pkg.HipDecoder()
    dec.set_wait_limit(1)
    with pytest.raises(pkg.HipError, match="waited too long"):
        dec.decode_wire(pkg.wire_set_slots(clip.wire(0), []), keep=True)
    dec.set_wait_limit(0)
    got = [pkg.picture_hash(planes) for _, _, planes in pkg.decode_clip(dec, clip)]
    want = [pkg.picture_hash(planes) for _, _, planes in pkg.decode_clip(OracleDecoder(), clip)]
    assert got == want
    dec.close()
