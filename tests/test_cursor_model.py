"""A model of the tile kernel's pixel hand-out (tile_loop.hpp): one cursor
per XCD over interleaved groups of slots, waves that draw from their own
XCD's cursor and then from the next ones in turn.  Checked on the CPU, with
waves interleaved at random and pixels of random length: every slot is
handed out exactly once, the binary64 second pass takes exactly the unsure
pixels, and every wave ends (a wave ends once it has found all eight cursors
drained and written every stopped pixel), and every pixel's results are
written exactly once -- for empty, tiny and ragged launches and grids that do not fill
the eight XCDs."""
import random

import pytest


def group_shift(npts):
    # tile_loop.hpp: clamp(floor(log2(max(npts / 8, 1))), 6, 14)
    return min(14, max(6, max(npts >> 3, 1).bit_length() - 1))


def run(npts, nblocks, waves_per_block, seed, unsure=None, refill_batch=60):
    rnd = random.Random(seed)
    cursors = [0] * 8
    gsh = group_shift(npts)
    started, skipped, written = [], [], []
    waves = [dict(xcd=b & 7, cur=0, active=[False] * 64, left=[0] * 64, exhausted=[False] * 64,
                  pending=[None] * 64)
             for b in range(nblocks) for _ in range(waves_per_block)]
    live = list(range(len(waves)))
    trips = 0
    while live:
        k = rnd.randrange(len(live))
        w = waves[live[k]]
        trips += 1
        assert trips < 5_000_000, "a wave does not end"
        act = sum(w["active"])
        want = [l for l in range(64) if not w["active"][l]
                and (w["pending"][l] is not None or not w["exhausted"][l])]
        if act == 0 and not want:
            live.pop(k)
            continue
        if want and (act == 0 or 64 - act >= refill_batch):
            drained_any = False
            for lane in range(64):
                if not w["active"][lane] and w["pending"][lane] is not None:
                    written.append(w["pending"][lane])
                    w["pending"][lane] = None
            idle = [l for l in range(64) if not w["active"][l] and not w["exhausted"][l]]
            if idle:
                cx = (w["xcd"] + w["cur"]) & 7
                base = cursors[cx]
                cursors[cx] += len(idle)
                for rank, lane in enumerate(idle):
                    loc = base + rank
                    idx = ((((loc >> gsh) << 3) + cx) << gsh) + (loc & ((1 << gsh) - 1))
                    if idx < npts:
                        if unsure is None or idx in unsure:
                            started.append(idx)
                            w["active"][lane] = True
                            w["left"][lane] = rnd.randint(1, 40)
                            w["pixel"] = w.get("pixel", {})
                            w["pixel"][lane] = idx
                        else:
                            skipped.append(idx)
                    else:
                        drained_any = True
                        w["exhausted"][lane] = w["cur"] >= 7
            if drained_any:
                w["cur"] += 1
        for lane in range(64):
            if w["active"][lane]:
                w["left"][lane] -= 1
                if w["left"][lane] == 0:
                    w["active"][lane] = False
                    w["pending"][lane] = w["pixel"][lane]
    assert sorted(written) == sorted(started), "a stopped pixel was not written exactly once"
    return started, skipped


@pytest.mark.parametrize("npts,nblocks,waves", [
    (0, 8, 4), (1, 3, 2), (63, 8, 2), (64 * 8 * 5 + 17, 16, 4), (5000, 5, 3), (777, 1, 1),
    (100000, 32, 4), (300000, 8, 8),
])
def test_every_slot_once_and_every_wave_ends(npts, nblocks, waves):
    for seed in range(2):
        started, skipped = run(npts, nblocks, waves, seed)
        assert sorted(started) == list(range(npts)) and not skipped


def test_second_pass_takes_the_unsure_pixels_only():
    rnd = random.Random(3)
    unsure = set(rnd.sample(range(20000), 6000))
    started, skipped = run(20000, 12, 4, 7, unsure=unsure)
    assert sorted(started) == sorted(unsure)
    assert sorted(started + skipped) == list(range(20000))


def test_group_shift():
    assert [group_shift(n) for n in (0, 511, 4096, 40000, 65536, 131072, 4194304)] == [6, 6, 9, 12, 13, 14, 14]
