"""CPU model of the batched kernel's per-workgroup stream (kami_amd/csrc/batched.hip):
queue batches fetched in steps at tile boundaries, a cursor that loads K stages up to
three ahead into two register slots and the LDS stage buffers, and the boundary repair of
a blocked cursor.  The model follows the kernel's control flow statement by statement
(stage loop, advance_slot, load_at_cursor, ensure_ready, the boundary block) and checks,
for random queues of tiles with 1 to 16 K stages, that every stage of every tile is
computed exactly once and in order and that every loop ends.  The last test keeps the
repair loop's former condition (without the blocked check) and shows it does not end."""
import random

import pytest


class Hang(Exception):
    pass


def run_stream(tiles_nch, qb, repair_checks_blocked=True, limit=200000):
    n = len(tiles_nch)
    lead = max(1, qb - 3)                    # the kernel's lead tile of a batch
    counter = [0]                            # the queue counter (one workgroup)
    recs = {}                                # stream tile -> stage count (None: invalid)
    st = dict(t0=0, tc=0, cc=0, blocked=False, x=0, loaded=0, fetched=0, phase=-1,
              fb=0, fc=0)
    slot = {0: None, 1: None}
    lds = {0: None, 1: None}
    ct = [None]
    computed = []
    steps = [0]

    def tick():
        steps[0] += 1
        if steps[0] > limit:
            raise Hang()

    def fetch_step():
        if st["phase"] == 0:                 # atomic add
            b = counter[0]
            counter[0] += qb
            st["fb"], st["fc"] = b, max(0, min(qb, n - b))
            st["phase"] = 1
        elif st["phase"] == 1:               # entry loads
            st["phase"] = 2
        else:                                # LDS record write
            for e in range(qb):
                recs[st["fetched"] + e] = tiles_nch[st["fb"] + e] if e < st["fc"] else None
            st["phase"] = 3

    def fetch_commit():
        if st["phase"] == 3:
            st["fetched"] += qb
            st["phase"] = -1

    def ensure_ready(s):
        while st["fetched"] <= s:
            tick()
            if st["phase"] < 0:
                st["phase"] = 0
            fetch_step()
            fetch_commit()

    def stream_tile(s):
        return recs.get(s) if s < st["fetched"] else None

    def load_at_cursor():
        if ct[0] is None or st["blocked"]:
            return None
        stage = (st["tc"], st["cc"])
        st["loaded"] += 1
        st["cc"] += 1
        if st["cc"] >= ct[0]:
            st["cc"] = 0
            st["tc"] += 1
            if st["tc"] < st["fetched"]:
                ct[0] = stream_tile(st["tc"])
            else:
                st["blocked"] = True
        return stage

    def load_next_into_slot():
        s = 1 if st["loaded"] & 1 else 0     # the slot is chosen before the load
        slot[s] = load_at_cursor()

    def stage(p):                            # compute from LDS p, then advance_slot(1 - p)
        computed.append(lds[p])
        q = 1 - p
        if slot[q] is not None:
            lds[q] = slot[q]
        slot[q] = load_at_cursor()
        st["x"] += 1

    ensure_ready(0)
    cur = stream_tile(0)
    if cur is None:
        return computed
    ct[0] = cur
    slot[0] = load_at_cursor()
    lds[0] = slot[0]
    slot[1] = load_at_cursor()
    slot[0] = load_at_cursor()
    while True:
        tick()
        c = 0
        if (st["x"] & 1) and c < cur:
            stage(1)
            c += 1
        while c + 1 < cur:
            stage(0)
            stage(1)
            c += 2
        if c < cur:
            stage(0)
        st["t0"] += 1
        t0 = st["t0"]
        if st["phase"] < 0 and (t0 & (qb - 1)) == lead and st["fetched"] == t0 - lead + qb:
            st["phase"] = 0
        if st["phase"] >= 0:
            fetch_step()
            fetch_commit()
        ensure_ready(t0)
        cur = stream_tile(t0)
        if cur is None:
            break
        if st["blocked"]:
            ensure_ready(st["tc"])
            ct[0] = stream_tile(st["tc"])
            st["blocked"] = False
            if st["loaded"] == st["x"]:
                load_next_into_slot()
                p = st["x"] & 1
                lds[p] = slot[p]
                slot[p] = None
            while (st["loaded"] < st["x"] + 3 and ct[0] is not None
                   and (not repair_checks_blocked or not st["blocked"])):
                tick()
                load_next_into_slot()
    return computed


def expected(tiles_nch):
    return [(t, c) for t, nch in enumerate(tiles_nch) for c in range(nch)]


@pytest.mark.parametrize("qb", [2, 4, 8])
def test_every_stage_once_in_order(qb):
    rng = random.Random(qb)
    for _ in range(300):
        tiles = [rng.choice([1, 1, 1, 2, 3, 8, 16]) for _ in range(rng.randint(1, 60))]
        assert run_stream(tiles, qb) == expected(tiles)


def test_repair_without_the_blocked_check_does_not_end():
    tiles = [1, 1, 1, 1, 1, 8, 1, 1, 1, 1, 3, 1, 1, 1]
    assert run_stream(tiles, 2) == expected(tiles)
    rng = random.Random(5)
    hung = 0
    for _ in range(100):
        tiles = [rng.choice([1, 1, 1, 2, 3, 8, 16]) for _ in range(rng.randint(10, 60))]
        try:
            run_stream(tiles, 2, repair_checks_blocked=False)
        except Hang:
            hung += 1
    assert hung > 0
