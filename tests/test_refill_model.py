"""CPU model of the register lane-refill buffer (BufferType.REG) of one
wavefront: the push (ds_permute to computed lanes) and the pop (ds_bpermute
from computed lanes) are simulated lane by lane with the same index
formulas the translator emits, over random activity patterns.  Checks that
the push permutation is total (every lane sends to a distinct lane, so no
two values collide), that the buffer never holds more than 64 entries, and
that every tuple that reaches the refill point runs the operators above it
exactly once, also when the input ends with entries buffered."""
import numpy as np
import pytest

W = 64


def lanesBelow(mask, lane):
    return sum(1 for j in range(lane) if mask[j])


def pushPermutation(active, cnt):
    """Destination lane of every lane, as emitted:
    (cnt + (active ? lb : rna + (lane - lb))) & 63."""
    rna = sum(active)
    dst = []
    for lane in range(W):
        lb = lanesBelow(active, lane)
        dst.append((cnt + (lb if active[lane] else rna + (lane - lb))) & 63)
    return dst


def runWave(trips, thr, rng):
    """trips: list of per-trip active vectors (tuple ids or None).  Returns
    the tuple ids processed by the operators above the refill point."""
    buf = [None] * W
    cnt = 0
    seen = []
    ntrips = len(trips)
    for t in range(ntrips + 1):
        flush = t == ntrips             # the trip past the end (more == false)
        regs = trips[t] if not flush else [None] * W
        active = [r is not None for r in regs]
        rna = sum(active)
        go = cnt + rna >= thr or (flush and cnt + rna > 0)
        if not go:
            assert cnt + rna < thr <= W
            dst = pushPermutation(active, cnt)
            assert sorted(dst) == list(range(W)), "push permutation not total"
            recv = [None] * W
            for lane in range(W):
                recv[dst[lane]] = regs[lane]
            for lane in range(W):
                if cnt <= lane < cnt + rna:
                    buf[lane] = recv[lane]
                    assert buf[lane] is not None
            cnt += rna
            assert cnt <= W
            continue
        while go:
            active = [r is not None for r in regs]
            rna = sum(active)
            take = min(cnt, W - rna)
            newregs = list(regs)
            for lane in range(W):
                rs = lane - lanesBelow(active, lane)
                dest = (not active[lane]) and rs < take
                src = cnt - take + rs if dest else lane
                assert 0 <= src < W
                if dest:
                    assert buf[src] is not None
                    newregs[lane] = buf[src]
            for k in range(cnt - take, cnt):
                buf[k] = None
            cnt -= take
            seen.extend(r for r in newregs if r is not None)
            regs = [None] * W
            go = cnt >= thr or (flush and cnt > 0)
    assert cnt == 0
    return seen


@pytest.mark.parametrize("thr", [1, 17, 57, 64])
@pytest.mark.parametrize("density", [0.02, 0.3, 0.9, 1.0])
def test_register_refill_model(thr, density):
    rng = np.random.default_rng(thr * 100 + int(density * 100))
    trips, nxt = [], 0
    for _ in range(40):
        act = rng.random(W) < density
        regs = []
        for a in act:
            regs.append(nxt if a else None)
            nxt += 1 if a else 0
        trips.append(regs)
    want = sorted(r for trip in trips for r in trip if r is not None)
    got = runWave(trips, thr, rng)
    assert sorted(got) == want
