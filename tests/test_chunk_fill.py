"""CPU model of chunked direct output (DirectSink::flush / direct_flush: rows reserved `chunk` at
a time, a wave's buffer split between the rest of its chunk and a new one, the unused tail of
each wave's last chunk recorded) and of the fill-up pass (chunk_fill_kernel in prims.hip: every
row found at or past the result count n moves, in order, to the r-th tail slot below n). Random
interleavings of waves and buffer sizes; the first n rows must then hold every result exactly
once. The index arithmetic is the device code's, line for line."""
import random

import pytest


def fill_up(mem, tail, chunk, n, reserved):
    """chunk_fill_kernel's mapping, applied to a dict position -> row."""
    nch = reserved // chunk
    tpre = [0] * nch
    for c in range(1, nch):
        tpre[c] = tpre[c - 1] + tail[c - 1]
    cn = n // chunk
    below = tpre[cn] if cn < nch else tpre[-1] + tail[-1]
    if cn < nch:
        keep, o = chunk - tail[cn], n % chunk
        below += o - keep if o > keep else 0
    for p in range(n, reserved):
        c, o = divmod(p, chunk)
        if o >= chunk - tail[c]:
            continue  # p is in a tail
        r = (p - n) - (tpre[c] - below)
        lo, hi = 0, (cn + 1 if cn < nch else nch)
        while hi - lo > 1:
            mid = (lo + hi) // 2
            if tpre[mid] <= r:
                lo = mid
            else:
                hi = mid
        dst = lo * chunk + (chunk - tail[lo]) + (r - tpre[lo])
        assert dst < n and dst not in mem
        mem[dst] = mem.pop(p)


def run(seed):
    rnd = random.Random(seed)
    chunk = rnd.choice([256, 512, 1024, 4096])
    waves = rnd.randint(1, 40)
    outs = [rnd.choice([0, rnd.randint(1, 300), rnd.randint(1, 5000)]) for _ in range(waves)]
    counter, mem, tails = 0, {}, {}
    state = [dict(cur=None, used=0, left=o, id=i, k=0) for i, o in enumerate(outs)]
    active = [s for s in state if s["left"] > 0]
    while active:
        s = rnd.choice(active)
        fill = min(s["left"], rnd.randint(1, 256))  # one buffer flush
        rem = 0 if s["cur"] is None else chunk - s["used"]
        na = min(fill, rem)
        base_a = (s["cur"] or 0) + s["used"]
        base_b = 0
        if fill > rem:  # one reservation of a whole chunk
            base_b = counter
            counter += chunk
            s["cur"], s["used"] = base_b, fill - na
        else:
            s["used"] += fill
        for i in range(fill):
            pos = base_a + i if i < na else base_b + (i - na)
            assert pos not in mem
            mem[pos] = (s["id"], s["k"])
            s["k"] += 1
        s["left"] -= fill
        if s["left"] == 0:
            active.remove(s)
            tails[s["cur"] // chunk] = chunk - s["used"]
    n, reserved = sum(outs), counter
    if reserved > n:
        fill_up(mem, [tails.get(c, 0) for c in range(reserved // chunk)], chunk, n, reserved)
    assert sorted(mem) == list(range(n))
    assert sorted(mem.values()) == sorted((i, k) for i, o in enumerate(outs) for k in range(o))


@pytest.mark.parametrize("block", range(4))
def test_chunked_output_and_fill_up_are_dense(block):
    for seed in range(block * 150, (block + 1) * 150):
        run(seed)
