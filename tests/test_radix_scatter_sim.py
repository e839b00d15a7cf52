"""CPU simulation of the binning radix passes' indexing (radix_hist_block and
radix_scatter_block in malevich_amd/csrc/mlv_kernels.hip).

The GPU kernels are followed step by step with the same block shape (kRadixItems = 4096 items
per block, NT threads, 64-lane waves, R = 4096 / NT rounds): the digit-major histogram and its
exclusive scan, the per-wave ballot ranks, the per-round prefix of the waves' digit counts (which
live in LDS between barriers), the LDS slot of every item and the output position the write-out
computes. The checks:
  - every LDS slot of a block is written exactly once and lies below the block's item count;
  - every output position lies in [0, n) and is written exactly once over all blocks;
  - the output is the stable sort of the input by the pass's digit.
The rounds are simulated in the order the kernel's barriers impose: a round's digit counts are
cleared only after every wave of the previous round has read its offsets. This is the order whose
loss (a dropped barrier) once let slower waves read cleared offsets and scatter out of range.
"""
import numpy as np
import pytest

ITEMS = 4096  # kRadixItems = 256 * kRadixRounds


def hist_digit_major(keys, n, shift):
    nb = (n + ITEMS - 1) // ITEMS
    h = np.zeros(256 * nb, np.int64)
    for b in range(nb):
        d = (keys[b * ITEMS:min(n, (b + 1) * ITEMS)] >> shift) & 255
        h[np.arange(256) * nb + b] = np.bincount(d, minlength=256)
    return h, nb


def scatter_block(kin, vin, n, nb, b, shift, hist, NT, kout, vout, written):
    NW, R = NT // 64, ITEMS // NT
    base = b * ITEMS
    # goff / loff / lbeg: thread t < 256 takes digit t; its count is the step of the scanned histogram
    idx = np.arange(256) * nb + b
    here = hist[idx]
    nxt = np.where(idx + 1 < 256 * nb, hist[np.minimum(idx + 1, 256 * nb - 1)], n)
    cnt_d = nxt - here
    ex = np.concatenate([[0], np.cumsum(cnt_d)[:-1]])
    goff, loff, lbeg = here.copy(), ex.copy(), ex.copy()
    cnt = min(ITEMS, n - base) if n > base else 0
    s_key = np.full(ITEMS, -1, np.int64)
    s_val = np.full(ITEMS, -1, np.int64)
    slot_used = np.zeros(ITEMS, bool)
    for r in range(R):
        wc = np.zeros((NW, 256), np.int64)  # cleared, then a barrier
        tids = np.arange(NT)
        i = base + r * NT + tids
        valid = i < n
        keys = np.where(valid, kin[np.minimum(i, n - 1)], 0)
        vals = np.where(valid, vin[np.minimum(i, n - 1)], 0)
        dg = (keys >> shift) & 255
        rank = np.zeros(NT, np.int64)
        for w in range(NW):
            lanes = slice(64 * w, 64 * w + 64)
            v, d = valid[lanes], dg[lanes]
            for lane in range(64):
                # m = ballot(valid) restricted to lanes whose 8 digit bits all equal this lane's
                m = v & (d == d[lane])
                rank[64 * w + lane] = int(m[:lane].sum())  # mbcnt
                if v[lane] and rank[64 * w + lane] == 0:
                    wc[w, d[lane]] = int(m.sum())  # the digit's first lane writes the wave's count
        # barrier; threads t < 256 turn the waves' counts into offsets in wave order
        for t in range(256):
            e = loff[t]
            for q in range(NW):
                c = wc[q, t]
                wc[q, t] = e
                e += c
            loff[t] = e
        # barrier; every valid item takes its LDS slot
        for t in np.nonzero(valid)[0]:
            lp = wc[t // 64, dg[t]] + rank[t]
            assert 0 <= lp < cnt, (b, r, t, lp, cnt)
            assert not slot_used[lp], (b, r, t, lp)
            slot_used[lp] = True
            s_key[lp] = keys[t]
            s_val[lp] = vals[t]
        # barrier before the next round clears wc
    assert slot_used[:cnt].all() and not slot_used[cnt:].any()
    for j in range(cnt):
        key = s_key[j]
        d = (key >> shift) & 255
        pos = goff[d] + (j - lbeg[d])
        assert 0 <= pos < n, (b, j, pos, n)
        assert not written[pos], (b, j, pos)
        written[pos] = True
        kout[pos] = key
        vout[pos] = s_val[j]


def run_pass(kin, vin, shift, NT):
    n = kin.size
    h, nb = hist_digit_major(kin, n, shift)
    hist = np.concatenate([[0], np.cumsum(h)[:-1]])  # the exclusive scan over all 256 * nb counters
    kout = np.full(n, -1, np.int64)
    vout = np.full(n, -1, np.int64)
    written = np.zeros(n, bool)
    for b in range(nb):
        scatter_block(kin, vin, n, nb, b, shift, hist, NT, kout, vout, written)
    assert written.all()
    return kout, vout


def _keys(kind, n, rng):
    if kind == "random":
        return rng.integers(0, 1 << 16, n)
    if kind == "one_digit":
        return np.full(n, 0x1234, np.int64)
    if kind == "sorted":
        return np.sort(rng.integers(0, 1 << 16, n))
    if kind == "reversed":
        return np.sort(rng.integers(0, 1 << 16, n))[::-1].copy()
    if kind == "few":
        return rng.choice(np.array([3, 700, 0xFFFF, 256]), n)
    raise ValueError(kind)


@pytest.mark.parametrize("NT", [256, 1024])
@pytest.mark.parametrize("kind,n", [("random", 2 * ITEMS + 777), ("one_digit", ITEMS + 1), ("sorted", 3 * ITEMS),
                                    ("reversed", ITEMS - 5), ("few", 2 * ITEMS + 64)])
def test_radix_scatter_indexing(NT, kind, n):
    rng = np.random.default_rng(n + NT)
    keys = _keys(kind, n, rng).astype(np.int64)
    vals = np.arange(n, dtype=np.int64)  # the pair's position: stability is visible in the values
    k, v = keys, vals
    for shift in (0, 8):  # two passes over 16-bit tile keys, as a 1080p flush runs
        k, v = run_pass(k, v, shift, NT)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(k, keys[order])
    assert np.array_equal(v, vals[order])


def test_radix_hist_counts_are_digit_major():
    rng = np.random.default_rng(3)
    n = 3 * ITEMS + 10
    keys = rng.integers(0, 1 << 16, n)
    h, nb = hist_digit_major(keys, n, 8)
    assert nb == 4 and h.sum() == n
    for b in range(nb):
        blk = (keys[b * ITEMS:(b + 1) * ITEMS] >> 8) & 255
        assert np.array_equal(h[np.arange(256) * nb + b], np.bincount(blk, minlength=256))
