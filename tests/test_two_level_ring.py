"""The two-level step keeps the inner level's output rows in a ring of
kRingRows (8) rows per wave (bmi_fused.hip, process_item2).  This emulates
the row schedule of process_item2 on the CPU, rows only: the final level's
pipeline reads its LL row one pair ahead (read_ll), the inner level runs
two steps (four rows) at a time until that row exists, odd segments run
bottom-up (5/3 ones only where neither pipeline's range holds the pair of
the reference's bottom-edge rule); 5/3 pipelines lag one pair (two warm-up
pairs where that rule applies).  For every item of every geometry tried,
the rows written since the row being read (inclusive) must fit in the
ring, with a margin: at most
4 here for 9/7 and 6 for 5/3 (its shorter lag lets two inner steps run
further ahead), against 8 slots (a wider sweep of heights up to 2160 and
segment heights up to 64 finds the same maxima)."""
import pathlib
import re

import pytest

RING_ROWS = 8
MAX_SEEN = {"97": 4, "53": 6}


def _level_span(t0, n, L, lev):
    s = L - 1 - lev
    return -((-t0) >> s), -((-(t0 + n)) >> s)


def _item_reach(b0, b1, ib0, ib1, seg_pairs, seg, rev, lag=2, warm=2, iwarm=2, quirk=0, iquirk=0):
    """Largest number of ring rows a read of this item depends on.  With
    rev None the direction is the kernel's rule for a 5/3 item."""
    K0, Klast = (b0 & ~1) >> 1, (b1 - 1) >> 1
    kout0 = K0 + seg * seg_pairs
    kout1 = min(Klast, kout0 + seg_pairs - 1)
    ks, ke = kout0 - warm, kout1 + lag
    nrefl = 1 if (b1 - b0) >= 8 else 3

    def reflect_row(r):
        for _ in range(nrefl):
            r = 2 * b0 - r if r < b0 else r
            r = 2 * (b1 - 1) - r if r >= b1 else r
        return r

    rl0 = -((-b0) // 2)

    def ll_row(k):
        return (reflect_row(2 * k) >> 1) - rl0

    rows = [ll_row(k) for k in range(ks, ke + 1)]
    rmin, rmax = min(rows), max(rows)
    qo0, qo1 = (ib0 + rmin) >> 1, (ib0 + rmax) >> 1
    iqs, iqe = qo0 - iwarm, qo1 + lag
    if rev is None:
        kq = K0 + ((b1 - b0) >> 1) if quirk else None
        ikq = ((ib0 & ~1) >> 1) + ((ib1 - ib0) >> 1) if iquirk else None
        rev = bool(seg & 1) and (kq is None or not ks <= kq <= ke) and \
            (ikq is None or not iqs <= ikq <= iqe)
    st = {"q": iqe if rev else iqs, "have": rmin, "low": rmax + 1}
    reach = 0

    def inner_step():
        po = st["q"] + lag if rev else st["q"] - lag
        r0 = 2 * po - ib0
        if qo0 <= po <= qo1:
            st["have"] = max(st["have"], r0 + 2)
            st["low"] = min(st["low"], r0)
        st["q"] += -1 if rev else 1

    def read_ll(k):
        nonlocal reach
        rk = ll_row(min(max(k, ks), ke))
        if rev:
            while st["low"] > rk and st["q"] >= iqs:
                inner_step()
                inner_step()
            assert st["low"] <= rk
            reach = max(reach, rk - st["low"] + 1)
        else:
            while st["have"] <= rk and st["q"] <= iqe:
                inner_step()
                inner_step()
            assert st["have"] > rk
            reach = max(reach, st["have"] - rk)

    read_ll(ke if rev else ks)
    D = 2
    for it in range((ke - ks + D) // D):
        for d in range(D):
            k = ke - (it * D + d) if rev else ks + it * D + d
            read_ll(k - 1 if rev else k + 1)
    return reach


@pytest.mark.parametrize("filt", ["97", "53"])
@pytest.mark.parametrize("L", [2, 3, 4, 5, 6])
def test_two_level_ring_holds_every_row_read(L, filt):
    worst = 0
    for lev in range(1, L):
        for n in range(2, 97):
            for t0 in range(0, 8):
                b0, b1 = _level_span(t0, n, L, lev)
                ib0, ib1 = _level_span(t0, n, L, lev - 1)
                if b1 - b0 < 2 or ib1 - ib0 < 2:
                    continue
                for sp in (1, 2, 3, 5, 16, 64):
                    K0, Klast = (b0 & ~1) >> 1, (b1 - 1) >> 1
                    for seg in range((Klast - K0 + sp) // sp):
                        if filt == "97":
                            worst = max(worst, _item_reach(b0, b1, ib0, ib1, sp, seg, bool(seg & 1)))
                        else:
                            for quirk, iquirk in ((0, 0), (1, 1), (1, 0), (0, 1)):
                                worst = max(worst, _item_reach(b0, b1, ib0, ib1, sp, seg, None, lag=1,
                                                               warm=1 + quirk, iwarm=1 + iquirk,
                                                               quirk=quirk, iquirk=iquirk))
    assert worst <= MAX_SEEN[filt] < RING_ROWS, worst


def test_ring_size_matches_the_kernel():
    src = (pathlib.Path(__file__).resolve().parents[1] / "jpeg2000-decoder_amd" / "csrc" /
           "bmi_fused.hip").read_text()
    m = re.search(r"constexpr int kRingRows = (\d+);", src)
    assert m and int(m.group(1)) == RING_ROWS
