"""Tier-1 pins on the CPU: the oracle's block decoder (restatement of
decode_code_block) and the project's encoder against ITU-T T.800 Tables
C.2, D.1 and D.3 typed here from the standard, and against each other:
encoder -> decoder round trips must return the quantisation indices, and
truncated streams the closed-form mid-point reconstruction."""
import numpy as np
import pytest

RNG = np.random.default_rng(1234)

# ITU-T T.800 Table C.2: (Qe, NMPS, NLPS, SWITCH) per index
T800_C2 = [
    (0x5601, 1, 1, 1), (0x3401, 2, 6, 0), (0x1801, 3, 9, 0), (0x0AC1, 4, 12, 0),
    (0x0521, 5, 29, 0), (0x0221, 38, 33, 0), (0x5601, 7, 6, 1), (0x5401, 8, 14, 0),
    (0x4801, 9, 14, 0), (0x3801, 10, 14, 0), (0x3001, 11, 17, 0), (0x2401, 12, 18, 0),
    (0x1C01, 13, 20, 0), (0x1601, 29, 21, 0), (0x5601, 15, 14, 1), (0x5401, 16, 14, 0),
    (0x5101, 17, 15, 0), (0x4801, 18, 16, 0), (0x3801, 19, 17, 0), (0x3401, 20, 18, 0),
    (0x3001, 21, 19, 0), (0x2801, 22, 19, 0), (0x2401, 23, 20, 0), (0x2201, 24, 21, 0),
    (0x1C01, 25, 22, 0), (0x1801, 26, 23, 0), (0x1601, 27, 24, 0), (0x1401, 28, 25, 0),
    (0x1201, 29, 26, 0), (0x1101, 30, 27, 0), (0x0AC1, 31, 28, 0), (0x09C1, 32, 29, 0),
    (0x08A1, 33, 30, 0), (0x0521, 34, 31, 0), (0x0441, 35, 32, 0), (0x02A1, 36, 33, 0),
    (0x0221, 37, 34, 0), (0x0141, 38, 35, 0), (0x0111, 39, 36, 0), (0x0085, 40, 37, 0),
    (0x0049, 41, 38, 0), (0x0025, 42, 39, 0), (0x0015, 43, 40, 0), (0x0009, 44, 41, 0),
    (0x0005, 45, 42, 0), (0x0001, 45, 43, 0), (0x5601, 46, 46, 0),
]


@pytest.fixture(scope="module")
def enc(workloads):
    return workloads._t1enc()


def _enc_tables(enc):
    import ctypes
    qe = np.zeros(47, np.uint16)
    nm, nl, sw = (np.zeros(47, np.uint8) for _ in range(3))
    enc.bmi_t1_encoder_tables(*(a.ctypes.data_as(ctypes.c_void_p) for a in (qe, nm, nl, sw)))
    return qe, nm, nl, sw


def test_mq_tables_equal_t800_c2(orc, enc):
    want = np.array(T800_C2)
    for name, tabs in (("oracle", orc.t1_tables()), ("encoder", _enc_tables(enc))):
        got = np.stack([t.astype(np.int64) for t in tabs], axis=1)
        assert np.array_equal(got, want), name


# T.800 Table D.1, as rows of the standard: (condition on (H, V, D)) -> label
def table_d1(h, v, d, orient):
    if orient == 1:  # HL: the table's H and V columns swap
        h, v = v, h
    if orient == 3:  # HH
        hv = h + v
        rows = [(d >= 3, 8), (d == 2 and hv >= 1, 7), (d == 2 and hv == 0, 6),
                (d == 1 and hv >= 2, 5), (d == 1 and hv == 1, 4), (d == 1 and hv == 0, 3),
                (d == 0 and hv >= 2, 2), (d == 0 and hv == 1, 1), (d == 0 and hv == 0, 0)]
    else:  # LL and LH
        rows = [(h == 2, 8), (h == 1 and v >= 1, 7), (h == 1 and v == 0 and d >= 1, 6),
                (h == 1 and v == 0 and d == 0, 5), (h == 0 and v == 2, 4),
                (h == 0 and v == 1, 3), (h == 0 and v == 0 and d >= 2, 2),
                (h == 0 and v == 0 and d == 1, 1), (h == 0 and v == 0 and d == 0, 0)]
    hits = [lab for cond, lab in rows if cond]
    assert len(hits) == 1
    return hits[0]


@pytest.mark.parametrize("orient", [0, 1, 2, 3])
def test_zero_coding_contexts_table_d1(orc, enc, orient):
    for idx in range(256):
        nb = [(idx >> b) & 1 for b in range(8)]
        nb9 = nb[:4] + [0] + nb[4:]  # NW N NE W (centre) E SW S SE
        h = nb9[3] + nb9[5]
        v = nb9[1] + nb9[7]
        d = nb9[0] + nb9[2] + nb9[6] + nb9[8]
        want = table_d1(h, v, d, orient)
        assert orc.t1_zc(nb9, orient) == want, (idx, orient)
        arr = np.asarray(nb9, np.uint8)
        assert enc.bmi_t1_encoder_zc(arr.ctypes.data, orient) == want, (idx, orient)


# T.800 Table D.3: (H, V) -> (context label, XOR bit)
TABLE_D3 = {(1, 1): (13, 0), (1, 0): (12, 0), (1, -1): (11, 0), (0, 1): (10, 0), (0, 0): (9, 0),
            (0, -1): (10, 1), (-1, 1): (11, 1), (-1, 0): (12, 1), (-1, -1): (13, 1)}


def test_sign_contexts_table_d3(orc, enc):
    # each of N, W, E, S: insignificant (0), positive (1), negative (3); a
    # negative flag without significance (2) must count as insignificant
    states = [0, 1, 2, 3]
    contrib = {0: 0, 1: 1, 2: 0, 3: -1}
    for n in states:
        for w in states:
            for e in states:
                for s in states:
                    nb9 = [0, n, 0, w, 0, e, 0, s, 0]
                    h = int(np.clip(contrib[w] + contrib[e], -1, 1))
                    v = int(np.clip(contrib[n] + contrib[s], -1, 1))
                    cx, xr = TABLE_D3[(h, v)]
                    assert orc.t1_sc(nb9) == 2 * cx + xr, nb9
                    arr = np.asarray(nb9, np.uint8)
                    assert enc.bmi_t1_encoder_sc(arr.ctypes.data) == 2 * cx + xr, nb9


def signmag(q):
    q = np.asarray(q, np.int64)
    return np.where(q < 0, np.abs(q) | 0x80000000, q).astype(np.uint32).view(np.int32)


def words_to_signed(wds):
    u = wds.view(np.uint32).astype(np.int64)
    mag = u & 0x7FFFFFFF
    return np.where(u >> 31, -mag, mag)


SIZES = [(1, 1), (1, 7), (7, 1), (2, 2), (3, 5), (4, 4), (5, 3), (8, 1), (1, 64), (64, 1),
         (17, 33), (63, 63), (64, 64), (32, 64), (64, 13)]


def _block(h, w, amp, sparsity=0.0):
    q = np.round(RNG.laplace(0, amp, (h, w))).astype(np.int32)
    if sparsity:
        q[RNG.random((h, w)) < sparsity] = 0
    return q


@pytest.mark.parametrize("orient", [0, 1, 2, 3])
@pytest.mark.parametrize("size", SIZES, ids=lambda s: f"{s[0]}x{s[1]}")
def test_round_trip_97_words(orc, workloads, orient, size):
    """9/7: numbps = exponent<<11 | mantissa (guard bits 1); every pass
    decoded -> the T1 words are the indices in sign-magnitude form."""
    h, w = size
    for amp, sp in ((6.0, 0.0), (40.0, 0.5), (0.7, 0.0)):
        q = _block(h, w, amp, sp)
        exp = max(int(np.abs(q).max()).bit_length(), 1) + int(RNG.integers(0, 4))
        data, npass, miss = workloads.t1_encode_block(q, exp, orient)
        got = orc.t1_decode_block(data, w, h, orient, (exp << 11) | 1234, miss, npass)
        K = int(np.abs(q).max()).bit_length()
        if K == 1:
            # bit-plane 0 is replaced by plane 1 (bmi_block_decoder.cpp:649):
            # a block whose largest magnitude is 1 decodes it as 3
            want = signmag(q * 3)
        else:
            want = signmag(q)
        assert np.array_equal(got, want), (size, amp, K)


@pytest.mark.parametrize("size", SIZES, ids=lambda s: f"{s[0]}x{s[1]}")
def test_round_trip_53_integers(orc, workloads, size):
    """5/3: numbps = exponent + guard bits - 1; magnitudes start at bit
    30 - missing and dec_cb_stripe's shift by 31 - numbps returns q."""
    h, w = size
    for amp in (3.0, 300.0):
        q = _block(h, w, amp, 0.3)
        mb = max(int(np.abs(q).max()).bit_length(), 1) + int(RNG.integers(0, 3))
        data, npass, miss = workloads.t1_encode_block(q, mb, 2)
        wds = orc.t1_decode_block(data, w, h, 2, mb, miss, npass)
        assert np.array_equal(orc.t1_words_to_int(wds, mb), q)


def test_truncated_passes_mid_point_closed_form(orc, workloads):
    """After the cleanup pass of plane p the decoded magnitude of every
    sample is its bits at and above p plus the mid-point 2^(p-1) when it
    is non-zero there (:90, :272-273)."""
    h, w = 37, 41
    q = _block(h, w, 25.0, 0.2)
    K = int(np.abs(q).max()).bit_length()
    exp = K + 2
    for k in range(K):
        npass_req = 3 * k + 1  # ends with the cleanup pass of plane K-1-k
        data, npass, miss = workloads.t1_encode_block(q, exp, 0, npass_req)
        assert npass == npass_req
        got = words_to_signed(orc.t1_decode_block(data, w, h, 0, exp << 11, miss, npass))
        p = K - 1 - k
        a = np.abs(q).astype(np.int64)
        top = (a >> p) << p
        mid = (1 << (p - 1)) if p >= 1 else 0
        want = np.where(top > 0, top + mid, 0) * np.sign(q)
        assert np.array_equal(got, want), k


def test_empty_and_full_range_blocks(orc, workloads):
    z = np.zeros((64, 64), np.int32)
    data, npass, miss = workloads.t1_encode_block(z, 9, 0)
    assert npass == 0 and len(data) == 0 and miss == 9
    assert not orc.t1_decode_block(data, 64, 64, 0, 9 << 11, miss, npass).any()
    # 5/3 magnitudes up to 2^30 - 1 (numbps 30: bit-plane 30 .. 1)
    q = np.array([[(1 << 30) - 1, -(1 << 29), 5, -7]], np.int32)
    data, npass, miss = workloads.t1_encode_block(q, 30, 3)
    wds = orc.t1_decode_block(data, 4, 1, 3, 30, miss, npass)
    assert np.array_equal(orc.t1_words_to_int(wds, 30), q)


def test_encoder_rejects_magnitude_overflow(workloads):
    with pytest.raises(ValueError):
        workloads.t1_encode_block(np.array([[8]], np.int32), 3, 0)


# ---------------------------------------------------- code-block styles --
# The style switches of T.800 Table A.19 (bypass 0x01, reset 0x02, termall
# 0x04, vertically causal 0x08, predictable termination 0x10, segmentation
# symbols 0x20).  OpenJPEG decoding whole streams from this encoder pins
# the encoder (tests/test_codestream.py); here the oracle's block decoder
# must invert it for every combination.

@pytest.mark.parametrize("style", range(64), ids=lambda s: f"{s:#04x}")
def test_round_trip_every_style(orc, workloads, style):
    for (h, w) in ((64, 64), (13, 17), (4, 4), (1, 9), (33, 64), (64, 1)):
        q = _block(h, w, 30.0, 0.3)
        mb = max(int(np.abs(q).max()).bit_length(), 1) + 1
        orient = int(RNG.integers(0, 4))
        data, npass, miss, segs = workloads.t1_encode_block(q, mb, orient, style=style,
                                                            segments=True)
        assert sum(segs) == len(data)
        assert len(segs) == (orc.t1_num_segments(npass, style) if style & 5 else (1 if npass else 0))
        wds = orc.t1_decode_block_ex(data, w, h, orient, mb, miss, npass, style, segs)
        assert np.array_equal(orc.t1_words_to_int(wds, mb), q), (style, h, w)


@pytest.mark.parametrize("style", [0x01, 0x05, 0x02, 0x08, 0x20, 0x3F])
def test_truncated_styles_mid_point_closed_form(orc, workloads, style):
    """Truncated after any pass, every style gives the closed-form
    reconstruction of the planes coded so far (raw passes included)."""
    h, w = 37, 41
    q = _block(h, w, 200.0, 0.2)
    K = int(np.abs(q).max()).bit_length()
    exp = K + 1
    for k in range(K):
        npass_req = 3 * k + 1
        data, npass, miss, segs = workloads.t1_encode_block(q, exp, 1, npass_req, style=style,
                                                            segments=True)
        got = words_to_signed(orc.t1_decode_block_ex(data, w, h, 1, exp << 11, miss, npass,
                                                     style, segs))
        p = K - 1 - k
        a = np.abs(q).astype(np.int64)
        top = (a >> p) << p
        mid = (1 << (p - 1)) if p >= 1 else 0
        want = np.where(top > 0, top + mid, 0) * np.sign(q)
        assert np.array_equal(got, want), (style, k)


@pytest.mark.parametrize("style", [0x01, 0x02, 0x08, 0x20])
def test_each_switch_changes_the_code_stream(orc, workloads, style):
    """Each switch alone changes the coded bytes of a busy block, and a
    stream decoded without its switch comes out wrong (the switches are
    not no-ops in the encoder or the decoder)."""
    q = _block(64, 64, 60.0, 0.0)
    mb = int(np.abs(q).max()).bit_length()
    d0, n0, m0, s0 = workloads.t1_encode_block(q, mb, 0, style=0, segments=True)
    d1, n1, m1, s1 = workloads.t1_encode_block(q, mb, 0, style=style, segments=True)
    assert d0 != d1
    if style & 5:
        return   # the segment table differs: decoding without it is refused
    wrong = orc.t1_decode_block_ex(d1, 64, 64, 0, mb, m1, n1, 0, s1)
    assert not np.array_equal(orc.t1_words_to_int(wrong, mb), q)


def test_segment_table_rejected_when_inconsistent(orc, workloads):
    q = _block(32, 32, 50.0, 0.0)
    mb = int(np.abs(q).max()).bit_length()
    data, npass, miss, segs = workloads.t1_encode_block(q, mb, 0, style=0x04, segments=True)
    with pytest.raises(ValueError):
        orc.t1_decode_block_ex(data, 32, 32, 0, mb, miss, npass, 0x04, segs[:-1])
    bad = list(segs)
    bad[0] += 1
    with pytest.raises(ValueError):
        orc.t1_decode_block_ex(data, 32, 32, 0, mb, miss, npass, 0x04, bad)
