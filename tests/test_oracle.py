"""Pins of the CPU oracle (no GPU).  Each check is chosen so a plausible
mistake in the restatement (dropped term, wrong sign or index, swapped
operand, wrong rounding) fails it; none retypes the oracle's own code."""
import itertools

import numpy as np
import pytest

RNG = np.random.default_rng(1234)


# ------------------------------------------------------------------ 5/3 --

def t800_53_line(y, i0):
    """ITU-T T.800 F.3.7 1D_SR with F.3.8.1 1D_FILTR_5-3R, written from the
    standard's formulas on an explicitly extended signal (PSE)."""
    n = len(y)
    if n == 1:
        return [y[0] if i0 % 2 == 0 else y[0] // 2]
    i1 = i0 + n
    ext = 2
    def pse(i):  # periodic symmetric extension index, T.800 eq. (F-4)
        period = 2 * (i1 - i0 - 1)
        j = (i - i0) % period
        return i0 + min(j, period - j)
    Y = {i: y[pse(i) - i0] for i in range(i0 - ext - 1, i1 + ext + 1)}
    X = dict(Y)
    for n2 in range((i0 - ext) // 2, (i1 + ext) // 2 + 1):
        i = 2 * n2
        if i0 - ext + 1 <= i <= i1 + ext - 2:
            X[i] = Y[i] - ((Y[i - 1] + Y[i + 1] + 2) // 4)
    for n2 in range((i0 - ext) // 2, (i1 + ext) // 2 + 1):
        i = 2 * n2 + 1
        if i0 - ext + 2 <= i <= i1 + ext - 3:
            X[i] = Y[i] + ((X[i - 1] + X[i + 1]) // 2)
    return [X[i] for i in range(i0, i1)]


@pytest.mark.parametrize("n", list(range(1, 34)))
def test_53_line_matches_t800_all_sizes(orc, n):
    for i0 in (0, 1, 6, 7):
        for _ in range(3):
            y = RNG.integers(-300, 300, n).tolist()
            got = orc.line53_textbook(np.array(y), i0 & 1).tolist()
            assert got == t800_53_line(y, i0), (n, i0)


@pytest.mark.parametrize("levels", [1, 2, 3, 5, 6])
def test_53_textbook_reversible(orc, levels):
    """forward 5/3 (T.800 F.4) then inverse is the identity, any size and
    parity (T.800 F.3 reversibility); the small sizes at odd offsets have
    empty coarse levels (no rows or no columns), which transform nothing."""
    for (w, h) in [(1, 1), (2, 1), (1, 2), (3, 2), (2, 3), (7, 5), (33, 17), (64, 64), (97, 131)]:
        for (x0, y0) in [(0, 0), (1, 0), (0, 3), (5, 7), (3, 1)]:
            img = RNG.integers(-2048, 2048, (h, w)).astype(np.int32)
            coef = orc.fdwt53(img, levels, x0=x0, y0=y0)
            back = orc.idwt53(coef, levels, x0=x0, y0=y0, mode=orc.R53_TEXTBOOK)
            np.testing.assert_array_equal(back, img, err_msg=f"{w}x{h}@{x0},{y0} L{levels}")


def test_53_constant_dc_gain_one(orc):
    for mode in (orc.R53_REFERENCE, orc.R53_TEXTBOOK):
        for (w, h, L) in [(64, 64, 5), (97, 131, 4), (512, 512, 5)]:
            plane = np.zeros((h, w), np.int32)
            import math
            lw, lh = w, h
            for _ in range(L):
                lw, lh = (lw + 1) // 2, (lh + 1) // 2
            plane[:lh, :lw] = 1234
            out = orc.idwt53(plane, L, mode=mode)
            assert (out == 1234).all()


def closed_form_tail(hres, lh):
    """Bottom rows of the reference's VDWT_R53_LINES from its hand trace:
    m = n//2, P = H[max(m-2,0)]; even n: X[n-1] = P + X[n-2];
    odd n: X[n-1] = L[m-1] - ((P+1)>>1), X[n-2] = P + ((X[n-1] + X[n-3])>>1)."""
    n = hres.shape[0]
    L, H = hres[:lh], hres[lh:]
    m = n // 2
    P = H[max(m - 2, 0)]
    return m, P, L


@pytest.mark.parametrize("n", list(range(2, 24)))
def test_r53_quirk_closed_form(orc, n):
    """Reference mode equals textbook mode except on the rows named by the
    closed form of the bottom edge; those rows follow the closed form."""
    w = 9
    lh = (n + 1) // 2
    hres = RNG.integers(-500, 500, (n, w)).astype(np.int32)
    ref = orc.v53_ref(hres, lh)
    # textbook vertical synthesis column by column
    tb = np.zeros_like(hres)
    for x in range(w):
        inter = np.empty(n, np.int32)
        inter[0::2] = hres[:lh, x]
        inter[1::2] = hres[lh:, x]
        tb[:, x] = orc.line53_textbook(inter, 0)
    m, P, L = closed_form_tail(hres, lh)
    if n % 2 == 0:
        quirk_rows = [n - 1] if n >= 4 else []
        exp_last = P + tb[n - 2]
        np.testing.assert_array_equal(ref[n - 1], exp_last)
    else:
        quirk_rows = [n - 2, n - 1]
        v1 = L[m - 1] - ((P + 1) >> 1)
        v2 = P + ((v1 + tb[n - 3]) >> 1)
        np.testing.assert_array_equal(ref[n - 1], v1)
        np.testing.assert_array_equal(ref[n - 2], v2)
    other = [r for r in range(n) if r not in quirk_rows]
    np.testing.assert_array_equal(ref[other], tb[other])


def test_h53_ref_equals_textbook_even_parity(orc):
    """HDWT_R53_INT is textbook for parity 0 (its deviation is parity only)."""
    for n in range(2, 40):
        lw = (n + 1) // 2
        lo = RNG.integers(-999, 999, lw).astype(np.int32)
        hi = RNG.integers(-999, 999, n - lw).astype(np.int32)
        inter = np.empty(n, np.int32)
        inter[0::2], inter[1::2] = lo, hi
        out = np.zeros(n, np.int32)
        import ctypes
        orc.lib().or_h53_line_ref(out.ctypes.data_as(ctypes.c_void_p), lo.ctypes.data_as(ctypes.c_void_p),
                                  hi.ctypes.data_as(ctypes.c_void_p), lw, n - lw)
        np.testing.assert_array_equal(out, orc.line53_textbook(inter, 0))


def test_r53_reference_full_tile_differs_from_textbook(orc, workloads):
    """On a multi-level tile the two modes must differ (the quirk is live),
    and both must keep the DC-free details' energy bounded."""
    planes = workloads.tile_planes(97, 131, 4, "53", 1, seed=3)
    ref = orc.idwt53(planes[0], 4, mode=orc.R53_REFERENCE)
    tb = orc.idwt53(planes[0], 4, mode=orc.R53_TEXTBOOK)
    assert (ref != tb).any()
    # the first rows are untouched by any level's bottom edge
    np.testing.assert_array_equal(ref[:8], tb[:8])


# ------------------------------------------------------------------ 9/7 --

@pytest.mark.parametrize("levels", [1, 2, 3, 5])
def test_97_f64_round_trip(orc, levels):
    """FP64 forward (analysis) then the FP64 oracle's synthesis is the
    identity: pins the lifting structure, constants, parities and the
    reference gain convention (low x K, high x 2/K)."""
    for (w, h) in [(1, 1), (2, 2), (3, 5), (6, 7), (33, 17), (97, 131)]:
        for (x0, y0) in [(0, 0), (1, 1), (3, 0)]:
            img = RNG.normal(0, 50, (h, w))
            coef = orc.fdwt97_f64(img, levels, x0=x0, y0=y0)
            back = orc.idwt97_f64(coef.astype(np.float32), levels, x0=x0, y0=y0)
            # coefficients were rounded to float32 on the way in
            assert np.abs(back - img).max() < 1e-3, (w, h, x0, y0)


def test_97_float_close_to_f64(orc, workloads):
    planes = workloads.tile_planes(256, 200, 5, "97", 1, seed=7)
    f32 = orc.idwt97(planes[0], 5)
    f64 = orc.idwt97_f64(planes[0], 5)
    err = np.abs(f32.astype(np.float64) - f64)
    assert err.max() < 2e-4 * max(1.0, np.abs(f64).max())
    rms = np.sqrt((err ** 2).mean())
    assert rms < 1e-4 * max(1.0, np.abs(f64).max())


@pytest.mark.parametrize("n", [3, 5, 7, 9, 17, 31])
def test_97_line_mirror_symmetry(orc, n):
    """Odd length, parity 0: reversing the line keeps every sample's band
    and, the operations being symmetric in (left, right), reverses the
    output bit for bit.  Fails on a shifted neighbour or a wrong mirror."""
    x = RNG.normal(0, 100, n).astype(np.float32)
    a = orc.line97(x, 0)
    b = orc.line97(x[::-1].copy(), 0)
    np.testing.assert_array_equal(a, b[::-1])


@pytest.mark.parametrize("w", [2, 3, 4, 5])
@pytest.mark.parametrize("h", [2, 3, 4, 5])
@pytest.mark.parametrize("x0,y0", [(0, 0), (1, 1), (1, 0), (0, 1)])
def test_97_small_resolutions_follow_t800(orc, w, h, x0, y0):
    """Widths and heights 2..5 are where the reference's 9/7 synthesis is
    undefined: HDWT_R97 writes nothing at odd parity
    (bmi_cpu_dwt_decoder.cpp:2307-2310); VDWT_R97 at even parity, heights 3
    and 5, reads a high row past the band and writes one row past the
    resolution (the `<=` loops at :3113, :3146); at odd parity, heights
    2..5, its step-4/6 loops update row 0 again from row -1 (:3237-3243,
    :3265-3270).  The oracle (and the GPU, which equals it bit for bit)
    produces the T.800 F.3.8.2 synthesis with symmetric extension there,
    pinned here by the FP64 analysis (T.800 F.4.8.2) followed by the
    oracle's float synthesis returning the image, at both parities, for one
    level (every line 2..5 samples) and two (the coarse level at 1..3)."""
    img = np.random.default_rng(w * 10 + h + 100 * x0 + 1000 * y0).normal(0, 60, (h, w))
    for levels in (1, 2):
        coef = orc.fdwt97_f64(img, levels, x0=x0, y0=y0)
        back = orc.idwt97(coef.astype(np.float32), levels, x0=x0, y0=y0)
        assert np.abs(back.astype(np.float64) - img).max() < 2e-3, (levels,)


def test_signmag_equals_float_input(orc, workloads):
    """Dequantising T1 sign-magnitude words equals feeding the same
    integers as floats (FETCHPIXEL_INT, bmi_cpu_dwt_decoder.cpp:2287)."""
    v = workloads.round_half_away(workloads.tile_planes(131, 97, 4, "97", 1, seed=9)[0].astype(np.float64))
    sm = workloads.to_signmag(v.astype(np.int64))
    steps = np.linspace(0.5, 2.0, 13).astype(np.float32)
    a = orc.idwt97(sm, 4, steps, coef_kind=orc.COEF_SIGNMAG)
    b = orc.idwt97(v.astype(np.float32), 4, steps, coef_kind=orc.COEF_FLOAT)
    np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))


def test_signmag_words(workloads):
    v = np.array([0, 1, -1, 5, -5, 2 ** 30, -(2 ** 30)], np.int64)
    w = workloads.to_signmag(v)
    assert w.tolist() == [0, 1, 0x80000001, 5, 0x80000005, 2 ** 30, 0x80000000 | 2 ** 30]


# --------------------------------------------------------------- colour --

def forward_rct(R, G, B, bd):
    off = 1 << (bd - 1)
    r, g, b = R - off, G - off, B - off
    return (r + 2 * g + b) >> 2, b - g, r - g


def test_rct_exhaustive_8bit(orc):
    """Forward RCT (T.800 G.2) then the oracle's inverse returns every 8-bit
    triple exactly, with the B,G,R byte order of RGB_24BITS."""
    v = np.arange(256, dtype=np.int32)
    R, G, B = np.meshgrid(v, v, v, indexing="ij")
    R, G, B = R.reshape(4096, 4096), G.reshape(4096, 4096), B.reshape(4096, 4096)
    y0, y1, y2 = forward_rct(R, G, B, 8)
    out = orc.rct_pack(y0, y1, y2, 8, 2).reshape(4096, 4096, 3)
    np.testing.assert_array_equal(out[..., 0], B)
    np.testing.assert_array_equal(out[..., 1], G)
    np.testing.assert_array_equal(out[..., 2], R)


def test_rct_12bit_into_48bit(orc):
    R, G, B = (RNG.integers(0, 4096, (512, 512)).astype(np.int32) for _ in range(3))
    y0, y1, y2 = forward_rct(R, G, B, 12)
    out = orc.rct_pack(y0, y1, y2, 12, 1).view(np.uint16).reshape(512, 512, 3)
    np.testing.assert_array_equal(out[..., 2], R << 4)
    np.testing.assert_array_equal(out[..., 1], G << 4)
    np.testing.assert_array_equal(out[..., 0], B << 4)


def test_rct_floor_not_truncation(orc):
    """(Y1+Y2)>>2 floors negative sums (the CPU's RRct_53), unlike the
    reference CUDA kernel's /4 (bmi_dwt_gpu_kernel.h:1115)."""
    y0 = np.array([[0]], np.int32)
    y1 = np.array([[-1]], np.int32)
    y2 = np.array([[-2]], np.int32)
    out = orc.rct_pack(y0, y1, y2, 8, 2)
    g = 0 - (-3 >> 2) + 128  # floor(-3/4) = -1 -> 129
    assert out[0, 1] == g


def test_ict_truncation_and_clamp_order(orc):
    """Cb = Cr = 0: R = G = B = (int)clamp(Y + 2^(B-1)) on a 1/64 grid
    around the clamp limits (truncation after the clamp, in float)."""
    for bd in (8, 12):
        top = 1 << bd
        y = (np.arange(-(top + 2) * 64, (top + 2) * 64) / 64.0).astype(np.float32) - (1 << (bd - 1))
        y = y.reshape(1, -1)
        z = np.zeros_like(y)
        fmt = 2 if bd == 8 else 1
        out = orc.ict_pack(y, z, z, bd, fmt)
        v = np.clip(y.astype(np.float64) + (1 << (bd - 1)), 0, top - 1)
        exp = np.trunc(v).astype(np.int64)
        if bd == 8:
            got = out.reshape(-1, 3)
        else:
            got = out.view(np.uint16).reshape(-1, 3).astype(np.int64) >> 4
        for ch in range(3):
            np.testing.assert_array_equal(got[:, ch], exp[0])


def test_ict_close_to_f64(orc):
    c = [RNG.normal(0, 60, (64, 64)).astype(np.float32) for _ in range(3)]
    r, g, b = orc.ict_float(*c, 8)
    y, cb, cr = (x.astype(np.float64) for x in c)
    np.testing.assert_allclose(r, y + 1.402 * cr + 128, atol=1e-3)
    np.testing.assert_allclose(g, y - 0.34413 * cb - 0.71414 * cr + 128, atol=1e-3)
    np.testing.assert_allclose(b, y + 1.772 * cb + 128, atol=1e-3)


def test_merge_truncates_before_shift(orc):
    """static_cast<int>(v) + 2^(B-1): -0.5 -> 0 + 128 (not floor -1)."""
    v = np.array([[-0.5, -1.5, 0.7, 200.0, -300.0]], np.float32)
    out = orc.merge_pack([v], 8, 6, 0, is_float=True).reshape(-1, 3)
    assert out[:, 0].tolist() == [128, 127, 128, 255, 0]


def test_merge_component_slots(orc):
    a = np.full((1, 2), 10, np.int32)
    b = np.full((1, 2), 20, np.int32)
    c = np.full((1, 2), 30, np.int32)
    d = np.full((1, 2), 40, np.int32)
    out4 = orc.merge_pack([a, b, c, d], 8, 3).reshape(-1, 4)
    # 4 components: byte 0 = Comp3, 1 = Comp2, 2 = Comp1, 3 = Comp0
    assert out4[0].tolist() == [168, 158, 148, 138]
    out3 = orc.merge_pack([a, b, c], 8, 6).reshape(-1, 3)
    assert out3[0].tolist() == [158, 148, 138]


def test_dc_shift(orc):
    v = np.array([[-200, -128, 0, 127, 300]], np.int32)
    assert orc.dc_shift(v, 8, False).tolist() == [[0, 0, 128, 255, 255]]
    f = np.array([[-0.5, 1.9, -128.9]], np.float32)
    assert orc.dc_shift(f, 8, True).tolist() == [[128, 129, 0]]


def test_one_component_from_result_extracts_the_byte(orc):
    """Mode 0 reads byte (2 - comp) of each packed 0x00RRGGBB word: against
    numpy bit operations on words built from known channels."""
    rng = np.random.default_rng(4)
    r, g, b = (rng.integers(0, 256, (7, 9)) for _ in range(3))
    words = ((r << 16) | (g << 8) | b | (rng.integers(0, 256, (7, 9)) << 24)).astype(np.uint32)
    for comp, ch in enumerate((r, g, b)):
        got = orc.one_component(words, 0, comp, 8)
        assert np.array_equal(got, ch | (ch << 8) | (ch << 16)), comp


def test_one_component_97_fixed_point_rounding(orc):
    """Mode 2 reads the word as fixed point with 13 fraction bits and rounds
    half away from zero: against np.floor(|w| / 2^13 + 1/2) with the sign,
    on a dense sample and at every half-way point, then level-shifted,
    clamped and reduced to 8 bits."""
    rng = np.random.default_rng(5)
    w = np.concatenate([rng.integers(-(1 << 24), 1 << 24, 20000),
                        np.arange(-40, 41) * 4096, np.arange(-40, 41) * 4096 - 1,
                        np.arange(-40, 41) * 4096 + 1, [-(1 << 31), (1 << 31) - 1]]).astype(np.int64)
    for bd in (8, 10, 12, 16):
        got = orc.one_component(w.astype(np.int32).reshape(1, -1), 2, 0, bd)[0]
        mag = np.floor(np.abs(w) / 8192.0 + 0.5)
        # the kernel keeps 15 integer bits of the rounded magnitude (mask 0xfffe000)
        mag = np.mod(mag, 1 << 15)
        v = (np.sign(w) * mag).astype(np.int64) + (1 << (bd - 1))
        o = np.clip(v, 0, (1 << bd) - 1) >> (bd - 8)
        assert np.array_equal(got, o | (o << 8) | (o << 16)), bd


def test_one_component_53_to_8_bits(orc):
    """Mode 1: component 0 is level-shifted, the others take (v + 2^B - 1) >> 1;
    then clamp and the 8-bit slot rule (>> B-8 above 8 bits, << 8-B below),
    at 4..16 bits, against a numpy statement of the same rule over the
    whole input range."""
    for bd in (4, 6, 8, 12, 16):
        v = np.arange(-(1 << bd) - 3, (1 << bd) + 4, dtype=np.int64)
        for comp in (0, 1):
            got = orc.one_component(v.astype(np.int32).reshape(1, -1), 1, comp, bd)[0]
            t = v + (1 << (bd - 1)) if comp == 0 else (v + (1 << bd) - 1) >> 1
            t = np.clip(t, 0, (1 << bd) - 1)
            t = (t >> (bd - 8)) if bd >= 8 else ((t << (8 - bd)) & 0xFF)
            assert np.array_equal(got, t | (t << 8) | (t << 16)), (bd, comp)
        # an 8-bit component 0 at its mid-grey and extremes
        if bd == 8:
            got = orc.one_component(np.array([[-128, 0, 127]], np.int32), 1, 0, 8)[0]
            assert got.tolist() == [0, 0x808080, 0xFFFFFF]


def test_97_error_budget_vs_f64(orc, workloads):
    """The stated 9/7 error bound: the float path (which the GPU matches
    bit for bit) against the FP64 lifting oracle on the image workload's
    distributions (LL ~ N(128,40), detail Laplace(4)*2^-l), 5 levels: the
    plane error and the reconstructed 8-bit samples (inverse ICT, +128,
    clamp, truncate) differ from FP64 by at most 1 LSB."""
    w, h, L = 512, 384, 5
    planes = workloads.tile_planes(w, h, L, "97", 3, seed=0)
    f32 = [orc.idwt97(planes[c], L) for c in range(3)]
    f64 = [orc.idwt97_f64(planes[c], L) for c in range(3)]
    err = np.concatenate([(a.astype(np.float64) - b).ravel() for a, b in zip(f32, f64)])
    max_abs, rms = float(np.abs(err).max()), float(np.sqrt((err ** 2).mean()))
    # 8-bit samples: float path exactly as the reference, FP64 path with the
    # same clamp-then-truncate rule
    bgr = orc.ict_pack(f32[0], f32[1], f32[2], 8, 2).reshape(h, w, 3).astype(np.int64)
    y, cb, cr = f64
    ref = np.stack([y + 1.772 * cb, y - 0.34413 * cb - 0.71414 * cr, y + 1.402 * cr], axis=-1)
    ref = np.trunc(np.clip(ref + 128.0, 0.0, 255.0)).astype(np.int64)
    d8 = np.abs(bgr - ref)
    print(f"\n9/7 float vs FP64: plane max-abs {max_abs:.3e}, RMS {rms:.3e}; "
          f"8-bit samples: max diff {d8.max()}, {int((d8 > 0).sum())} of {d8.size} differ")
    assert max_abs < 2e-3 and rms < 2e-4
    assert d8.max() <= 1


@pytest.mark.parametrize("threads", [1, 3, 8])
@pytest.mark.parametrize("shape", [(1003, 517, 5), (64, 40, 3), (97, 33, 4), (7, 5, 2)])
def test_cpu_baseline_equals_oracle(orc, workloads, threads, shape):
    """The multithreaded, vectorised CPU baseline (row-band / column-band
    threading, -O3 x86-64-v3, no FP contraction) gives the scalar oracle's
    planes and pixels bit for bit at any thread count, so the time it
    reports is the time of the same decode."""
    w, h, L = shape
    planes = workloads.tile_planes(w, h, L, "97", 3, seed=w + h)
    steps = np.stack([np.linspace(0.5 + 0.2 * c, 1.5, 3 * L + 1) for c in range(3)]).astype(np.float32)
    got_planes, got = orc.cpu_decode97_ict(planes, L, threads, steps=steps)
    exp_planes, exp = orc.decode_tile(planes, levels=L, filt="97", out_format=2, steps=steps)
    for c in range(3):
        assert np.array_equal(got_planes[c].view(np.uint32), exp_planes[c].view(np.uint32)), c
    assert np.array_equal(got, exp)


def test_ignore_components_slot_order(orc):
    """mIgnoreComponents: the remaining components move down to the lower
    output slots in order (bmi_cpu_dwt_decoder.cpp:1180-1290); a four-
    component 8-bit merge with component 1 ignored packs components 0, 2, 3
    into a three-component pixel exactly as a three-component tile made of
    them; an ignored check-out component means none; nothing changes for a
    mask of 0."""
    rng = np.random.default_rng(12)
    planes = rng.integers(-50, 50, (4, 6, 7)).astype(np.int32)
    _, a = orc.decode_tile(planes, levels=0, filt="53", out_format=6, ignore_mask=0b0010)
    _, b = orc.decode_tile(planes[[0, 2, 3]], levels=0, filt="53", out_format=6)
    assert np.array_equal(a, b)
    # byte slots of the 24-bit merge: comp2 | comp1 | comp0 of the remaining
    # list, i.e. planes 3, 2, 0 (hand check of one pixel)
    px = a.reshape(6, 7, 3)[0, 0]
    assert px.tolist() == [planes[3, 0, 0] + 128, planes[2, 0, 0] + 128, planes[0, 0, 0] + 128]
    # check-out of an ignored component: no check-out
    _, c = orc.decode_tile(planes, levels=0, filt="53", out_format=6, check_out=1, ignore_mask=0b0010)
    assert np.array_equal(c, a)
    # check-out of a kept one is renumbered: plane 2 is slot 1
    _, d = orc.decode_tile(planes, levels=0, filt="53", out_format=6, check_out=2, ignore_mask=0b0010)
    _, e = orc.decode_tile(planes[[0, 2, 3]], levels=0, filt="53", out_format=6, check_out=1)
    assert np.array_equal(d, e)
    _, f = orc.decode_tile(planes, levels=0, filt="53", out_format=3)
    _, g = orc.decode_tile(planes, levels=0, filt="53", out_format=3, ignore_mask=0)
    assert np.array_equal(f, g)
    with pytest.raises(ValueError):
        orc.decode_tile(planes, levels=0, filt="53", out_format=3, ignore_mask=0b1111)
