"""Deterministic synthetic workloads (coefficient planes in Mallat layout).

The reference ships no sample codestreams, so every input is generated
here, identically on any machine:

* PRNG: splitmix64; one sub-stream per (image, tile, component, band) with
  seed ^ (img << 40 | tile << 20 | comp << 8 | band).
* uniform u in [0,1) from the top 53 bits; Laplacian(b) by the inverse CDF
  x = -b*sgn(u-1/2)*ln(1-2|u-1/2|); Normal by Box-Muller in float64;
  integers by rounding half away from zero.

Band layout per tile-component (as the reference GPU "orig" buffer,
decoder/cudainterface/src/cudainterface.cpp:674-703): at synthesis level
lev (0 = coarsest) LL is at [0,lw)x[0,lh) (level 0 only), HL at
[lw,nw)x[0,lh), LH at [0,lw)x[lh,nh), HH at [lw,nw)x[lh,nh).
Subband ids: 0 = LL, 3*lev+1 = HL, 3*lev+2 = LH, 3*lev+3 = HH.
"""
from __future__ import annotations

import math

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int) -> np.ndarray:
    """First n outputs of splitmix64 started at `seed`."""
    with np.errstate(over="ignore"):
        i = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + i * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, n: int) -> np.ndarray:
    return (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def laplace(seed: int, n: int, b: float) -> np.ndarray:
    u = uniform(seed, n)
    d = u - 0.5
    a = np.minimum(2.0 * np.abs(d), 1.0 - 2.0 ** -53)
    return -b * np.sign(d) * np.log1p(-a)


def normal(seed: int, n: int, mu: float, sigma: float) -> np.ndarray:
    u1 = uniform(seed, n)
    u2 = uniform(seed ^ 0x5851F42D4C957F2D, n)
    r = np.sqrt(-2.0 * np.log1p(-u1))
    return mu + sigma * r * np.cos(2.0 * math.pi * u2)


def round_half_away(x: np.ndarray) -> np.ndarray:
    return np.where(x >= 0, np.floor(x + 0.5), -np.floor(-x + 0.5))


def substream(seed: int, img: int, tile: int, comp: int, band: int) -> int:
    return (seed ^ ((img << 40) | (tile << 20) | (comp << 8) | band)) & 0xFFFFFFFFFFFFFFFF


def level_span(t0: int, n: int, levels: int, lev: int):
    s = levels - 1 - lev
    return -(-t0 // (1 << s)), -(-(t0 + n) // (1 << s))


def band_rects(w: int, h: int, x0: int, y0: int, levels: int):
    """[(band_id, decomposition_level, x, y, bw, bh)] in Mallat layout;
    decomposition level 1 is the finest."""
    rects = []
    for lev in range(levels):
        a0, a1 = level_span(x0, w, levels, lev)
        b0, b1 = level_span(y0, h, levels, lev)
        nw, nh = a1 - a0, b1 - b0
        lw = (a1 + 1) // 2 - (a0 + 1) // 2
        lh = (b1 + 1) // 2 - (b0 + 1) // 2
        dl = levels - lev
        if lev == 0:
            rects.append((0, dl, 0, 0, lw, lh))
        rects.append((3 * lev + 1, dl, lw, 0, nw - lw, lh))
        rects.append((3 * lev + 2, dl, 0, lh, lw, nh - lh))
        rects.append((3 * lev + 3, dl, lw, lh, nw - lw, nh - lh))
    if levels == 0:
        rects.append((0, 0, 0, 0, w, h))
    return rects


def tile_planes(w: int, h: int, levels: int, filt: str, ncomp: int, seed: int = 0, *,
                x0: int = 0, y0: int = 0, img: int = 0, tile: int = 0,
                ll53=(0, 4095), lap53: float = 6.0, ll97=(128.0, 40.0), lap97: float = 4.0,
                signmag: bool = False) -> np.ndarray:
    """Coefficient planes (ncomp, h, w).

    5/3: int32; LL uniform integers in [ll53[0], ll53[1]], detail bands
    round(Laplace(lap53)).
    9/7: float32; LL ~ N(ll97), detail at decomposition level l ~
    Laplace(lap97) * 2^-l.  With signmag=True the 9/7 values are rounded to
    integers and stored as T1 sign-magnitude words (uint32).
    """
    if filt not in ("53", "97"):
        raise ValueError(filt)
    dtype = np.int32 if filt == "53" else np.float32
    out = np.zeros((ncomp, h, w), dtype=dtype)
    for c in range(ncomp):
        for (bid, dl, bx, by, bw, bh) in band_rects(w, h, x0, y0, levels):
            n = bw * bh
            if n == 0:
                continue
            s = substream(seed, img, tile, c, bid)
            if filt == "53":
                if bid == 0:
                    lo, hi = ll53
                    v = np.floor(uniform(s, n) * (hi - lo + 1)) + lo
                else:
                    v = round_half_away(laplace(s, n, lap53))
            else:
                if bid == 0:
                    v = normal(s, n, ll97[0], ll97[1])
                else:
                    v = laplace(s, n, lap97) * (2.0 ** -dl)
            out[c, by:by + bh, bx:bx + bw] = v.reshape(bh, bw).astype(dtype)
    if filt == "97" and signmag:
        return to_signmag(round_half_away(out.astype(np.float64)).astype(np.int64))
    return out


def to_signmag(v: np.ndarray) -> np.ndarray:
    """int -> T1 sign-magnitude words (bit 31 = sign)."""
    v = np.asarray(v, dtype=np.int64)
    mag = np.minimum(np.abs(v), 0x7FFFFFFF).astype(np.uint32)
    return np.where(v < 0, mag | np.uint32(0x80000000), mag).astype(np.uint32)


# --------------------------------------------------------------- configs --

def config1(seed: int = 0):
    """512x512 single component, 5 levels, 5/3 int32 (grey output)."""
    return dict(w=512, h=512, levels=5, filt="53", ncomp=1,
                planes=tile_planes(512, 512, 5, "53", 1, seed))


def config2(seed: int = 0, w: int = 4096, h: int = 2160):
    """4096x2160 3 components, 5 levels, 9/7 float32, ICT to 8-bit BGR."""
    return dict(w=w, h=h, levels=5, filt="97", ncomp=3,
                planes=tile_planes(w, h, 5, "97", 3, seed))


def config3_params(seed: int = 0, ntiles: int = 256, lo: int = 64, hi: int = 1024):
    """Tile parameters of the batched workload: sides U{lo..hi} (odd sizes
    included), levels U{3..6}, 5/3 or 9/7 (p=1/2), 1 or 3 components."""
    u = uniform(seed ^ 0xC3C3C3C3, ntiles * 5).reshape(ntiles, 5)
    tiles = []
    for i in range(ntiles):
        w = lo + int(u[i, 0] * (hi - lo + 1))
        h = lo + int(u[i, 1] * (hi - lo + 1))
        levels = 3 + int(u[i, 2] * 4)
        filt = "53" if u[i, 3] < 0.5 else "97"
        ncomp = 1 if u[i, 4] < 0.5 else 3
        tiles.append(dict(w=w, h=h, levels=levels, filt=filt, ncomp=ncomp, tile=i))
    return tiles


def config4_requests(seed: int, nreq: int, rate: float, lo: int = 256, hi: int = 8192):
    """Poisson arrivals at `rate` req/s; square sides log-uniform in
    [lo, hi]; 8 or 12 bits; 5/3 or 9/7; 3 components; 5 levels."""
    u = uniform(seed ^ 0x4A4A4A4A, nreq * 4).reshape(nreq, 4)
    t = np.cumsum(-np.log1p(-u[:, 0]) / rate)
    reqs = []
    for i in range(nreq):
        side = int(round(math.exp(math.log(lo) + u[i, 1] * (math.log(hi) - math.log(lo)))))
        reqs.append(dict(arrival=float(t[i]), w=side, h=side, bits=8 if u[i, 2] < 0.5 else 12,
                         filt="53" if u[i, 3] < 0.5 else "97", ncomp=3, levels=5, img=i))
    return reqs


# ----------------------------------------------------------- T1 streams --
# Code streams for the GPU tier-1 decoder, made by the project's own EBCOT
# encoder (csrc/bmi_t1_encoder.c, built into libbmi_t1enc.so).

_T1ENC = None


def _t1enc():
    global _T1ENC
    if _T1ENC is None:
        import ctypes
        import pathlib
        path = pathlib.Path(__file__).resolve().parent / "libbmi_t1enc.so"
        if not path.exists():
            raise RuntimeError(f"{path} missing: build it with make (or __graft_entry__.build())")
        L = ctypes.CDLL(str(path))
        P, i = ctypes.c_void_p, ctypes.c_int
        L.bmi_t1_encode_block.argtypes = [P, i, i, i, i, i, P, ctypes.c_long,
                                          ctypes.POINTER(ctypes.c_long), ctypes.POINTER(i),
                                          ctypes.POINTER(i)]
        L.bmi_t1_encode_block.restype = i
        L.bmi_t1_encode_block_ex.argtypes = [P, i, i, i, i, i, i, P, ctypes.c_long,
                                             ctypes.POINTER(ctypes.c_long), ctypes.POINTER(i),
                                             ctypes.POINTER(i), P, i, ctypes.POINTER(i)]
        L.bmi_t1_encode_block_ex.restype = i
        L.bmi_t1_encoder_tables.argtypes = [P, P, P, P]
        L.bmi_t1_encoder_zc.argtypes = [P, i]
        L.bmi_t1_encoder_sc.argtypes = [P]
        _T1ENC = L
    return _T1ENC


def t1_encode_block(q: np.ndarray, mb: int, orient: int, pass_limit: int = 0, style: int = 0,
                    segments: bool = False):
    """Encode one code-block of quantisation indices q (h, w) with mb
    magnitude bit-planes and code-block style `style` (ITU-T T.800 Table
    A.19 bits) -> (bytes, num_passes, missing), plus the byte length of
    every codeword segment when `segments`."""
    import ctypes
    a = np.ascontiguousarray(q, dtype=np.int32)
    h, w = a.shape
    cap = 64 + 4 * w * h * max(1, mb) // 3 + 4 * w * h + 8 * 3 * max(1, mb)
    out = np.zeros(cap, dtype=np.uint8)
    segs = np.zeros(3 * 32, dtype=np.int64)
    n, npass, miss, nseg = ctypes.c_long(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    rc = _t1enc().bmi_t1_encode_block_ex(a.ctypes.data_as(ctypes.c_void_p), w, h, mb, orient,
                                         pass_limit, style, out.ctypes.data_as(ctypes.c_void_p),
                                         cap, ctypes.byref(n), ctypes.byref(npass),
                                         ctypes.byref(miss), segs.ctypes.data_as(ctypes.c_void_p),
                                         len(segs), ctypes.byref(nseg))
    if rc != 0:
        raise ValueError(f"t1 encode failed ({rc}): magnitude >= 2^{mb} or buffer too small")
    data = out[:n.value].tobytes()
    if segments:
        return data, npass.value, miss.value, [int(x) for x in segs[:nseg.value]]
    return data, npass.value, miss.value


def band_orient(band_id: int) -> int:
    """Subband id -> orientation (0 LL, 1 HL, 2 LH, 3 HH)."""
    return 0 if band_id == 0 else (band_id - 1) % 3 + 1


def abs_step_norm(exp: int, mant: int, bit_depth: int = 8) -> float:
    """mfAbsStepNorm of a 9/7 band from its QCD exponent / mantissa
    (decoder/cudainterface/src/cudainterface.cpp:712-720)."""
    step = ((mant & 0x7FF) << 2) | 0x2000
    n = bit_depth - exp
    step = step << n if n > 0 else step >> (-n)
    return float(np.float32(step) * np.float32(2.0 ** -13))


def t1_image(w: int, h: int, levels: int, ncomp: int, seed: int = 0, cb: int = 64,
             filt: str = "97", threads: int = 8, pass_limit: int = 0):
    """Quantise a synthetic image's subbands and encode every code-block.

    9/7: coefficients as tile_planes (LL ~ N(128,40), detail Laplace(4)*2^-l),
    deadzone-quantised with the step of a QCD (exponent, mantissa) chosen per
    band: LL (10, 2047) -> step ~0.5, detail level l (9+l, 0) -> 2^-(l+1);
    guard bits 1, so numbps = exponent<<11 | mantissa and a band has
    `exponent` magnitude bit-planes.  5/3: integer coefficients of
    tile_planes, numbps = 12 (guard bits 1, exponent 12).
    Returns dict(q=(ncomp,h,w) int32 indices, blocks=[dict], data=bytes,
    steps=(ncomp, 3L+1) float32)."""
    from concurrent.futures import ThreadPoolExecutor
    planes = tile_planes(w, h, levels, filt, ncomp, seed)
    rects = band_rects(w, h, 0, 0, levels)
    q = np.zeros((ncomp, h, w), dtype=np.int32)
    steps = np.ones((ncomp, 3 * levels + 1), dtype=np.float32)
    qcd = {}
    for (bid, dl, bx, by, bw, bh) in rects:
        if filt == "97":
            exp, mant = (10, 2047) if bid == 0 else (9 + dl, 0)
            numbps = (exp << 11) | mant
            mb = exp
            st = abs_step_norm(exp, mant)
            qcd[bid] = (numbps, mb)
            for c in range(ncomp):
                steps[c, bid] = st
                v = planes[c, by:by + bh, bx:bx + bw].astype(np.float64) / st
                q[c, by:by + bh, bx:bx + bw] = (np.sign(v) * np.floor(np.abs(v))).astype(np.int32)
        else:
            qcd[bid] = (12, 12)
            q[:, by:by + bh, bx:bx + bw] = planes[:, by:by + bh, bx:bx + bw]
    jobs = []
    for c in range(ncomp):
        for (bid, dl, bx, by, bw, bh) in rects:
            for y0 in range(0, bh, cb):
                for x0 in range(0, bw, cb):
                    jobs.append((c, bid, bx + x0, by + y0, min(cb, bw - x0), min(cb, bh - y0)))

    def enc(j):
        c, bid, x, y, bw, bh = j
        numbps, mb = qcd[bid]
        data, npass, miss = t1_encode_block(q[c, y:y + bh, x:x + bw], mb, band_orient(bid),
                                            pass_limit)
        return data, npass, miss

    with ThreadPoolExecutor(threads) as ex:
        enc_out = list(ex.map(enc, jobs))
    blocks, chunks, off = [], [], 0
    for (c, bid, x, y, bw, bh), (data, npass, miss) in zip(jobs, enc_out):
        blocks.append(dict(comp=c, band=bid, x=x, y=y, w=bw, h=bh, orient=band_orient(bid),
                           numbps=qcd[bid][0], missing=miss, num_pass=npass, offset=off,
                           len=len(data)))
        chunks.append(data)
        off += len(data)
    return dict(q=q, blocks=blocks, data=b"".join(chunks) + b"\xff" * 16, steps=steps,
                filt=filt, levels=levels, w=w, h=h, ncomp=ncomp)


def _fdwt53_axis0(x: np.ndarray) -> np.ndarray:
    """One forward reversible 5/3 decomposition (ITU-T T.800 F.4.8.2,
    signal starting at an even index, symmetric extension) along axis 0;
    returns the low samples followed by the high samples."""
    n = x.shape[0]
    if n == 1:
        return x.copy()
    ev, od = x[0::2], x[1::2]
    right = np.concatenate([ev[1:], ev[-1:] if n % 2 == 0 else ev[-2:-1]], axis=0)[:od.shape[0]]
    d = od - np.floor_divide(ev[:od.shape[0]] + right, 2)
    dl = np.concatenate([d[:1], d], axis=0)[:ev.shape[0]]           # Y(2k-1), Y(-1) = Y(1)
    dr = np.concatenate([d, d[-1:]], axis=0)[:ev.shape[0]]          # Y(2k+1), Y(N) = Y(N-2)
    s = ev + np.floor_divide(dl + dr + 2, 4)
    return np.concatenate([s, d], axis=0)


def forward_53(plane: np.ndarray, levels: int) -> np.ndarray:
    """Forward 5/3 DWT of a tile at the origin into Mallat layout: per
    level the vertical decomposition, then the horizontal one (T.800
    2D_SD, F.4.2), repeated on the LL quadrant."""
    a = np.asarray(plane, dtype=np.int64).copy()
    h, w = a.shape
    for _ in range(levels):
        if h == 0 or w == 0:
            break
        sub = a[:h, :w]
        sub = _fdwt53_axis0(sub)
        sub = _fdwt53_axis0(sub.T).T
        a[:h, :w] = sub
        h, w = (h + 1) // 2, (w + 1) // 2
    return a


def forward_rct(rgb: np.ndarray) -> np.ndarray:
    """Reversible colour transform (T.800 G.2) of (..., 3) level-shifted
    samples -> (Y0, Y1, Y2) = (floor((R+2G+B)/4), B-G, R-G)."""
    r, g, b = (rgb[..., i].astype(np.int64) for i in range(3))
    return np.stack([np.floor_divide(r + 2 * g + b, 4), b - g, r - g], axis=-1)


def lossless_image(pixels: np.ndarray, levels: int, *, bit_depth: int = 8, cb: int = 64,
                   style: int = 0, mct: bool | None = None, guard: int = 2,
                   threads: int = 8) -> dict:
    """A lossless (reversible 5/3) encode of an image with the project's own
    tier-1 encoder: level shift, RCT for 3+ components (when `mct`),
    forward 5/3, code-blocks of cb x cb coded with code-block style
    `style`.  Band exponents are bit_depth + log2 of the band gain (LL 0,
    HL / LH 1, HH 2), guard bits `guard`, so a band has
    guard + exponent - 1 magnitude bit-planes.  Returns the dict of
    t1_image (filt "53") plus seg_lens per block, style, mct, guard and the
    band exponents, ready for t1_codestream."""
    from concurrent.futures import ThreadPoolExecutor
    px = np.asarray(pixels)
    if px.ndim == 2:
        px = px[..., None]
    h, w, nc = px.shape
    if mct is None:
        mct = nc >= 3
    v = px.astype(np.int64) - (1 << (bit_depth - 1))
    if mct:
        v = np.concatenate([forward_rct(v[..., :3]), v[..., 3:]], axis=-1)
    q = np.stack([forward_53(v[..., c], levels) for c in range(nc)]).astype(np.int32)
    rects = band_rects(w, h, 0, 0, levels)
    gain = {0: 0, 1: 1, 2: 1, 3: 2}
    expn = {bid: bit_depth + gain[band_orient(bid)] for (bid, *_r) in rects}
    jobs = []
    for c in range(nc):
        for (bid, dl, bx, by, bw, bh) in rects:
            for y0 in range(0, bh, cb):
                for x0 in range(0, bw, cb):
                    jobs.append((c, bid, bx + x0, by + y0, min(cb, bw - x0), min(cb, bh - y0)))

    def enc(j):
        c, bid, x, y, bw, bh = j
        return t1_encode_block(q[c, y:y + bh, x:x + bw], guard + expn[bid] - 1, band_orient(bid),
                               style=style, segments=True)

    with ThreadPoolExecutor(threads) as ex:
        enc_out = list(ex.map(enc, jobs))
    blocks, chunks, off = [], [], 0
    for (c, bid, x, y, bw, bh), (data, npass, miss, segs) in zip(jobs, enc_out):
        blocks.append(dict(comp=c, band=bid, x=x, y=y, w=bw, h=bh, orient=band_orient(bid),
                           numbps=guard + expn[bid] - 1, missing=miss, num_pass=npass, offset=off,
                           len=len(data), seg_lens=segs))
        chunks.append(data)
        off += len(data)
    return dict(q=q, blocks=blocks, data=b"".join(chunks) + b"\xff" * 16,
                steps=np.ones((nc, 3 * levels + 1), np.float32), filt="53", levels=levels, w=w,
                h=h, ncomp=nc, style=style, mct=bool(mct), guard=guard, expn=expn,
                bit_depth=bit_depth)


# A JPEG 2000 codestream around the code-blocks of t1_image: one tile, one
# layer, LRCP, default precincts, guard bits 1 (ITU-T T.800 Annex A markers,
# B.10 packet headers), so the codestream front end decodes exactly the
# blocks the tier-1 workload holds.

class _HeaderBits:
    """Packet-header bit packing: a byte after 0xFF carries 7 bits."""

    def __init__(self):
        self.buf, self.acc, self.n, self.cap = bytearray(), 0, 0, 8

    def put(self, bit: int):
        self.acc = (self.acc << 1) | (bit & 1)
        self.n += 1
        if self.n == self.cap:
            self.buf.append(self.acc)
            self.cap = 7 if self.acc == 0xFF else 8
            self.acc, self.n = 0, 0

    def put_bits(self, v: int, n: int):
        for i in range(n - 1, -1, -1):
            self.put(v >> i)

    def finish(self) -> bytes:
        if self.n:
            self.buf.append(self.acc << (self.cap - self.n))
        if self.buf and self.buf[-1] == 0xFF:
            self.buf.append(0)
        return bytes(self.buf)


class _TagTreeWriter:
    """Tag-tree encoder over a leaf grid (node value = minimum of its
    children; per node the current lower bound and whether it is known)."""

    def __init__(self, vals: np.ndarray):
        self.lv = [np.asarray(vals, dtype=np.int64)]
        while self.lv[-1].shape != (1, 1):
            v = self.lv[-1]
            h, w = v.shape
            p = np.full(((h + 1) // 2, (w + 1) // 2), 1 << 40, dtype=np.int64)
            for dy in (0, 1):
                for dx in (0, 1):
                    s = v[dy::2, dx::2]
                    p[:s.shape[0], :s.shape[1]] = np.minimum(p[:s.shape[0], :s.shape[1]], s)
            self.lv.append(p)
        self.low = [np.zeros(v.shape, np.int64) for v in self.lv]
        self.known = [np.zeros(v.shape, bool) for v in self.lv]

    def encode(self, bits: _HeaderBits, i: int, j: int, threshold: int):
        low = 0
        for k in range(len(self.lv) - 1, -1, -1):
            a, b = i >> k, j >> k
            if low > self.low[k][a, b]:
                self.low[k][a, b] = low
            else:
                low = int(self.low[k][a, b])
            val = int(self.lv[k][a, b])
            while low < threshold:
                if low >= val:
                    if not self.known[k][a, b]:
                        bits.put(1)
                        self.known[k][a, b] = True
                    break
                bits.put(0)
                low += 1
            self.low[k][a, b] = low


def _pass_codeword(bits: _HeaderBits, n: int):
    if n == 1:
        bits.put(0)
    elif n == 2:
        bits.put_bits(0b10, 2)
    elif n <= 5:
        bits.put_bits(0b1100 | (n - 3), 4)
    elif n <= 36:
        bits.put_bits(0b1111, 4)
        bits.put_bits(n - 6, 5)
    else:
        bits.put_bits(0b111111111, 9)
        bits.put_bits(n - 37, 7)


def _pass_new_segment(i: int, style: int) -> bool:
    """Pass i of a code-block opens a codeword segment (T.800 Table A.19
    switches: every pass with termination on each pass; with bypass, the
    first raw pass of each bit-plane and each cleanup from pass 10 on)."""
    if i == 0 or style & 0x04:
        return True
    return bool(style & 0x01) and i >= 10 and i % 3 != 2


def _layer_parts(b: dict, style: int, a: int, e: int):
    """Codeword-segment parts of passes [a, e) of block b: [(passes, bytes)],
    the bytes of a segment shared out over its passes."""
    segs = b.get("seg_lens") or [b["len"]]
    starts = [i for i in range(b["num_pass"]) if _pass_new_segment(i, style)]
    ends = starts[1:] + [b["num_pass"]]
    parts, pos = [], 0
    for s0, s1, ln in zip(starts, ends, segs):
        lo, hi = max(a, s0), min(e, s1)
        if lo < hi:
            c0 = ln * (lo - s0) // (s1 - s0)
            c1 = ln * (hi - s0) // (s1 - s0)
            parts.append((hi - lo, pos + c0, c1 - c0))
        pos += ln
    return parts


def t1_codestream(img: dict, cb: int = 64, layers: int = 1) -> bytes:
    """Raw codestream (.j2k) holding the code-blocks of t1_image or
    lossless_image: one tile, LRCP, default precincts, code-block style
    img["style"] (0 when absent).  With `layers` > 1 every code-block's
    passes are shared out over the quality layers (the bytes of a codeword
    segment split between the layers that hold its passes), so packets
    carry several segment lengths and segments span layers."""
    import struct
    w, h, L, nc, filt = img["w"], img["h"], img["levels"], img["ncomp"], img["filt"]
    style = img.get("style", 0)
    data = img["data"]
    by_band = {}
    for b in img["blocks"]:
        by_band.setdefault((b["comp"], b["band"]), []).append(b)
    numbps = {b["band"]: b["numbps"] for b in img["blocks"]}
    bit_depth = img.get("bit_depth", 8)
    mct = img.get("mct", nc == 3)
    out = bytearray(struct.pack(">H", 0xFF4F))
    siz = struct.pack(">HIIIIIIIIH", 0, w, h, 0, 0, w, h, 0, 0, nc) + bytes([bit_depth - 1, 1, 1]) * nc
    out += struct.pack(">HH", 0xFF51, 2 + len(siz)) + siz
    lg = cb.bit_length() - 1
    cod = struct.pack(">BBHBBBBBB", 0, 0, layers, 1 if mct else 0, L, lg - 2, lg - 2, style,
                      0 if filt == "97" else 1)
    out += struct.pack(">HH", 0xFF52, 2 + len(cod)) + cod
    if filt == "97":        # guard bits 1: QCD value = numbps
        qcd = bytes([(1 << 5) | 2]) + b"".join(struct.pack(">H", numbps[k]) for k in range(3 * L + 1))
    elif "expn" in img:     # reversible: exponents and guard bits of the encode
        qcd = bytes([img["guard"] << 5]) + bytes(img["expn"][k] << 3 for k in range(3 * L + 1))
    else:                   # exponent = numbps with guard bits 1
        qcd = bytes([(1 << 5) | 0]) + bytes(numbps[k] << 3 for k in range(3 * L + 1))
    out += struct.pack(">HH", 0xFF5C, 2 + len(qcd)) + qcd
    # per block: passes of each layer, first layer with passes, Lblock
    split = {}
    for b in img["blocks"]:
        n = b["num_pass"]
        cuts = [n * k // layers for k in range(layers + 1)]
        first = next((k for k in range(layers) if cuts[k + 1] > cuts[k]), 1 << 20)
        split[id(b)] = dict(cuts=cuts, first=first, lblock=3)
    trees = {}
    for key, blks in by_band.items():
        blks.sort(key=lambda b: (b["y"], b["x"]))
        xs = sorted({b["x"] for b in blks})
        ys = sorted({b["y"] for b in blks})
        first = np.array([split[id(b)]["first"] for b in blks]).reshape(len(ys), len(xs))
        zbp = np.array([b["missing"] if b["num_pass"] else 1 << 20 for b in blks]).reshape(first.shape)
        trees[key] = (len(xs), _TagTreeWriter(first), _TagTreeWriter(zbp))
    packets = bytearray()
    for lay in range(layers):
        for r in range(L + 1):
            for c in range(nc):
                bands = [0] if r == 0 else [3 * (r - 1) + o for o in (1, 2, 3)]
                hb, body = _HeaderBits(), bytearray()
                contrib = any(split[id(b)]["cuts"][lay + 1] > split[id(b)]["cuts"][lay]
                              for bid in bands for b in by_band.get((c, bid), []))
                if not contrib:
                    hb.put(0)
                    packets += hb.finish()
                    continue
                hb.put(1)
                for bid in bands:
                    if (c, bid) not in trees:
                        continue
                    gw, incl, zbp = trees[(c, bid)]
                    for k, b in enumerate(by_band[(c, bid)]):
                        i, j = divmod(k, gw)
                        st = split[id(b)]
                        a, e = st["cuts"][lay], st["cuts"][lay + 1]
                        if st["first"] >= lay:   # not included by an earlier layer
                            incl.encode(hb, i, j, lay + 1)
                            if st["first"] != lay:
                                continue
                            zbp.encode(hb, i, j, 1 << 30)
                        else:
                            hb.put(1 if e > a else 0)
                            if e == a:
                                continue
                        _pass_codeword(hb, e - a)
                        parts = _layer_parts(b, style, a, e)
                        need = max(max(1, ln.bit_length()) - (np_.bit_length() - 1)
                                   for np_, _o, ln in parts)
                        extra = max(0, need - st["lblock"])
                        hb.put_bits((1 << (extra + 1)) - 2, extra + 1)   # extra ones, then 0
                        st["lblock"] += extra
                        for np_, o, ln in parts:
                            hb.put_bits(ln, st["lblock"] + np_.bit_length() - 1)
                            body += data[b["offset"] + o:b["offset"] + o + ln]
                packets += hb.finish() + body
    sot_len = 12 + 2 + len(packets)
    out += struct.pack(">HHHIBB", 0xFF90, 10, 0, sot_len, 0, 1)
    out += struct.pack(">H", 0xFF93) + packets + struct.pack(">H", 0xFFD9)
    return bytes(out)
