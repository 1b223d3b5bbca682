"""CPU restatement of the reference's codestream parser and tier-2 packet
decoder (pure Python, small images only).

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product.  It
restates, function by function:
  - JP2 box walk ............ decoder/codec/src/dec_syntax.cpp:46-128
  - main header markers ..... dec_syntax.cpp:135-415 (SIZ, COD, COC, QCD, QCC)
  - tile-part headers ....... dec_syntax.cpp:423-757, tile loop dec_tsk.cpp:107-185
  - tile / band / precinct /
    code-block geometry ..... decoder/codec/src/dec_init.cpp:645-1056
  - packet header bit reader  decoder/codec/src/dec_io.cpp:99-177
  - tag trees ............... decoder/codec/src/dec_tagtree.cpp:40-142
  - packets ................. decoder/codec/src/dec_packet.cpp:37-380
  - progression orders ...... dec_packet.cpp:384-730
  - numbps of a band ........ decoder/codec/src/dec_jp2.cpp:177-178
  - multi-layer segments .... decoder/codec/src/bmi_block_decoder.cpp:742-760
  - code-block styles ....... ITU-T T.800 Table A.19, B.10.7.2 (codeword
                              segments with bypass / termination on every
                              pass; the reference asserts on them,
                              bmi_block_decoder.cpp:792-797, 818-822)
  - 9/7 step (mfAbsStepNorm)  decoder/cudainterface/src/cudainterface.cpp:712-720

Where the reference is narrower than ITU-T T.800 this restatement follows
T.800, which reduces to the reference on every stream the reference
decodes correctly:
  - precinct sizes per resolution (the reference uses the highest
    resolution's size for every resolution, dec_packet.cpp:812-813; equal
    whenever band-space precincts are uniform, e.g. default precincts);
  - image / tile origins (the reference reads and drops XOsiz, YOsiz,
    XTOsiz, YTOsiz, dec_syntax.cpp:158-164: equal when they are 0);
  - QCD/QCC by marker length, scalar-derived steps per T.800 E.1.1.1 (the
    reference reads 3L+1 values for every style, dec_syntax.cpp:300-318);
  - SOP / EPH markers accepted when present (the reference requires them
    when Scod flags them, dec_packet.cpp:339-366).
The tag tree keeps the reference's (low, value) node representation; the
product's C++ parser uses the T.800 (lower bound, known) formulation, so the
two are independent restatements of one decoding rule.
"""
from __future__ import annotations

import math

import numpy as np

MK_SOC, MK_SOT, MK_SOD, MK_EOC = 0xFF4F, 0xFF90, 0xFF93, 0xFFD9
MK_SIZ, MK_COD, MK_COC, MK_TLM, MK_PLM, MK_PLT = 0xFF51, 0xFF52, 0xFF53, 0xFF55, 0xFF57, 0xFF58
MK_QCD, MK_QCC, MK_RGN, MK_POC, MK_PPM, MK_PPT = 0xFF5C, 0xFF5D, 0xFF5E, 0xFF5F, 0xFF60, 0xFF61
MK_CRG, MK_COM, MK_SOP, MK_EPH = 0xFF63, 0xFF64, 0xFF91, 0xFF92

LRCP, RLCP, RPCL, PCRL, CPRL = range(5)


class Unsupported(Exception):
    """A stream feature the back end rejects (as the reference does)."""


def _ceil(a: int, b: int) -> int:
    return -(-a // b)


def _u(buf, pos, n):
    return int.from_bytes(buf[pos:pos + n], "big")


# ------------------------------------------------------------- JP2 boxes --

def codestream_of(buf: bytes) -> bytes:
    """Raw codestream, or the contents of a JP2 file's jp2c box
    (dec_read_jp2_header, dec_syntax.cpp:46-128: colr with method 1 must be
    sRGB 16, grey 17 or sYCC 18)."""
    if buf[:2] == b"\xff\x4f":
        return buf
    if buf[:12] != b"\x00\x00\x00\x0cjP  \r\n\x87\n":
        raise ValueError("neither a codestream nor a JP2 file")
    pos = 0
    while pos + 8 <= len(buf):
        ln, typ = _u(buf, pos, 4), buf[pos + 4:pos + 8]
        hdr = 8
        if ln == 1:
            ln, hdr = _u(buf, pos + 8, 8), 16
        elif ln == 0:
            ln = len(buf) - pos
        if typ == b"jp2h":
            sub = pos + hdr
            while sub < pos + ln:
                sl, st = _u(buf, sub, 4), buf[sub + 4:sub + 8]
                if st == b"colr" and buf[sub + 8] == 1:
                    cs = _u(buf, sub + 11, 4)
                    if cs not in (16, 17, 18):
                        raise Unsupported(f"colour space {cs}")
                sub += sl
        if typ == b"jp2c":
            return buf[pos + hdr:pos + ln]
        pos += ln
    raise ValueError("no jp2c box")


# ---------------------------------------------------------------- markers --

def _parse_cod_sp(buf, p, scod_prec, out):
    """SPcod / SPcoc fields (dec_syntax.cpp:211-250)."""
    out["levels"] = buf[p]
    out["xcb"] = (buf[p + 1] & 0xF) + 2
    out["ycb"] = (buf[p + 2] & 0xF) + 2
    out["cbstyle"] = buf[p + 3]
    out["transform"] = buf[p + 4]
    p += 5
    if scod_prec:
        out["ppx"] = [buf[p + i] & 0xF for i in range(out["levels"] + 1)]
        out["ppy"] = [buf[p + i] >> 4 for i in range(out["levels"] + 1)]
        if 0 in out["ppx"][1:] + out["ppy"][1:]:
            raise ValueError("precinct exponent 0 above resolution 0 (T.800 A.6.1)")
    else:
        out["ppx"] = [15] * (out["levels"] + 1)
        out["ppy"] = [15] * (out["levels"] + 1)


def _parse_qcd_sp(buf, p, end):
    """Sqcd + SPqcd (dec_syntax.cpp:294-318): returns (guard, style, values)
    with values[n] = exponent (style 0) or exponent<<11 | mantissa."""
    sq = buf[p]
    guard, style = sq >> 5, sq & 0x1F
    p += 1
    if style == 0:
        vals = [buf[i] >> 3 for i in range(p, end)]
    else:
        vals = [_u(buf, i, 2) for i in range(p, end - 1, 2)]
        if style == 1:
            vals = vals[:1]
    return dict(guard=guard, style=style, vals=vals)


def _band_quant(q, band, levels):
    """Quantisation value of band index `band` (0 = LL, 3k+1.. = HL, LH, HH
    of level k from the coarsest) in the reference's numbps encoding."""
    if q["style"] == 1:      # scalar derived: eps_b = eps_0 - NL + n_b, same mantissa
        e0, m0 = q["vals"][0] >> 11, q["vals"][0] & 0x7FF
        nb = levels if band == 0 else levels - (band - 1) // 3
        return ((e0 - levels + nb) << 11) | m0
    if band >= len(q["vals"]):
        raise ValueError("QCD too short")
    return q["vals"][band]


def parse(buf: bytes, guard_t800: bool = False) -> dict:
    """Parse a codestream (or JP2 file) and decode every packet header.
    guard_t800: 9/7 numbps as ((exponent + guard bits - 1) << 11) |
    mantissa (T.800 E.1) instead of the reference's exponent << 11 |
    mantissa + guard bits - 1.

    Returns dict(siz=..., tiles=[dict(tx0, ty0, tx1, ty1, levels, filt, mct,
    ncomp, steps, blocks=[...]) ...]).  Each block: comp, band, x, y (Mallat
    position in the tile-component plane), w, h, orient, numbps, missing,
    num_pass, style (code-block style bits), data (its codeword segments
    concatenated, each gathered over the layers) and seg_lens."""
    cs = codestream_of(buf)
    if _u(cs, 0, 2) != MK_SOC:
        raise ValueError("no SOC")
    pos = 2
    siz = None
    main_cod, main_coc, main_qcd, main_qcc = None, {}, None, {}
    while True:
        mk, ln = _u(cs, pos, 2), _u(cs, pos + 2, 2)
        p, end = pos + 4, pos + 2 + ln
        if mk == MK_SIZ:
            siz = dict(Xsiz=_u(cs, p + 2, 4), Ysiz=_u(cs, p + 6, 4), XOsiz=_u(cs, p + 10, 4),
                       YOsiz=_u(cs, p + 14, 4), XTsiz=_u(cs, p + 18, 4), YTsiz=_u(cs, p + 22, 4),
                       XTOsiz=_u(cs, p + 26, 4), YTOsiz=_u(cs, p + 30, 4), Csiz=_u(cs, p + 34, 2))
            comps = []
            for c in range(siz["Csiz"]):
                s = cs[p + 36 + 3 * c]
                comps.append(dict(depth=(s & 0x7F) + 1, signed=s >> 7, dx=cs[p + 37 + 3 * c],
                                  dy=cs[p + 38 + 3 * c]))
            siz["comps"] = comps
        elif mk == MK_COD:
            main_cod = _parse_cod(cs, p)
        elif mk == MK_COC:
            c, cp = _coc_comp(cs, p, siz)
            main_coc[c] = cp
        elif mk == MK_QCD:
            main_qcd = _parse_qcd_sp(cs, p, end)
        elif mk == MK_QCC:
            c, p2 = (cs[p], p + 1) if siz["Csiz"] < 257 else (_u(cs, p, 2), p + 2)
            main_qcc[c] = _parse_qcd_sp(cs, p2, end)
        elif mk in (MK_RGN, MK_POC, MK_PPM):
            raise Unsupported(f"marker {mk:04X}")
        elif mk in (MK_TLM, MK_PLM, MK_CRG, MK_COM):
            pass
        elif mk == MK_SOT:
            break
        else:
            raise ValueError(f"unexpected marker {mk:04X} in main header")
        pos = end
    if siz is None or main_cod is None or main_qcd is None:
        raise ValueError("main header lacks SIZ/COD/QCD")
    for cp in siz["comps"]:
        if cp["dx"] != 1 or cp["dy"] != 1 or cp["signed"]:
            raise Unsupported("sub-sampled or signed components")
    ntx = _ceil(siz["Xsiz"] - siz["XTOsiz"], siz["XTsiz"])
    nty = _ceil(siz["Ysiz"] - siz["YTOsiz"], siz["YTsiz"])
    tiles = [None] * (ntx * nty)
    # tile-parts (dec_read_tile_header + dec_t2's loop)
    while pos + 2 <= len(cs) and _u(cs, pos, 2) == MK_SOT:
        sot = pos
        t, psot, tpsot = _u(cs, pos + 4, 2), _u(cs, pos + 6, 4), cs[pos + 10]
        pos += 12
        if t >= len(tiles):
            raise ValueError("tile index out of range")
        if tiles[t] is None:
            tiles[t] = dict(cod=None, coc={}, qcd=None, qcc={}, data=bytearray())
        T = tiles[t]
        while True:
            mk, ln = _u(cs, pos, 2), _u(cs, pos + 2, 2)
            if mk == MK_SOD:
                pos += 2
                break
            p, end = pos + 4, pos + 2 + ln
            if mk == MK_COD and tpsot == 0:
                T["cod"] = _parse_cod(cs, p)
            elif mk == MK_COC and tpsot == 0:
                c, cp = _coc_comp(cs, p, siz)
                T["coc"][c] = cp
            elif mk == MK_QCD and tpsot == 0:
                T["qcd"] = _parse_qcd_sp(cs, p, end)
            elif mk == MK_QCC and tpsot == 0:
                c, p2 = (cs[p], p + 1) if siz["Csiz"] < 257 else (_u(cs, p, 2), p + 2)
                T["qcc"][c] = _parse_qcd_sp(cs, p2, end)
            elif mk in (MK_COM, MK_PLT):
                pass
            elif mk in (MK_RGN, MK_POC, MK_PPT):
                raise Unsupported(f"marker {mk:04X}")
            else:
                raise ValueError(f"unexpected marker {mk:04X} in tile-part header")
            pos = end
        stop = sot + psot if psot else (len(cs) - 2 if cs[-2:] == b"\xff\xd9" else len(cs))
        T["data"] += cs[pos:stop]
        pos = stop
    out = []
    for t, T in enumerate(tiles):
        if T is None:
            T = dict(cod=None, coc={}, qcd=None, qcc={}, data=bytearray())
        out.append(_decode_tile(siz, t, ntx, main_cod, main_coc, main_qcd, main_qcc, T, guard_t800))
    return dict(siz=siz, tiles=out, ntx=ntx, nty=nty)


def _parse_cod(cs, p):
    scod = cs[p]
    cod = dict(scod=scod, prog=cs[p + 1], layers=_u(cs, p + 2, 2), mct=cs[p + 4])
    _parse_cod_sp(cs, p + 5, scod & 1, cod)
    return cod


def _coc_comp(cs, p, siz):
    c, p = (cs[p], p + 1) if siz["Csiz"] < 257 else (_u(cs, p, 2), p + 2)
    cp = {}
    _parse_cod_sp(cs, p + 1, cs[p] & 1, cp)
    return c, cp


# --------------------------------------------------------- bit reader --

class _Bits:
    """Packet-header bit reader with 0xFF bit stuffing (dec_io.cpp:99-177):
    a byte following 0xFF carries 7 bits."""

    def __init__(self, data: bytes, pos: int):
        self.d, self.cur, self.bits = data, pos, 0

    def bit(self) -> int:
        if self.bits == 8:
            self.bits = 1 if self.d[self.cur] == 0xFF else 0
            self.cur += 1
        if self.cur >= len(self.d):
            raise ValueError("packet header past the end of the tile data")
        self.bits += 1
        return (self.d[self.cur] >> (8 - self.bits)) & 1

    def int(self, n: int) -> int:
        v = 0
        for _ in range(n):
            v = (v << 1) | self.bit()
        return v

    def align(self):
        if self.bits:
            if self.d[self.cur] == 0xFF:
                self.cur += 1
            self.bits = 0
            self.cur += 1


# ------------------------------------------------------------ tag trees --

class _TagTree:
    """Tag tree over a w x h leaf grid (dec_tagtree.cpp:40-76 build,
    :112-142 decode), nodes as (low, value): `value` counts the zero bits
    read at the node, `low` all bits read; the value is known once a 1 bit
    was read (low == value + 1)."""

    def __init__(self, w: int, h: int):
        self.parent, self.low, self.value = [], [], []
        base = 0
        while True:
            nw, nh = (w + 1) >> 1, (h + 1) >> 1
            nxt = base + w * h
            root = (w == 1 and h == 1)
            for i in range(h):
                for j in range(w):
                    self.parent.append(-1 if root else nxt + (i >> 1) * nw + (j >> 1))
            base = nxt
            if root:
                break
            w, h = nw, nh
        self.low = [0] * len(self.parent)
        self.value = [0] * len(self.parent)

    def decode(self, bits: _Bits, leaf: int, threshold: int) -> bool:
        path = []
        n = leaf
        while self.parent[n] >= 0:
            path.append(n)
            n = self.parent[n]
        low = 0
        while True:
            if self.low[n] < low:
                self.low[n] = self.value[n] = low
            while self.low[n] == self.value[n] and self.low[n] < threshold:
                if not bits.bit():
                    self.value[n] += 1
                self.low[n] += 1
            low = self.value[n]
            if not path:
                break
            n = path.pop()
        return self.low[n] != self.value[n]


def _num_passes(bits: _Bits) -> int:
    """Coding-pass count codeword (dec_packet.cpp:37-60, T.800 Table B.4)."""
    n = 1 + bits.bit()
    if n == 2:
        n += bits.bit()
        if n == 3:
            n += bits.int(2)
            if n >= 6:
                n += bits.int(5)
                if n >= 37:
                    n += bits.int(7)
    return n


# ------------------------------------------------------------- geometry --

def _tile_params(siz, T, main_cod, main_coc, main_qcd, main_qcc, c):
    """Coding / quantisation parameters of component c in a tile, by T.800
    precedence: tile COC > tile COD > main COC > main COD (and the same
    for QCC / QCD)."""
    cod = dict(T["cod"] or main_cod)
    sp = T["coc"].get(c) or (None if T["cod"] else main_coc.get(c))
    if sp:
        cod.update(sp)
    q = T["qcc"].get(c) or T["qcd"] or main_qcc.get(c) or main_qcd
    return cod, q


def _decode_tile(siz, t, ntx, main_cod, main_coc, main_qcd, main_qcc, T, guard_t800):
    p, q = t % ntx, t // ntx
    tx0 = max(siz["XTOsiz"] + p * siz["XTsiz"], siz["XOsiz"])
    ty0 = max(siz["YTOsiz"] + q * siz["YTsiz"], siz["YOsiz"])
    tx1 = min(siz["XTOsiz"] + (p + 1) * siz["XTsiz"], siz["Xsiz"])
    ty1 = min(siz["YTOsiz"] + (q + 1) * siz["YTsiz"], siz["Ysiz"])
    ncomp = siz["Csiz"]
    comps = []
    for c in range(ncomp):
        cod, qc = _tile_params(siz, T, main_cod, main_coc, main_qcd, main_qcc, c)
        if cod["cbstyle"] & ~0x3F:
            raise Unsupported(f"code-block style {cod['cbstyle']:#x}")
        if cod["xcb"] > 6 or cod["ycb"] > 6:
            raise Unsupported("code-blocks over 64 samples a side")
        comps.append(_comp_geometry(tx0, ty0, tx1, ty1, cod, qc, siz["comps"][c]["depth"], c,
                                    guard_t800))
    cod0 = comps[0]["cod"]
    tile = dict(tx0=tx0, ty0=ty0, tx1=tx1, ty1=ty1, ncomp=ncomp, levels=cod0["levels"],
                filt="53" if cod0["transform"] == 1 else "97", mct=cod0["mct"],
                layers=cod0["layers"], prog=cod0["prog"], comps=comps)
    _read_packets(tile, bytes(T["data"]), cod0)
    blocks = []
    for cp in comps:
        for b in cp["blocks"]:
            blocks.append(dict(comp=cp["c"], band=b["band"], x=b["x"], y=b["y"], w=b["w"],
                               h=b["h"], orient=b["orient"], numbps=b["numbps"],
                               missing=b["missing"], num_pass=b["num_pass"],
                               style=b["style"], data=b"".join(b["segs"]),
                               seg_lens=[len(x) for x in b["segs"]]))
    tile["blocks"] = blocks
    tile["steps"] = [cp["steps"] for cp in comps]
    return tile


def _comp_geometry(tx0, ty0, tx1, ty1, cod, qc, depth, c, guard_t800=False):
    """Resolutions, bands, precincts and code-blocks of one tile-component
    (dec_init.cpp:717-1050, per-resolution precinct sizes per T.800)."""
    NL = cod["levels"]
    res, blocks = [], []
    steps = []
    for r in range(NL + 1):
        s = NL - r
        trx0, try0 = _ceil(tx0, 1 << s), _ceil(ty0, 1 << s)
        trx1, try1 = _ceil(tx1, 1 << s), _ceil(ty1, 1 << s)
        ppx, ppy = cod["ppx"][r], cod["ppy"][r]
        npx = (_ceil(trx1, 1 << ppx) - (trx0 >> ppx)) if trx1 > trx0 else 0
        npy = (_ceil(try1, 1 << ppy) - (try0 >> ppy)) if try1 > try0 else 0
        sh = 1 if r else 0
        xcb = min(cod["xcb"], ppx - sh)
        ycb = min(cod["ycb"], ppy - sh)
        bands = []
        for o in ((0,) if r == 0 else (1, 2, 3)):
            nb = NL if r == 0 else NL - r + 1
            xob, yob = o & 1, o >> 1
            tbx0 = _ceil(tx0 - xob * (1 << nb >> 1), 1 << nb)
            tby0 = _ceil(ty0 - yob * (1 << nb >> 1), 1 << nb)
            tbx1 = _ceil(tx1 - xob * (1 << nb >> 1), 1 << nb)
            tby1 = _ceil(ty1 - yob * (1 << nb >> 1), 1 << nb)
            # Mallat position: HL right of, LH below the LL of its level
            llw = _ceil(tx1, 1 << nb) - _ceil(tx0, 1 << nb)
            llh = _ceil(ty1, 1 << nb) - _ceil(ty0, 1 << nb)
            bid = 0 if r == 0 else 3 * (r - 1) + o
            qv = _band_quant(qc, bid, NL)
            numbps = qv + qc["guard"] - 1
            if guard_t800 and cod["transform"] != 1:
                numbps = (((qv >> 11) + qc["guard"] - 1) << 11) | (qv & 0x7FF)
            if cod["transform"] == 1:
                steps.append(1.0)
            else:
                steps.append(_abs_step_norm(qv >> 11, qv & 0x7FF, depth))
            band = dict(o=o, bid=bid, tbx0=tbx0, tby0=tby0, tbx1=tbx1, tby1=tby1, prec=[])
            gx0, gy0 = tbx0 >> xcb, tby0 >> ycb
            ncbx = (_ceil(tbx1, 1 << xcb) - gx0) if tbx1 > tbx0 else 0
            ncby = (_ceil(tby1, 1 << ycb) - gy0) if tby1 > tby0 else 0
            grid = {}
            for j in range(ncby):
                for i in range(ncbx):
                    x0 = max((gx0 + i) << xcb, tbx0)
                    y0 = max((gy0 + j) << ycb, tby0)
                    x1 = min((gx0 + i + 1) << xcb, tbx1)
                    y1 = min((gy0 + j + 1) << ycb, tby1)
                    b = dict(band=bid, orient=o, x=x0 - tbx0 + llw * xob, y=y0 - tby0 + llh * yob,
                             w=x1 - x0, h=y1 - y0, numbps=numbps, missing=0, num_pass=0,
                             lblock=3, incl=False, segs=[], style=cod["cbstyle"])
                    grid[(i, j)] = b
                    blocks.append(b)
            # precincts of this band (dec_init.cpp:963-1008)
            for py in range(npy):
                for px in range(npx):
                    ppx0 = (((trx0 >> ppx) + px) << ppx) >> sh
                    ppy0 = (((try0 >> ppy) + py) << ppy) >> sh
                    ppx1, ppy1 = ppx0 + (1 << ppx >> sh), ppy0 + (1 << ppy >> sh)
                    ppx0, ppy0 = max(ppx0, tbx0), max(ppy0, tby0)
                    ppx1, ppy1 = min(ppx1, tbx1), min(ppy1, tby1)
                    if ppx1 <= ppx0 or ppy1 <= ppy0:
                        band["prec"].append(None)
                        continue
                    cx0, cy0 = ppx0 >> xcb, ppy0 >> ycb
                    cw = _ceil(ppx1, 1 << xcb) - cx0
                    ch = _ceil(ppy1, 1 << ycb) - cy0
                    cbs = [grid[(cx0 - gx0 + i, cy0 - gy0 + j)] for j in range(ch) for i in range(cw)]
                    band["prec"].append(dict(cbs=cbs, incl=_TagTree(cw, ch), zbp=_TagTree(cw, ch)))
            bands.append(band)
        res.append(dict(r=r, trx0=trx0, try0=try0, trx1=trx1, try1=try1, ppx=ppx, ppy=ppy,
                        npx=npx, npy=npy, bands=bands))
    return dict(c=c, cod=cod, res=res, blocks=blocks, steps=steps)


def _abs_step_norm(exp, mant, depth):
    """mfAbsStepNorm (cudainterface.cpp:712-720)."""
    step = (mant << 2) | 0x2000
    n = depth - exp
    step = step << n if n > 0 else step >> (-n)
    return float(np.float32(step) * np.float32(2.0 ** -13))


# -------------------------------------------------------------- packets --

def pass_new_segment(i: int, style: int) -> bool:
    """Coding pass i of a code-block (0 = the first cleanup pass) opens a
    codeword segment (T.800 Table A.19, D.6): the first pass; every pass
    when each pass is terminated; with bypass, from pass 10 on each raw
    significance pass and each cleanup pass (the raw segment of a
    bit-plane holds its significance and refinement passes)."""
    if i == 0 or style & 0x04:
        return True
    return bool(style & 0x01) and i >= 10 and i % 3 != 2


def _segment_parts(p0: int, n: int, style: int):
    """[(segment index, passes)] of the passes [p0, p0 + n)."""
    k = sum(pass_new_segment(i, style) for i in range(p0 + 1)) - 1
    parts = []
    for i in range(p0, p0 + n):
        if i > p0 and pass_new_segment(i, style):
            k += 1
        if parts and parts[-1][0] == k:
            parts[-1][1] += 1
        else:
            parts.append([k, 1])
    return [tuple(x) for x in parts]



def _read_packet(bits: _Bits, data: bytes, res: dict, pi: int, layer: int, sop: bool, eph: bool,
                 log: list):
    """dec_read_packet + dec_prec + dec_restore_coding_pass
    (dec_packet.cpp:66-170, 172-220, 223-380)."""
    if bits.cur + 1 < len(data) and _u(data, bits.cur, 2) == MK_SOP:
        bits.cur += 6
    nonempty = bits.bit()
    precs = [b["prec"][pi] for b in res["bands"]]
    if nonempty:
        for pr in precs:
            if pr is None:
                continue
            for k, cb in enumerate(pr["cbs"]):
                if not cb["incl"]:
                    inc = pr["incl"].decode(bits, k, layer + 1)
                else:
                    inc = bits.bit()
                if not inc:
                    continue
                if not cb["incl"]:
                    pr["zbp"].decode(bits, k, 1 << 62)
                    cb["missing"] = pr["zbp"].value[k]
                    if cb["missing"] > 74:
                        raise ValueError("missing bit-planes > 74")
                cb["incl"] = True
                n = _num_passes(bits)
                p0 = cb["num_pass"]
                cb["num_pass"] += n
                if cb["num_pass"] > 128:
                    raise ValueError("more than 128 coding passes")
                while bits.bit():
                    cb["lblock"] += 1
                # one length per codeword segment the new passes touch,
                # Lblock + floor(log2(passes of that segment)) bits each
                cb["_new"] = [(k, bits.int(cb["lblock"] + int(math.floor(math.log2(m)))))
                              for k, m in _segment_parts(p0, n, cb["style"])]
    bits.align()
    hdr_end = bits.cur
    if eph and bits.cur + 1 < len(data) and _u(data, bits.cur, 2) == MK_EPH:
        bits.cur += 2
    if not nonempty:
        log.append((hdr_end, bits.cur))
        return
    for pr in precs:
        if pr is None:
            continue
        for cb in pr["cbs"]:
            for k, n in cb.pop("_new", ()):
                if bits.cur + n > len(data):
                    raise ValueError("code-block data past the end of the tile")
                while len(cb["segs"]) <= k:
                    cb["segs"].append(b"")
                cb["segs"][k] += data[bits.cur:bits.cur + n]
                bits.cur += n
    log.append((hdr_end, bits.cur))


def _packet_order(tile):
    """Packets of a tile in progression order (T.800 B.12; the reference's
    dec_lrcp .. dec_cprl, dec_packet.cpp:384-730): (layer, comp, res, k)."""
    comps, L = tile["comps"], tile["layers"]
    prog = tile["prog"]
    maxr = max(cp["cod"]["levels"] for cp in comps)
    if prog in (LRCP, RLCP):
        order = []
        outer = [(l, r) for l in range(L) for r in range(maxr + 1)] if prog == LRCP else \
                [(l, r) for r in range(maxr + 1) for l in range(L)]
        for l, r in outer:
            for cp in comps:
                if r >= len(cp["res"]):
                    continue
                R = cp["res"][r]
                for k in range(R["npx"] * R["npy"]):
                    order.append((l, cp["c"], r, k))
        return order
    tx0, ty0, tx1, ty1 = tile["tx0"], tile["ty0"], tile["tx1"], tile["ty1"]
    dx = min(1 << (R["ppx"] + len(cp["res"]) - 1 - R["r"]) for cp in comps for R in cp["res"])
    dy = min(1 << (R["ppy"] + len(cp["res"]) - 1 - R["r"]) for cp in comps for R in cp["res"])

    def at(cp, r, x, y):
        R = cp["res"][r]
        if R["npx"] == 0 or R["npy"] == 0:
            return None
        s = len(cp["res"]) - 1 - r
        okx = x % (1 << (R["ppx"] + s)) == 0 or (x == tx0 and (R["trx0"] << s) % (1 << (R["ppx"] + s)))
        oky = y % (1 << (R["ppy"] + s)) == 0 or (y == ty0 and (R["try0"] << s) % (1 << (R["ppy"] + s)))
        if not (okx and oky):
            return None
        i = (_ceil(x, 1 << s) >> R["ppx"]) - (R["trx0"] >> R["ppx"])
        j = (_ceil(y, 1 << s) >> R["ppy"]) - (R["try0"] >> R["ppy"])
        return i + j * R["npx"]

    def ys():
        y = ty0
        while y < ty1:
            yield y
            y += dy - (y % dy)

    def xs():
        x = tx0
        while x < tx1:
            yield x
            x += dx - (x % dx)

    order = []
    if prog == RPCL:
        for r in range(maxr + 1):
            for y in ys():
                for x in xs():
                    for cp in comps:
                        if r < len(cp["res"]):
                            k = at(cp, r, x, y)
                            if k is not None:
                                order += [(l, cp["c"], r, k) for l in range(L)]
    elif prog == PCRL:
        for y in ys():
            for x in xs():
                for cp in comps:
                    for r in range(len(cp["res"])):
                        k = at(cp, r, x, y)
                        if k is not None:
                            order += [(l, cp["c"], r, k) for l in range(L)]
    elif prog == CPRL:
        for cp in comps:
            for y in ys():
                for x in xs():
                    for r in range(len(cp["res"])):
                        k = at(cp, r, x, y)
                        if k is not None:
                            order += [(l, cp["c"], r, k) for l in range(L)]
    else:
        raise Unsupported(f"progression {prog}")
    return order


def _read_packets(tile, data: bytes, cod0):
    """Decode the tile's packets in progression order; tile["packets"]
    records (start, header end, end) of each packet read."""
    bits = _Bits(data, 0)
    sop, eph = bool(cod0["scod"] & 2), bool(cod0["scod"] & 4)
    log = []
    for (l, c, r, k) in _packet_order(tile):
        if bits.cur >= len(data):
            break          # truncated tile: remaining packets absent
        start = bits.cur
        _read_packet(bits, data, tile["comps"][c]["res"][r], k, l, sop, eph, log)
        log[-1] = (start,) + log[-1]
    tile["packets"] = log
    tile["data"] = data


# --------------------------------------------------- decode to pixels --

def tile_planes(tile, orc):
    """Run the oracle's T1 on every block of a parsed tile: returns
    (ncomp, h, w) int32 Mallat planes (T1 words for 9/7, integers for 5/3)
    and the (ncomp, 3L+1) steps."""
    w, h = tile["tx1"] - tile["tx0"], tile["ty1"] - tile["ty0"]
    planes = np.zeros((tile["ncomp"], h, w), dtype=np.int32)
    for b in tile["blocks"]:
        words = orc.t1_decode_block_ex(b["data"], b["w"], b["h"], b["orient"], b["numbps"],
                                       b["missing"], b["num_pass"], b["style"], b["seg_lens"])
        if tile["filt"] == "53":
            words = orc.t1_words_to_int(words, b["numbps"])
        planes[b["comp"], b["y"]:b["y"] + b["h"], b["x"]:b["x"] + b["w"]] = words
    return planes, np.asarray(tile["steps"], dtype=np.float32)


def decode_image(buf: bytes, orc, *, out_format: int, check_out: int = -1, r53_mode: int = 0,
                 guard_t800: bool = False, toplevel: int = -1, ignore_mask: int = 0):
    """Whole-image oracle decode: T2 (this module), T1 and the reference
    pipeline (oracle.decode_tile) per tile, tiles placed in one packed
    image (rows of width * pixel bytes), or per-component planes for
    out_format 7 (NONE)."""
    P = parse(buf, guard_t800)
    siz = P["siz"]
    L0 = P["tiles"][0]["levels"]
    sh = L0 - toplevel if 0 <= toplevel < L0 else 0

    def cd(a):
        return -(-a // (1 << sh))
    W, H = cd(siz["Xsiz"]) - cd(siz["XOsiz"]), cd(siz["Ysiz"]) - cd(siz["YOsiz"])
    depth = siz["comps"][0]["depth"]
    nc = siz["Csiz"]
    if out_format == 7:
        out = [np.zeros((H, W), np.int32) for _ in range(nc)]
    else:
        pb = orc.lib().or_pixel_bytes(out_format)
        out = np.zeros((H, W * pb), np.uint8)
    for T in P["tiles"]:
        planes, steps = tile_planes(T, orc)
        rec, packed = orc.decode_tile(planes, levels=T["levels"], filt=T["filt"],
                                      out_format=out_format, bit_depth=depth, check_out=check_out,
                                      x0=T["tx0"], y0=T["ty0"], steps=steps,
                                      coef_kind=orc.COEF_SIGNMAG, r53_mode=r53_mode,
                                      toplevel=toplevel if sh else -1, ignore_mask=ignore_mask)
        ox, oy = cd(T["tx0"]) - cd(siz["XOsiz"]), cd(T["ty0"]) - cd(siz["YOsiz"])
        h, w = cd(T["ty1"]) - cd(T["ty0"]), cd(T["tx1"]) - cd(T["tx0"])
        if out_format == 7:
            kept = [c for c in range(nc) if not (ignore_mask >> c) & 1]
            for i, c in enumerate(kept):
                out[c][oy:oy + h, ox:ox + w] = np.asarray(rec[i]).view(np.int32)
        else:
            out[oy:oy + h, ox * pb:(ox + w) * pb] = packed
    return out
