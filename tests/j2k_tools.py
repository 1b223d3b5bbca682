"""Codestream rewriting for the tier-2 tests: re-emit a parsed codestream
with SOP / EPH markers, several tile-parts per tile, truncated tiles or a
foreign marker, so the parsers meet stream features the fixture encoder
does not write.  Packet boundaries come from the oracle's parse
(tile["packets"] = (start, header end, end) per packet)."""
from __future__ import annotations

import struct

import t2_oracle as t2


def main_header(cs: bytes) -> bytearray:
    """Bytes from SOC up to the first SOT."""
    pos = 2
    while t2._u(cs, pos, 2) != t2.MK_SOT:
        pos += 2 + t2._u(cs, pos + 2, 2)
    return bytearray(cs[:pos])


def patch_scod(hdr: bytearray, set_bits: int) -> None:
    pos = 2
    while pos < len(hdr):
        mk, ln = t2._u(hdr, pos, 2), t2._u(hdr, pos + 2, 2)
        if mk == t2.MK_COD:
            hdr[pos + 4] |= set_bits
        pos += 2 + ln


def insert_marker(hdr: bytearray, marker: int, payload: bytes) -> bytearray:
    """Append a marker segment at the end of the main header."""
    return hdr + struct.pack(">HH", marker, len(payload) + 2) + payload


def rewrite(buf: bytes, *, sop=False, eph=False, parts=1, keep_packets=None, extra_main=b"",
            comment_in_tile=False) -> bytes:
    """Re-emit `buf` (a codestream or JP2 file; the result is a raw
    codestream).  sop / eph: add the markers (and set Scod's flags);
    parts: tile-parts per tile, split at packet boundaries; keep_packets:
    keep only the first n packets of every tile (a truncated stream);
    extra_main: raw marker segments appended to the main header;
    comment_in_tile: a COM segment in every tile-part header."""
    cs = t2.codestream_of(buf)
    P = t2.parse(cs)
    hdr = main_header(cs)
    patch_scod(hdr, (2 if sop else 0) | (4 if eph else 0))
    out = bytearray(hdr) + extra_main
    for t, T in enumerate(P["tiles"]):
        data = T["data"]
        pk = T["packets"] if keep_packets is None else T["packets"][:keep_packets]
        bodies = []
        for i, (start, hend, end) in enumerate(pk):
            h, b = data[start:hend], data[hend:end]
            seg = bytearray()
            if sop:
                seg += struct.pack(">HHH", t2.MK_SOP, 4, i & 0xFFFF)
            seg += h
            if eph:
                seg += struct.pack(">H", t2.MK_EPH)
            seg += b
            bodies.append(bytes(seg))
        n = max(1, min(parts, len(bodies)))
        cuts = [len(bodies) * i // n for i in range(n + 1)]
        for i in range(n):
            body = b"".join(bodies[cuts[i]:cuts[i + 1]])
            extra = struct.pack(">HH", t2.MK_COM, 6) + b"\x00\x01hi" if comment_in_tile else b""
            psot = 12 + len(extra) + 2 + len(body)
            out += struct.pack(">HHHIBB", t2.MK_SOT, 10, t, psot, i, n) + extra
            out += struct.pack(">H", t2.MK_SOD) + body
    out += struct.pack(">H", t2.MK_EOC)
    return bytes(out)
