// bmi_codestream.cpp -- host-side JPEG 2000 codestream parser and tier-2
// (packet header) decoder feeding the GPU tier-1 decoder and the fused
// IDWT/colour plan.
//
// Replaces the reference's T2 stage: marker parsing
// (decoder/codec/src/dec_syntax.cpp:46-757), tile / band / precinct /
// code-block geometry (decoder/codec/src/dec_init.cpp:645-1056), the
// packet loop (decoder/codec/src/dec_tsk.cpp:107-185,
// decoder/codec/src/dec_packet.cpp:37-730), tag trees
// (decoder/codec/src/dec_tagtree.cpp:40-142) and the stuffed bit reader
// (decoder/codec/src/dec_io.cpp:99-177).  Written from ITU-T T.800 Annex A
// (markers), B (geometry, packets, progression orders) and E (quantisation):
//  - tag trees keep a lower bound and a "known" flag per node (T.800 B.10.2);
//  - position-driven progressions are a sort of the precincts by their
//    anchor on the reference grid (the first (x, y) at which T.800 B.12's
//    scan reaches them), not a scan of the tile;
//  - a tile's tile-parts are concatenated and its packets decoded in one
//    pass; each code-block keeps the (offset, length) of its chunk in every
//    packet, and its layers are concatenated when the table is emitted (the
//    reference concatenates them in decode_code_block,
//    bmi_block_decoder.cpp:742-760).
// numbps per band is the reference's (dec_jp2.cpp:177-178) unless
// BMI_CS_GUARD_T800 asks for the T.800 placement of 9/7 bit-planes.

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "bmi_codestream.h"
#include "bmi_t1_style.h"

int bmi_fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));

namespace {

struct ParseError {
    int code;
    std::string what;
};

[[noreturn]] void fail(int code, const std::string &what) { throw ParseError{code, what}; }

// ceil(a / b) for b > 0 and any sign of a (band origins of T.800 B-15 can
// be negative before the division)
inline int ceil_div(long long a, long long b) { return (int)(a >= 0 ? (a + b - 1) / b : -((-a) / b)); }
inline int floor_log2(unsigned v)
{
    int n = -1;
    while (v) {
        ++n;
        v >>= 1;
    }
    return n;
}

// big-endian reader over a byte range
struct Reader {
    const uint8_t *d;
    size_t end;
    size_t pos;
    void need(size_t n) const
    {
        if (pos + n > end) fail(BMI_E_INVALID, "truncated marker segment");
    }
    unsigned u8() { need(1); return d[pos++]; }
    unsigned u16() { need(2); unsigned v = (d[pos] << 8) | d[pos + 1]; pos += 2; return v; }
    unsigned u32() { need(4); unsigned v = ((unsigned)d[pos] << 24) | (d[pos + 1] << 16) | (d[pos + 2] << 8) | d[pos + 3]; pos += 4; return v; }
};

// COD / COC coding style of one tile-component
struct CodingStyle {
    int scod = 0;         // bit 0 explicit precincts, bit 1 SOP, bit 2 EPH
    int prog = 0, layers = 1, mct = 0;
    int levels = 0, xcb = 6, ycb = 6, cbstyle = 0, transform = 0;
    uint8_t ppx[33], ppy[33];
};

// QCD / QCC
struct QuantStyle {
    int guard = 1, style = 0;
    std::vector<unsigned> v;   // exponent (style 0) or exponent<<11 | mantissa
};

enum { MK_SOC = 0xFF4F, MK_SOT = 0xFF90, MK_SOD = 0xFF93, MK_EOC = 0xFFD9, MK_SIZ = 0xFF51,
       MK_COD = 0xFF52, MK_COC = 0xFF53, MK_TLM = 0xFF55, MK_PLM = 0xFF57, MK_PLT = 0xFF58,
       MK_QCD = 0xFF5C, MK_QCC = 0xFF5D, MK_RGN = 0xFF5E, MK_POC = 0xFF5F, MK_PPM = 0xFF60,
       MK_PPT = 0xFF61, MK_CRG = 0xFF63, MK_COM = 0xFF64, MK_SOP = 0xFF91, MK_EPH = 0xFF92 };

void read_spcod(Reader &r, int precincts, CodingStyle &cs, unsigned seg_end)
{
    cs.levels = r.u8();
    cs.xcb = (r.u8() & 0xF) + 2;
    cs.ycb = (r.u8() & 0xF) + 2;
    cs.cbstyle = r.u8();
    cs.transform = r.u8();
    if (cs.levels > 32) fail(BMI_E_INVALID, "more than 32 decomposition levels");
    if (cs.xcb + cs.ycb > 12) fail(BMI_E_INVALID, "code-block area over 4096");
    for (int i = 0; i <= cs.levels; ++i) {
        if (precincts) {
            if (r.pos >= seg_end) fail(BMI_E_INVALID, "COD precinct sizes missing");
            unsigned b = r.u8();
            cs.ppx[i] = b & 0xF;
            cs.ppy[i] = b >> 4;
            if (i && (cs.ppx[i] == 0 || cs.ppy[i] == 0))
                fail(BMI_E_INVALID, "precinct exponent 0 above resolution 0");
        } else {
            cs.ppx[i] = cs.ppy[i] = 15;
        }
    }
}

QuantStyle read_sqcd(Reader &r, unsigned seg_end)
{
    QuantStyle q;
    unsigned s = r.u8();
    q.guard = s >> 5;
    q.style = s & 0x1F;
    if (q.style > 2) fail(BMI_E_INVALID, "unknown quantisation style");
    while (r.pos < seg_end) {
        q.v.push_back(q.style == 0 ? (r.u8() >> 3) : r.u16());
        if (q.style == 1) break;
    }
    if (q.v.empty()) fail(BMI_E_INVALID, "empty QCD");
    return q;
}

// ---------------------------------------------------------------- tiers --

// T.800 B.10.2 tag tree: each node holds a lower bound of its value and
// whether the value is known
struct TagTree {
    std::vector<int> parent;
    std::vector<int> lb;
    std::vector<uint8_t> known;
    void init(int w, int h)
    {
        parent.clear();
        int base = 0;
        for (;;) {
            int nw = (w + 1) >> 1, nh = (h + 1) >> 1;
            bool root = (w == 1 && h == 1);
            int next = base + w * h;
            for (int i = 0; i < h; ++i)
                for (int j = 0; j < w; ++j)
                    parent.push_back(root ? -1 : next + (i >> 1) * nw + (j >> 1));
            if (root) break;
            base = next;
            w = nw;
            h = nh;
        }
        lb.assign(parent.size(), 0);
        known.assign(parent.size(), 0);
    }
};

// packet-header bit reader: after a 0xFF byte the next byte carries 7 bits
struct BitReader {
    const uint8_t *d;
    size_t end, pos;
    unsigned cur = 0;
    int left = 0;
    bool last_ff = false;
    int bit()
    {
        if (!left) {
            if (pos >= end) fail(BMI_E_INVALID, "packet header past the end of the tile");
            cur = d[pos++];
            left = last_ff ? 7 : 8;
            last_ff = (cur == 0xFF);
        }
        return (cur >> --left) & 1;
    }
    unsigned bits(int n)
    {
        unsigned v = 0;
        while (n--) v = (v << 1) | bit();
        return v;
    }
    // end of a packet header: drop the rest of the byte, and the stuffed
    // byte after a final 0xFF
    void align()
    {
        if (last_ff) ++pos;
        left = 0;
        last_ff = false;
    }
};

// value of `leaf` if it is below `threshold` (true), else false with the
// tree advanced to that threshold
bool tag_decode(TagTree &t, BitReader &br, int leaf, int threshold, int *value)
{
    int path[40], n = 0;
    for (int v = leaf; v >= 0; v = t.parent[v]) {
        if (n == 40) fail(BMI_E_INVALID, "tag tree too deep");
        path[n++] = v;
    }
    int low = 0;
    for (int i = n - 1; i >= 0; --i) {
        int v = path[i];
        if (!t.known[v] && t.lb[v] < low) t.lb[v] = low;
        while (!t.known[v] && t.lb[v] < threshold) {
            if (br.bit())
                t.known[v] = 1;
            else
                ++t.lb[v];
        }
        low = t.lb[v];
    }
    *value = t.lb[leaf];
    return t.known[leaf] != 0;
}

// T.800 Table B.4
int read_num_passes(BitReader &br)
{
    if (!br.bit()) return 1;
    if (!br.bit()) return 2;
    unsigned v = br.bits(2);
    if (v < 3) return 3 + v;
    v = br.bits(5);
    if (v < 31) return 6 + v;
    return 37 + br.bits(7);
}

// bytes of codeword segment `seg` of a code-block, found in one packet
struct Chunk {
    uint32_t off, len;
    int seg;
};

struct Block {
    int comp, band, x, y, w, h, orient, numbps, style = 0;
    int missing = 0, npass = 0, lblock = 3;
    bool seen = false;
    std::vector<Chunk> chunks;   // in stream order; segments never interleave
    std::vector<Chunk> pending;  // lengths read from the current packet header
    size_t bytes = 0;
    int nseg() const { return bmi_style_segmented(style) ? BMI_t1_num_segments(npass, style) : 1; }
};

struct Precinct {
    int cw = 0, ch = 0;
    std::vector<int> blocks;     // tile block indices, raster order
    TagTree incl, zbp;
};

struct Band {
    std::vector<Precinct> prec;  // per precinct of the resolution (may be empty)
};

struct Resolution {
    int trx0, try0, trx1, try1, ppx, ppy, npx = 0, npy = 0;
    std::vector<Band> bands;
};

struct TileComp {
    CodingStyle cod;
    QuantStyle q;
    std::vector<Resolution> res;
};

struct Tile {
    int tx0 = 0, ty0 = 0, tx1 = 0, ty1 = 0;
    bool have_cod = false, have_qcd = false;
    CodingStyle cod;
    QuantStyle qcd;
    std::vector<std::pair<int, CodingStyle>> coc;
    std::vector<std::pair<int, QuantStyle>> qcc;
    std::vector<uint8_t> data;
    std::vector<TileComp> comps;
    std::vector<Block> blocks;
    float steps[MAX_COMPONENT][3 * MAX_DWT_LEVEL + 1];
    size_t seg_bytes = 0;     // segment-length tables, then the block data
    size_t table_bytes = 0;   // of which the tables (4 bytes per segment)
};

}  // namespace

struct BmiCodestream {
    BmiCodestreamInfo info;
    int flags = 0;
    int Xsiz, Ysiz, XOsiz, YOsiz, XTsiz, YTsiz, XTOsiz, YTOsiz, Csiz;
    int depth[MAX_COMPONENT];
    CodingStyle cod;
    QuantStyle qcd;
    std::vector<std::pair<int, CodingStyle>> coc;
    std::vector<std::pair<int, QuantStyle>> qcc;
    std::vector<Tile> tiles;
};

namespace {

// codestream inside a JP2 file (dec_read_jp2_header, dec_syntax.cpp:46-128)
void locate_codestream(const uint8_t *d, size_t n, size_t *cs_off, size_t *cs_len, int *colour)
{
    *colour = 0;
    if (n >= 2 && d[0] == 0xFF && d[1] == 0x4F) {
        *cs_off = 0;
        *cs_len = n;
        return;
    }
    static const uint8_t sig[12] = {0, 0, 0, 12, 'j', 'P', ' ', ' ', 13, 10, 0x87, 10};
    if (n < 12 || memcmp(d, sig, 12) != 0) fail(BMI_E_INVALID, "neither a codestream nor a JP2 file");
    auto box = [&](size_t pos, size_t lim, uint64_t *len, int *hdr, uint32_t *type) {
        if (pos + 8 > lim) fail(BMI_E_INVALID, "truncated JP2 box");
        uint64_t l = ((uint64_t)d[pos] << 24) | (d[pos + 1] << 16) | (d[pos + 2] << 8) | d[pos + 3];
        *type = ((uint32_t)d[pos + 4] << 24) | (d[pos + 5] << 16) | (d[pos + 6] << 8) | d[pos + 7];
        *hdr = 8;
        if (l == 1) {
            if (pos + 16 > lim) fail(BMI_E_INVALID, "truncated JP2 box");
            l = 0;
            for (int i = 0; i < 8; ++i) l = (l << 8) | d[pos + 8 + i];
            *hdr = 16;
        } else if (l == 0) {
            l = lim - pos;
        }
        if (l < (uint64_t)*hdr || pos + l > lim) fail(BMI_E_INVALID, "bad JP2 box length");
        *len = l;
    };
    size_t pos = 0;
    while (pos < n) {
        uint64_t len;
        int hdr;
        uint32_t type;
        box(pos, n, &len, &hdr, &type);
        if (type == 0x6A703268) {          // jp2h
            size_t sub = pos + hdr, lim = pos + len;
            while (sub < lim) {
                uint64_t sl;
                int sh;
                uint32_t st;
                box(sub, lim, &sl, &sh, &st);
                if (st == 0x636F6C72 && sl >= (uint64_t)sh + 7 && d[sub + sh] == 1) {   // colr, enumerated
                    size_t p = sub + sh + 3;
                    int space = (d[p] << 24) | (d[p + 1] << 16) | (d[p + 2] << 8) | d[p + 3];
                    if (space != 16 && space != 17 && space != 18)
                        fail(BMI_E_UNSUPPORTED, "JP2 colour space not sRGB / grey / sYCC");
                    *colour = space;
                }
                sub += sl;
            }
        } else if (type == 0x6A703263) {   // jp2c
            *cs_off = pos + hdr;
            *cs_len = len - hdr;
            return;
        }
        pos += len;
    }
    fail(BMI_E_INVALID, "JP2 file without a codestream box");
}

void parse_marker_segments(BmiCodestream *S, const uint8_t *d, size_t n)
{
    Reader r{d, n, 0};
    if (r.u16() != MK_SOC) fail(BMI_E_INVALID, "no SOC marker");
    bool have_siz = false, have_cod = false, have_qcd = false;
    // main header
    for (;;) {
        unsigned mk = r.u16();
        if (mk == MK_SOT) {
            r.pos -= 2;
            break;
        }
        if (mk < 0xFF30) fail(BMI_E_INVALID, "marker expected in the main header");
        unsigned len = r.u16();
        if (len < 2) fail(BMI_E_INVALID, "bad marker segment length");
        size_t seg_end = r.pos + len - 2;
        if (seg_end > n) fail(BMI_E_INVALID, "truncated main header");
        switch (mk) {
        case MK_SIZ: {
            r.u16();
            // the eight 32-bit fields are unsigned (T.800 A.5.1); values
            // above INT_MAX are refused so every later sum fits in 64 bits
            // and every coordinate in an int
            uint32_t f[8];
            for (uint32_t &v : f) v = r.u32();
            for (uint32_t v : f)
                if (v > 0x7FFFFFFFu) fail(BMI_E_UNSUPPORTED, "SIZ value above 2^31 - 1");
            S->Xsiz = (int)f[0]; S->Ysiz = (int)f[1]; S->XOsiz = (int)f[2]; S->YOsiz = (int)f[3];
            S->XTsiz = (int)f[4]; S->YTsiz = (int)f[5]; S->XTOsiz = (int)f[6]; S->YTOsiz = (int)f[7];
            S->Csiz = r.u16();
            if (S->Csiz < 1 || S->Csiz > MAX_COMPONENT)
                fail(BMI_E_UNSUPPORTED, "1..4 components supported");
            // T.800 A.5.1: XOsiz < Xsiz, 0 < XTsiz, XTOsiz <= XOsiz,
            // XTsiz + XTOsiz > XOsiz (and the same vertically)
            if (S->Xsiz <= S->XOsiz || S->Ysiz <= S->YOsiz || S->XTsiz <= 0 || S->YTsiz <= 0 ||
                S->XTOsiz > S->XOsiz || S->YTOsiz > S->YOsiz ||
                (long long)S->XTOsiz + S->XTsiz <= S->XOsiz ||
                (long long)S->YTOsiz + S->YTsiz <= S->YOsiz)
                fail(BMI_E_INVALID, "inconsistent SIZ geometry");
            for (int c = 0; c < S->Csiz; ++c) {
                unsigned s = r.u8(), dx = r.u8(), dy = r.u8();
                if ((s & 0x80) || dx != 1 || dy != 1)
                    fail(BMI_E_UNSUPPORTED, "signed or sub-sampled component");
                S->depth[c] = (s & 0x7F) + 1;
                if (S->depth[c] > 16) fail(BMI_E_UNSUPPORTED, "bit depth over 16");
                if (S->depth[c] != S->depth[0]) fail(BMI_E_UNSUPPORTED, "components of different bit depths");
            }
            have_siz = true;
            break;
        }
        case MK_COD: {
            CodingStyle cs;
            cs.scod = r.u8();
            cs.prog = r.u8();
            cs.layers = r.u16();
            cs.mct = r.u8();
            read_spcod(r, cs.scod & 1, cs, seg_end);
            S->cod = cs;
            have_cod = true;
            break;
        }
        case MK_COC: {
            if (!have_siz || !have_cod) fail(BMI_E_INVALID, "COC before SIZ / COD");
            int c = S->Csiz < 257 ? r.u8() : r.u16();
            CodingStyle cs = S->cod;
            int scoc = r.u8();
            read_spcod(r, scoc & 1, cs, seg_end);
            if (c >= S->Csiz) fail(BMI_E_INVALID, "COC component index");
            S->coc.emplace_back(c, cs);
            break;
        }
        case MK_QCD:
            S->qcd = read_sqcd(r, seg_end);
            have_qcd = true;
            break;
        case MK_QCC: {
            if (!have_siz) fail(BMI_E_INVALID, "QCC before SIZ");
            int c = S->Csiz < 257 ? r.u8() : r.u16();
            if (c >= S->Csiz) fail(BMI_E_INVALID, "QCC component index");
            S->qcc.emplace_back(c, read_sqcd(r, seg_end));
            break;
        }
        case MK_RGN: case MK_POC: case MK_PPM:
            fail(BMI_E_UNSUPPORTED, "RGN / POC / PPM marker");
        case MK_TLM: case MK_PLM: case MK_CRG: case MK_COM:
            break;
        default:
            fail(BMI_E_INVALID, "unexpected marker in the main header");
        }
        r.pos = seg_end;
    }
    if (!have_siz || !have_cod || !have_qcd) fail(BMI_E_INVALID, "main header lacks SIZ / COD / QCD");
    int ntx = ceil_div((long long)S->Xsiz - S->XTOsiz, S->XTsiz);
    int nty = ceil_div((long long)S->Ysiz - S->YTOsiz, S->YTsiz);
    if ((long long)ntx * nty > 65535) fail(BMI_E_INVALID, "more than 65535 tiles");
    // a tile is at most XTsiz wide and lies inside the image; the back end
    // (like the reference's 16-bit tile sizes) takes tiles up to 32767
    if (std::min(S->XTsiz, S->Xsiz - S->XOsiz) > 32767 || std::min(S->YTsiz, S->Ysiz - S->YOsiz) > 32767)
        fail(BMI_E_UNSUPPORTED, "tiles larger than 32767 samples");
    S->tiles.resize((size_t)ntx * nty);
    S->info.ntiles_x = ntx;
    S->info.ntiles_y = nty;
    // tile-parts
    while (r.pos + 2 <= n) {
        size_t sot = r.pos;
        unsigned mk = r.u16();
        if (mk == MK_EOC) break;
        if (mk != MK_SOT) fail(BMI_E_INVALID, "SOT expected");
        if (r.u16() != 10) fail(BMI_E_INVALID, "bad SOT length");
        unsigned t = r.u16();
        uint32_t psot = r.u32();
        unsigned tpsot = r.u8();
        r.u8();
        if (t >= S->tiles.size()) fail(BMI_E_INVALID, "tile index out of range");
        size_t stop = psot ? sot + psot : n;
        if (!psot && n >= 2 && d[n - 2] == 0xFF && d[n - 1] == 0xD9) stop = n - 2;
        if (stop > n || stop < sot + 14) fail(BMI_E_INVALID, "tile-part length out of range");
        Tile &T = S->tiles[t];
        for (;;) {
            unsigned m = r.u16();
            if (m == MK_SOD) break;
            unsigned len = r.u16();
            size_t seg_end = r.pos + len - 2;
            if (len < 2 || seg_end > stop) fail(BMI_E_INVALID, "truncated tile-part header");
            bool first = (tpsot == 0);
            switch (m) {
            case MK_COD:
                if (first) {
                    CodingStyle cs;
                    cs.scod = r.u8();
                    cs.prog = r.u8();
                    cs.layers = r.u16();
                    cs.mct = r.u8();
                    read_spcod(r, cs.scod & 1, cs, seg_end);
                    T.cod = cs;
                    T.have_cod = true;
                }
                break;
            case MK_COC:
                if (first) {
                    int c = S->Csiz < 257 ? r.u8() : r.u16();
                    if (c >= S->Csiz) fail(BMI_E_INVALID, "COC component index");
                    CodingStyle cs;
                    int scoc = r.u8();
                    read_spcod(r, scoc & 1, cs, seg_end);
                    T.coc.emplace_back(c, cs);
                }
                break;
            case MK_QCD:
                if (first) {
                    T.qcd = read_sqcd(r, seg_end);
                    T.have_qcd = true;
                }
                break;
            case MK_QCC:
                if (first) {
                    int c = S->Csiz < 257 ? r.u8() : r.u16();
                    if (c >= S->Csiz) fail(BMI_E_INVALID, "QCC component index");
                    T.qcc.emplace_back(c, read_sqcd(r, seg_end));
                }
                break;
            case MK_RGN: case MK_POC: case MK_PPT:
                fail(BMI_E_UNSUPPORTED, "RGN / POC / PPT marker");
            case MK_COM: case MK_PLT:
                break;
            default:
                fail(BMI_E_INVALID, "unexpected marker in a tile-part header");
            }
            r.pos = seg_end;
        }
        if (r.pos > stop) fail(BMI_E_INVALID, "tile-part header overruns the tile-part");
        T.data.insert(T.data.end(), d + r.pos, d + stop);
        r.pos = stop;
    }
}

// numbps and mfAbsStepNorm of band `bid` (0 LL, 3k+1.. HL, LH, HH of level k
// from the coarsest) of a component with `levels` decompositions
void band_quant(const QuantStyle &q, int bid, int levels, int depth, int reversible, int flags,
                int *numbps, float *step)
{
    unsigned v;
    if (q.style == 1) {
        // scalar derived (T.800 E.1.1.1): eps_b = eps_0 - NL + n_b, mu_b = mu_0
        int nb = bid == 0 ? levels : levels - (bid - 1) / 3;
        int e = (int)(q.v[0] >> 11) - levels + nb;
        if (e < 0) fail(BMI_E_INVALID, "derived exponent below zero");
        v = ((unsigned)e << 11) | (q.v[0] & 0x7FF);
    } else {
        if ((size_t)bid >= q.v.size()) fail(BMI_E_INVALID, "QCD has fewer bands than the tile");
        v = q.v[bid];
    }
    if (reversible) {
        *numbps = (int)v + q.guard - 1;            // dec_jp2.cpp:177-178
        *step = 1.0f;
        if (*numbps > 31) fail(BMI_E_INVALID, "more than 31 magnitude bit-planes");
        return;
    }
    int exp = (int)(v >> 11), mant = (int)(v & 0x7FF);
    if (flags & BMI_CS_GUARD_T800)
        *numbps = ((exp + q.guard - 1) << 11) | mant;
    else
        *numbps = (int)v + q.guard - 1;            // dec_jp2.cpp:177-178
    // the tier-1 words hold bit-planes 0..30 below the sign bit
    if ((*numbps >> 11) > 31) fail(BMI_E_UNSUPPORTED, "more than 31 magnitude bit-planes");
    // mfAbsStepNorm (cudainterface.cpp:712-720)
    int s = (mant << 2) | 0x2000;
    int sh = depth - exp;
    long long abs_step = sh > 0 ? ((long long)s << std::min(sh, 40)) : (long long)(s >> std::min(-sh, 31));
    *step = (float)abs_step * 0.0001220703125f;
}

void build_tile(BmiCodestream *S, int t, Tile &T)
{
    int p = t % S->info.ntiles_x, q = t / S->info.ntiles_x;
    T.tx0 = (int)std::max((long long)S->XTOsiz + (long long)p * S->XTsiz, (long long)S->XOsiz);
    T.ty0 = (int)std::max((long long)S->YTOsiz + (long long)q * S->YTsiz, (long long)S->YOsiz);
    T.tx1 = std::min((long long)S->XTOsiz + (long long)(p + 1) * S->XTsiz, (long long)S->Xsiz);
    T.ty1 = std::min((long long)S->YTOsiz + (long long)(q + 1) * S->YTsiz, (long long)S->Ysiz);
    T.comps.resize(S->Csiz);
    for (int c = 0; c < S->Csiz; ++c) {
        TileComp &tc = T.comps[c];
        // precedence (T.800 A.6): tile COC > tile COD > main COC > main COD
        tc.cod = T.have_cod ? T.cod : S->cod;
        auto take_sp = [&](const CodingStyle &src) {
            tc.cod.levels = src.levels; tc.cod.xcb = src.xcb; tc.cod.ycb = src.ycb;
            tc.cod.cbstyle = src.cbstyle; tc.cod.transform = src.transform;
            memcpy(tc.cod.ppx, src.ppx, sizeof src.ppx);
            memcpy(tc.cod.ppy, src.ppy, sizeof src.ppy);
        };
        if (!T.have_cod)
            for (auto &e : S->coc) if (e.first == c) take_sp(e.second);
        for (auto &e : T.coc) if (e.first == c) take_sp(e.second);
        // tile QCC > tile QCD > main QCC > main QCD
        tc.q = S->qcd;
        for (auto &e : S->qcc) if (e.first == c) tc.q = e.second;
        if (T.have_qcd) tc.q = T.qcd;
        for (auto &e : T.qcc) if (e.first == c) tc.q = e.second;
        const CodingStyle &cs = tc.cod;
        if (cs.cbstyle & ~BMI_T1_STYLES)
            fail(BMI_E_UNSUPPORTED, "code-block style " + std::to_string(cs.cbstyle));
        if (cs.xcb > 6 || cs.ycb > 6) fail(BMI_E_UNSUPPORTED, "code-blocks over 64 samples a side");
        if (cs.transform > 1) fail(BMI_E_UNSUPPORTED, "wavelet transform other than 9/7 / 5/3");
        if (cs.levels > MAX_DWT_LEVEL) fail(BMI_E_UNSUPPORTED, "more than 6 decomposition levels");
        if (cs.prog > 4) fail(BMI_E_INVALID, "unknown progression order");
        if (cs.layers < 1) fail(BMI_E_INVALID, "zero layers");
        if (c && (cs.levels != T.comps[0].cod.levels || cs.transform != T.comps[0].cod.transform))
            fail(BMI_E_UNSUPPORTED, "components of one tile with different levels or filters");
        const int NL = cs.levels;
        tc.res.resize(NL + 1);
        for (int r = 0; r <= NL; ++r) {
            Resolution &R = tc.res[r];
            const int s = NL - r, sh = r ? 1 : 0;
            R.trx0 = ceil_div(T.tx0, 1ll << s); R.try0 = ceil_div(T.ty0, 1ll << s);
            R.trx1 = ceil_div(T.tx1, 1ll << s); R.try1 = ceil_div(T.ty1, 1ll << s);
            R.ppx = cs.ppx[r]; R.ppy = cs.ppy[r];
            if (R.trx1 > R.trx0 && R.try1 > R.try0) {
                R.npx = ceil_div(R.trx1, 1ll << R.ppx) - (R.trx0 >> R.ppx);
                R.npy = ceil_div(R.try1, 1ll << R.ppy) - (R.try0 >> R.ppy);
            }
            const int xcb = std::min(cs.xcb, R.ppx - sh), ycb = std::min(cs.ycb, R.ppy - sh);
            const int nbands = r ? 3 : 1;
            R.bands.resize(nbands);
            for (int bi = 0; bi < nbands; ++bi) {
                const int o = r ? bi + 1 : 0;             // orientation: 0 LL, 1 HL, 2 LH, 3 HH
                const int nb = r ? NL - r + 1 : NL;       // decompositions to this band
                const int xo = o & 1, yo = o >> 1;
                const long long half = (1ll << nb) >> 1;
                const int bx0 = ceil_div(T.tx0 - xo * half, 1ll << nb), by0 = ceil_div(T.ty0 - yo * half, 1ll << nb);
                const int bx1 = ceil_div(T.tx1 - xo * half, 1ll << nb), by1 = ceil_div(T.ty1 - yo * half, 1ll << nb);
                // Mallat position: HL / HH right of, LH / HH below this level's LL
                const int llw = ceil_div(T.tx1, 1ll << nb) - ceil_div(T.tx0, 1ll << nb);
                const int llh = ceil_div(T.ty1, 1ll << nb) - ceil_div(T.ty0, 1ll << nb);
                const int bid = r ? 3 * (r - 1) + o : 0;
                int numbps;
                float step;
                band_quant(tc.q, bid, NL, S->depth[c], cs.transform == 1, S->flags, &numbps, &step);
                T.steps[c][bid] = step;
                Band &B = R.bands[bi];
                B.prec.resize((size_t)R.npx * R.npy);
                if (bx1 <= bx0 || by1 <= by0) continue;
                // code-block grid of the band, anchored at multiples of the block size
                const int gx0 = bx0 >> xcb, gy0 = by0 >> ycb;
                const int gw = ceil_div(bx1, 1ll << xcb) - gx0, gh = ceil_div(by1, 1ll << ycb) - gy0;
                const int first = (int)T.blocks.size();
                for (int j = 0; j < gh; ++j)
                    for (int i = 0; i < gw; ++i) {
                        Block b;
                        const int x0 = std::max((gx0 + i) << xcb, bx0), y0 = std::max((gy0 + j) << ycb, by0);
                        const int x1 = std::min((gx0 + i + 1) << xcb, bx1), y1 = std::min((gy0 + j + 1) << ycb, by1);
                        b.comp = c; b.band = bid; b.orient = o; b.numbps = numbps;
                        b.style = tc.cod.cbstyle;
                        b.x = x0 - bx0 + llw * xo; b.y = y0 - by0 + llh * yo;
                        b.w = x1 - x0; b.h = y1 - y0;
                        T.blocks.push_back(std::move(b));
                    }
                // precincts: band-space cells of 2^(PP - sh) aligned on the resolution's grid
                for (int py = 0; py < R.npy; ++py)
                    for (int px = 0; px < R.npx; ++px) {
                        long long cx0 = ((long long)(R.trx0 >> R.ppx) + px) << R.ppx >> sh;
                        long long cy0 = ((long long)(R.try0 >> R.ppy) + py) << R.ppy >> sh;
                        long long cx1 = cx0 + ((1ll << R.ppx) >> sh), cy1 = cy0 + ((1ll << R.ppy) >> sh);
                        cx0 = std::max<long long>(cx0, bx0); cy0 = std::max<long long>(cy0, by0);
                        cx1 = std::min<long long>(cx1, bx1); cy1 = std::min<long long>(cy1, by1);
                        if (cx1 <= cx0 || cy1 <= cy0) continue;
                        Precinct &P = B.prec[(size_t)py * R.npx + px];
                        const int i0 = (int)(cx0 >> xcb) - gx0, j0 = (int)(cy0 >> ycb) - gy0;
                        P.cw = ceil_div(cx1, 1ll << xcb) - (int)(cx0 >> xcb);
                        P.ch = ceil_div(cy1, 1ll << ycb) - (int)(cy0 >> ycb);
                        for (int j = 0; j < P.ch; ++j)
                            for (int i = 0; i < P.cw; ++i) P.blocks.push_back(first + (j0 + j) * gw + i0 + i);
                        P.incl.init(P.cw, P.ch);
                        P.zbp.init(P.cw, P.ch);
                    }
            }
        }
    }
}

struct PacketId {
    int layer, comp, res, prec;
};

// packets in progression order (T.800 B.12)
std::vector<PacketId> packet_order(const Tile &T)
{
    const int L = T.comps[0].cod.layers, prog = T.comps[0].cod.prog;
    int maxr = 0;
    for (auto &tc : T.comps) maxr = std::max(maxr, (int)tc.res.size() - 1);
    std::vector<PacketId> out;
    const int nc = (int)T.comps.size();
    if (prog == 0 || prog == 1) {          // LRCP, RLCP: raster precincts
        for (int a = 0; a < (prog == 0 ? L : maxr + 1); ++a)
            for (int b = 0; b < (prog == 0 ? maxr + 1 : L); ++b) {
                const int l = prog == 0 ? a : b, r = prog == 0 ? b : a;
                for (int c = 0; c < nc; ++c) {
                    if (r >= (int)T.comps[c].res.size()) continue;
                    const Resolution &R = T.comps[c].res[r];
                    for (int k = 0; k < R.npx * R.npy; ++k) out.push_back({l, c, r, k});
                }
            }
        return out;
    }
    // position-driven: each precinct is reached at its anchor, the first
    // reference-grid point of the tile inside it
    struct Anchor {
        long long y, x;
        int c, r, k;
    };
    std::vector<Anchor> a;
    for (int c = 0; c < nc; ++c) {
        const TileComp &tc = T.comps[c];
        for (int r = 0; r < (int)tc.res.size(); ++r) {
            const Resolution &R = tc.res[r];
            const int s = (int)tc.res.size() - 1 - r;
            for (int j = 0; j < R.npy; ++j)
                for (int i = 0; i < R.npx; ++i) {
                    long long x = (((long long)(R.trx0 >> R.ppx) + i) << R.ppx) << s;
                    long long y = (((long long)(R.try0 >> R.ppy) + j) << R.ppy) << s;
                    a.push_back({std::max<long long>(y, T.ty0), std::max<long long>(x, T.tx0), c, r, j * R.npx + i});
                }
        }
    }
    auto key = [&](const Anchor &e) {
        switch (prog) {
        case 2: return std::make_tuple((long long)e.r, e.y, e.x, (long long)e.c);   // RPCL
        case 3: return std::make_tuple(e.y, e.x, (long long)e.c, (long long)e.r);   // PCRL
        default: return std::make_tuple((long long)e.c, e.y, e.x, (long long)e.r);  // CPRL
        }
    };
    std::stable_sort(a.begin(), a.end(), [&](const Anchor &u, const Anchor &v) { return key(u) < key(v); });
    for (auto &e : a)
        for (int l = 0; l < L; ++l) out.push_back({l, e.c, e.r, e.k});
    return out;
}

// one packet (T.800 B.9-B.10; the reference's dec_read_packet / dec_prec /
// dec_restore_coding_pass, dec_packet.cpp:66-380)
void read_packet(Tile &T, BitReader &br, const PacketId &id, bool eph_flag)
{
    const uint8_t *d = T.data.data();
    const size_t n = T.data.size();
    if (br.pos + 6 <= n && d[br.pos] == 0xFF && d[br.pos + 1] == 0x91) br.pos += 6;   // SOP
    Resolution &R = T.comps[id.comp].res[id.res];
    br.left = 0;
    br.last_ff = false;
    const bool nonempty = br.bit();
    if (nonempty) {
        for (Band &B : R.bands) {
            if (B.prec.empty()) continue;
            Precinct &P = B.prec[id.prec];
            for (size_t k = 0; k < P.blocks.size(); ++k) {
                Block &cb = T.blocks[P.blocks[k]];
                int v;
                bool inc = cb.seen ? br.bit() != 0 : tag_decode(P.incl, br, (int)k, id.layer + 1, &v);
                if (!inc) continue;
                if (!cb.seen) {
                    tag_decode(P.zbp, br, (int)k, 1 << 30, &v);
                    if (v > 74) fail(BMI_E_INVALID, "more than 74 missing bit-planes");
                    cb.missing = v;
                    cb.seen = true;
                }
                const int np = read_num_passes(br);
                const int p0 = cb.npass;
                cb.npass += np;
                if (cb.npass > 128) fail(BMI_E_INVALID, "more than 128 coding passes");
                while (br.bit()) {
                    if (++cb.lblock > 32) fail(BMI_E_INVALID, "Lblock over 32");
                }
                // one length per codeword segment the new passes reach
                // (T.800 B.10.7.2): Lblock + floor(log2(its passes here)) bits
                int seg = -1;
                for (int i = 0; i <= p0; ++i) seg += bmi_pass_new_segment(i, cb.style) ? 1 : 0;
                for (int i = p0; i < p0 + np;) {
                    if (i > p0 && bmi_pass_new_segment(i, cb.style)) ++seg;
                    int m = 1;
                    while (i + m < p0 + np && !bmi_pass_new_segment(i + m, cb.style)) ++m;
                    const int nbits = cb.lblock + floor_log2((unsigned)m);
                    if (nbits > 32) fail(BMI_E_INVALID, "segment length field over 32 bits");
                    const int64_t ln = (int64_t)br.bits(nbits);
                    if (ln < 0 || ln > 0x7FFFFFFF) fail(BMI_E_INVALID, "segment length");
                    cb.pending.push_back({0, (uint32_t)ln, seg});
                    i += m;
                }
            }
        }
    }
    br.align();
    if (eph_flag && br.pos + 2 <= n && d[br.pos] == 0xFF && d[br.pos + 1] == 0x92) br.pos += 2;
    if (!nonempty) return;
    for (Band &B : R.bands) {
        if (B.prec.empty()) continue;
        for (int bi : B.prec[id.prec].blocks) {
            Block &cb = T.blocks[bi];
            for (const Chunk &c : cb.pending) {
                if ((size_t)c.len > n - br.pos) fail(BMI_E_INVALID, "code-block data past the end of the tile");
                cb.chunks.push_back({(uint32_t)br.pos, c.len, c.seg});
                cb.bytes += c.len;
                br.pos += c.len;
            }
            cb.pending.clear();
        }
    }
}

void decode_tile(BmiCodestream *S, int t)
{
    Tile &T = S->tiles[t];
    build_tile(S, t, T);
    if (T.data.size() > 0xFFFFFFF0u) fail(BMI_E_UNSUPPORTED, "tile data over 4 GiB");
    BitReader br{T.data.data(), T.data.size(), 0};
    const bool eph = (T.comps[0].cod.scod & 4) != 0;
    for (const PacketId &id : packet_order(T)) {
        if (br.pos >= T.data.size()) break;   // truncated after a packet: the rest are absent
        read_packet(T, br, id, eph);
    }
    size_t data = 0, nsegs = 0;
    for (const Block &b : T.blocks) {
        data += b.bytes;
        if (bmi_style_segmented(b.style)) nsegs += (size_t)b.nseg();
    }
    T.table_bytes = (4 * nsegs + 15) & ~(size_t)15;
    T.seg_bytes = T.table_bytes + data;
}

}  // namespace

extern "C" int BMI_codestream_open(const void *data, size_t len, int flags, BmiCodestream **out,
                                   BmiCodestreamInfo *info)
{
    if (!data || !out || (flags & ~BMI_CS_GUARD_T800)) return bmi_fail(BMI_E_INVALID, "BMI_codestream_open arguments");
    *out = nullptr;
    BmiCodestream *S = new (std::nothrow) BmiCodestream();
    if (!S) return bmi_fail(BMI_E_NOMEM, "BMI_codestream_open");
    try {
        memset(&S->info, 0, sizeof S->info);
        S->flags = flags;
        const uint8_t *d = static_cast<const uint8_t *>(data);
        size_t off, n;
        int colour;
        locate_codestream(d, len, &off, &n, &colour);
        parse_marker_segments(S, d + off, n);
        for (int t = 0; t < (int)S->tiles.size(); ++t) decode_tile(S, t);
        BmiCodestreamInfo &I = S->info;
        I.width = S->Xsiz - S->XOsiz;
        I.height = S->Ysiz - S->YOsiz;
        I.x0 = S->XOsiz;
        I.y0 = S->YOsiz;
        I.tile_w = S->XTsiz;
        I.tile_h = S->YTsiz;
        I.tile_x0 = S->XTOsiz;
        I.tile_y0 = S->YTOsiz;
        I.ncomp = S->Csiz;
        I.bit_depth = S->depth[0];
        I.levels = S->cod.levels;
        I.reversible = S->cod.transform == 1;
        I.mct = S->cod.mct;
        I.layers = S->cod.layers;
        I.progression = S->cod.prog;
        I.cb_w = 1 << S->cod.xcb;
        I.cb_h = 1 << S->cod.ycb;
        I.guard_bits = S->qcd.guard;
        I.colour_space = colour;
        long long total = 0;
        for (auto &T : S->tiles) total += (long long)T.blocks.size();
        I.total_blocks = (int)std::min<long long>(total, 0x7FFFFFFF);
        for (auto &T : S->tiles) T.data.shrink_to_fit();
    } catch (const ParseError &e) {
        delete S;
        return bmi_fail(e.code, "codestream: %s", e.what.c_str());
    } catch (const std::bad_alloc &) {
        delete S;
        return bmi_fail(BMI_E_NOMEM, "codestream: out of host memory");
    }
    if (info) *info = S->info;
    *out = S;
    return BMI_OK;
}

extern "C" int BMI_codestream_tile_info(const BmiCodestream *S, int t, int *x0, int *y0, int *w, int *h,
                                        int *nblocks, size_t *seg_bytes)
{
    if (!S || t < 0 || t >= (int)S->tiles.size()) return bmi_fail(BMI_E_INVALID, "BMI_codestream_tile_info");
    const Tile &T = S->tiles[t];
    if (x0) *x0 = T.tx0;
    if (y0) *y0 = T.ty0;
    if (w) *w = T.tx1 - T.tx0;
    if (h) *h = T.ty1 - T.ty0;
    if (nblocks) *nblocks = (int)T.blocks.size();
    if (seg_bytes) *seg_bytes = T.seg_bytes;
    return BMI_OK;
}

extern "C" int BMI_codestream_tile_blocks(const BmiCodestream *S, int t, void *const *planes, int plane_pitch,
                                          unsigned long long seg_base, BmiT1Block *blocks, void *segs,
                                          size_t seg_capacity, BmiTileDesc *desc)
{
    if (!S || t < 0 || t >= (int)S->tiles.size()) return bmi_fail(BMI_E_INVALID, "BMI_codestream_tile_blocks");
    const Tile &T = S->tiles[t];
    const int w = T.tx1 - T.tx0, h = T.ty1 - T.ty0;
    if (plane_pitch < w) return bmi_fail(BMI_E_INVALID, "plane pitch %d below tile width %d", plane_pitch, w);
    if (seg_base & 15) return bmi_fail(BMI_E_INVALID, "seg_base %llu not a multiple of 16", seg_base);
    if (blocks && (!planes || (segs == nullptr && T.seg_bytes)))
        return bmi_fail(BMI_E_INVALID, "BMI_codestream_tile_blocks: planes / segs missing");
    if (blocks) {
        if (seg_capacity < T.seg_bytes + 16) return bmi_fail(BMI_E_INVALID, "segment buffer below seg_bytes + 16");
        for (int c = 0; c < S->Csiz; ++c)
            if (!planes[c]) return bmi_fail(BMI_E_INVALID, "plane %d missing", c);
        // every block's rectangle lies in the tile-sized planes: the geometry
        // guarantees it, and the check keeps a wrong table from ever becoming
        // an out-of-bounds store of the decoder
        for (const Block &b : T.blocks)
            if (b.x < 0 || b.y < 0 || b.w < 1 || b.h < 1 || b.w > 64 || b.h > 64 ||
                (long long)b.x + b.w > w || (long long)b.y + b.h > h || b.comp < 0 || b.comp >= S->Csiz)
                return bmi_fail(BMI_E_INVALID, "code-block outside its tile");
        // layout: the segment-length tables of segmented blocks, then
        // every block's data (its segments in order), then 16 bytes of 0xFF
        uint8_t *dst = static_cast<uint8_t *>(segs);
        int32_t *table = reinterpret_cast<int32_t *>(dst);
        size_t off = T.table_bytes, ntab = 0;
        memset(dst, 0, T.table_bytes);
        const int reversible = T.comps[0].cod.transform == 1;
        for (size_t i = 0; i < T.blocks.size(); ++i) {
            const Block &b = T.blocks[i];
            BmiT1Block &o = blocks[i];
            o.data_offset = seg_base + off;
            o.len = (int)b.bytes;
            o.style = b.style;
            o.nseg = b.nseg();
            o.seg_offset = 0;
            if (bmi_style_segmented(b.style)) {
                o.seg_offset = seg_base + 4 * ntab;
                for (const Chunk &ch : b.chunks) table[ntab + ch.seg] += (int32_t)ch.len;
                ntab += (size_t)o.nseg;
            }
            for (const Chunk &ch : b.chunks) {
                memcpy(dst + off, T.data.data() + ch.off, ch.len);
                off += ch.len;
            }
            o.width = b.w;
            o.height = b.h;
            o.orient = b.orient;
            o.numbps = b.numbps;
            o.missing = b.missing;
            o.num_pass = b.npass;
            o.kind = reversible ? BMI_T1_INT : BMI_T1_WORDS;
            o.out = static_cast<char *>(planes[b.comp]) + 4 * ((size_t)b.y * plane_pitch + b.x);
            o.out_pitch = plane_pitch;
        }
        memset(dst + off, 0xFF, 16);
    }
    if (desc) {
        memset(desc, 0, sizeof *desc);
        for (int c = 0; c < S->Csiz; ++c) desc->coef[c] = planes ? planes[c] : nullptr;
        desc->coef_pitch = plane_pitch;
        desc->width = w;
        desc->height = h;
        desc->x0 = T.tx0;
        desc->y0 = T.ty0;
        desc->levels = T.comps[0].cod.levels;
        desc->top_level = -1;
        desc->ncomp = S->Csiz;
        desc->method = T.comps[0].cod.transform == 1 ? Ckernels_W5X3 : Ckernels_W9X7;
        desc->coef_kind = BMI_COEF_SIGNMAG;
        desc->bit_depth = S->depth[0];
        desc->out_format = IMAGE_FOTMAT_NONE;
        desc->check_out = -1;
        desc->mct = (T.comps[0].cod.mct && S->Csiz >= 3) ? 1 : 0;
        desc->r53_mode = BMI_R53_REFERENCE;
        for (int c = 0; c < S->Csiz; ++c)
            for (int b = 0; b < 3 * desc->levels + 1; ++b) desc->steps[c][b] = T.steps[c][b];
    }
    return BMI_OK;
}

extern "C" int BMI_codestream_close(BmiCodestream *S)
{
    delete S;
    return BMI_OK;
}
