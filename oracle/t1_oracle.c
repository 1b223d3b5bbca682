/*
 * t1_oracle.c -- CPU restatement of the reference's EBCOT tier-1 block
 * decoder.  TEST INFRASTRUCTURE ONLY (the checker of the GPU T1 kernel;
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * it).
 *
 * Follows decode_code_block (decoder/codec/src/bmi_block_decoder.cpp:589-853)
 * and its three passes (:10-215 significance propagation, :217-349
 * magnitude refinement, :351-586 cleanup), the MQ decoder
 * (decoder/codec/src/bmi_mq_decoder.cpp:186-355, bmi_mq_decoder.h:86-169)
 * and the stripe copy of dec_cb_stripe (decoder/codec/src/dec_t1_comp.cpp:
 * 283-426).  The reference keeps its state in packed "context words"
 * (bmi_t1_macro.h) and skips runs of empty columns; this restatement keeps
 * one flag byte per sample and visits every column, which decodes the same
 * symbols: a skipped column is one where no sample can be coded in that
 * pass.  Behaviour kept from the reference:
 *   - the passes start with a cleanup pass at bit-plane
 *     (numbps >> 11) - 1 - missing, or 30 - missing when that is negative,
 *     and plane 0 is replaced by plane 1 (:647-649);
 *   - one codeword segment for all passes, the data followed by two 0xFF
 *     bytes (bmi_mq_decoder.cpp:195);
 *   - a sample found significant at plane p becomes sign<<31 | 3<<(p-1)
 *     (1.5 * 2^p, :90), a refinement bit s at plane p clears bit p when s
 *     is 0 and sets bit p-1 (:272-273);
 *   - the refinement context is "first refinement" while the magnitude
 *     above plane p is below 4 (:260-270);
 *   - run mode only in columns of a complete four-row stripe (the partial
 *     last stripe carries out-of-bounds markers, :636-650);
 *   - the context states start at index 0 / MPS 0, except the all-zero
 *     zero-coding context at index 4 and the run context at index 3
 *     (bmi_mq_decoder.cpp:60-72); the two run-position bits use a fixed
 *     uniform state (MQ_DECODE_START, bmi_mq_decoder.h:131-168).
 * The MQ procedure is written as ITU-T T.800 C.3 (INITDEC, BYTEIN, DECODE,
 * RENORMD); the reference's scaled-register form computes the same symbols.
 *
 * or_t1_decode_block_ex adds the code-block style switches of T.800 Table
 * A.19, which the reference asserts on (bmi_block_decoder.cpp:792-797,
 * 818-822; its segment loop :729-760 already restarts the MQ decoder per
 * codeword segment): raw significance / refinement passes from pass 10 on
 * with bypass (D.6), contexts reset at every pass boundary, a codeword
 * segment per pass with termall (else, with bypass, segments after pass 9,
 * per raw pair and per cleanup pass), vertically causal contexts (D.7) and
 * the segmentation symbol after every cleanup pass (D.5).  These follow
 * the standard; they are pinned by OpenJPEG decoding the project encoder's
 * streams to their source images.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "j2k_oracle.h"

static const uint16_t kQe[47] = {
    0x5601, 0x3401, 0x1801, 0x0AC1, 0x0521, 0x0221, 0x5601, 0x5401, 0x4801, 0x3801, 0x3001, 0x2401,
    0x1C01, 0x1601, 0x5601, 0x5401, 0x5101, 0x4801, 0x3801, 0x3401, 0x3001, 0x2801, 0x2401, 0x2201,
    0x1C01, 0x1801, 0x1601, 0x1401, 0x1201, 0x1101, 0x0AC1, 0x09C1, 0x08A1, 0x0521, 0x0441, 0x02A1,
    0x0221, 0x0141, 0x0111, 0x0085, 0x0049, 0x0025, 0x0015, 0x0009, 0x0005, 0x0001, 0x5601};
/* bmi_mq_decoder.cpp:37-42 */
static const uint8_t kNmps[47] = {1,  2,  3,  4,  5,  38, 7,  8,  9,  10, 11, 12, 13, 29, 15, 16,
                                  17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29, 30, 31, 32,
                                  33, 34, 35, 36, 37, 38, 39, 40, 41, 42, 43, 44, 45, 45, 46};
static const uint8_t kNlps[47] = {1,  6,  9,  12, 29, 33, 6,  14, 14, 14, 17, 18, 20, 21, 14, 14,
                                  15, 16, 17, 18, 19, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29,
                                  30, 31, 32, 33, 34, 35, 36, 37, 38, 39, 40, 41, 42, 43, 46};

enum { CX_SC = 9, CX_MR = 14, CX_RL = 17, CX_UNI = 18, NCX = 19 };

typedef struct {
    const uint8_t *d;
    long len, bp;
    uint32_t A, C;
    int CT;
    uint8_t idx[NCX], mps[NCX];
} MqDec;

/* byte at position i of the segment; past its end the two appended 0xFF
   bytes, then 0xFF (a marker, never consumed) */
static int seg_byte(const MqDec *m, long i) { return i < m->len ? m->d[i] : 0xFF; }

static void bytein(MqDec *m)
{
    if (seg_byte(m, m->bp) == 0xFF) {
        if (seg_byte(m, m->bp + 1) > 0x8F) {
            m->C += 0xFF00;
            m->CT = 8;
        } else {
            m->bp++;
            m->C += (uint32_t)seg_byte(m, m->bp) << 9;
            m->CT = 7;
        }
    } else {
        m->bp++;
        m->C += (uint32_t)seg_byte(m, m->bp) << 8;
        m->CT = 8;
    }
}

/* INITDEC (C.3.5) on a codeword segment; context states are kept */
static void mq_start(MqDec *m, const uint8_t *d, long len)
{
    m->d = d;
    m->len = len;
    m->bp = 0;
    m->C = (uint32_t)seg_byte(m, 0) << 16;
    bytein(m);
    m->C <<= 7;
    m->CT -= 7;
    m->A = 0x8000;
}

/* initial context states (bmi_mq_decoder.cpp:60-72) */
static void mq_reset(MqDec *m)
{
    for (int i = 0; i < NCX; ++i) m->idx[i] = 0, m->mps[i] = 0;
    m->idx[0] = 4;
    m->idx[CX_RL] = 3;
    m->idx[CX_UNI] = 46;
}

static void renormd(MqDec *m)
{
    do {
        if (m->CT == 0) bytein(m);
        m->A <<= 1;
        m->C <<= 1;
        m->CT--;
    } while (!(m->A & 0x8000));
}

static int mq_decode(MqDec *m, int cx)
{
    const int i = m->idx[cx];
    const uint32_t qe = kQe[i];
    int d;
    m->A -= qe;
    if ((m->C >> 16) < qe) {
        /* LPS exchange */
        if (m->A < qe) {
            d = m->mps[cx];
            m->idx[cx] = kNmps[i];
        } else {
            d = 1 - m->mps[cx];
            if (kQe[i] == 0x5601 && kNlps[i] != 46) m->mps[cx] ^= 1;
            m->idx[cx] = kNlps[i];
        }
        m->A = qe;
        renormd(m);
    } else {
        m->C -= qe << 16;
        if (!(m->A & 0x8000)) {
            /* MPS exchange */
            if (m->A < qe) {
                d = 1 - m->mps[cx];
                if (kQe[i] == 0x5601 && kNlps[i] != 46) m->mps[cx] ^= 1;
                m->idx[cx] = kNlps[i];
            } else {
                d = m->mps[cx];
                m->idx[cx] = kNmps[i];
            }
            renormd(m);
        } else {
            d = m->mps[cx];
        }
    }
    return d;
}

/* ---- contexts ------------------------------------------------------------ */

enum { F_SIG = 1, F_NEG = 2, F_VISIT = 4, F_MU = 8 };

/* zero coding (dec_t1.h lut_ctx_lllh / lut_ctx_hl / lut_ctx_hh chosen per
   band, dec_t1_comp.cpp:510-512) */
static int zc_ctx(const uint8_t *f, int s, int orient)
{
    const int h = (f[-1] & F_SIG) + (f[1] & F_SIG);
    const int v = (f[-s] & F_SIG) + (f[s] & F_SIG);
    const int d = (f[-s - 1] & F_SIG) + (f[-s + 1] & F_SIG) + (f[s - 1] & F_SIG) + (f[s + 1] & F_SIG);
    static const uint8_t hh[4][5] = {{0, 1, 2, 2, 2}, {3, 4, 5, 5, 5}, {6, 7, 7, 7, 7}, {8, 8, 8, 8, 8}};
    if (orient == 3) return hh[d > 3 ? 3 : d][h + v];
    const int a = (orient == 1) ? v : h; /* HL: vertical sum leads */
    const int b = (orient == 1) ? h : v;
    static const uint8_t lead1[2][2] = {{5, 6}, {7, 7}}; /* [b>=1][d>=1] */
    if (a == 2) return 8;
    if (a == 1) return lead1[b >= 1][d >= 1];
    if (b) return b == 2 ? 4 : 3;
    return d >= 2 ? 2 : d;
}

/* sign coding: context 9..13 and XOR bit (dec_t1.h lut_ctx_sign) */
static int sc_ctx(const uint8_t *f, int s, int *xr)
{
    static const int8_t c3[4] = {0, 1, 0, -1}; /* flags & 3 -> contribution */
    int h = c3[f[-1] & 3] + c3[f[1] & 3];
    int v = c3[f[-s] & 3] + c3[f[s] & 3];
    h = h > 0 ? 1 : (h < 0 ? -1 : 0);
    v = v > 0 ? 1 : (v < 0 ? -1 : 0);
    /* rows: H = -1, 0, 1; columns: V = -1, 0, 1 -> (context, xor) */
    static const uint8_t cx[3][3] = {{13, 12, 11}, {10, 9, 10}, {11, 12, 13}};
    static const uint8_t xb[3][3] = {{1, 1, 1}, {1, 0, 0}, {0, 0, 0}};
    *xr = xb[h + 1][v + 1];
    return cx[h + 1][v + 1];
}

static int nb_sig(const uint8_t *f, int s)
{
    return ((f[-s - 1] | f[-s] | f[-s + 1] | f[-1] | f[1] | f[s - 1] | f[s] | f[s + 1]) & F_SIG) != 0;
}

/* ---- raw (bypass) segments --------------------------------------------------
 * T.800 D.6: raw bits most significant first, a byte after 0xFF holds seven
 * bits (its top bit is stuffed).  Past the end of the segment 0xFF bytes are
 * read; 0xFF followed by a byte above 0x8F is not consumed (a marker, or
 * the end), so the reader then returns 1 bits. */
typedef struct {
    const uint8_t *d;
    long len, bp;
    uint32_t c;
    int ct;
} RawDec;

static void raw_start(RawDec *r, const uint8_t *d, long len)
{
    r->d = d;
    r->len = len;
    r->bp = 0;
    r->c = 0;
    r->ct = 0;
}

static int raw_bit(RawDec *r)
{
    if (r->ct == 0) {
        const uint32_t next = r->bp < r->len ? r->d[r->bp] : 0xFF;
        if (r->c == 0xFF) {
            if (next > 0x8F) {
                r->c = 0xFF;
                r->ct = 8;
            } else {
                r->c = next;
                r->bp++;
                r->ct = 7;
            }
        } else {
            r->c = next;
            r->bp++;
            r->ct = 8;
        }
    }
    r->ct--;
    return (int)((r->c >> r->ct) & 1);
}

/* ---- code-block style (T.800 Table A.19) ------------------------------------ */

enum { CB_BYPASS = 0x01, CB_RESET = 0x02, CB_TERMALL = 0x04, CB_VSC = 0x08, CB_SEGSYM = 0x20 };

/* pass i (0 = first cleanup; type (i + 2) % 3) is raw: bypass, from the
   fifth coded bit-plane on, significance and refinement passes (D.6) */
static int pass_raw(int i, int style) { return (style & CB_BYPASS) && i >= 10 && i % 3 != 0; }

/* pass i opens a codeword segment: the first pass; every pass with
   termination on each pass; with bypass, each raw significance pass and
   each cleanup pass from pass 10 on (a raw segment holds the significance
   and refinement passes of one bit-plane) */
static int pass_new_segment(int i, int style)
{
    if (i == 0 || (style & CB_TERMALL)) return 1;
    return (style & CB_BYPASS) && i >= 10 && i % 3 != 2;
}

int or_t1_num_segments(int num_pass, int style)
{
    int n = 0;
    for (int i = 0; i < num_pass; ++i) n += pass_new_segment(i, style);
    return n;
}

/* 3x3 neighbourhood of the flag grid at f as contexts see it: with
   vertically causal contexts (D.7) the row below the last row of a stripe
   counts as insignificant */
static const uint8_t *view(const uint8_t *f, int s, int causal, uint8_t g[9])
{
    for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) g[(dy + 1) * 3 + dx + 1] = (causal && dy > 0) ? 0 : f[dy * s + dx];
    return g + 4;
}

/* ---- block decode ---------------------------------------------------------- */

int or_t1_decode_block_ex(const uint8_t *data, long len, int w, int h, int orient, int numbps,
                          int missing, int num_pass, int style, const int32_t *seg_len, int nseg,
                          int32_t *out, int out_pitch)
{
    if (w < 1 || h < 1 || w > 1024 || h > 1024 || orient < 0 || orient > 3 || !out || (style & ~0x3F))
        return -1;
    /* codeword segments: one of `len` bytes unless passes are terminated */
    const int multi = (style & (CB_BYPASS | CB_TERMALL)) != 0;
    if (multi) {
        long sum = 0;
        if (num_pass < 0 || nseg != or_t1_num_segments(num_pass, style) || (nseg && !seg_len)) return -1;
        for (int k = 0; k < nseg; ++k) {
            if (seg_len[k] < 0) return -1;
            sum += seg_len[k];
        }
        if (sum != len) return -1;
    }
    for (int y = 0; y < h; ++y) memset(out + (long)y * out_pitch, 0, sizeof(int32_t) * w);
    /* :610-624 */
    const int max_original_passes = (numbps - missing) * 3 - 2;
    if (max_original_passes <= 0) return 0;
    const int p_max = 30 - missing;
    int num_passes = 3 * p_max - 2;
    if (num_passes > num_pass) num_passes = num_pass;
    if (num_passes <= 0) return 0;
    /* :645-649 */
    int plane = (numbps >> 11) - 1 - missing;
    if (plane < 0) plane = p_max;
    if (plane == 0) plane = 1;

    const int s = w + 2;
    const int vsc = (style & CB_VSC) != 0;
    uint8_t *fl = (uint8_t *)calloc((size_t)s * (h + 2), 1);
    uint32_t *val = (uint32_t *)calloc((size_t)w * h, sizeof(uint32_t));
    if (!fl || !val) {
        free(fl);
        free(val);
        return -1;
    }
    MqDec m;
    RawDec r;
    mq_reset(&m);
    long seg_pos = 0;
    int seg = 0, raw = 0;
    int type = 2;
    for (int pi = 0; pi < num_passes; ++pi, ++type) {
        if (type == 3) {
            type = 0;
            --plane;
        }
        /* a pass below bit-plane 0 (more passes than the band has planes)
           would shift by a negative count, undefined in the reference:
           such passes are not decoded */
        if (plane < 0) break;
        if (pass_new_segment(pi, style)) {
            const long n = multi ? seg_len[seg] : len;
            raw = pass_raw(pi, style);
            if (raw)
                raw_start(&r, data + seg_pos, n);
            else
                mq_start(&m, data + seg_pos, n);
            seg_pos += n;
            ++seg;
        }
        if ((style & CB_RESET) && pi > 0) mq_reset(&m);
        const uint32_t one_half = (plane >= 0 && plane < 31)
                                      ? ((uint32_t)1 << plane) + (((uint32_t)1 << plane) >> 1)
                                      : 0;
        const uint32_t half = (plane >= 1 && plane < 32) ? ((uint32_t)1 << plane) >> 1 : 0;
        for (int y0 = 0; y0 < h; y0 += 4) {
            const int rows = (h - y0) < 4 ? (h - y0) : 4;
            for (int x = 0; x < w; ++x) {
                int y = y0;
                if (type == 2 && rows == 4) {
                    int run = 1;
                    for (int k = 0; k < 4 && run; ++k) {
                        uint8_t g[9];
                        const uint8_t *f = &fl[(y0 + k + 1) * s + x + 1];
                        if ((*f & (F_SIG | F_VISIT)) || nb_sig(view(f, s, vsc && k == 3, g), 3)) run = 0;
                    }
                    if (run) {
                        if (!mq_decode(&m, CX_RL)) continue;
                        int k = mq_decode(&m, CX_UNI) << 1;
                        k |= mq_decode(&m, CX_UNI);
                        uint8_t *f = &fl[(y0 + k + 1) * s + x + 1];
                        uint8_t g[9];
                        int xr;
                        const int cx = sc_ctx(view(f, s, vsc && k == 3, g), 3, &xr);
                        const int neg = mq_decode(&m, cx) ^ xr;
                        *f |= F_SIG | (neg ? F_NEG : 0);
                        val[(long)(y0 + k) * w + x] = ((uint32_t)neg << 31) + one_half;
                        y = y0 + k + 1;
                    }
                }
                for (; y < y0 + rows; ++y) {
                    uint8_t *f = &fl[(y + 1) * s + x + 1];
                    uint8_t g[9];
                    const uint8_t *n = view(f, s, vsc && (y & 3) == 3, g);
                    uint32_t *v = &val[(long)y * w + x];
                    if (type == 1) {
                        if (!(*f & F_MU)) continue;
                        const uint32_t mag = (*v & 0x7FFFFFFFu) >> plane;
                        const int bit = raw ? raw_bit(&r) : mq_decode(&m, mag < 4 ? CX_MR + nb_sig(n, 3) : CX_MR + 2);
                        *v ^= (uint32_t)(1 - bit) << plane;
                        *v |= half;
                        continue;
                    }
                    if (type == 0) {
                        if ((*f & F_SIG) || !nb_sig(n, 3)) continue;
                        *f |= F_VISIT;
                    } else if (*f & (F_SIG | F_VISIT)) {
                        continue;
                    }
                    int neg;
                    if (raw) {
                        /* raw significance bit, then the sign bit itself */
                        if (!raw_bit(&r)) continue;
                        neg = raw_bit(&r);
                    } else {
                        if (!mq_decode(&m, zc_ctx(n, 3, orient))) continue;
                        int xr;
                        const int cx = sc_ctx(n, 3, &xr);
                        neg = mq_decode(&m, cx) ^ xr;
                    }
                    *f |= F_SIG | (neg ? F_NEG : 0);
                    *v = ((uint32_t)neg << 31) + one_half;
                }
            }
        }
        if (type == 2) {
            /* segmentation symbol (D.5): four decisions in the uniform
               context after every cleanup pass (1010 when intact) */
            if (style & CB_SEGSYM)
                for (int k = 0; k < 4; ++k) (void)mq_decode(&m, CX_UNI);
            /* end of cleanup: significant samples become refinable */
            for (long i = 0; i < (long)s * (h + 2); ++i)
                fl[i] = (uint8_t)((fl[i] & ~F_VISIT) | ((fl[i] & F_SIG) ? F_MU : 0));
        }
    }
    for (int y = 0; y < h; ++y)
        memcpy(out + (long)y * out_pitch, val + (long)y * w, sizeof(int32_t) * w);
    free(fl);
    free(val);
    return 0;
}

/* code-block style 0: the reference's decode_code_block */
int or_t1_decode_block(const uint8_t *data, long len, int w, int h, int orient, int numbps,
                       int missing, int num_pass, int32_t *out, int out_pitch)
{
    return or_t1_decode_block_ex(data, len, w, h, orient, numbps, missing, num_pass, 0, NULL, 0, out,
                                 out_pitch);
}

/* dec_t1_comp.cpp:343-367: T1 words of a 5/3 band to two's complement */
void or_t1_words_to_int(const int32_t *words, int w, int h, int pitch, int numbps, int32_t *out,
                        int out_pitch)
{
    const int sh = 31 - numbps;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const int32_t t = words[(long)y * pitch + x];
            out[(long)y * out_pitch + x] =
                t < 0 ? -(int32_t)(((uint32_t)t & 0x7FFFFFFFu) >> sh) : (t >> sh);
        }
}

void or_t1_tables(uint16_t *qe, uint8_t *nmps, uint8_t *nlps, uint8_t *sw)
{
    memcpy(qe, kQe, sizeof(kQe));
    memcpy(nmps, kNmps, sizeof(kNmps));
    memcpy(nlps, kNlps, sizeof(kNlps));
    for (int i = 0; i < 47; ++i) sw[i] = (kQe[i] == 0x5601 && kNlps[i] != 46) ? 1 : 0;
}

int or_t1_zc(const uint8_t *nb9, int orient)
{
    uint8_t g[9];
    for (int i = 0; i < 9; ++i) g[i] = (nb9[i] & 1 ? F_SIG : 0) | (nb9[i] & 2 ? F_NEG : 0);
    return zc_ctx(&g[4], 3, orient);
}

int or_t1_sc(const uint8_t *nb9)
{
    uint8_t g[9];
    for (int i = 0; i < 9; ++i) g[i] = (nb9[i] & 1 ? F_SIG : 0) | (nb9[i] & 2 ? F_NEG : 0);
    int xr;
    const int cx = sc_ctx(&g[4], 3, &xr);
    return cx * 2 + xr;
}
