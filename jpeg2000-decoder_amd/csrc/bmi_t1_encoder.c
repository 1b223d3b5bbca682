/*
 * bmi_t1_encoder.c -- the project's own EBCOT tier-1 encoder, used to make
 * the code streams that the GPU block decoder is tested and measured on.
 *
 * Written from ITU-T T.800 (JPEG 2000 part 1):
 *   Annex C.2  MQ arithmetic encoder (INITENC, CODEMPS/CODELPS, RENORME,
 *              BYTEOUT with bit stuffing, FLUSH with SETBITS);
 *   Annex D    coding passes per bit-plane (significance propagation,
 *              magnitude refinement, cleanup with run mode), stripes of four
 *              rows, context tables D.1 (zero coding), D.3 (sign coding),
 *              D.4 (magnitude refinement).
 * bmi_t1_encode_block writes what the reference block decoder supports
 * (decoder/codec/src/bmi_block_decoder.cpp:589-853): one codeword segment
 * per code-block, code-block style 0.  bmi_t1_encode_block_ex adds the
 * code-block style switches of T.800 Table A.19 that the reference asserts
 * on (:792-797, 818-822): selective arithmetic coding bypass (raw
 * significance and refinement passes from the 11th pass on, D.6), context
 * reset on every pass boundary, termination of every pass, vertically
 * causal context formation (D.7) and segmentation symbols (D.5); it
 * returns the byte length of every codeword segment.
 *
 * This file produces inputs; it decodes nothing.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* T.800 Table C.2: Qe, next index after MPS / LPS, MPS switch */
static const uint16_t kQe[47] = {
    0x5601, 0x3401, 0x1801, 0x0AC1, 0x0521, 0x0221, 0x5601, 0x5401, 0x4801, 0x3801, 0x3001, 0x2401,
    0x1C01, 0x1601, 0x5601, 0x5401, 0x5101, 0x4801, 0x3801, 0x3401, 0x3001, 0x2801, 0x2401, 0x2201,
    0x1C01, 0x1801, 0x1601, 0x1401, 0x1201, 0x1101, 0x0AC1, 0x09C1, 0x08A1, 0x0521, 0x0441, 0x02A1,
    0x0221, 0x0141, 0x0111, 0x0085, 0x0049, 0x0025, 0x0015, 0x0009, 0x0005, 0x0001, 0x5601};
static const uint8_t kNmps[47] = {1,  2,  3,  4,  5,  38, 7,  8,  9,  10, 11, 12, 13, 29, 15, 16,
                                  17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29, 30, 31, 32,
                                  33, 34, 35, 36, 37, 38, 39, 40, 41, 42, 43, 44, 45, 45, 46};
static const uint8_t kNlps[47] = {1,  6,  9,  12, 29, 33, 6,  14, 14, 14, 17, 18, 20, 21, 14, 14,
                                  15, 16, 17, 18, 19, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29,
                                  30, 31, 32, 33, 34, 35, 36, 37, 38, 39, 40, 41, 42, 43, 46};
static const uint8_t kSwitch[47] = {1, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 1, 0,
                                    0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,
                                    0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};

/* contexts: 0..8 zero coding, 9..13 sign, 14..16 refinement, 17 run, 18 uniform */
enum { CX_SC = 9, CX_MR = 14, CX_RL = 17, CX_UNI = 18, NCX = 19 };

typedef struct {
    uint32_t A, C;
    int CT;
    uint8_t *buf; /* buf[-1] is a scratch byte (the byte before the stream) */
    long bp;      /* index of the current byte B, -1 initially */
    long cap;
    int overflow;
    uint8_t idx[NCX], mps[NCX];
} MqEnc;

static void mq_reset_contexts(MqEnc *e)
{
    for (int i = 0; i < NCX; ++i) e->idx[i] = 0, e->mps[i] = 0;
    e->idx[0] = 4;       /* all-zero neighbourhood zero-coding context */
    e->idx[CX_RL] = 3;   /* run-length context */
    e->idx[CX_UNI] = 46; /* uniform context */
}

static uint8_t *mq_byte(MqEnc *e) { return e->bp < 0 ? e->buf - 1 : e->buf + e->bp; }

static void mq_byteout(MqEnc *e)
{
    if (e->bp + 1 >= e->cap) {
        e->overflow = 1;
        e->bp = e->cap - 2;
    }
    if (*mq_byte(e) == 0xFF) {
        e->bp++;
        *mq_byte(e) = (uint8_t)(e->C >> 20);
        e->C &= 0xFFFFF;
        e->CT = 7;
    } else if (e->C < 0x8000000) {
        e->bp++;
        *mq_byte(e) = (uint8_t)(e->C >> 19);
        e->C &= 0x7FFFF;
        e->CT = 8;
    } else {
        (*mq_byte(e))++;
        if (*mq_byte(e) == 0xFF) {
            e->C &= 0x7FFFFFF;
            e->bp++;
            *mq_byte(e) = (uint8_t)(e->C >> 20);
            e->C &= 0xFFFFF;
            e->CT = 7;
        } else {
            e->bp++;
            *mq_byte(e) = (uint8_t)(e->C >> 19);
            e->C &= 0x7FFFF;
            e->CT = 8;
        }
    }
}

static void mq_renorme(MqEnc *e)
{
    do {
        e->A <<= 1;
        e->C <<= 1;
        if (--e->CT == 0) mq_byteout(e);
    } while (!(e->A & 0x8000));
}

static void mq_encode(MqEnc *e, int cx, int d)
{
    const int i = e->idx[cx];
    const uint32_t qe = kQe[i];
    if (d == e->mps[cx]) { /* CODEMPS */
        e->A -= qe;
        if (!(e->A & 0x8000)) {
            if (e->A < qe)
                e->A = qe;
            else
                e->C += qe;
            e->idx[cx] = kNmps[i];
            mq_renorme(e);
        } else {
            e->C += qe;
        }
    } else { /* CODELPS */
        e->A -= qe;
        if (e->A < qe)
            e->C += qe;
        else
            e->A = qe;
        if (kSwitch[i]) e->mps[cx] ^= 1;
        e->idx[cx] = kNlps[i];
        mq_renorme(e);
    }
}

/* INITENC (C.2.8) on a new codeword segment; the context states are kept */
static void mq_start(MqEnc *e, uint8_t *buf, long cap)
{
    e->A = 0x8000;
    e->C = 0;
    e->CT = 12;
    e->buf = buf;
    e->bp = -1;
    e->cap = cap;
    e->overflow = 0;
    buf[-1] = 0;
}

/* FLUSH (C.2.9): SETBITS, two BYTEOUTs; a final 0xFF byte is not emitted */
static long mq_flush(MqEnc *e)
{
    const uint32_t tempc = e->C + e->A;
    e->C |= 0xFFFF;
    if (e->C >= tempc) e->C -= 0x8000;
    e->C <<= e->CT;
    mq_byteout(e);
    e->C <<= e->CT;
    mq_byteout(e);
    long len = e->bp + 1;
    if (len > 0 && e->buf[len - 1] == 0xFF) --len;
    return len;
}

/* ---- raw (bypass) segments (D.6) ----------------------------------------
 * Bits are packed most significant first; a byte that follows 0xFF carries
 * seven bits (its top bit is a stuffed 0).  Termination pads the last byte
 * with 0 bits and drops a final 0xFF (a decoder reads 1 bits past the end
 * of a segment, which is what that byte held). */
typedef struct {
    uint8_t *buf;
    long n, cap;
    uint32_t acc;
    int nbits, room, overflow;
} RawEnc;

static void raw_start(RawEnc *r, uint8_t *buf, long cap)
{
    r->buf = buf;
    r->n = 0;
    r->cap = cap;
    r->acc = 0;
    r->nbits = 0;
    r->room = 8;
    r->overflow = 0;
}

static void raw_emit(RawEnc *r, uint32_t byte)
{
    if (r->n >= r->cap) {
        r->overflow = 1;
        return;
    }
    r->buf[r->n++] = (uint8_t)byte;
    r->room = (byte == 0xFF) ? 7 : 8;
}

static void raw_put(RawEnc *r, int bit)
{
    r->acc = (r->acc << 1) | (uint32_t)(bit & 1);
    if (++r->nbits == r->room) {
        raw_emit(r, r->acc);
        r->acc = 0;
        r->nbits = 0;
    }
}

static long raw_flush(RawEnc *r)
{
    if (r->nbits) raw_emit(r, r->acc << (r->room - r->nbits));
    r->acc = 0;
    r->nbits = 0;
    if (r->n > 0 && r->buf[r->n - 1] == 0xFF) --r->n;
    return r->n;
}

/* ---- contexts (T.800 Tables D.1, D.3, D.4) ---------------------------- */

/* state flags per sample, on a grid with a one-sample border */
enum { F_SIG = 1, F_NEG = 2, F_VISIT = 4, F_REFINED = 8 };

static int zc_context(const uint8_t *f, int stride, int orient)
{
    const int h = (f[-1] & F_SIG) + (f[1] & F_SIG);
    const int v = (f[-stride] & F_SIG) + (f[stride] & F_SIG);
    const int d = (f[-stride - 1] & F_SIG) + (f[-stride + 1] & F_SIG) + (f[stride - 1] & F_SIG) +
                  (f[stride + 1] & F_SIG);
    if (orient == 3) { /* HH */
        const int hv = h + v;
        if (d >= 3) return 8;
        if (d == 2) return hv >= 1 ? 7 : 6;
        if (d == 1) return hv >= 2 ? 5 : (hv == 1 ? 4 : 3);
        return hv >= 2 ? 2 : hv;
    }
    int a = h, b = v; /* LL, LH: horizontal sum leads */
    if (orient == 1) a = v, b = h; /* HL: vertical sum leads */
    if (a == 2) return 8;
    if (a == 1) return b >= 1 ? 7 : (d >= 1 ? 6 : 5);
    if (b == 2) return 4;
    if (b == 1) return 3;
    return d >= 2 ? 2 : d;
}

static int contrib(uint8_t n) { return (n & F_SIG) ? ((n & F_NEG) ? -1 : 1) : 0; }

/* sign context and XOR bit */
static void sc_context(const uint8_t *f, int stride, int *cx, int *xr)
{
    int h = contrib(f[-1]) + contrib(f[1]);
    int v = contrib(f[-stride]) + contrib(f[stride]);
    h = h > 1 ? 1 : (h < -1 ? -1 : h);
    v = v > 1 ? 1 : (v < -1 ? -1 : v);
    *xr = 0;
    if (h < 0 || (h == 0 && v < 0)) {
        h = -h;
        v = -v;
        *xr = 1;
    }
    /* h = 0: v in {0,1}; h = 1: v in {-1,0,1} */
    *cx = CX_SC + (h == 0 ? v : 3 + v);
}

static int any_sig_neighbour(const uint8_t *f, int stride)
{
    return ((f[-stride - 1] | f[-stride] | f[-stride + 1] | f[-1] | f[1] | f[stride - 1] |
             f[stride] | f[stride + 1]) &
            F_SIG) != 0;
}

/* ---- the coder ----------------------------------------------------------- */

/* code-block style switches (T.800 Table A.19) */
enum { CB_BYPASS = 0x01, CB_RESET = 0x02, CB_TERMALL = 0x04, CB_VSC = 0x08, CB_PTERM = 0x10,
       CB_SEGSYM = 0x20 };

/* Coding pass i of a code-block (0 = the first cleanup pass, then
   significance, refinement, cleanup per bit-plane: type (i + 2) % 3).
   With bypass, passes 10, 11, 13, 14, ... (significance and refinement
   from the fifth coded bit-plane on) are raw; a codeword segment ends at
   every pass (termall), or with bypass after pass 9, after each raw pair
   and after each cleanup pass from pass 12 on. */
static int pass_is_raw(int i, int style) { return (style & CB_BYPASS) && i >= 10 && (i % 3) != 0; }
static int pass_starts_segment(int i, int style)
{
    if (i == 0 || (style & CB_TERMALL)) return 1;
    return (style & CB_BYPASS) && i >= 10 && (i % 3) != 2;
}

/* the 3x3 neighbourhood of f as the context rules see it: with vertically
   causal contexts the row below a stripe's last row reads as insignificant */
static const uint8_t *nbhood(const uint8_t *f, int stride, int causal, uint8_t g[9])
{
    for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx)
            g[(dy + 1) * 3 + dx + 1] = (causal && dy == 1) ? 0 : f[dy * stride + dx];
    return &g[4];
}

/*
 * coef   w*h quantisation indices, two's complement, row pitch w
 * mb     magnitude bit-planes of the band (numbps): |q| < 2^mb required
 * orient 0 LL, 1 HL, 2 LH, 3 HH
 * pass_limit  0 = every pass, else at most this many passes
 * style  code-block style bits (CB_*)
 * out/cap     output buffer; *len bytes written (all segments in order)
 * *num_passes, *missing   passes coded and all-zero leading bit-planes
 * seg_len, max_segs, nseg   byte length of each codeword segment
 * Returns 0, -1 on bad arguments or magnitude overflow, -2 if cap or
 * max_segs is short.
 */
int bmi_t1_encode_block_ex(const int32_t *coef, int w, int h, int mb, int orient, int pass_limit,
                           int style, uint8_t *out, long cap, long *len, int *num_passes,
                           int *missing, long *seg_len, int max_segs, int *nseg)
{
    if (!coef || w < 1 || h < 1 || w > 1024 || h > 1024 || mb < 0 || mb > 30 || orient < 0 ||
        orient > 3 || !out || cap < 4 || !len || !num_passes || !missing || !seg_len ||
        max_segs < 1 || !nseg || (style & ~0x3F))
        return -1;
    uint32_t maxmag = 0;
    for (long i = 0; i < (long)w * h; ++i) {
        const int64_t v = coef[i];
        const uint32_t m = (uint32_t)(v < 0 ? -v : v);
        if (m > maxmag) maxmag = m;
    }
    if (mb < 31 && (uint64_t)maxmag >= ((uint64_t)1 << mb)) return -1;
    int K = 0;
    while (K < 31 && (maxmag >> K)) ++K;
    *missing = mb - K;
    *len = 0;
    *num_passes = 0;
    *nseg = 0;
    if (K == 0) return 0;
    int total = 3 * K - 2;
    if (pass_limit > 0 && pass_limit < total) total = pass_limit;

    const int stride = w + 2;
    const int vsc = (style & CB_VSC) != 0;
    uint8_t *flags = (uint8_t *)calloc((size_t)stride * (h + 2), 1);
    uint32_t *mag = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)w * h);
    uint8_t *tmp = (uint8_t *)malloc((size_t)cap + 1);
    if (!flags || !mag || !tmp) {
        free(flags);
        free(mag);
        free(tmp);
        return -1;
    }
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const int32_t v = coef[(long)y * w + x];
            mag[(long)y * w + x] = (uint32_t)(v < 0 ? -(int64_t)v : v);
            if (v < 0) flags[(y + 1) * stride + x + 1] |= F_NEG;
        }
    MqEnc e;
    RawEnc r;
    mq_reset_contexts(&e);
    long used = 0;  /* bytes of finished segments in out */
    int segs = 0, rc = 0, raw = 0;
    /* end the open segment: terminate its coder and append its bytes */
    #define END_SEGMENT()                                                              \
        do {                                                                           \
            const long n_ = raw ? raw_flush(&r) : mq_flush(&e);                        \
            if ((raw ? r.overflow : e.overflow) || used + n_ > cap || segs >= max_segs) \
                rc = -2;                                                               \
            else {                                                                     \
                memcpy(out + used, tmp + 1, (size_t)n_);                               \
                used += n_;                                                            \
                seg_len[segs++] = n_;                                                  \
            }                                                                          \
        } while (0)

    int passes = 0;
    for (int p = K - 1; p >= 0 && passes < total && rc == 0; --p) {
        for (int type = (p == K - 1) ? 2 : 0; type < 3 && passes < total && rc == 0; ++type, ++passes) {
            if (pass_starts_segment(passes, style)) {
                if (passes > 0) END_SEGMENT();
                raw = pass_is_raw(passes, style);
                if (raw)
                    raw_start(&r, tmp + 1, cap);
                else
                    mq_start(&e, tmp + 1, cap);
            }
            if ((style & CB_RESET) && passes > 0) mq_reset_contexts(&e);
            for (int y0 = 0; y0 < h; y0 += 4) {
                const int rows = (h - y0) < 4 ? (h - y0) : 4;
                for (int x = 0; x < w; ++x) {
                    int y = y0;
                    if (type == 2 && rows == 4) {
                        /* run mode: four insignificant, unvisited samples with
                           insignificant neighbourhoods */
                        int run = 1;
                        for (int k = 0; k < 4 && run; ++k) {
                            uint8_t g[9];
                            const uint8_t *f = &flags[(y0 + k + 1) * stride + x + 1];
                            if ((*f & (F_SIG | F_VISIT)) ||
                                any_sig_neighbour(nbhood(f, stride, vsc && k == 3, g), 3))
                                run = 0;
                        }
                        if (run) {
                            int first = -1;
                            for (int k = 0; k < 4; ++k)
                                if ((mag[(long)(y0 + k) * w + x] >> p) & 1) {
                                    first = k;
                                    break;
                                }
                            if (first < 0) {
                                mq_encode(&e, CX_RL, 0);
                                continue;
                            }
                            mq_encode(&e, CX_RL, 1);
                            mq_encode(&e, CX_UNI, first >> 1);
                            mq_encode(&e, CX_UNI, first & 1);
                            uint8_t *f = &flags[(y0 + first + 1) * stride + x + 1];
                            uint8_t g[9];
                            int cx, xr;
                            sc_context(nbhood(f, stride, vsc && first == 3, g), 3, &cx, &xr);
                            mq_encode(&e, cx, ((*f & F_NEG) ? 1 : 0) ^ xr);
                            *f |= F_SIG;
                            y = y0 + first + 1;
                        }
                    }
                    for (; y < y0 + rows; ++y) {
                        uint8_t *f = &flags[(y + 1) * stride + x + 1];
                        uint8_t g[9];
                        const uint8_t *n = nbhood(f, stride, vsc && (y & 3) == 3, g);
                        const int bit = (mag[(long)y * w + x] >> p) & 1;
                        if (type == 0) {
                            if ((*f & F_SIG) || !any_sig_neighbour(n, 3)) continue;
                            *f |= F_VISIT;
                            if (raw) {
                                raw_put(&r, bit);
                                if (bit) {
                                    raw_put(&r, (*f & F_NEG) ? 1 : 0);
                                    *f |= F_SIG;
                                }
                                continue;
                            }
                            mq_encode(&e, zc_context(n, 3, orient), bit);
                        } else if (type == 1) {
                            if (!(*f & F_SIG) || (*f & F_VISIT)) continue;
                            if (raw)
                                raw_put(&r, bit);
                            else
                                mq_encode(&e, (*f & F_REFINED) ? CX_MR + 2 : CX_MR + any_sig_neighbour(n, 3),
                                          bit);
                            *f |= F_REFINED;
                            continue;
                        } else {
                            if (*f & (F_SIG | F_VISIT)) continue;
                            mq_encode(&e, zc_context(n, 3, orient), bit);
                        }
                        if (bit) {
                            int cx, xr;
                            sc_context(n, 3, &cx, &xr);
                            mq_encode(&e, cx, ((*f & F_NEG) ? 1 : 0) ^ xr);
                            *f |= F_SIG;
                        }
                    }
                }
            }
            if (type == 2) {
                if (style & CB_SEGSYM) { /* 1010 in the uniform context */
                    mq_encode(&e, CX_UNI, 1);
                    mq_encode(&e, CX_UNI, 0);
                    mq_encode(&e, CX_UNI, 1);
                    mq_encode(&e, CX_UNI, 0);
                }
                for (long i = 0; i < (long)stride * (h + 2); ++i) flags[i] &= (uint8_t)~F_VISIT;
            }
        }
    }
    if (rc == 0) END_SEGMENT();
    #undef END_SEGMENT
    if (rc == 0) {
        *len = used;
        *num_passes = passes;
        *nseg = segs;
    }
    free(flags);
    free(mag);
    free(tmp);
    return rc;
}

/* style 0: one codeword segment (the form the reference decodes) */
int bmi_t1_encode_block(const int32_t *coef, int w, int h, int mb, int orient, int pass_limit,
                        uint8_t *out, long cap, long *len, int *num_passes, int *missing)
{
    long seg[1];
    int nseg;
    return bmi_t1_encode_block_ex(coef, w, h, mb, orient, pass_limit, 0, out, cap, len, num_passes,
                                  missing, seg, 1, &nseg);
}

/* the encoder's Table C.2, for the test that compares it with the standard */
void bmi_t1_encoder_tables(uint16_t *qe, uint8_t *nmps, uint8_t *nlps, uint8_t *sw)
{
    memcpy(qe, kQe, sizeof(kQe));
    memcpy(nmps, kNmps, sizeof(kNmps));
    memcpy(nlps, kNlps, sizeof(kNlps));
    memcpy(sw, kSwitch, sizeof(kSwitch));
}

/* the encoder's zero-coding and sign contexts for a 3x3 neighbourhood
   given as 9 flags (row-major, centre at 4; bit0 significant, bit1
   negative), for the table tests */
int bmi_t1_encoder_zc(const uint8_t *nb9, int orient)
{
    uint8_t g[9];
    for (int i = 0; i < 9; ++i) g[i] = (nb9[i] & 1 ? F_SIG : 0) | (nb9[i] & 2 ? F_NEG : 0);
    return zc_context(&g[4], 3, orient);
}

int bmi_t1_encoder_sc(const uint8_t *nb9)
{
    uint8_t g[9];
    for (int i = 0; i < 9; ++i) g[i] = (nb9[i] & 1 ? F_SIG : 0) | (nb9[i] & 2 ? F_NEG : 0);
    int cx, xr;
    sc_context(&g[4], 3, &cx, &xr);
    return cx * 2 + xr;
}
