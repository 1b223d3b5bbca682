// bmi_t1.hip -- EBCOT tier-1 code-block decoding on gfx950, one code-block
// per wavefront.
//
// The MQ decoder is one serial chain of decisions per code-block, so a
// wave runs it as wave-uniform code: the MQ registers, context states and
// the current column's state are uniform and live in scalar registers, and
// the branches are scalar.  The 64 lanes do the parallel parts, one lane
// per column: clearing the block state, building per-stripe state, and
// reading and writing the decoded block (coalesced rows).
//
// Two kernels run over the same block table, each skipping the other's
// blocks:
//  - t1_style0_kernel, for code-block style 0 (the style the reference
//    decodes): per stripe and column a flag word in LDS, 3x3 significance
//    windows built by the lanes at each stripe, and the MQ decoder as one
//    block of scalar assembly (its comments below);
//  - t1_styled_kernel, for every other switch of T.800 Table A.19
//    (bypass with raw passes, context reset, termination on every pass,
//    vertically causal contexts, segmentation symbols; D.5-D.7), which the
//    reference asserts on.  There the stripe being decoded lives in
//    registers, one column per lane: four sample words and four 16-bit
//    state words; across stripes the state words are kept in LDS (8 KB per
//    wave) and the sample words in the block's own output rows in global
//    memory, loaded one stripe ahead of use and stored back after it.  The
//    state word carries the significance of the eight neighbours, so a
//    context is formed from one lookup:
//      bits 0..7   significance of N, S, W, E, NW, NE, SW, SE
//      bit 8       significant        bit 9   negative
//      bit 10      coded in this plane's significance pass
//      bit 11      refinable (significant before the current plane)
//      bits 12..15 sign of the N, S, W, E neighbours
//
// Semantics follow decode_code_block (decoder/codec/src/bmi_block_decoder.cpp:
// 589-853) and its passes (:10-586), the MQ decoder
// (decoder/codec/src/bmi_mq_decoder.cpp:186-355) and dec_cb_stripe's copy
// (decoder/codec/src/dec_t1_comp.cpp:283-426); the procedures are written
// from ITU-T T.800 C.3 and D.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <type_traits>

#include "bmi_device.h"
#include "bmi_internal.h"
#include "bmi_t1.h"
#include "bmi_t1_style.h"

namespace bmi {
namespace {

constexpr int kMaxSide = 64;
constexpr int kWavesPerGroup = 4;

// T.800 Table C.2 packed: Qe | NMPS << 16 | NLPS << 22 | SWITCH << 28
#define MQE(q, nm, nl, sw) ((uint32_t)(q) | ((uint32_t)(nm) << 16) | ((uint32_t)(nl) << 22) | ((uint32_t)(sw) << 28))
__constant__ uint32_t kMqTable[47] = {
    MQE(0x5601, 1, 1, 1),    MQE(0x3401, 2, 6, 0),    MQE(0x1801, 3, 9, 0),
    MQE(0x0AC1, 4, 12, 0),   MQE(0x0521, 5, 29, 0),   MQE(0x0221, 38, 33, 0),
    MQE(0x5601, 7, 6, 1),    MQE(0x5401, 8, 14, 0),   MQE(0x4801, 9, 14, 0),
    MQE(0x3801, 10, 14, 0),  MQE(0x3001, 11, 17, 0),  MQE(0x2401, 12, 18, 0),
    MQE(0x1C01, 13, 20, 0),  MQE(0x1601, 29, 21, 0),  MQE(0x5601, 15, 14, 1),
    MQE(0x5401, 16, 14, 0),  MQE(0x5101, 17, 15, 0),  MQE(0x4801, 18, 16, 0),
    MQE(0x3801, 19, 17, 0),  MQE(0x3401, 20, 18, 0),  MQE(0x3001, 21, 19, 0),
    MQE(0x2801, 22, 19, 0),  MQE(0x2401, 23, 20, 0),  MQE(0x2201, 24, 21, 0),
    MQE(0x1C01, 25, 22, 0),  MQE(0x1801, 26, 23, 0),  MQE(0x1601, 27, 24, 0),
    MQE(0x1401, 28, 25, 0),  MQE(0x1201, 29, 26, 0),  MQE(0x1101, 30, 27, 0),
    MQE(0x0AC1, 31, 28, 0),  MQE(0x09C1, 32, 29, 0),  MQE(0x08A1, 33, 30, 0),
    MQE(0x0521, 34, 31, 0),  MQE(0x0441, 35, 32, 0),  MQE(0x02A1, 36, 33, 0),
    MQE(0x0221, 37, 34, 0),  MQE(0x0141, 38, 35, 0),  MQE(0x0111, 39, 36, 0),
    MQE(0x0085, 40, 37, 0),  MQE(0x0049, 41, 38, 0),  MQE(0x0025, 42, 39, 0),
    MQE(0x0015, 43, 40, 0),  MQE(0x0009, 44, 41, 0),  MQE(0x0005, 45, 42, 0),
    MQE(0x0001, 45, 43, 0),  MQE(0x5601, 46, 46, 0)};
#undef MQE

// state word bits (the four direct neighbours first, so that the sign
// context is indexed by bits 0..3 and 12..15)
constexpr uint32_t S_N = 1u << 0, S_S = 1u << 1, S_W = 1u << 2, S_E = 1u << 3, S_NW = 1u << 4,
                   S_NE = 1u << 5, S_SW = 1u << 6, S_SE = 1u << 7;
constexpr uint32_t S_NB = 0xFFu;
constexpr uint32_t S_SIG = 1u << 8, S_NEG = 1u << 9, S_VISIT = 1u << 10, S_MU = 1u << 11;
constexpr uint32_t S_NNEG = 1u << 12, S_SNEG = 1u << 13, S_WNEG = 1u << 14, S_ENEG = 1u << 15;

// contexts: 0..8 zero coding, 9..13 sign, 14..16 refinement, 17 run, 18 uniform
constexpr int CX_SC = 9, CX_MR = 14, CX_RL = 17, CX_UNI = 18;

__device__ __forceinline__ uint32_t uni(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// zero-coding context from the neighbour bits (T.800 Table D.1)
__device__ __forceinline__ uint32_t zc_context(uint32_t s, int orient)
{
    const uint32_t h = __builtin_popcount(s & (S_W | S_E));
    const uint32_t v = __builtin_popcount(s & (S_N | S_S));
    const uint32_t d = __builtin_popcount(s & (S_NW | S_NE | S_SW | S_SE));
    if (orient == 3) {
        const uint32_t hv = h + v;
        if (d >= 3) return 8;
        if (d == 2) return hv ? 7 : 6;
        if (d == 1) return hv >= 2 ? 5 : 3 + hv;
        return hv >= 2 ? 2 : hv;
    }
    const uint32_t a = orient == 1 ? v : h;
    const uint32_t b = orient == 1 ? h : v;
    if (a == 2) return 8;
    if (a == 1) return b ? 7 : (d ? 6 : 5);
    if (b) return b == 2 ? 4 : 3;
    return d >= 2 ? 2 : d;
}

// sign context (T.800 Table D.3): returns context | xor << 8
__device__ __forceinline__ uint32_t sc_context(uint32_t s)
{
    const int w = (s & S_W) ? ((s & S_WNEG) ? -1 : 1) : 0;
    const int e = (s & S_E) ? ((s & S_ENEG) ? -1 : 1) : 0;
    const int n = (s & S_N) ? ((s & S_NNEG) ? -1 : 1) : 0;
    const int so = (s & S_S) ? ((s & S_SNEG) ? -1 : 1) : 0;
    int h = w + e, v = n + so;
    h = h > 0 ? 1 : (h < 0 ? -1 : 0);
    v = v > 0 ? 1 : (v < 0 ? -1 : 0);
    uint32_t xr = 0;
    if (h < 0 || (h == 0 && v < 0)) {
        h = -h;
        v = -v;
        xr = 1;
    }
    return (uint32_t)(CX_SC + (h == 0 ? v : 3 + v)) | (xr << 8);
}

constexpr int kMaxStripes = kMaxSide / 4;

// Per wave, LDS holds the state words of the whole block, stripe-major
// and column-major inside a stripe (one lane loads one column's four rows
// with one access).  The sample words live in the block's own output
// rows in global memory: each stripe is loaded one stripe ahead of its
// use and stored back after it, which keeps LDS at 8 KB per wave and the
// number of waves per CU high (the decoder is latency-bound).
struct WaveLds {
    uint32_t st[kMaxStripes][kMaxSide][2];  // 16-bit state of rows 0,1 | 2,3
};

__device__ __forceinline__ uint32_t rdl(uint32_t v, uint32_t lane)
{
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)lane);
}
// value v with lane `lane` replaced by the uniform s (a compare and a
// select per lane)
__device__ __forceinline__ uint32_t wrl(uint32_t s, uint32_t lane, uint32_t v)
{
    return ((threadIdx.x & 63) == lane) ? s : v;
}

// Context tables of one code-block, 256 nibbles each, eight per lane in
// lanes 0..31 of a vector register: zero coding indexed by the eight
// neighbour bits (for the block's orientation), sign coding indexed by
// the four direct neighbours' significance and sign bits, holding
// (context - 9) | xor << 3.
__device__ __forceinline__ uint32_t build_lut(uint32_t lane, int orient, bool sign)
{
    uint32_t word = 0;
    if (lane < 32)
        for (int j = 0; j < 8; ++j) {
            const uint32_t i = lane * 8 + j;
            uint32_t e;
            if (sign) {
                const uint32_t st = (i & 0xF) | ((i & 0xF0) << 8);
                const uint32_t sc = sc_context(st);
                e = ((sc & 0xFF) - CX_SC) | ((sc >> 8) << 3);
            } else {
                e = zc_context(i, orient);
            }
            word |= e << (4 * j);
        }
    return word;
}
__device__ __forceinline__ uint32_t lut_get(uint32_t lut, uint32_t i)
{
    return (rdl(lut, i >> 3) >> (4 * (i & 7))) & 0xF;
}
__device__ __forceinline__ uint32_t zc_lut(uint32_t zl, uint32_t s) { return lut_get(zl, s & S_NB); }
// sign context | xor << 8, as sc_context
__device__ __forceinline__ uint32_t sc_lut(uint32_t sl, uint32_t s)
{
    const uint32_t e = lut_get(sl, (s & 0xF) | ((s >> 8) & 0xF0));
    return (CX_SC + (e & 7)) | ((e >> 3) << 8);
}

// MQ decoder (T.800 C.3).  Registers A, C, CT and the byte position are
// wave-uniform (scalar).  The code stream is read through a 64-bit buffer
// holding the bytes from the current one on, refilled with scalar loads
// (the stream is read-only for the kernel).  The Qe table and the 19
// context states live one entry per lane in vector registers, read with
// readlane and written with a lane select.
struct Mq {
    uint32_t A, C;
    int CT;
    int bp;        // segment index of the current byte
    int len;
    const uint8_t *seg;
    uint64_t buf;  // bytes bp, bp+1, ... (little endian), nbuf of them valid
    int nbuf;
    uint32_t qtab;  // per lane: Table C.2 entry `lane`
    uint32_t ctx;   // per lane: Table C.2 entry of context `lane`'s index | MPS << 29
    uint32_t zl, sl;  // zero-coding and sign context tables (build_lut)

    // keep at least 5 bytes from bp on in the buffer; bytes past the
    // segment are masked by byte_at, the loads stay within the 16 bytes
    // of slack the caller guarantees after every segment
    __device__ __forceinline__ void fill()
    {
        while (nbuf <= 4) {
            const uintptr_t a = (uintptr_t)(seg + bp + nbuf);
            const auto *p = (const __attribute__((address_space(4))) uint32_t *)(a & ~(uintptr_t)3);
            const int sh = (int)(a & 3);
            buf |= (uint64_t)(*p >> (8 * sh)) << (8 * nbuf);
            nbuf += 4 - sh;
        }
    }
    // byte bp + i (i = 0 or 1); 0xFF past the segment, as the reference's
    // decoder appends 0xFF 0xFF (bmi_mq_decoder.cpp:195)
    __device__ __forceinline__ uint32_t byte_at(int i) const
    {
        return (bp + i >= len) ? 0xFFu : (uint32_t)(buf >> (8 * i)) & 0xFF;
    }
    __device__ __forceinline__ void advance()
    {
        ++bp;
        buf >>= 8;
        --nbuf;
        fill();
    }
    __device__ __forceinline__ void bytein()
    {
        if (byte_at(0) == 0xFF) {
            const uint32_t b1 = byte_at(1);
            if (b1 > 0x8F) {
                C += 0xFF00;
                CT = 8;
            } else {
                advance();
                C += b1 << 9;
                CT = 7;
            }
        } else {
            advance();
            C += byte_at(0) << 8;
            CT = 8;
        }
    }
    // per code-block: context tables, Table C.2 and the initial states
    __device__ __forceinline__ void init_block(uint32_t lane, int orient)
    {
        zl = build_lut(lane, orient, false);
        sl = build_lut(lane, orient, true);
        qtab = lane < 47 ? kMqTable[lane] : 0u;
        reset_states(lane);
    }
    // all contexts index 0 / MPS 0, except zero coding 0 (index 4), run
    // (index 3) and uniform (index 46)
    __device__ __forceinline__ void reset_states(uint32_t lane)
    {
        ctx = kMqTable[lane == 0 ? 4 : (lane == CX_RL ? 3 : (lane == CX_UNI ? 46 : 0))];
    }
    __device__ __forceinline__ void open(const uint8_t *sg, int n)
    {
        len = n;
        seg = sg;
        buf = 0;
        nbuf = 0;
        bp = 0;
        fill();
    }
    // INITDEC on a codeword segment (the context states carry over)
    __device__ __forceinline__ void start(const uint8_t *sg, int n)
    {
        open(sg, n);
        C = byte_at(0) << 16;
        bytein();
        C <<= 7;
        CT -= 7;
        A = 0x8000;
    }
    // Raw segment (bypass passes, T.800 D.6): bits most significant first,
    // seven bits in a byte that follows 0xFF.  bp is the next unread byte;
    // 0xFF followed by a byte above 0x8F (a marker, or the 0xFF bytes past
    // the segment) is not consumed and reads as 1 bits.  C holds the
    // current byte and CT its unread bits.
    __device__ __forceinline__ void start_raw(const uint8_t *sg, int n)
    {
        open(sg, n);
        C = 0;
        CT = 0;
    }
    __device__ __forceinline__ uint32_t raw_bit()
    {
        if (CT == 0) {
            const uint32_t nb = byte_at(0);
            if (C == 0xFF && nb > 0x8F) {
                CT = 8;
            } else {
                CT = (C == 0xFF) ? 7 : 8;
                C = nb;
                advance();
            }
        }
        --CT;
        return (C >> CT) & 1;
    }
    // RENORMD: shift A and C until bit 15 of A is set, reading a byte each
    // time CT runs out (several shifts at once between reads)
    __device__ __forceinline__ void renorm()
    {
        int n = __builtin_clz(A) - 16;
        while (n > 0) {
            if (CT == 0) bytein();
            const int k = n < CT ? n : CT;
            A <<= k;
            C <<= k;
            CT -= k;
            n -= k;
        }
    }
    // One decision in context cx.  Lane cx of `ctx` holds the Table C.2
    // entry of the context's current index with the MPS in bit 29, so the
    // common MPS path needs one readlane; the table is read again only
    // when the index changes.
    __device__ __forceinline__ uint32_t decode(uint32_t cx, uint32_t lane)
    {
        const uint32_t e = rdl(ctx, cx);
        const uint32_t qe = e & 0xFFFF, mps = (e >> 29) & 1;
        uint32_t d;
        A -= qe;
        if ((C >> 16) < qe) {
            if (A < qe) {
                d = mps;
                ctx = wrl(rdl(qtab, (e >> 16) & 63) | (mps << 29), cx, ctx);
            } else {
                d = mps ^ 1;
                ctx = wrl(rdl(qtab, (e >> 22) & 63) | ((mps ^ ((e >> 28) & 1)) << 29), cx, ctx);
            }
            A = qe;
            renorm();
        } else {
            C -= qe << 16;
            if (!(A & 0x8000)) {
                if (A < qe) {
                    d = mps ^ 1;
                    ctx = wrl(rdl(qtab, (e >> 22) & 63) | ((mps ^ ((e >> 28) & 1)) << 29), cx, ctx);
                } else {
                    d = mps;
                    ctx = wrl(rdl(qtab, (e >> 16) & 63) | (mps << 29), cx, ctx);
                }
                renorm();
            } else {
                d = mps;
            }
        }
        return d;
    }
};

// State of the stripe being decoded: lane x holds column x (rows 0,1 in
// lo, rows 2,3 in hi, 16 bits each) and its four sample words; pending
// neighbour bits for row 3 of the stripe above and row 0 of the stripe
// below, applied when those stripes are next touched.
struct Stripe {
    uint32_t lo, hi;
    uint32_t v0, v1, v2, v3;
    uint32_t pend_prev;  // hi-word bits (row 3) of the stripe above
    uint32_t pend_next;  // lo-word bits (row 0) of the stripe below
};

template <int R>
__device__ __forceinline__ uint32_t row_state(uint32_t lo, uint32_t hi)
{
    return ((R < 2 ? lo : hi) >> (16 * (R & 1))) & 0xFFFF;
}

template <int R>
__device__ __forceinline__ uint32_t &vrow(Stripe &S)
{
    return R == 0 ? S.v0 : (R == 1 ? S.v1 : (R == 2 ? S.v2 : S.v3));
}

// OR 16-bit row bits `m` into row `r` of the (lo, hi) pair
__device__ __forceinline__ void or_row(uint32_t &lo, uint32_t &hi, int r, uint32_t m)
{
    if (r < 2)
        lo |= m << (16 * (r & 1));
    else
        hi |= m << (16 * (r & 1));
}

// sample (x, R) became significant with sign `neg`: update the state of
// its own column (uniform lo/hi), of columns x-1 / x+1 (in the vectors)
// and the pending bits of the stripes above / below
template <int R>
__device__ __forceinline__ void significant(Stripe &S, uint32_t &lo, uint32_t &hi, int x, int w,
                                            uint32_t neg, bool vsc)
{
    or_row(lo, hi, R, S_SIG | (neg ? S_NEG : 0));
    const uint32_t vn = S_N | (neg ? S_NNEG : 0), vs = S_S | (neg ? S_SNEG : 0);
    if (R > 0) or_row(lo, hi, R - 1, vs);
    if (R < 3) or_row(lo, hi, R + 1, vn);
    // with vertically causal contexts the last row of the stripe above
    // does not see this stripe
    if (R == 0 && !vsc) S.pend_prev = wrl(rdl(S.pend_prev, x) | (vs << 16), x, S.pend_prev);
    if (R == 3) S.pend_next = wrl(rdl(S.pend_next, x) | vn, x, S.pend_next);
    // the columns to the left and right
#pragma unroll
    for (int side = 0; side < 2; ++side) {
        const int c = side ? x + 1 : x - 1;
        if (c < 0 || c >= w) continue;
        const uint32_t hbit = side ? (S_W | (neg ? S_WNEG : 0)) : (S_E | (neg ? S_ENEG : 0));
        const uint32_t up = side ? S_SW : S_SE;    // for the row above
        const uint32_t down = side ? S_NW : S_NE;  // for the row below
        uint32_t clo = 0, chi = 0;
        or_row(clo, chi, R, hbit);
        if (R > 0) or_row(clo, chi, R - 1, up);
        if (R < 3) or_row(clo, chi, R + 1, down);
        if (clo) S.lo = wrl(rdl(S.lo, c) | clo, c, S.lo);
        if (chi) S.hi = wrl(rdl(S.hi, c) | chi, c, S.hi);
        if (R == 0 && !vsc) S.pend_prev = wrl(rdl(S.pend_prev, c) | (up << 16), c, S.pend_prev);
        if (R == 3) S.pend_next = wrl(rdl(S.pend_next, c) | down, c, S.pend_next);
    }
}

// significance propagation of row R of column x
// (raw: the significance bit and then the sign bit itself, uncoded)
template <int R>
__device__ __forceinline__ void sigprop_row(Mq &m, Stripe &S, uint32_t &lo, uint32_t &hi, int x,
                                            int w, int rows, int orient, uint32_t one_half,
                                            uint32_t lane, bool raw, bool vsc)
{
    if (R >= rows) return;
    const uint32_t s = row_state<R>(lo, hi);
    if ((s & S_SIG) || !(s & S_NB)) return;
    or_row(lo, hi, R, S_VISIT);
    uint32_t neg;
    if (raw) {
        if (!m.raw_bit()) return;
        neg = m.raw_bit();
    } else {
        if (!m.decode(zc_lut(m.zl, s), lane)) return;
        const uint32_t sc = sc_lut(m.sl, s);
        neg = m.decode(sc & 0xFF, lane) ^ (sc >> 8);
    }
    significant<R>(S, lo, hi, x, w, neg, vsc);
    vrow<R>(S) = wrl((neg << 31) + one_half, x, vrow<R>(S));
}

template <int R>
__device__ __forceinline__ void magref_row(Mq &m, Stripe &S, uint32_t lo, uint32_t hi, int x,
                                           int rows, int plane, uint32_t half, uint32_t lane,
                                           bool raw)
{
    if (R >= rows) return;
    const uint32_t s = row_state<R>(lo, hi);
    if (!(s & S_MU)) return;
    const uint32_t v = rdl(vrow<R>(S), x);
    const uint32_t mag = (v & 0x7FFFFFFFu) >> plane;
    const uint32_t cx = mag < 4 ? CX_MR + ((s & S_NB) ? 1 : 0) : CX_MR + 2;
    const uint32_t bit = raw ? m.raw_bit() : m.decode(cx, lane);
    vrow<R>(S) = wrl((v ^ ((1u - bit) << plane)) | half, x, vrow<R>(S));
}

// cleanup of row R of column x; rows before `first` are skipped (run mode)
template <int R>
__device__ __forceinline__ void cleanup_row(Mq &m, Stripe &S, uint32_t &lo, uint32_t &hi, int x,
                                            int w, int rows, int orient, uint32_t one_half,
                                            int first, uint32_t lane, bool vsc)
{
    if (R >= rows || R < first) return;
    const uint32_t s = row_state<R>(lo, hi);
    if (s & (S_SIG | S_VISIT)) return;
    if (!m.decode(zc_lut(m.zl, s), lane)) return;
    const uint32_t sc = sc_lut(m.sl, s);
    const uint32_t neg = m.decode(sc & 0xFF, lane) ^ (sc >> 8);
    significant<R>(S, lo, hi, x, w, neg, vsc);
    vrow<R>(S) = wrl((neg << 31) + one_half, x, vrow<R>(S));
}

template <int R>
__device__ __forceinline__ void run_hit(Mq &m, Stripe &S, uint32_t &lo, uint32_t &hi, int x, int w,
                                        uint32_t one_half, uint32_t lane, bool vsc)
{
    const uint32_t s = row_state<R>(lo, hi);
    const uint32_t sc = sc_lut(m.sl, s);
    const uint32_t neg = m.decode(sc & 0xFF, lane) ^ (sc >> 8);
    significant<R>(S, lo, hi, x, w, neg, vsc);
    vrow<R>(S) = wrl((neg << 31) + one_half, x, vrow<R>(S));
}

// mask of the 16-bit rows [0, rows) in the (lo, hi) pair, as (mlo, mhi)
__device__ __forceinline__ void rows_mask(int rows, uint32_t bits, uint32_t &mlo, uint32_t &mhi)
{
    mlo = (rows > 0 ? bits : 0) | (rows > 1 ? bits << 16 : 0);
    mhi = (rows > 2 ? bits : 0) | (rows > 3 ? bits << 16 : 0);
}

// The blocks with any switch of T.800 Table A.19 other than predictable
// termination; t1_style0_kernel, launched over the same table, decodes the
// others, and each skips the other's blocks.  Kept apart so that the style
// machinery costs the common kernel no registers.
__global__ void __launch_bounds__(64 * kWavesPerGroup)
    t1_styled_kernel(const BmiT1Block *__restrict__ blocks, int nblocks,
                     const uint8_t *__restrict__ data)
{
    constexpr bool STYLED = true;
    __shared__ WaveLds lds_all[kWavesPerGroup];
    const uint32_t lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int bi = (int)uni(blockIdx.x * kWavesPerGroup + wave);
    if (bi >= nblocks) return;
    WaveLds &L = lds_all[wave];
    const BmiT1Block *B = blocks + bi;
    const int w = (int)uni(B->width), h = (int)uni(B->height);
    if (w < 1 || h < 1 || w > kMaxSide || h > kMaxSide) return;
    const int orient = (int)uni(B->orient);
    const int numbps = (int)uni(B->numbps);
    const int missing = (int)uni(B->missing);
    const int num_pass = (int)uni(B->num_pass);
    const int kind = (int)uni(B->kind);
    const int len = (int)uni(B->len);
    const uint64_t off = ((uint64_t)uni((uint32_t)(B->data_offset >> 32)) << 32) |
                         uni((uint32_t)B->data_offset);
    const uintptr_t outp = ((uintptr_t)uni((uint32_t)((uintptr_t)B->out >> 32)) << 32) |
                           uni((uint32_t)(uintptr_t)B->out);
    const int out_pitch = (int)uni(B->out_pitch);
    const int block_style = (int)uni(B->style) & BMI_T1_STYLES;
    if (((block_style & ~BMI_T1_PTERM) != 0) != STYLED) return;
    const int nseg = bmi_style_segmented(block_style) ? (int)uni(B->nseg) : 1;
    const uint64_t segl_off = ((uint64_t)uni((uint32_t)(B->seg_offset >> 32)) << 32) |
                              uni((uint32_t)B->seg_offset);
    const auto *segl = (const __attribute__((address_space(4))) int *)(data + segl_off);
    const int nstripes = (h + 3) >> 2;
    const uint64_t colmask = (w == 64) ? ~0ull : ((1ull << w) - 1);

    auto *o = gptr_w<uint32_t>((void *)outp);
    const bool col_ok = (int)lane < w;
    for (int s = 0; s < nstripes; ++s) L.st[s][lane][0] = L.st[s][lane][1] = 0;
    if (col_ok)
        for (int y = 0; y < h; ++y) o[(size_t)y * out_pitch + lane] = 0;
    // sample word of row y of this lane's column (0 past the block)
    auto vload = [&](int y) -> uint32_t {
        return (col_ok && y < h) ? o[(size_t)y * out_pitch + lane] : 0u;
    };
    auto vstore = [&](int y, uint32_t v) {
        if (col_ok && y < h) o[(size_t)y * out_pitch + lane] = v;
    };

    // decode_code_block :610-649
    int num_passes = 0;
    int plane = 0;
    if ((numbps - missing) * 3 - 2 > 0) {
        const int p_max = 30 - missing;
        num_passes = 3 * p_max - 2;
        if (num_passes > num_pass) num_passes = num_pass;
        plane = (numbps >> 11) - 1 - missing;
        if (plane < 0) plane = p_max;
        if (plane == 0) plane = 1;
    }
    // the passes (with STYLED false every style switch folds away)
    {
        const int style = STYLED ? block_style : 0;
        const bool segmented = STYLED && bmi_style_segmented(style);
        const bool vsc = STYLED && (style & BMI_T1_VSC);
        Mq m;
        if (num_passes > 0) m.init_block(lane, orient);
        int type = 2;
        int seg = 0, seg_pos = 0;
        bool raw = false;
        for (int pi = 0; pi < num_passes; ++pi, ++type) {
            if (type == 3) {
                type = 0;
                --plane;
            }
            // passes below bit-plane 0 are not decoded (a negative shift
            // count, undefined in the reference)
            if (plane < 0) break;
            if (bmi_pass_new_segment(pi, style)) {
                // next codeword segment, clamped to the block's bytes (a
                // missing or oversized length reads nothing beyond them)
                int n = segmented ? (seg < nseg ? (int)uni((uint32_t)segl[seg]) : 0) : len;
                n = max(0, min(n, len - seg_pos));
                raw = bmi_pass_raw(pi, style);
                if (raw)
                    m.start_raw(data + off + seg_pos, n);
                else
                    m.start(data + off + seg_pos, n);
                seg_pos += n;
                ++seg;
            }
            if ((style & BMI_T1_RESET) && pi > 0) m.reset_states(lane);
            const uint32_t one_half = (plane >= 0 && plane < 31) ? (1u << plane) + ((1u << plane) >> 1) : 0u;
            const uint32_t half = (plane >= 1 && plane < 32) ? (1u << plane) >> 1 : 0u;
            Stripe S;
            S.pend_next = 0;
            uint32_t n0 = vload(0), n1 = vload(1), n2 = vload(2), n3 = vload(3);
            for (int s = 0; s < nstripes; ++s) {
                const int rows = min(4, h - 4 * s);
                S.lo = L.st[s][lane][0] | S.pend_next;
                S.hi = L.st[s][lane][1];
                S.v0 = n0;
                S.v1 = n1;
                S.v2 = n2;
                S.v3 = n3;
                if (s + 1 < nstripes) {
                    const int y = 4 * (s + 1);
                    n0 = vload(y);
                    n1 = vload(y + 1);
                    n2 = vload(y + 2);
                    n3 = vload(y + 3);
                }
                S.pend_next = 0;
                S.pend_prev = 0;
                uint32_t mlo, mhi;
                uint64_t todo;
                if (type == 0) {
                    // a row that is insignificant with a significant neighbour
                    rows_mask(rows, 0xFFFF, mlo, mhi);
                    const uint32_t lo = S.lo & mlo, hi = S.hi & mhi;
                    bool any = false;
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const uint32_t st = ((r < 2 ? lo : hi) >> (16 * (r & 1))) & 0xFFFF;
                        any = any || (r < rows && !(st & S_SIG) && (st & S_NB));
                    }
                    todo = __builtin_amdgcn_ballot_w64(any) & colmask;
                } else if (type == 1) {
                    rows_mask(rows, S_MU, mlo, mhi);
                    todo = __builtin_amdgcn_ballot_w64(((S.lo & mlo) | (S.hi & mhi)) != 0) & colmask;
                } else {
                    // a row not yet significant or coded
                    bool any = false;
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const uint32_t st = ((r < 2 ? S.lo : S.hi) >> (16 * (r & 1))) & 0xFFFF;
                        any = any || (r < rows && !(st & (S_SIG | S_VISIT)));
                    }
                    todo = __builtin_amdgcn_ballot_w64(any) & colmask;
                }
                while (todo) {
                    const int x = (int)__builtin_ctzll(todo);
                    todo &= todo - 1;
                    uint32_t lo = rdl(S.lo, x), hi = rdl(S.hi, x);
                    const uint32_t sig0 = (lo | (hi << 1)) & (S_SIG | (S_SIG << 16) | (S_SIG << 1) | (S_SIG << 17));
                    if (type == 0) {
                        sigprop_row<0>(m, S, lo, hi, x, w, rows, orient, one_half, lane, raw, vsc);
                        sigprop_row<1>(m, S, lo, hi, x, w, rows, orient, one_half, lane, raw, vsc);
                        sigprop_row<2>(m, S, lo, hi, x, w, rows, orient, one_half, lane, raw, vsc);
                        sigprop_row<3>(m, S, lo, hi, x, w, rows, orient, one_half, lane, raw, vsc);
                        // a sample that became significant makes column x+1 a candidate
                        const uint32_t sig1 = (lo | (hi << 1)) & (S_SIG | (S_SIG << 16) | (S_SIG << 1) | (S_SIG << 17));
                        if (x + 1 < w && sig1 != sig0) todo |= 1ull << (x + 1);
                    } else if (type == 1) {
                        magref_row<0>(m, S, lo, hi, x, rows, plane, half, lane, raw);
                        magref_row<1>(m, S, lo, hi, x, rows, plane, half, lane, raw);
                        magref_row<2>(m, S, lo, hi, x, rows, plane, half, lane, raw);
                        magref_row<3>(m, S, lo, hi, x, rows, plane, half, lane, raw);
                        continue;  // refinement leaves the state words unchanged
                    } else {
                        int first = 0;
                        if (rows == 4) {
                            const uint32_t busy = (S_NB | S_SIG | S_VISIT) | ((S_NB | S_SIG | S_VISIT) << 16);
                            if (((lo | hi) & busy) == 0) {
                                if (!m.decode(CX_RL, lane)) continue;  // four zeros
                                first = (int)m.decode(CX_UNI, lane) << 1;
                                first |= (int)m.decode(CX_UNI, lane);
                                switch (first) {
                                case 0: run_hit<0>(m, S, lo, hi, x, w, one_half, lane, vsc); break;
                                case 1: run_hit<1>(m, S, lo, hi, x, w, one_half, lane, vsc); break;
                                case 2: run_hit<2>(m, S, lo, hi, x, w, one_half, lane, vsc); break;
                                default: run_hit<3>(m, S, lo, hi, x, w, one_half, lane, vsc); break;
                                }
                                ++first;
                            }
                        }
                        cleanup_row<0>(m, S, lo, hi, x, w, rows, orient, one_half, first, lane, vsc);
                        cleanup_row<1>(m, S, lo, hi, x, w, rows, orient, one_half, first, lane, vsc);
                        cleanup_row<2>(m, S, lo, hi, x, w, rows, orient, one_half, first, lane, vsc);
                        cleanup_row<3>(m, S, lo, hi, x, w, rows, orient, one_half, first, lane, vsc);
                    }
                    S.lo = wrl(lo, x, S.lo);
                    S.hi = wrl(hi, x, S.hi);
                }
                // write the stripe back; after a cleanup pass "coded" is
                // cleared and significant samples become refinable
                uint32_t slo = S.lo, shi = S.hi;
                if (type == 2) {
                    const uint32_t vis = S_VISIT | (S_VISIT << 16);
                    slo = (slo & ~vis) | ((slo & (S_SIG | (S_SIG << 16))) << 3);
                    shi = (shi & ~vis) | ((shi & (S_SIG | (S_SIG << 16))) << 3);
                }
                L.st[s][lane][0] = slo;
                L.st[s][lane][1] = shi;
                vstore(4 * s, S.v0);
                vstore(4 * s + 1, S.v1);
                vstore(4 * s + 2, S.v2);
                vstore(4 * s + 3, S.v3);
                if (s > 0) L.st[s - 1][lane][1] |= S.pend_prev;
            }
            // segmentation symbol (D.5): four uniform-context decisions
            // after each cleanup pass
            if (type == 2 && (style & BMI_T1_SEGSYM))
                for (int k = 0; k < 4; ++k) (void)m.decode(CX_UNI, lane);
        }
    }

    // 5/3: convert the words in place (dec_cb_stripe, dec_t1_comp.cpp:343-367);
    // 9/7 words are copied as they are (:329-341)
    if (kind == BMI_T1_INT && col_ok) {
        const int sh = 31 - numbps;
        for (int y = 0; y < h; ++y) {
            const uint32_t v = o[(size_t)y * out_pitch + lane];
            const int32_t r = (v & 0x80000000u) ? -(int32_t)((v & 0x7FFFFFFFu) >> sh) : (int32_t)(v >> sh);
            o[(size_t)y * out_pitch + lane] = (uint32_t)r;
        }
    }
}


// ---------------------------------------------------------------------------
// Code-block style 0 (and predictable termination, which changes nothing
// for the decoder): the blocks the reference decodes, in a kernel of their
// own that keeps the scalar work per decision small.
//
// State.  Per stripe and column one 32-bit flag word, in LDS between passes
// (4 KB per wave):
//   bits 3r, 3r+1, 3r+2 (row r = 0..3): significant, negative, coded in
//     this bit-plane's significance pass ("visited");
//   bits 12+r: refinable (significant before this bit-plane);
//   bits 16+r: refinable before the previous bit-plane as well, i.e. not
//     the first refinement (T.800 Table D.4), which the reference decides
//     from the magnitude above the plane being below 4: a sample found
//     significant at plane p has magnitude 3 << (p-1), so the two rules
//     agree.
// At the start of each stripe of each pass the lanes build, with vector
// instructions, every column's 3x3 windows of significance and of signs
// over rows -1..4 of the stripe: bit 3j + dx + 1 holds row j-1, column
// x+dx.  Row r's neighbourhood is then the 9-bit field at bit 3r
// (0 NW, 1 N, 2 NE, 3 W, 4 the sample, 5 E, 6 SW, 7 S, 8 SE), which indexes
// the zero-coding table directly.  A sample found significant in column x
// updates the scalar copy of its own window (the rows below it in the
// column) and the window of column x+1 (one lane write); column x-1 and
// the other stripes rebuild theirs from the flag words when next visited.
// The sample words of a stripe are updated at its end, with vector
// instructions, from 64-bit masks of the columns whose row r became
// significant (or decoded a refinement bit of 0).
//
// MQ registers.  A and the Qe values are kept shifted left by 16 (A's bit
// 31 is T.800's 0x8000), so C is compared and reduced without shifts and
// RENORMD shifts by the leading zero count at once.  Lane i of `nm` / `nl`
// holds the context entry that follows index i after an MPS / LPS
// decision; an entry is Qe << 16 | index | MPS << 6, and `nl` carries the
// SWITCH flag in bit 6, so the new entry is one select and one XOR.

namespace s0 {

constexpr uint32_t FS = 0x249u;   // significant, rows 0..3
constexpr uint32_t FSN = 0x6DBu;  // significant and negative, rows 0..3

struct Lds {
    uint32_t fl[kMaxStripes][kMaxSide];
};

__device__ __forceinline__ uint32_t wrlane(uint32_t v, uint32_t lane, uint32_t old) { return wrl(v, lane, old); }

// zero-coding table over the 9-bit window, 4-bit entries, eight per lane
__device__ __forceinline__ uint32_t build_zc(uint32_t lane, int orient)
{
    uint32_t word = 0;
    for (int j = 0; j < 8; ++j) {
        const uint32_t i = lane * 8 + j;
        const uint32_t st = ((i >> 1) & 1 ? S_N : 0) | ((i >> 7) & 1 ? S_S : 0) | ((i >> 3) & 1 ? S_W : 0) |
                            ((i >> 5) & 1 ? S_E : 0) | (i & 1 ? S_NW : 0) | ((i >> 2) & 1 ? S_NE : 0) |
                            ((i >> 6) & 1 ? S_SW : 0) | ((i >> 8) & 1 ? S_SE : 0);
        word |= zc_context(st, orient) << (4 * j);
    }
    return word;
}

// sign-coding table, 8-bit entries (context | xor << 4), four per lane,
// indexed by the significance of N, W, E, S at bits 1, 3, 5, 7 and their
// signs at bits 0, 2, 4, 6
__device__ __forceinline__ uint32_t build_sc(uint32_t lane)
{
    uint32_t word = 0;
    for (int j = 0; j < 4; ++j) {
        const uint32_t i = lane * 4 + j;
        const uint32_t st = ((i >> 1) & 1 ? S_N : 0) | ((i >> 3) & 1 ? S_W : 0) | ((i >> 5) & 1 ? S_E : 0) |
                            ((i >> 7) & 1 ? S_S : 0) | (i & 1 ? S_NNEG : 0) | ((i >> 2) & 1 ? S_WNEG : 0) |
                            ((i >> 4) & 1 ? S_ENEG : 0) | ((i >> 6) & 1 ? S_SNEG : 0);
        const uint32_t sc = sc_context(st);
        word |= ((sc & 0xFF) | ((sc >> 8) << 4)) << (8 * j);
    }
    return word;
}

struct Mq {
    // MQ registers, A and Qe shifted left by 16
    uint32_t A, C;
    int CT;
    // code stream reader: bytes from the current one on in lo | hi << 32
    // (nbuf >= 2 of them valid); the next dword is loaded from base + pos
    uint32_t lo, hi;
    int nbuf, pos;
    const uint8_t *base;  // the segment's first byte rounded down to 4
    int lenx;             // segment length + (segment start - base)
    uint32_t ctx, nm, nl, zl, sl;

    __device__ __forceinline__ void init(uint32_t lane, int orient)
    {
        const uint32_t t = kMqTable[lane < 47 ? lane : 46];
        const uint32_t ni = (t >> 16) & 63, li = (t >> 22) & 63;
        nm = ((kMqTable[ni] & 0xFFFFu) << 16) | ni;
        nl = ((kMqTable[li] & 0xFFFFu) << 16) | li | (((t >> 28) & 1) << 6);
        const uint32_t k = lane == 0 ? 4u : (lane == CX_RL ? 3u : (lane == CX_UNI ? 46u : 0u));
        ctx = ((kMqTable[k] & 0xFFFFu) << 16) | k;
        zl = build_zc(lane, orient);
        sl = build_sc(lane);
    }
    // One DECODE (C.3.2) in context cx with RENORMD and BYTEIN (C.3.3,
    // C.3.4), in one block of scalar code: the decision chain is wave-
    // uniform, and as straight-line assembly it keeps every register in
    // place (compiled as C++, each rare path made the compiler copy the
    // whole decoder state on the common ones).
    //   e = the context's entry, qe = Qe << 16;
    //   C >= qe (no borrow from C -= qe): A -= qe; A's bit 31 still set
    //   (no signed overflow): MPS, done; else MPS exchange (LPS if A < qe);
    //   C < qe: C restored, A -= qe, LPS exchange (LPS if A >= qe), A = qe;
    //   an exchange makes e the next entry (nl on LPS, with the MPS switch
    //   in bit 6, else nm) and RENORMD shifts by A's leading zeros, reading
    //   a byte whenever CT runs out (two reads cover any shift: after a
    //   byte that follows 0xFF, which holds 7 bits, comes one of 8).
    // The entry is written back after the block (unchanged on the MPS
    // path).  BYTEIN's temporaries are t, qe and l.
    // Lane selects are scalar-ALU results (cx, and e & 63 here), so no
    // vector-ALU-to-lane-select wait states arise; the one scalar load is
    // waited for inside the block.
#define S0_BYTEIN(K)                                              \
    "s_and_b32 %[t], %[lo], 0xff\n\t"                            \
    "s_bfe_u32 %[qe], %[lo], 0x80008\n\t"                         \
    "s_cmp_eq_u32 %[t], 0xff\n\t"                                \
    "s_cbranch_scc0 .Lb8" K "_%=\n\t"                            \
    "s_cmp_gt_u32 %[qe], 0x8f\n\t"                                \
    "s_cbranch_scc0 .Lb9" K "_%=\n\t"                            \
    "s_add_u32 %[C], %[C], 0xff00\n\t"                           \
    "s_mov_b32 %[CT], 8\n\t"                                     \
    "s_branch .Lbe" K "_%=\n"                                    \
    ".Lb9" K "_%=:\n\t"                                          \
    "s_lshl_b32 %[qe], %[qe], 9\n\t"                               \
    "s_add_u32 %[C], %[C], %[qe]\n\t"                             \
    "s_mov_b32 %[CT], 7\n\t"                                     \
    "s_branch .Lba" K "_%=\n"                                    \
    ".Lb8" K "_%=:\n\t"                                          \
    "s_lshl_b32 %[qe], %[qe], 8\n\t"                               \
    "s_add_u32 %[C], %[C], %[qe]\n\t"                             \
    "s_mov_b32 %[CT], 8\n"                                       \
    ".Lba" K "_%=:\n\t"                                          \
    "s_lshr_b32 %[lo], %[lo], 8\n\t"                             \
    "s_lshl_b32 %[t], %[hi], 24\n\t"                             \
    "s_or_b32 %[lo], %[lo], %[t]\n\t"                            \
    "s_lshr_b32 %[hi], %[hi], 8\n\t"                             \
    "s_sub_i32 %[nbuf], %[nbuf], 1\n\t"                          \
    "s_cmp_gt_i32 %[nbuf], 1\n\t"                                \
    "s_cbranch_scc1 .Lbe" K "_%=\n\t"                            \
    "s_load_dword %[t], %[base], %[pos]\n\t"                     \
    "s_sub_i32 %[qe], %[lenx], %[pos]\n\t"                        \
    "s_lshl_b32 %[l], %[qe], 3\n\t"                               \
    "s_lshl_b32 %[l], -1, %[l]\n\t"                              \
    "s_cmp_lt_i32 %[qe], 4\n\t"                                   \
    "s_cselect_b32 %[l], %[l], 0\n\t"                            \
    "s_cmp_le_i32 %[qe], 0\n\t"                                   \
    "s_cselect_b32 %[l], -1, %[l]\n\t"                           \
    "s_waitcnt lgkmcnt(0)\n\t"                                   \
    "s_or_b32 %[t], %[t], %[l]\n\t"                              \
    "s_lshl_b32 %[l], %[t], 8\n\t"                               \
    "s_or_b32 %[lo], %[lo], %[l]\n\t"                            \
    "s_lshr_b32 %[hi], %[t], 24\n\t"                             \
    "s_add_i32 %[nbuf], %[nbuf], 4\n\t"                          \
    "s_add_i32 %[pos], %[pos], 4\n"                              \
    ".Lbe" K "_%=:\n\t"

    // INITDEC (C.3.5) on the block's one codeword segment: the first two
    // dwords from the segment's start rounded down to 4 (inside the data
    // buffer, which is 16-byte aligned), bytes past the segment as 0xFF (the
    // reference appends 0xFF bytes), C = byte 0 << 16, BYTEIN, C <<= 7,
    // CT -= 7, A = 0x8000.  Later loads reach at most 8 bytes past the
    // segment's end.
    __device__ __forceinline__ void start(const uint8_t *sg, int n)
    {
        const uintptr_t a = (uintptr_t)sg;
        const uint32_t sh = (uint32_t)(a & 3);
        base = (const uint8_t *)(a - sh);
        lenx = n + (int)sh;
        uint32_t t, m, l, qe;
        asm volatile(
            "s_load_dword %[lo], %[base], 0x0\n\t"
            "s_load_dword %[hi], %[base], 0x4\n\t"
            "s_lshl_b32 %[l], %[lenx], 3\n\t"
            "s_lshl_b32 %[m], -1, %[l]\n\t"
            "s_cmp_lt_i32 %[lenx], 4\n\t"
            "s_cselect_b32 %[m], %[m], 0\n\t"
            "s_cmp_le_i32 %[lenx], 0\n\t"
            "s_cselect_b32 %[m], -1, %[m]\n\t"
            "s_sub_i32 %[t], %[lenx], 4\n\t"
            "s_lshl_b32 %[l], %[t], 3\n\t"
            "s_lshl_b32 %[l], -1, %[l]\n\t"
            "s_cmp_lt_i32 %[t], 4\n\t"
            "s_cselect_b32 %[l], %[l], 0\n\t"
            "s_cmp_le_i32 %[t], 0\n\t"
            "s_cselect_b32 %[l], -1, %[l]\n\t"
            "s_waitcnt lgkmcnt(0)\n\t"
            "s_or_b32 %[lo], %[lo], %[m]\n\t"
            "s_or_b32 %[hi], %[hi], %[l]\n\t"
            "s_lshl_b32 %[l], %[sh], 3\n\t"
            "s_lshr_b32 %[lo], %[lo], %[l]\n\t"
            "s_sub_u32 %[m], 32, %[l]\n\t"
            "s_lshl_b32 %[t], %[hi], %[m]\n\t"
            "s_cmp_eq_u32 %[sh], 0\n\t"
            "s_cselect_b32 %[t], 0, %[t]\n\t"
            "s_or_b32 %[lo], %[lo], %[t]\n\t"
            "s_lshr_b32 %[hi], %[hi], %[l]\n\t"
            "s_sub_u32 %[nbuf], 8, %[sh]\n\t"
            "s_mov_b32 %[pos], 8\n\t"
            "s_and_b32 %[C], %[lo], 0xff\n\t"
            "s_lshl_b32 %[C], %[C], 16\n\t"
            S0_BYTEIN("0")
            "s_lshl_b32 %[C], %[C], 7\n\t"
            "s_sub_i32 %[CT], %[CT], 7\n\t"
            "s_mov_b32 %[A], 0x80000000"
            : [A] "=&s"(A), [C] "=&s"(C), [CT] "=&s"(CT), [lo] "=&s"(lo), [hi] "=&s"(hi),
              [nbuf] "=&s"(nbuf), [pos] "=&s"(pos), [t] "=&s"(t), [m] "=&s"(m), [l] "=&s"(l),
              [qe] "=&s"(qe)
            : [base] "s"(base), [lenx] "s"(lenx), [sh] "s"(sh)
            : "scc");
    }

    __device__ __forceinline__ uint32_t decode(uint32_t cx)
    {
        uint32_t d, e, qe, t, l;
        int n;
        asm volatile(
            "v_readlane_b32 %[e], %[ctx], %[cx]\n\t"
            "s_and_b32 %[qe], %[e], 0xffff0000\n\t"
            "s_bfe_u32 %[d], %[e], 0x10006\n\t"
            // SCC: borrow, C < qe
            "s_sub_u32 %[C], %[C], %[qe]\n\t"
            "s_cbranch_scc1 .Llps_%=\n\t"
            // SCC: signed overflow, A's bit 31 now clear
            "s_sub_i32 %[A], %[A], %[qe]\n\t"
            "s_cbranch_scc0 .Lend_%=\n\t"
            // MPS exchange: the LPS if A < qe
            "s_and_b32 %[t], %[e], 63\n\t"
            "s_cmp_lt_u32 %[A], %[qe]\n\t"
            "s_cbranch_scc1 .Lxl_%=\n"
            ".Lxm_%=:\n\t"
            "v_readlane_b32 %[l], %[nm], %[t]\n\t"
            "s_and_b32 %[t], %[e], 64\n\t"
            "s_or_b32 %[e], %[l], %[t]\n\t"
            "s_branch .Lrn_%=\n"
            ".Llps_%=:\n\t"
            "s_add_u32 %[C], %[C], %[qe]\n\t"
            "s_sub_u32 %[A], %[A], %[qe]\n\t"
            "s_and_b32 %[t], %[e], 63\n\t"
            // LPS exchange: the LPS if A >= qe
            "s_cmp_ge_u32 %[A], %[qe]\n\t"
            "s_mov_b32 %[A], %[qe]\n\t"
            "s_cbranch_scc0 .Lxm_%=\n"
            ".Lxl_%=:\n\t"
            "v_readlane_b32 %[l], %[nl], %[t]\n\t"
            "s_and_b32 %[t], %[e], 64\n\t"
            "s_xor_b32 %[e], %[l], %[t]\n\t"
            "s_xor_b32 %[d], %[d], 1\n"
            ".Lrn_%=:\n\t"
            "s_flbit_i32_b32 %[n], %[A]\n\t"
            "s_cmp_gt_i32 %[n], %[CT]\n\t"
            "s_cbranch_scc1 .Lrd_%=\n"
            ".Lsh_%=:\n\t"
            "s_lshl_b32 %[A], %[A], %[n]\n\t"
            "s_lshl_b32 %[C], %[C], %[n]\n\t"
            "s_sub_i32 %[CT], %[CT], %[n]\n\t"
            "s_branch .Lend_%=\n"
            ".Lrd_%=:\n\t"
            "s_lshl_b32 %[A], %[A], %[CT]\n\t"
            "s_lshl_b32 %[C], %[C], %[CT]\n\t"
            "s_sub_i32 %[n], %[n], %[CT]\n\t"
            S0_BYTEIN("1")
            "s_cmp_gt_i32 %[n], %[CT]\n\t"
            "s_cbranch_scc0 .Lsh_%=\n\t"
            "s_lshl_b32 %[A], %[A], %[CT]\n\t"
            "s_lshl_b32 %[C], %[C], %[CT]\n\t"
            "s_sub_i32 %[n], %[n], %[CT]\n\t"
            S0_BYTEIN("2")
            "s_branch .Lsh_%=\n"
            ".Lend_%=:"
            : [A] "+s"(A), [C] "+s"(C), [CT] "+s"(CT), [lo] "+s"(lo), [hi] "+s"(hi),
              [nbuf] "+s"(nbuf), [pos] "+s"(pos), [d] "=&s"(d), [e] "=&s"(e), [qe] "=&s"(qe),
              [t] "=&s"(t), [l] "=&s"(l), [n] "=&s"(n)
            : [cx] "s"(cx), [ctx] "v"(ctx), [nm] "v"(nm), [nl] "v"(nl), [base] "s"(base),
              [lenx] "s"(lenx)
            : "scc");
        // the context's entry (unchanged unless an exchange happened),
        // written outside the block: a vector result among the scalar ones
        // would make the compiler take them all for vector values, and
        // v_writelane with a scalar value needs its lane in M0 (one more
        // scalar instruction; the compare and select are vector ones)
        ctx = wrlane(e, cx, ctx);
        return d;
    }
#undef S0_BYTEIN
    __device__ __forceinline__ uint32_t zc(uint32_t win)
    {
        return (rdl(zl, win >> 3) >> ((win << 2) & 31)) & 0xF;
    }
    // sign of a sample whose window of significance is `win` and of signs
    // (shifted down by one) `nwin`
    __device__ __forceinline__ uint32_t sign(uint32_t win, uint32_t nwin)
    {
        const uint32_t i = (win & 0xAA) | (nwin & 0x55);
        const uint32_t e = rdl(sl, i >> 2) >> ((i << 3) & 31);
        return decode(e & 0xF) ^ ((e >> 4) & 1);
    }
};

// every column's 3x3 window from `col`, the column's own rows -1..4 at bits
// 0, 3, ..., 15
__device__ __forceinline__ uint32_t window(uint32_t col)
{
    return (uint32_t)wave_from_left((int)col) | (col << 1) | ((uint32_t)wave_from_right((int)col) << 2);
}
__device__ __forceinline__ uint32_t sig_col(uint32_t up, uint32_t cur, uint32_t dn)
{
    return ((up >> 9) & 1) | ((cur & FS) << 3) | ((dn & 1) << 15);
}
__device__ __forceinline__ uint32_t neg_col(uint32_t up, uint32_t cur, uint32_t dn)
{
    return ((up >> 10) & 1) | (((cur >> 1) & FS) << 3) | (((dn >> 1) & 1) << 15);
}
// rows of a window's column that are insignificant with a significant
// neighbour (bit 3r)
__device__ __forceinline__ uint32_t sig_candidates(uint32_t sw)
{
    const uint32_t n3 = sw | (sw >> 1) | (sw >> 2);
    return (n3 | (n3 >> 3) | (n3 >> 6)) & ~(sw >> 4) & FS;
}

struct Stripe {
    uint32_t v0, v1, v2, v3;  // sample words of rows 0..3 (lane = column)
};

// the sample words of the stripe's new significant samples: sign << 31 |
// 3 << (plane-1)
__device__ __forceinline__ void apply_sig(Stripe &v, uint32_t cur, uint32_t lane, uint64_t n0,
                                          uint64_t n1, uint64_t n2, uint64_t n3, uint32_t one_half)
{
    if ((n0 >> lane) & 1) v.v0 = ((cur << 30) & 0x80000000u) + one_half;
    if ((n1 >> lane) & 1) v.v1 = ((cur << 27) & 0x80000000u) + one_half;
    if ((n2 >> lane) & 1) v.v2 = ((cur << 24) & 0x80000000u) + one_half;
    if ((n3 >> lane) & 1) v.v3 = ((cur << 21) & 0x80000000u) + one_half;
}

// a sample of row R of column x found significant with sign `neg`
#define S0_SIGNIFICANT(R)                                        \
    do {                                                         \
        f |= (1u | (neg << 1)) << (3 * (R));                     \
        sw |= 1u << (3 * (R) + 4);                               \
        nw |= neg << (3 * (R) + 4);                              \
        cs |= 1u << (3 * (R) + 3);                               \
        cn |= neg << (3 * (R) + 3);                              \
        ns##R |= 1ull << x;                                      \
    } while (0)

// significance propagation pass over one stripe
__device__ __forceinline__ void sigprop(Mq &m, uint32_t &cur, uint32_t up, uint32_t dn, Stripe &v,
                                        uint32_t rm, int w, uint64_t colmask, uint32_t lane,
                                        uint32_t one_half)
{
    uint32_t SW = window(sig_col(up, cur, dn));
    uint32_t CD = sig_candidates(SW) & rm;
    uint64_t todo = __builtin_amdgcn_ballot_w64(CD != 0) & colmask;
    if (!todo) return;
    uint32_t NW = window(neg_col(up, cur, dn));
    uint64_t ns0 = 0, ns1 = 0, ns2 = 0, ns3 = 0;
    do {
        const uint32_t x = (uint32_t)__builtin_ctzll(todo);
        todo &= ~(1ull << x);
        uint32_t sw = rdl(SW, x), nw = rdl(NW, x), f = rdl(cur, x), c = rdl(CD, x);
        uint32_t cs = 0, cn = 0;
#define S0_SP_ROW(R)                                                             \
        if (c & (1u << (3 * (R)))) {                                             \
            const uint32_t win = (sw >> (3 * (R))) & 0x1FF;                      \
            if (m.decode(m.zc(win))) {                                           \
                const uint32_t neg = m.sign(win, (nw >> (3 * (R) + 1)) & 0xFF);  \
                S0_SIGNIFICANT(R);                                               \
                if ((R) < 3) c |= (1u << (3 * (R) + 3)) & ~(sw >> 4) & rm;       \
            }                                                                    \
        }
        S0_SP_ROW(0)
        S0_SP_ROW(1)
        S0_SP_ROW(2)
        S0_SP_ROW(3)
#undef S0_SP_ROW
        // every candidate row was coded: "visited"
        cur = wrlane(f | (c << 2), x, cur);
        if (cs && (int)x + 1 < w) {
            // column x+1 sees the new significance to its west; its rows
            // next to it may become candidates (recomputed by the lanes:
            // the columns up to x are done, the others unchanged)
            const uint32_t x1 = x + 1;
            SW = wrlane(rdl(SW, x1) | cs, x1, SW);
            NW = wrlane(rdl(NW, x1) | cn, x1, NW);
            CD = sig_candidates(SW) & rm;
            todo |= 1ull << x1;
        }
    } while (todo);
    apply_sig(v, cur, lane, ns0, ns1, ns2, ns3, one_half);
}

// magnitude refinement pass over one stripe.  The stripe's significance
// does not change in this pass, so each refinable sample's context is
// known at the start: the lanes compute it (16 after the first
// refinement, else 15 with a significant neighbour, 14 without; 0 where
// the sample is not refinable) for the four rows, five bits each.
__device__ __forceinline__ void refine(Mq &m, uint32_t cur, uint32_t up, uint32_t dn, Stripe &v,
                                       uint64_t colmask, uint32_t lane, int plane, uint32_t half)
{
    const uint32_t mu = (cur >> 12) & 0xF;
    uint64_t todo = __builtin_amdgcn_ballot_w64(mu != 0) & colmask;
    if (!todo) return;
    const uint32_t SW = window(sig_col(up, cur, dn));
    uint32_t MC = 0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const uint32_t c = ((cur >> (16 + r)) & 1) ? CX_MR + 2u
                                                   : (((SW >> (3 * r)) & 0x1EF) ? CX_MR + 1u : CX_MR + 0u);
        MC |= (((mu >> r) & 1) ? c : 0u) << (5 * r);
    }
    uint64_t z0 = 0, z1 = 0, z2 = 0, z3 = 0;
    do {
        const uint32_t x = (uint32_t)__builtin_ctzll(todo);
        todo &= ~(1ull << x);
        const uint32_t mc = rdl(MC, x);
#define S0_MR_ROW(R)                                                  \
        if (const uint32_t cx = (mc >> (5 * (R))) & 31) {             \
            if (!m.decode(cx)) z##R |= 1ull << x;                     \
        }
        S0_MR_ROW(0)
        S0_MR_ROW(1)
        S0_MR_ROW(2)
        S0_MR_ROW(3)
#undef S0_MR_ROW
    } while (todo);
    // bit `plane` cleared where the refinement bit was 0, bit plane-1 set
    const uint32_t pb = 1u << plane;
    if (mu & 1) v.v0 = (v.v0 ^ (((z0 >> lane) & 1) ? pb : 0u)) | half;
    if (mu & 2) v.v1 = (v.v1 ^ (((z1 >> lane) & 1) ? pb : 0u)) | half;
    if (mu & 4) v.v2 = (v.v2 ^ (((z2 >> lane) & 1) ? pb : 0u)) | half;
    if (mu & 8) v.v3 = (v.v3 ^ (((z3 >> lane) & 1) ? pb : 0u)) | half;
}

// cleanup pass over one stripe, then the end-of-plane flag update
__device__ __forceinline__ void cleanup(Mq &m, uint32_t &cur, uint32_t up, uint32_t dn, Stripe &v,
                                        uint32_t rm, int w, uint64_t colmask,
                                        uint32_t lane, uint32_t one_half)
{
    // rows not significant and not visited
    const uint32_t CD = ~(cur | (cur >> 2)) & rm;
    uint64_t todo = __builtin_amdgcn_ballot_w64(CD != 0) & colmask;
    if (todo) {
        uint32_t SW = window(sig_col(up, cur, dn));
        uint32_t NW = window(neg_col(up, cur, dn));
        uint64_t ns0 = 0, ns1 = 0, ns2 = 0, ns3 = 0;
        do {
            const uint32_t x = (uint32_t)__builtin_ctzll(todo);
            todo &= ~(1ull << x);
            uint32_t sw = rdl(SW, x), nw = rdl(NW, x), f = rdl(cur, x), c = rdl(CD, x);
            uint32_t cs = 0, cn = 0;
            if ((sw | (c ^ FS)) == 0) {
                // run mode: four codable samples without significant
                // neighbours (c == FS needs all four rows, so never in the
                // short last stripe)
                if (!m.decode(CX_RL)) continue;
                uint32_t k = m.decode(CX_UNI) << 1;
                k |= m.decode(CX_UNI);
                // its neighbours are all insignificant: context 9, no XOR
                const uint32_t neg = m.decode(CX_SC);
                switch (k) {
                case 0: S0_SIGNIFICANT(0); break;
                case 1: S0_SIGNIFICANT(1); break;
                case 2: S0_SIGNIFICANT(2); break;
                default: S0_SIGNIFICANT(3); break;
                }
                c &= ~0u << (3 * k + 3);
            }
#define S0_CU_ROW(R)                                                               \
            if (c & (1u << (3 * (R)))) {                                           \
                const uint32_t win = (sw >> (3 * (R))) & 0x1FF;                    \
                if (m.decode(m.zc(win))) {                                         \
                    const uint32_t neg = m.sign(win, (nw >> (3 * (R) + 1)) & 0xFF); \
                    S0_SIGNIFICANT(R);                                             \
                }                                                                  \
            }
            S0_CU_ROW(0)
            S0_CU_ROW(1)
            S0_CU_ROW(2)
            S0_CU_ROW(3)
#undef S0_CU_ROW
            if (cs) {
                cur = wrlane(f, x, cur);
                if ((int)x + 1 < w) {
                    SW = wrlane(rdl(SW, x + 1) | cs, x + 1, SW);
                    NW = wrlane(rdl(NW, x + 1) | cn, x + 1, NW);
                }
            }
        } while (todo);
        apply_sig(v, cur, lane, ns0, ns1, ns2, ns3, one_half);
    }
    // end of the bit-plane: "visited" cleared, significant samples become
    // refinable, refinable ones become "refined before"
    const uint32_t sig4 = (cur & 1) | ((cur >> 2) & 2) | ((cur >> 4) & 4) | ((cur >> 6) & 8);
    cur = (cur & FSN) | (sig4 << 12) | (((cur >> 12) & 0xF) << 16);
}
#undef S0_SIGNIFICANT

}  // namespace s0

__global__ void __launch_bounds__(64 * kWavesPerGroup)
    t1_style0_kernel(const BmiT1Block *__restrict__ blocks, int nblocks, const uint8_t *__restrict__ data)
{
    __shared__ s0::Lds lds_all[kWavesPerGroup];
    const uint32_t lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int bi = (int)uni(blockIdx.x * kWavesPerGroup + wave);
    if (bi >= nblocks) return;
    s0::Lds &L = lds_all[wave];
    const BmiT1Block *B = blocks + bi;
    const int w = (int)uni(B->width), h = (int)uni(B->height);
    if (w < 1 || h < 1 || w > kMaxSide || h > kMaxSide) return;
    if (((int)uni(B->style) & BMI_T1_STYLES & ~BMI_T1_PTERM) != 0) return;  // the styled kernel's
    const int orient = (int)uni(B->orient);
    const int numbps = (int)uni(B->numbps);
    const int missing = (int)uni(B->missing);
    const int num_pass = (int)uni(B->num_pass);
    const int kind = (int)uni(B->kind);
    const int len = (int)uni(B->len);
    const uint64_t off = ((uint64_t)uni((uint32_t)(B->data_offset >> 32)) << 32) |
                         uni((uint32_t)B->data_offset);
    const uintptr_t outp = ((uintptr_t)uni((uint32_t)((uintptr_t)B->out >> 32)) << 32) |
                           uni((uint32_t)(uintptr_t)B->out);
    const int out_pitch = (int)uni(B->out_pitch);
    const int nstripes = (h + 3) >> 2;
    const uint64_t colmask = (w == 64) ? ~0ull : ((1ull << w) - 1);

    auto *o = gptr_w<uint32_t>((void *)outp);
    const bool col_ok = (int)lane < w;
    for (int s = 0; s < nstripes; ++s) L.fl[s][lane] = 0;
    if (col_ok)
        for (int y = 0; y < h; ++y) o[(size_t)y * out_pitch + lane] = 0;
    auto vload = [&](int y) -> uint32_t { return (col_ok && y < h) ? o[(size_t)y * out_pitch + lane] : 0u; };
    auto vstore = [&](int y, uint32_t v) {
        if (col_ok && y < h) o[(size_t)y * out_pitch + lane] = v;
    };

    // decode_code_block :610-649 (as t1_styled_kernel)
    int num_passes = 0;
    int plane = 0;
    if ((numbps - missing) * 3 - 2 > 0) {
        const int p_max = 30 - missing;
        num_passes = 3 * p_max - 2;
        if (num_passes > num_pass) num_passes = num_pass;
        plane = (numbps >> 11) - 1 - missing;
        if (plane < 0) plane = p_max;
        if (plane == 0) plane = 1;
    }
    // passes below bit-plane 0 are not decoded (a negative shift count,
    // undefined in the reference): the first cleanup pass is at `plane`,
    // so at most 3 * plane + 1 passes
    num_passes = min(num_passes, 3 * plane + 1);
    if (num_passes > 0) {
        s0::Mq m;
        m.init(lane, orient);
        m.start(data + off, len);
        // one pass (TYPE 0 significance propagation, 1 refinement, 2
        // cleanup) over the stripes at bit-plane `plane`
        auto pass = [&](auto type_tag, int plane) {
            constexpr int TYPE = decltype(type_tag)::value;
            const uint32_t one_half = (plane < 31) ? (1u << plane) + ((1u << plane) >> 1) : 0u;
            const uint32_t half = (plane >= 1 && plane < 32) ? (1u << plane) >> 1 : 0u;
            uint32_t up = 0, cur = L.fl[0][lane];
            s0::Stripe v{vload(0), vload(1), vload(2), vload(3)};
            for (int s = 0; s < nstripes; ++s) {
                const int rows = min(4, h - 4 * s);
                const uint32_t rm = rows == 4 ? s0::FS : (s0::FS & ((1u << (3 * rows)) - 1));
                const uint32_t dn = (s + 1 < nstripes) ? L.fl[s + 1][lane] : 0u;
                s0::Stripe nx{0, 0, 0, 0};
                if (s + 1 < nstripes) {
                    const int y = 4 * (s + 1);
                    nx = s0::Stripe{vload(y), vload(y + 1), vload(y + 2), vload(y + 3)};
                }
                if constexpr (TYPE == 0) s0::sigprop(m, cur, up, dn, v, rm, w, colmask, lane, one_half);
                if constexpr (TYPE == 1) s0::refine(m, cur, up, dn, v, colmask, lane, plane, half);
                if constexpr (TYPE == 2) s0::cleanup(m, cur, up, dn, v, rm, w, colmask, lane, one_half);
                if constexpr (TYPE != 1) L.fl[s][lane] = cur;
                vstore(4 * s, v.v0);
                vstore(4 * s + 1, v.v1);
                vstore(4 * s + 2, v.v2);
                vstore(4 * s + 3, v.v3);
                up = cur;
                cur = dn;
                v = nx;
            }
        };
        // a cleanup pass at the start plane, then significance propagation,
        // refinement and cleanup a plane lower each time (written without
        // a three-way branch per pass, which the compiler handles badly
        // around the scalar decoder state)
        int rest = num_passes;
        for (int it = 0; rest > 0; ++it) {
            if (it > 0) {
                --plane;
                pass(std::integral_constant<int, 0>{}, plane);
                if (--rest > 0) {
                    pass(std::integral_constant<int, 1>{}, plane);
                    --rest;
                }
            }
            if (rest > 0) {
                pass(std::integral_constant<int, 2>{}, plane);
                --rest;
            }
        }
    }
    // 5/3: convert the words in place (dec_cb_stripe, dec_t1_comp.cpp:343-367)
    if (kind == BMI_T1_INT && col_ok) {
        const int sh = 31 - numbps;
        for (int y = 0; y < h; ++y) {
            const uint32_t v = o[(size_t)y * out_pitch + lane];
            const int32_t r = (v & 0x80000000u) ? -(int32_t)((v & 0x7FFFFFFFu) >> sh) : (int32_t)(v >> sh);
            o[(size_t)y * out_pitch + lane] = (uint32_t)r;
        }
    }
}

}  // namespace
}  // namespace bmi

using namespace bmi;

extern "C" int BMI_t1_check_blocks(const BmiT1Block *blocks, int nblocks, int *bad)
{
    if (bad) *bad = -1;
    if (!blocks || nblocks < 0) return bmi_fail(BMI_E_INVALID, "t1_check_blocks: bad args");
    for (int i = 0; i < nblocks; ++i) {
        const BmiT1Block &b = blocks[i];
        int rc = BMI_OK;
        if (b.width < 1 || b.height < 1 || b.len < 0 || b.num_pass < 0 || b.missing < 0 ||
            b.orient < 0 || b.orient > 3 || (b.kind != BMI_T1_WORDS && b.kind != BMI_T1_INT) ||
            !b.out || b.out_pitch < b.width || (b.style & ~BMI_T1_STYLES) ||
            (bmi_style_segmented(b.style) &&
             (b.nseg != BMI_t1_num_segments(b.num_pass, b.style) || (b.seg_offset & 3))))
            rc = BMI_E_INVALID;
        else if (b.width > kMaxSide || b.height > kMaxSide)
            rc = BMI_E_UNSUPPORTED;
        else if (b.kind == BMI_T1_INT && (b.numbps < 1 || b.numbps > 31))
            rc = BMI_E_INVALID;
        if (rc) {
            if (bad) *bad = i;
            return bmi_fail(rc, "t1_check_blocks: block %d rejected", i);
        }
    }
    return BMI_OK;
}

extern "C" int BMI_t1_order_blocks(BmiT1Block *blocks, int nblocks)
{
    if (nblocks < 0 || (nblocks > 0 && !blocks)) return bmi_fail(BMI_E_INVALID, "t1_order_blocks: bad args");
    std::stable_sort(blocks, blocks + nblocks,
                     [](const BmiT1Block &a, const BmiT1Block &b) { return a.len > b.len; });
    return BMI_OK;
}

extern "C" int BMI_t1_num_segments(int num_pass, int style)
{
    int n = 0;
    for (int i = 0; i < num_pass; ++i) n += bmi_pass_new_segment(i, style) ? 1 : 0;
    return n;
}

extern "C" int BMI_t1_decode_blocks(const BmiT1Block *blocks, int nblocks, const void *data,
                                    StreamId stream)
{
    if (!blocks || !data || nblocks < 0 || ((uintptr_t)data & 15))
        return bmi_fail(BMI_E_INVALID, "t1_decode_blocks: bad args (data must be 16-byte aligned)");
    if (nblocks == 0) return BMI_OK;
    const dim3 grid((nblocks + kWavesPerGroup - 1) / kWavesPerGroup), block(64 * kWavesPerGroup);
    // the check below must see this launch, not an error left by an
    // earlier, unrelated runtime call of the caller
    (void)hipGetLastError();
    hipLaunchKernelGGL(t1_style0_kernel, grid, block, 0, (hipStream_t)stream, blocks, nblocks,
                       (const uint8_t *)data);
    hipLaunchKernelGGL(t1_styled_kernel, grid, block, 0, (hipStream_t)stream, blocks, nblocks,
                       (const uint8_t *)data);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return bmi_fail(BMI_E_HIP, "t1 launch: %s", hipGetErrorString(e));
    return BMI_OK;
}
