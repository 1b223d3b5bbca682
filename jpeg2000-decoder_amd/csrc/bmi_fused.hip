// bmi_fused.hip -- fused multi-level IDWT + colour transform + pack for
// gfx950, batched over any number of tiles in one persistent launch.
//
// Replaces the reference's per-level row/column launches
// (decoder/cudainterface/src/gpufunc.cu:419-546, 953-1089) and its separate
// colour kernels (gpufunc.cu:1093-1401) with one engine:
//
//  * A work item is one wave64 sweeping a vertical strip of one synthesis
//    level: lane j owns the column pair (2j, 2j+1) of a 128-column window
//    (60 output pairs for 9/7, 62 for 5/3; the rest is symmetric-extension
//    or neighbour halo).  Each input row pair (low row from LL|HL, high row
//    from LH|HH) is loaded coalesced, lifted horizontally with wave
//    shuffles, and pushed into a per-lane register pipeline that does the
//    vertical lifting, so a level makes one pass over HBM and uses no LDS.
//  * The final level of a tile runs all components in lockstep and applies
//    the inverse RCT/ICT or the merge level shift before packing pixels, so
//    the last resolution never goes to HBM as float planes.
//  * Items of all tiles and levels sit in one device work queue; waves take
//    items in order and wait (bounded spin, acquire) only on the per-job
//    completion counters of the level below in the same tile.
//
// Arithmetic follows the reference CPU back end: gains folded into the row
// pass as in bmi_cpu_dwt_decoder.cpp:1433-1473, lifting updates
// x - c*(l + r) rounded per operation, ICT in float
// (bmi_cpu_dwt_decoder.cpp:3570-3841), RCT with floor (:3936-4130), merge
// with truncation before the level shift (:4282-4596).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <cstdio>
#include <cstring>
#include <type_traits>
#include <vector>

#include "bmi_device.h"
#include "bmi_fused.h"
#include "bmi_internal.h"

namespace bmi {

template <int FILT>
struct SampleT;
template <>
struct SampleT<97> {
    using T = float;
};
template <>
struct SampleT<53> {
    using T = int;
};

template <int FILT>
__device__ inline typename SampleT<FILT>::T load_sample(const void *base, size_t idx, int kind,
                                                        float gain)
{
    if constexpr (FILT == 97) {
        uint32_t w = gptr<uint32_t>(base)[idx];
        float v = (kind == SRC_SIGNMAG) ? (float)signmag_to_int(w) : __uint_as_float(w);
        return __fmul_rn(v, gain);
    } else {
        return gptr<int>(base)[idx];
    }
}

// one horizontal synthesis across the wave: lane holds (low, high) of its
// column pair; T.800 lifting with the reference's operation order
template <int FILT, typename T>
__device__ inline void hlift(T &lo, T &hi)
{
    if constexpr (FILT == 97) {
        float hl = wave_from_left(hi);
        lo = lift97(lo, kDelta, hl, hi);
        float lr = wave_from_right(lo);
        hi = lift97(hi, kGamma, lo, lr);
        hl = wave_from_left(hi);
        lo = lift97(lo, kBeta, hl, hi);
        lr = wave_from_right(lo);
        hi = lift97(hi, kAlpha, lo, lr);
    } else {
        int hl = wave_from_left(hi);
        lo = lo - ((hl + hi + 2) >> 2);
        int lr = wave_from_right(lo);
        hi = hi + ((lo + lr) >> 1);
    }
}

// single-sample line: T.800 F.3.7 (X = Y, or Y/2 for an odd position)
template <int FILT, typename T>
__device__ inline T half_sample(T v)
{
    if constexpr (FILT == 97)
        return __fmul_rn(v, 0.5f);
    else
        return v >> 1;
}

// Output kinds of a work item (kernel template parameter)
constexpr int K_PLANE = 0;  // float / int32 plane (intermediate or final)
constexpr int K_RGB = 1;    // inverse RCT / ICT + pack
constexpr int K_MERGE = 2;  // level shift + pack of 1..4 components

// Pixel layouts (element order B,G,R[,A], codec_base.h:51-143)
constexpr int PL_BGR8 = 0, PL_BGRA8 = 1, PL_BGR16 = 2, PL_BGRA16 = 3, PL_ANY = 4;

__host__ __device__ inline int pixel_layout(int fmt)
{
    switch (fmt) {
    case IMAGE_FOTMAT_24BITS_RGB:
    case IMAGE_FOTMAT_24BITS_3COMPONENTS: return PL_BGR8;
    case IMAGE_FOTMAT_32BITS_ARGB:
    case IMAGE_FOTMAT_32BITS_4COMPONENTS: return PL_BGRA8;
    case IMAGE_FOTMAT_48BITS_RGB:
    case IMAGE_FOTMAT_48BITS_3COMPONENTS: return PL_BGR16;
    default: return PL_BGRA16;
    }
}

// Output parameters of one item, read from the descriptor once.  SIMPLE:
// the caller has checked that the tile's columns start and end on an even
// column (no lane holds just one pixel of its pair inside the tile) and
// that a packed output's pitch is even (every row starts on an even byte),
// so the stores need neither case.
template <int FILT, int NC, int OUTK, int PL, bool SIMPLE = false>
struct Emitter {
    using T = typename SampleT<FILT>::T;
    __attribute__((address_space(1))) T *dst[NC];
    __attribute__((address_space(1))) uint8_t *out;
    int pitch, a0, b0, ox, oy, bd, od, ls, rs, mx, check_out, fmt, nc;
    float fmx, foff;
    bool pair_fast;      // both pixels of a lane in one aligned store sequence
    bool uniform_pairs;  // no lane of the item has just one of its pixels in the tile

    __device__ inline void init(const LevelJob &__restrict__ J)
    {
        uniform_pairs = false;
        a0 = J.a0;
        b0 = J.b0;
        nc = J.nc;
        if constexpr (OUTK == K_PLANE) {
#pragma unroll
            for (int c = 0; c < NC; ++c) dst[c] = gptr_w<T>(J.dst[c]);
            pitch = J.dst_pitch;
            bool al = true;
#pragma unroll
            for (int c = 0; c < NC; ++c) al = al && !((uintptr_t)J.dst[c] & 7);
            pair_fast = al && !(a0 & 1) && !(pitch & 1);
        } else {
            out = gptr_w<uint8_t>(J.out);
            pitch = J.out_pitch_px;
            ox = J.out_x - J.a0;
            oy = J.out_y - J.b0;
            bd = J.bit_depth;
            fmt = J.fmt;
            od = fmt_depth(fmt);
            ls = od > bd ? od - bd : 0;
            rs = bd > od ? bd - od : 0;
            mx = (1 << bd) - 1;
            fmx = (float)mx;
            foff = (float)(1 << (bd - 1));
            check_out = J.check_out;
            // 24-bit layout (fixed by the variant, or the merge's format)
            const bool bgr8 = (PL == PL_BGR8) || (PL == PL_ANY && pixel_layout(fmt) == PL_BGR8);
            pair_fast = bgr8 && !((uintptr_t)J.out & 3) && !(ox & 1);
            // 48-bit layout of a specialised variant: a lane's two pixels
            // are 12 bytes at an offset that is a multiple of 4 on rows
            // whose start is
            if constexpr (PL == PL_BGR16) pair_fast = !((uintptr_t)J.out & 3) && !(ox & 1);
        }
    }

    template <bool SHIFT = true>
    __device__ inline int lift(int v) const
    {
        if constexpr (SHIFT) return (v << ls) >> rs;
        else return v;
    }

    // the lanes' output masks of an item: whether any lane stores just one
    // of its two pixels (at an odd edge of the tile), tested once per store
    // with a uniform branch
    __device__ inline void set_lanes(bool e_ok, bool o_ok)
    {
        if constexpr (SIMPLE) uniform_pairs = true;
        else uniform_pairs = __builtin_amdgcn_ballot_w64(e_ok != o_ok) == 0;
    }

    // put2_full applies (rows inside the level, paired layout, no element
    // shift, no grey check-out of a colour transform; a merge keeps its
    // grey slot choice): the common case of an item, decided once.  Lanes
    // that hold one pixel of their pair inside the tile (at an odd edge)
    // stay on the same path and narrow their stores.
    __device__ inline bool full_ok() const
    {
        if constexpr (OUTK == K_PLANE) return pair_fast;
        else if constexpr (PL == PL_BGR8 || PL == PL_ANY)
            return pair_fast && ls == 0 && rs == 0 && (OUTK == K_MERGE || check_out < 0);
        else if constexpr (PL == PL_BGR16)
            return pair_fast && (OUTK == K_MERGE || check_out < 0);
        else return false;
    }

    // both samples of a lane to the planes, or the one inside the tile
    __device__ inline void store_plane_pair(int row, int ce, const T *ve, const T *vo, bool e_ok,
                                            bool o_ok) const
    {
        const uint32_t off = ((uint32_t)(row - b0) * (uint32_t)pitch + (uint32_t)(ce - a0)) * 4u;
        if (!SIMPLE && !uniform_pairs) {
            if (e_ok && !o_ok) {
#pragma unroll
                for (int c = 0; c < NC; ++c) *at(dst[c], off) = ve[c];
            } else if (o_ok && !e_ok) {
#pragma unroll
                for (int c = 0; c < NC; ++c) *at(dst[c], off + 4) = vo[c];
            }
        }
        if (e_ok && o_ok) {
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                typedef T v2 __attribute__((ext_vector_type(2)));
                v2 pairv;
                pairv.x = ve[c];
                pairv.y = vo[c];
                *(__attribute__((address_space(1))) v2 *)at(dst[c], off) = pairv;
            }
        }
    }

    // put2 for items where full_ok() holds: no branch but the store masks
    __device__ inline void put2_full(int row, int ce, const T *ve, const T *vo, bool e_ok, bool o_ok) const
    {
        if constexpr (OUTK == K_PLANE) {
            store_plane_pair(row, ce, ve, vo, e_ok, o_ok);
        } else if constexpr (FILT == 97 && OUTK == K_RGB && NC >= 3 && (PL == PL_BGR8 || PL == PL_ANY)) {
            // full_ok implies the 24-bit BGR layout (pair_fast) and 8-bit
            // samples with no shift: a channel is
            // (int)clamp(x, 0, 255), i.e. the clamped truncation, which
            // v_cvt_pk_u8_f32 of trunc(x) computes and packs in one step
            float r0, g0, b0, r1, g1, b1;
            ict97(ve, r0, g0, b0);
            ict97(vo, r1, g1, b1);
            auto u8 = [](float x, uint32_t k, uint32_t acc) {
                return __builtin_amdgcn_cvt_pk_u8_f32(__builtin_truncf(x), k, acc);
            };
            // w0 = B0 G0 R0 B1, w1 = G1 R1 (bytes, lowest first)
            const uint32_t w0 = u8(b1, 3, u8(r0, 2, u8(g0, 1, u8(b0, 0, 0u))));
            const uint32_t w1 = u8(r1, 1, u8(g1, 0, 0u));
            store_pair_words(row, ce, w0, w1, e_ok, o_ok);
        } else if constexpr (PL == PL_BGR8 || PL == PL_ANY) {
            int R0, G0, B0, A0, R1, G1, B1, A1;
            colour<false, false>(ve, R0, G0, B0, A0);
            colour<false, false>(vo, R1, G1, B1, A1);
            store_pair_bgr8(row, ce, R0, G0, B0, R1, G1, B1, e_ok, o_ok);
        } else if constexpr (PL == PL_BGR16) {
            int R0, G0, B0, A0, R1, G1, B1, A1;
            colour<true, false>(ve, R0, G0, B0, A0);
            colour<true, false>(vo, R1, G1, B1, A1);
            store_pair_bgr16(row, ce, R0, G0, B0, R1, G1, B1, e_ok, o_ok);
        }
    }

    // B0 G0 R0 B1 G1 R1 of an even output column as 16-bit elements (12
    // bytes).  With an even row start the byte offset is a multiple of 4:
    // three aligned dwords; otherwise six aligned shorts.  A lane with one
    // pixel inside the tile writes that pixel's six bytes only (a dword and
    // a short, or a short and a dword).
    __device__ inline void store_pair_bgr16(int row, int ce, int R0, int G0, int B0, int R1, int G1, int B1,
                                            bool e_ok, bool o_ok) const
    {
        typedef __attribute__((address_space(1))) uint32_t gu32;
        typedef __attribute__((address_space(1))) uint16_t gu16;
        const uint32_t off = ((uint32_t)(oy + row) * (uint32_t)pitch + (uint32_t)(ox + ce)) * 6u;
        const uint32_t w0 = (uint32_t)(B0 & 0xffff) | ((uint32_t)G0 << 16);
        const uint32_t w1 = (uint32_t)(R0 & 0xffff) | ((uint32_t)B1 << 16);
        const uint32_t w2 = (uint32_t)(G1 & 0xffff) | ((uint32_t)R1 << 16);
        const bool odd_row = !SIMPLE && (((uint32_t)(oy + row) * (uint32_t)pitch) & 1u) != 0;
        if (!SIMPLE && !uniform_pairs) {
            if (e_ok && !o_ok) {
                if (odd_row) {
                    *(gu16 *)at(out, off) = (uint16_t)w0;
                    *(gu16 *)at(out, off + 2) = (uint16_t)(w0 >> 16);
                } else {
                    *(gu32 *)at(out, off) = w0;
                }
                *(gu16 *)at(out, off + 4) = (uint16_t)w1;
            } else if (o_ok && !e_ok) {
                *(gu16 *)at(out, off + 6) = (uint16_t)(w1 >> 16);
                if (odd_row) {
                    *(gu16 *)at(out, off + 8) = (uint16_t)w2;
                    *(gu16 *)at(out, off + 10) = (uint16_t)(w2 >> 16);
                } else {
                    *(gu32 *)at(out, off + 8) = w2;
                }
            }
        }
        if (e_ok && o_ok) {
            if (odd_row) {
                *(gu16 *)at(out, off) = (uint16_t)w0;
                *(gu16 *)at(out, off + 2) = (uint16_t)(w0 >> 16);
                *(gu16 *)at(out, off + 4) = (uint16_t)w1;
                *(gu16 *)at(out, off + 6) = (uint16_t)(w1 >> 16);
                *(gu16 *)at(out, off + 8) = (uint16_t)w2;
                *(gu16 *)at(out, off + 10) = (uint16_t)(w2 >> 16);
            } else {
                *(gu32 *)at(out, off) = w0;
                *(gu32 *)at(out, off + 4) = w1;
                *(gu32 *)at(out, off + 8) = w2;
            }
        }
    }

    // B0 G0 R0 B1 G1 R1 of an even output column, packed as w0 = B0 G0 R0
    // B1, w1 = G1 R1 (lowest byte first).  With an even row start (an even
    // pitch, or an even row of an odd one) the byte offset is even: one
    // aligned dword and one aligned short; otherwise three shorts at odd
    // offsets.  The parity is the row's (ox + ce is even), so uniform.  A
    // lane with one pixel inside the tile (its odd pixel past the right
    // edge, or its even one before an odd left edge) writes that pixel's
    // three bytes only, as a short and a byte.
    __device__ inline void store_pair_words(int row, int ce, uint32_t w0, uint32_t w1, bool e_ok,
                                            bool o_ok) const
    {
        const uint32_t off = ((uint32_t)(oy + row) * (uint32_t)pitch + (uint32_t)(ox + ce)) * 3u;
        if (!SIMPLE && !uniform_pairs) {
            if (e_ok && !o_ok) {
                *(__attribute__((address_space(1))) uint16_t *)at(out, off) = (uint16_t)w0;
                *(__attribute__((address_space(1))) uint8_t *)at(out, off + 2) = (uint8_t)(w0 >> 16);
            } else if (o_ok && !e_ok) {
                *(__attribute__((address_space(1))) uint8_t *)at(out, off + 3) = (uint8_t)(w0 >> 24);
                *(__attribute__((address_space(1))) uint16_t *)at(out, off + 4) = (uint16_t)w1;
            }
        }
        const bool both = e_ok && o_ok;
        if (!SIMPLE && (((uint32_t)(oy + row) * (uint32_t)pitch) & 1u)) {
            if (both) {
                *(__attribute__((address_space(1))) uint16_t *)at(out, off) = (uint16_t)w0;
                *(__attribute__((address_space(1))) uint16_t *)at(out, off + 2) = (uint16_t)(w0 >> 16);
                *(__attribute__((address_space(1))) uint16_t *)at(out, off + 4) = (uint16_t)w1;
            }
            return;
        }
        const bool mid = (off & 2) != 0;
        const uint32_t hi4 = __builtin_amdgcn_alignbit(w1, w0, 16);  // R0 B1 G1 R1
        const uint32_t dw_off = mid ? off + 2 : off;
        const uint32_t sh_off = mid ? off : off + 4;
        if (both) {
            *(__attribute__((address_space(1))) uint32_t *)at(out, dw_off) = mid ? hi4 : w0;
            *(__attribute__((address_space(1))) uint16_t *)at(out, sh_off) = (uint16_t)(mid ? w0 : w1);
        }
    }

    // RICT_97_FLOAT of one pixel before the clamp (bmi_cpu_dwt_decoder.cpp:3570-3841)
    __device__ inline void ict97(const T *v, float &r, float &g, float &b) const
    {
        r = __fadd_rn(__fadd_rn(v[0], __fmul_rn(kCrR, v[2])), foff);
        g = __fadd_rn(__fadd_rn(__fadd_rn(v[0], __fmul_rn(kCbG, v[1])), __fmul_rn(kCrG, v[2])), foff);
        b = __fadd_rn(__fadd_rn(v[0], __fmul_rn(kCbB, v[1])), foff);
    }

    // the six bytes of store_pair_words from unpacked channels
    __device__ inline void store_pair_bgr8(int row, int ce, int R0, int G0, int B0, int R1, int G1, int B1,
                                           bool e_ok, bool o_ok) const
    {
        const uint32_t w0 = (uint32_t)(B0 & 0xff) | ((uint32_t)(G0 & 0xff) << 8) |
                            ((uint32_t)(R0 & 0xff) << 16) | ((uint32_t)(B1 & 0xff) << 24);
        const uint32_t w1 = (uint32_t)(G1 & 0xff) | ((uint32_t)(R1 & 0xff) << 8);
        store_pair_words(row, ce, w0, w1, e_ok, o_ok);
    }

    // colour of one pixel (before packing); SHIFT false when the output
    // element depth equals the bit depth (ls = rs = 0)
    template <bool SHIFT = true, bool CHECK = true>
    __device__ inline void colour(const T *v, int &R, int &G, int &B, int &A) const
    {
        A = 0;
        if constexpr (OUTK == K_RGB) {
            if constexpr (NC >= 3) {
                if constexpr (FILT == 97) {
                    // RICT_97_FLOAT, bmi_cpu_dwt_decoder.cpp:3570-3841
                    const float r = __fadd_rn(__fadd_rn(v[0], __fmul_rn(kCrR, v[2])), foff);
                    const float g = __fadd_rn(
                        __fadd_rn(__fadd_rn(v[0], __fmul_rn(kCbG, v[1])), __fmul_rn(kCrG, v[2])), foff);
                    const float b = __fadd_rn(__fadd_rn(v[0], __fmul_rn(kCbB, v[1])), foff);
                    R = lift<SHIFT>((int)__builtin_amdgcn_fmed3f(r, 0.0f, fmx));
                    G = lift<SHIFT>((int)__builtin_amdgcn_fmed3f(g, 0.0f, fmx));
                    B = lift<SHIFT>((int)__builtin_amdgcn_fmed3f(b, 0.0f, fmx));
                } else {
                    // RRct_53, bmi_cpu_dwt_decoder.cpp:3936-4130
                    const int g = v[0] - ((v[1] + v[2]) >> 2) + (1 << (bd - 1));
                    R = lift<SHIFT>(min(max(v[2] + g, 0), mx));
                    G = lift<SHIFT>(min(max(g, 0), mx));
                    B = lift<SHIFT>(min(max(v[1] + g, 0), mx));
                }
                // grey preview of one component: one uniform test in the
                // common case (check_out < 0), selects otherwise
                if (CHECK && check_out >= 0) {
                    const int g1 = check_out == 0 ? R : (check_out == 1 ? G : B);
                    R = G = B = g1;
                }
            } else {
                R = G = B = 0;
            }
        } else {
            int t[4] = {0, 0, 0, 0};
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                if (c < nc) {
                    int sv;
                    if constexpr (FILT == 97) sv = (int)v[c];
                    else sv = v[c];
                    t[c] = lift<SHIFT>(min(max(sv + (1 << (bd - 1)), 0), mx));
                }
            }
            merge_slots(t, fmt, check_out == 0, R, G, B, A);
        }
    }

    __device__ inline void store(__attribute__((address_space(1))) uint8_t *p, int R, int G, int B,
                                 int A) const
    {
        const int lay = (PL == PL_ANY) ? pixel_layout(fmt) : PL;
        if (lay == PL_BGR8) {
            p[0] = (uint8_t)B;
            p[1] = (uint8_t)G;
            p[2] = (uint8_t)R;
        } else if (lay == PL_BGRA8) {
            *(__attribute__((address_space(1))) uint32_t *)p =
                (uint32_t)(B & 0xff) | ((uint32_t)(G & 0xff) << 8) | ((uint32_t)(R & 0xff) << 16) |
                ((uint32_t)(A & 0xff) << 24);
        } else if (lay == PL_BGR16) {
            auto q = (__attribute__((address_space(1))) uint16_t *)p;
            q[0] = (uint16_t)B;
            q[1] = (uint16_t)G;
            q[2] = (uint16_t)R;
        } else {
            auto q = (__attribute__((address_space(1))) uint16_t *)p;
            q[0] = (uint16_t)B;
            q[1] = (uint16_t)G;
            q[2] = (uint16_t)R;
            q[3] = (uint16_t)A;
        }
    }

    __device__ inline int layout_bytes() const
    {
        const int lay = (PL == PL_ANY) ? pixel_layout(fmt) : PL;
        return lay == PL_BGR8 ? 3 : lay == PL_BGRA8 ? 4 : lay == PL_BGR16 ? 6 : 8;
    }

    // 32-bit byte offsets from the uniform bases (destinations are limited
    // to 2^31 bytes by validate_tile), so stores take a scalar base
    template <typename U>
    __device__ static inline __attribute__((address_space(1))) U *at(__attribute__((address_space(1))) U *base,
                                                                      uint32_t byte_off)
    {
        return (__attribute__((address_space(1))) U *)((__attribute__((address_space(1))) char *)base + byte_off);
    }

    __device__ inline void put(int row, int col, const T *v) const
    {
        if constexpr (OUTK == K_PLANE) {
            const uint32_t off = ((uint32_t)(row - b0) * (uint32_t)pitch + (uint32_t)(col - a0)) * 4u;
#pragma unroll
            for (int c = 0; c < NC; ++c) *at(dst[c], off) = v[c];
        } else {
            int R, G, B, A;
            colour(v, R, G, B, A);
            store(at(out, ((uint32_t)(oy + row) * (uint32_t)pitch + (uint32_t)(ox + col)) *
                              (uint32_t)layout_bytes()),
                  R, G, B, A);
        }
    }

    // the two pixels of a lane, (row, ce) and (row, ce + 1).  The paired
    // path is chosen per item (uniform), so a wave runs one copy of the
    // colour code and masks only its stores.
    __device__ inline void put2(int row, int ce, const T *ve, const T *vo, bool e_ok, bool o_ok) const
    {
        if (pair_fast) {
            if constexpr (OUTK == K_PLANE) {
                store_plane_pair(row, ce, ve, vo, e_ok, o_ok);
                return;
            } else if constexpr (PL == PL_BGR8 || PL == PL_ANY) {
                // 6 bytes B0 G0 R0 B1 G1 R1 (pair_fast: a 24-bit layout)
                int R0, G0, B0, A0, R1, G1, B1, A1;
                colour(ve, R0, G0, B0, A0);
                colour(vo, R1, G1, B1, A1);
                store_pair_bgr8(row, ce, R0, G0, B0, R1, G1, B1, e_ok, o_ok);
                return;
            } else if constexpr (PL == PL_BGR16) {
                int R0, G0, B0, A0, R1, G1, B1, A1;
                colour(ve, R0, G0, B0, A0);
                colour(vo, R1, G1, B1, A1);
                store_pair_bgr16(row, ce, R0, G0, B0, R1, G1, B1, e_ok, o_ok);
                return;
            }
        }
        if (e_ok) put(row, ce, ve);
        if (o_ok) put(row, ce + 1, vo);
    }
};

// Prefetch depth of the row-pair pipeline: loads for pair k+D are issued
// while pair k is lifted, so each wave keeps D pairs of loads in flight
// (deeper for single-component items, which need fewer registers).
template <int NC>
struct Prefetch {
    // deeper rings for several components (3 or 4 pairs in flight) made the
    // final level slower with its data in HBM (37.6 -> 38.8 / 39.3 us)
    static constexpr int D = (NC == 1) ? 4 : 2;
};

// Coefficient encodings of a 9/7 item, fixed per kernel variant: LL and
// detail bands float (KD_FF), LL float and details sign-magnitude (KD_FS,
// levels above 0 of T1 data), both sign-magnitude (KD_SS), or chosen at
// run time (KD_ANY)
constexpr int KD_ANY = 0, KD_FF = 1, KD_FS = 2, KD_SS = 3;
constexpr int SK_RUNTIME = 0, SK_FLOAT = 1, SK_SIGNMAG = 2;

// convert a raw 32-bit word to a sample: sign-magnitude or float words
// times the folded gain (9/7), two's complement (5/3)
template <int FILT, int SK>
__device__ inline typename SampleT<FILT>::T to_sample(uint32_t w, int is_float, float gain)
{
    if constexpr (FILT == 97) {
        float v;
        if constexpr (SK == SK_FLOAT) v = __uint_as_float(w);
        else if constexpr (SK == SK_SIGNMAG) v = (float)signmag_to_int(w);
        else v = is_float ? __uint_as_float(w) : (float)signmag_to_int(w);
        return __fmul_rn(v, gain);
    } else {
        return (int)w;
    }
}

// Converted samples fixed in registers before their load buffer is
// refilled.  Without this the compiler sinks the conversions (the gain
// multiplications) to the samples' first use, after the refill, keeps the
// raw words alive by copying them out of the buffer at the loop's back
// edge, and waits for every load in flight to do so (s_waitcnt vmcnt(0)
// once per row pair), which undoes the prefetching.
template <typename T, int NC, int N>
__device__ inline void pin_samples(T (&smp)[NC][N])
{
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int j = 0; j < N; ++j) asm volatile("" : "+v"(smp[c][j]));
}

// Degenerate levels (a single row or a single column): no prefetching.
template <int FILT, int NC, int OUTK, int PL>
__device__ __attribute__((always_inline)) inline void process_item_small(const LevelJob &__restrict__ J,
                                                             int strip, int seg, int lane);

// One work item: strip `strip`, row segment `seg` of level job J, run by
// one wave64 (lane = 0..63).
template <int FILT, int NC, int OUTK, int PL, int KD>
__device__ __attribute__((always_inline)) inline void process_item(
    const LevelJob &__restrict__ J, int strip, int seg, int lane)
{
    constexpr int SK_LL = (KD == KD_FF || KD == KD_FS) ? SK_FLOAT : (KD == KD_SS ? SK_SIGNMAG : SK_RUNTIME);
    constexpr int SK_DET = (KD == KD_FF) ? SK_FLOAT : ((KD == KD_FS || KD == KD_SS) ? SK_SIGNMAG : SK_RUNTIME);
    using T = typename SampleT<FILT>::T;
    constexpr int HP = (FILT == 97) ? 2 : 1;  // halo pairs on each side
    constexpr int OWP = 64 - 2 * HP;          // output pairs per strip
    constexpr int D = Prefetch<NC>::D;

    const int a0 = J.a0, a1 = J.a1, b0 = J.b0, b1 = J.b1;
    if ((a1 - a0) == 1 || (b1 - b0) == 1) {
        process_item_small<FILT, NC, OUTK, PL>(J, strip, seg, lane);
        return;
    }
    // every field the loop needs, hoisted out of the descriptor once
    const int lw = J.lw, lh = J.lh, rl0 = J.rl0, rh0 = J.rh0;
    const int ll_pitch = J.ll_pitch, det_pitch = J.det_pitch;
    const int ll_float = J.ll_kind == SRC_FLOAT, det_float = J.det_kind == SRC_FLOAT;
    const int seg_pairs = J.seg_pairs, quirk = J.quirk;
    const __attribute__((address_space(1))) uint32_t *llp[NC];
    const __attribute__((address_space(1))) uint32_t *detp[NC];
    float gn[NC][4];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        llp[c] = gptr<uint32_t>(J.ll[c]);
        detp[c] = gptr<uint32_t>(J.det[c]);
#pragma unroll
        for (int j = 0; j < 4; ++j) gn[c][j] = J.gain[c][j];
    }

    // absolute columns of this lane (even = low sample, odd = high sample)
    const int wpair0 = ((a0 & ~1) >> 1) + strip * OWP - HP;
    const int ce = 2 * (wpair0 + lane), co = ce + 1;
    const bool lane_out = (lane >= HP) && (lane < 64 - HP);
    const int lo_idx = (reflect(ce, a0, a1) >> 1) - J.cl0;
    const int hi_off = lw + (reflect(co, a0, a1) >> 1) - J.ch0;

    // row pairs of this segment
    const int K0 = (b0 & ~1) >> 1;
    const int Klast = (b1 - 1) >> 1;
    const int kout0 = K0 + seg * seg_pairs;
    const int kout1 = min(Klast, kout0 + seg_pairs - 1);

    // raw words of pair k: [comp][top-low, top-high, bottom-low, bottom-high];
    // a segment fetches rows in [b0 - 5, b1 + 4]; for a level of two or
    // more rows three single reflections bring every one of them into
    // range (one for a level of 8 rows or more), equal to the periodic
    // symmetric extension (reflect)
    const int nrefl = (b1 - b0) >= 8 ? 1 : 3;
    auto reflect_row = [&](int r) {
#pragma unroll 1
        for (int i = 0; i < nrefl; ++i) {
            r = r < b0 ? 2 * b0 - r : r;
            r = r >= b1 ? 2 * (b1 - 1) - r : r;
        }
        return r;
    };
    auto fetch = [&](int k, uint32_t (&r)[NC][4]) {
        const int rre = reflect_row(2 * k), rro = reflect_row(2 * k + 1);
        // 32-bit element offsets from the (uniform) plane bases: the loads
        // take a scalar base and a 32-bit vector offset (planes are
        // limited to 2^31 bytes by validate_tile)
        const uint32_t rt = (uint32_t)((rre >> 1) - rl0);
        const uint32_t q = (uint32_t)((rro >> 1) - rh0 + lh);
        const uint32_t llr = rt * (uint32_t)ll_pitch, dtr = rt * (uint32_t)det_pitch;
        const uint32_t dbr = q * (uint32_t)det_pitch;
        auto ld = [](const __attribute__((address_space(1))) uint32_t *base, uint32_t elem) {
            return *(const __attribute__((address_space(1))) uint32_t *)(
                (const __attribute__((address_space(1))) char *)base + (elem << 2));
        };
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            r[c][0] = ld(llp[c], llr + (uint32_t)lo_idx);
            r[c][1] = ld(detp[c], dtr + (uint32_t)hi_off);
            r[c][2] = ld(detp[c], dbr + (uint32_t)lo_idx);
            r[c][3] = ld(detp[c], dbr + (uint32_t)hi_off);
        }
    };

    Emitter<FILT, NC, OUTK, PL> em;
    em.init(J);
    const bool ce_ok = lane_out && ce >= a0 && ce < a1;
    const bool co_ok = lane_out && co >= a0 && co < a1;
    em.set_lanes(ce_ok, co_ok);
    const bool full = em.full_ok();
    auto emit_row = [&](int row, const T *ve, const T *vo) {
        if (row < b0 || row >= b1) return;
        em.put2(row, ce, ve, vo, ce_ok, co_ok);
    };

    const int nh = b1 - b0;
    // iteration at which the reference's VDWT_R53_LINES tail applies:
    // local pair nh/2 (bmi_cpu_dwt_decoder.cpp:3359-3378)
    const int kq = (FILT == 53 && quirk) ? (K0 + (nh >> 1)) : -0x7fffffff;
    const bool nh_even = !(nh & 1);
    constexpr int LAG = (FILT == 97) ? 2 : 1;  // pair k in, pair k-LAG out
    const int ks = kout0 - ((FILT == 97) ? 2 : (quirk ? 2 : 1));
    const int ke = kout1 + LAG;

    // Segments alternate direction: odd segments run bottom-up.  The halo
    // rows two vertically neighbouring segments both read are then read at
    // the same time by both (at the end of each, or at the start of each),
    // so the second read is served by L2 instead of HBM.  Every lifting
    // step keeps its operands and their order, so the results are the same
    // bit for bit in either direction.  A 5/3 segment whose pipeline range
    // holds the pair of the reference's bottom-edge rule runs top-down (the
    // rule is stated on the top-down pipeline's state).
    const bool rev = (seg & 1) && (FILT == 97 || kq < ks || kq > ke);

    // vertical pipeline state per component and column (0 = ce, 1 = co)
    // 9/7 top-down:  s0 = H_{k-1}, s1 = e3_{k-1}, s2 = o4_{k-2}, s3 = e5_{k-2}
    // 9/7 bottom-up: s0 = L_{k+1}, s1 = H_{k+1}, s2 = e3_{k+2}, s3 = o4_{k+2},
    //                s4 = e5_{k+3}
    // 5/3 top-down:  s0 = H_{k-1}, s1 = e1_{k-1}, s2 = L_{k-1}, s3 = H_{k-2}
    // 5/3 bottom-up: s0 = L_{k+1}, s1 = H_{k+1}, s2 = e1_{k+2}
    T s0[NC][2], s1[NC][2], s2[NC][2], s3[NC][2], s4[NC][2];
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int e = 0; e < 2; ++e) s0[c][e] = s1[c][e] = s2[c][e] = s3[c][e] = s4[c][e] = (T)0;

    // one pipeline step with pair k (samples already converted)
    auto consume = [&](auto dir, int k, T (&smp)[NC][4]) {
        constexpr bool REV = decltype(dir)::value;
        T evn[2][NC], odd[2][NC];  // [column][component]
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            T tl = smp[c][0], th = smp[c][1], bl = smp[c][2], bh = smp[c][3];
            hlift<FILT>(tl, th);
            hlift<FILT>(bl, bh);
            const T Lk[2] = {tl, th}, Hk[2] = {bl, bh};
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                if constexpr (FILT == 97) {
                    if constexpr (!REV) {
                        // pair k in, pair k - 2 out
                        const float e3 = lift97(Lk[e], kDelta, s0[c][e], Hk[e]);
                        const float o4 = lift97(s0[c][e], kGamma, s1[c][e], e3);
                        const float e5 = lift97(s1[c][e], kBeta, s2[c][e], o4);
                        const float o6 = lift97(s2[c][e], kAlpha, s3[c][e], e5);
                        evn[e][c] = s3[c][e];
                        odd[e][c] = o6;
                        s0[c][e] = Hk[e];
                        s1[c][e] = e3;
                        s2[c][e] = o4;
                        s3[c][e] = e5;
                    } else {
                        // pair k in, pair k + 2 out; the same steps with the
                        // upper neighbour first, as above
                        const float e3 = lift97(s0[c][e], kDelta, Hk[e], s1[c][e]);  // e3_{k+1}
                        const float o4 = lift97(s1[c][e], kGamma, e3, s2[c][e]);     // o4_{k+1}
                        const float e5 = lift97(s2[c][e], kBeta, o4, s3[c][e]);      // e5_{k+2}
                        const float o6 = lift97(s3[c][e], kAlpha, e5, s4[c][e]);     // o6_{k+2}
                        evn[e][c] = e5;
                        odd[e][c] = o6;
                        s0[c][e] = Lk[e];
                        s1[c][e] = Hk[e];
                        s2[c][e] = e3;
                        s3[c][e] = o4;
                        s4[c][e] = e5;
                    }
                } else if constexpr (REV) {
                    // pair k in, pair k + 1 out: e1_{k+1}, then o2_{k+1}
                    const int e1 = s0[c][e] - ((Hk[e] + s1[c][e] + 2) >> 2);
                    evn[e][c] = e1;
                    odd[e][c] = s1[c][e] + ((e1 + s2[c][e]) >> 1);
                    s0[c][e] = Lk[e];
                    s1[c][e] = Hk[e];
                    s2[c][e] = e1;
                } else {
                    int e1 = Lk[e] - ((s0[c][e] + Hk[e] + 2) >> 2);
                    int o2 = s0[c][e] + ((s1[c][e] + e1) >> 1);
                    if (k == kq) {
                        if (nh_even) {
                            // last odd row uses the stale high row H_{m-2}
                            o2 = Hk[e] + s1[c][e];
                        } else {
                            // last two rows use L_{m-1} and H_{m-2}
                            e1 = s2[c][e] - ((s3[c][e] + s3[c][e] + 2) >> 2);
                            o2 = s3[c][e] + ((s1[c][e] + e1) >> 1);
                        }
                    }
                    evn[e][c] = s1[c][e];
                    odd[e][c] = o2;
                    s3[c][e] = s0[c][e];
                    s0[c][e] = Hk[e];
                    s2[c][e] = Lk[e];
                    s1[c][e] = e1;
                }
            }
        }
        const int ko = REV ? k + LAG : k - LAG;
        if (ko >= kout0 && ko <= kout1) {
            if (full && 2 * ko >= b0 && 2 * ko + 1 < b1) {
                em.put2_full(2 * ko, ce, evn[0], evn[1], ce_ok, co_ok);
                em.put2_full(2 * ko + 1, ce, odd[0], odd[1], ce_ok, co_ok);
            } else {
                emit_row(2 * ko, evn[0], evn[1]);
                emit_row(2 * ko + 1, odd[0], odd[1]);
            }
        }
    };

    // software pipeline, D pairs of loads in flight; loads are issued
    // unconditionally (pairs past the range [ks, ke] are clamped to it and
    // the pipeline steps that use them emit nothing); bottom-up segments
    // walk the same pairs from ke down to ks
    // (one copy of the loop per direction, chosen once per item)
    auto run = [&](auto dir) {
        constexpr bool REV = decltype(dir)::value;
        uint32_t ring[D][NC][4];
#pragma unroll
        for (int d = 0; d < D; ++d) fetch(REV ? max(ke - d, ks) : min(ks + d, ke), ring[d]);
        const int rounds = (ke - ks + D) / D;
        for (int it = 0; it < rounds; ++it) {
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const int k = REV ? ke - (it * D + d) : ks + it * D + d;
                // convert slot d (its loads were issued D pairs ago), then reuse
                // its registers for pair k + D before doing the arithmetic
                T smp[NC][4];
#pragma unroll
                for (int c = 0; c < NC; ++c) {
                    smp[c][0] = to_sample<FILT, SK_LL>(ring[d][c][0], ll_float, gn[c][0]);
                    smp[c][1] = to_sample<FILT, SK_DET>(ring[d][c][1], det_float, gn[c][1]);
                    smp[c][2] = to_sample<FILT, SK_DET>(ring[d][c][2], det_float, gn[c][2]);
                    smp[c][3] = to_sample<FILT, SK_DET>(ring[d][c][3], det_float, gn[c][3]);
                }
                fetch(REV ? max(k - D, ks) : min(k + D, ke), ring[d]);
                consume(dir, k, smp);
                __builtin_amdgcn_sched_barrier(0);
            }
        }
    };
    if (rev)
        run(std::true_type{});
    else
        run(std::false_type{});
}

template <int FILT, int NC, int OUTK, int PL>
__device__ __attribute__((always_inline)) inline void process_item_small(const LevelJob &__restrict__ J,
                                                             int strip, int seg, int lane)
{
    using T = typename SampleT<FILT>::T;
    constexpr int HP = (FILT == 97) ? 2 : 1;
    constexpr int OWP = 64 - 2 * HP;
    const int a0 = J.a0, a1 = J.a1, b0 = J.b0, b1 = J.b1;
    const bool single_col = (a1 - a0) == 1;
    const bool single_row = (b1 - b0) == 1;
    const int nc = J.nc;
    const int wpair0 = ((a0 & ~1) >> 1) + strip * OWP - HP;
    const int ce = 2 * (wpair0 + lane), co = ce + 1;
    const bool lane_out = (lane >= HP) && (lane < 64 - HP);
    int lo_idx = 0, hi_idx = 0;
    if (!single_col) {
        lo_idx = (reflect(ce, a0, a1) >> 1) - J.cl0;
        hi_idx = (reflect(co, a0, a1) >> 1) - J.ch0;
    }
    const bool col_is_low = single_col && !(a0 & 1);
    const int hi_off = single_col ? J.lw : J.lw + hi_idx;
    const int K0 = (b0 & ~1) >> 1;
    const int Klast = (b1 - 1) >> 1;
    const int kout0 = K0 + seg * J.seg_pairs;
    const int kout1 = min(Klast, kout0 + J.seg_pairs - 1);
    const bool has_top = !single_row || !(b0 & 1);
    const bool has_bot = !single_row || (b0 & 1);
    const bool load_lo = !single_col || col_is_low;
    const bool load_hi = !single_col || !col_is_low;

    auto load_rows = [&](int k, int c, T &tl, T &th, T &bl, T &bh) {
        int rre = b0, rro = b0;
        if (!single_row) {
            rre = reflect(2 * k, b0, b1);
            rro = reflect(2 * k + 1, b0, b1);
        }
        const size_t rt = (size_t)((rre >> 1) - J.rl0);
        const size_t q = (size_t)((rro >> 1) - J.rh0 + J.lh);
        const float *g = J.gain[c];
        tl = th = bl = bh = (T)0;
        if (has_top) {
            if (load_lo) tl = load_sample<FILT>(J.ll[c], rt * J.ll_pitch + lo_idx, J.ll_kind, g[0]);
            if (load_hi) th = load_sample<FILT>(J.det[c], rt * J.det_pitch + hi_off, J.det_kind, g[1]);
        }
        if (has_bot) {
            if (load_lo) bl = load_sample<FILT>(J.det[c], q * J.det_pitch + lo_idx, J.det_kind, g[2]);
            if (load_hi) bh = load_sample<FILT>(J.det[c], q * J.det_pitch + hi_off, J.det_kind, g[3]);
        }
        if (!single_col) {
            hlift<FILT>(tl, th);
            hlift<FILT>(bl, bh);
        } else if (col_is_low) {
            th = tl;
            bh = bl;
        } else {
            tl = th = half_sample<FILT>(th);
            bl = bh = half_sample<FILT>(bh);
        }
    };
    Emitter<FILT, NC, OUTK, PL> em;
    em.init(J);
    const bool ce_ok = lane_out && ce >= a0 && ce < a1;
    const bool co_ok = lane_out && co >= a0 && co < a1;
    em.set_lanes(ce_ok, co_ok);
    auto emit_row = [&](int row, const T *ve, const T *vo) {
        if (row < b0 || row >= b1) return;
        em.put2(row, ce, ve, vo, ce_ok, co_ok);
    };

    if (single_row) {
        T ve[NC], vo[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            T tl = 0, th = 0, bl = 0, bh = 0;
            if (c < nc) load_rows(b0 >> 1, c, tl, th, bl, bh);
            if (!(b0 & 1)) {
                ve[c] = tl;
                vo[c] = th;
            } else {
                ve[c] = half_sample<FILT>(bl);
                vo[c] = half_sample<FILT>(bh);
            }
        }
        emit_row(b0, ve, vo);
        return;
    }
    // single column, several rows: plain (unpipelined) vertical lifting
    T s0[NC][2], s1[NC][2], s2[NC][2], s3[NC][2];
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int e = 0; e < 2; ++e) s0[c][e] = s1[c][e] = s2[c][e] = s3[c][e] = (T)0;
    const int nh = b1 - b0;
    const int kq = (FILT == 53 && J.quirk) ? (K0 + (nh >> 1)) : -0x7fffffff;
    const bool nh_even = !(nh & 1);
    constexpr int LAG = (FILT == 97) ? 2 : 1;
    const int ks = kout0 - ((FILT == 97) ? 2 : (J.quirk ? 2 : 1));
    const int ke = kout1 + LAG;
    for (int k = ks; k <= ke; ++k) {
        T evn[2][NC], odd[2][NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            T tl = 0, th = 0, bl = 0, bh = 0;
            if (c < nc) load_rows(k, c, tl, th, bl, bh);
            const T Lk[2] = {tl, th}, Hk[2] = {bl, bh};
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                if constexpr (FILT == 97) {
                    const float e3 = lift97(Lk[e], kDelta, s0[c][e], Hk[e]);
                    const float o4 = lift97(s0[c][e], kGamma, s1[c][e], e3);
                    const float e5 = lift97(s1[c][e], kBeta, s2[c][e], o4);
                    const float o6 = lift97(s2[c][e], kAlpha, s3[c][e], e5);
                    evn[e][c] = s3[c][e];
                    odd[e][c] = o6;
                    s0[c][e] = Hk[e];
                    s1[c][e] = e3;
                    s2[c][e] = o4;
                    s3[c][e] = e5;
                } else {
                    int e1 = Lk[e] - ((s0[c][e] + Hk[e] + 2) >> 2);
                    int o2 = s0[c][e] + ((s1[c][e] + e1) >> 1);
                    if (k == kq) {
                        if (nh_even) {
                            o2 = Hk[e] + s1[c][e];
                        } else {
                            e1 = s2[c][e] - ((s3[c][e] + s3[c][e] + 2) >> 2);
                            o2 = s3[c][e] + ((s1[c][e] + e1) >> 1);
                        }
                    }
                    evn[e][c] = s1[c][e];
                    odd[e][c] = o2;
                    s3[c][e] = s0[c][e];
                    s0[c][e] = Hk[e];
                    s2[c][e] = Lk[e];
                    s1[c][e] = e1;
                }
            }
        }
        const int ko = k - LAG;
        if (ko >= kout0) {
            emit_row(2 * ko, evn[0], evn[1]);
            emit_row(2 * ko + 1, odd[0], odd[1]);
        }
    }
}

// Thumbnail (a tile decoded with zero synthesis levels, top_level 0): the
// LL band alone, colour-converted and packed like a final level.  9/7
// samples are dequantised with the LL step only, (float)twos(w) * step:
// with the gains folded into the synthesis passes (K / 2/K per direction)
// every synthesis level has DC gain 1, so this is the scale the full
// decode gives a flat region.  The reference's GPU thumbnail
// (cudainterface.cpp:1278-1293, 1496-1510) converts the LL band of a
// 5/3 tile the same way; one item is 64 pixels of one LL row.
template <int FILT, int NC, int OUTK, int PL>
__device__ __attribute__((always_inline)) inline void thumb_item(const LevelJob &__restrict__ J,
                                                                  int strip, int seg, int lane)
{
    using T = typename SampleT<FILT>::T;
    const int x = J.a0 + strip * 64 + lane;
    const int y = J.b0 + seg;
    if (x >= J.a1) return;
    Emitter<FILT, NC, OUTK, PL> em;
    em.init(J);
    T v[NC];
    const size_t idx = (size_t)seg * (size_t)J.ll_pitch + (size_t)(x - J.a0);
#pragma unroll
    for (int c = 0; c < NC; ++c)
        v[c] = (c < J.nc) ? load_sample<FILT>(J.ll[c], idx, J.ll_kind, J.gain[c][0]) : (T)0;
    em.put(y, x, v);
}

// Rows of the inner level kept per wave by a two-level item: the final
// level reads them in order except near the tile edges, where the
// symmetric extension revisits rows already read.  When a row is read, at
// most four rows (9/7) or six (5/3, whose shorter pipeline lag lets two
// inner steps run further ahead) have been written since it (inclusive;
// an emulation of this schedule, two inner steps at a time, over a million
// items of every level count, offset, height and segment height finds no
// more), so eight are enough.  A small ring keeps LDS from limiting the
// waves per CU.
constexpr int kRingRows = 8;

// Strips of a two-level item (LevelJob::fuse2).  One strip of the final
// level needs 64 columns of the inner level's output (its LL band) but the
// inner level's pipeline, like every pipeline here, lifts 128 columns per
// wave, so with one strip per wave the inner level keeps 28 of its 64
// lanes idle.  A two-strip item runs two strips of the final level side by
// side (each lane owns one column pair of each), 58 + 57 output pairs,
// whose 119 LL columns (with the final level's lifting halo) the inner
// level makes in one 128-column pass.
__host__ __device__ constexpr int fused2_out_pairs(int ns) { return ns == 2 ? 115 : 60; }

// Two-level work item (9/7): item `strip` (NS strips side by side), row
// segment `seg` of the final level of J, with the level below it
// synthesised in the same wave.  The inner level runs the same pipeline as
// process_item over the inner columns the item's LL band needs and writes
// each finished row to a ring of kRingRows rows in this wave's LDS; the
// final level takes its LL samples from that ring instead of a scratch
// plane.  The inner level runs two steps (two row pairs in, four rows out)
// whenever the final level needs a row it has not produced yet: an even
// number of steps leaves the pipeline's state in the registers it started
// in, so no step ends in register copies.  Every sample gets the
// operations of the two-launch form in the same order, so the output is
// the same bit for bit; the inner level's output never goes to HBM.
template <int FILT, int NC, int OUTK, int PL, int KD, int NS, bool SIMPLE>
__device__ __attribute__((always_inline)) inline void process_item2(
    const LevelJob &__restrict__ J, int strip, int seg, int lane, float *__restrict__ ring_f)
{
    constexpr int SK_ILL = (KD == KD_FF || KD == KD_FS) ? SK_FLOAT : (KD == KD_SS ? SK_SIGNMAG : SK_RUNTIME);
    constexpr int SK_DET = (KD == KD_FF) ? SK_FLOAT : ((KD == KD_FS || KD == KD_SS) ? SK_SIGNMAG : SK_RUNTIME);
    using T = typename SampleT<FILT>::T;
    T *__restrict__ ring = (T *)ring_f;
    constexpr int LAG = (FILT == 97) ? 2 : 1;  // pair k in, pair k - LAG out
    // the 5/3 items use the 9/7 geometry (two halo pairs a side; 5/3 needs one)
    constexpr int HP = 2;
    constexpr int OWP = fused2_out_pairs(NS);
    constexpr int RW = 64 * NS;  // ring row: floats per component
    constexpr int D = 2;         // final-level pairs in flight

    // ---- final level: as process_item, without the LL loads; strip h of
    // the item holds window pairs wbase + 58 h + lane
    const int a0 = J.a0, a1 = J.a1, b0 = J.b0, b1 = J.b1;
    const int lw = J.lw, lh = J.lh, rl0 = J.rl0, rh0 = J.rh0;
    const int det_pitch = J.det_pitch;
    const int det_float = J.det_kind == SRC_FLOAT;
    const __attribute__((address_space(1))) uint32_t *detp[NC];
    float gn[NC][4];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        detp[c] = gptr<uint32_t>(J.det[c]);
#pragma unroll
        for (int j = 0; j < 4; ++j) gn[c][j] = J.gain[c][j];
    }
    const int wbase = ((a0 & ~1) >> 1) + strip * OWP - HP;
    const int mbase = wbase - J.cl0;  // LL column (band-relative) of ring slot 0
    int ce[NS], lo_idx[NS], hi_off[NS], xl[NS];
    bool ce_ok[NS], co_ok[NS];
#pragma unroll
    for (int h = 0; h < NS; ++h) {
        ce[h] = 2 * (wbase + 58 * h + lane);
        const int co = ce[h] + 1;
        // output lanes: 2..61 of a single strip; 2..59 and 2..58 of two
        const int lane_end = NS == 1 ? 64 - HP : (h == 0 ? 60 : 59);
        const bool lane_out = lane >= HP && lane < lane_end;
        lo_idx[h] = (reflect(ce[h], a0, a1) >> 1) - J.cl0;
        hi_off[h] = lw + (reflect(co, a0, a1) >> 1) - J.ch0;
        ce_ok[h] = lane_out && ce[h] >= a0 && ce[h] < a1;
        co_ok[h] = lane_out && co >= a0 && co < a1;
        // ring slot of this lane's LL sample (its unreflected column is
        // mbase + 58 h + lane; a reflected one of a lane that matters stays
        // inside the window, and lanes far past the tile edge only feed
        // halo lanes that are never stored, so their slot is clamped)
        xl[h] = min(max(lo_idx[h] - mbase, 0), RW - 1);
    }
    const int K0 = (b0 & ~1) >> 1;
    const int Klast = (b1 - 1) >> 1;
    const int kout0 = K0 + seg * J.seg_pairs;
    const int kout1 = min(Klast, kout0 + J.seg_pairs - 1);
    // 5/3: the pair at which the reference's bottom-edge rule applies (as
    // in process_item), one more warm-up pair for it
    const int nh = b1 - b0;
    const int kq = (FILT == 53 && J.quirk) ? (K0 + (nh >> 1)) : -0x7fffffff;
    const bool nh_even = !(nh & 1);
    const int ks = kout0 - ((FILT == 97) ? 2 : (J.quirk ? 2 : 1)), ke = kout1 + LAG;
    const int nrefl = (b1 - b0) >= 8 ? 1 : 3;
    auto reflect_row = [&](int r) {
#pragma unroll 1
        for (int i = 0; i < nrefl; ++i) {
            r = r < b0 ? 2 * b0 - r : r;
            r = r >= b1 ? 2 * (b1 - 1) - r : r;
        }
        return r;
    };
    auto ld = [](const __attribute__((address_space(1))) uint32_t *base, uint32_t elem) {
        return *(const __attribute__((address_space(1))) uint32_t *)(
            (const __attribute__((address_space(1))) char *)base + (elem << 2));
    };
    // LL row (relative to the inner level's first row) of final pair k
    auto ll_row = [&](int k) { return (reflect_row(2 * k) >> 1) - rl0; };
    auto fetch = [&](int k, uint32_t (&r)[NS][NC][3]) {
        const int rre = reflect_row(2 * k), rro = reflect_row(2 * k + 1);
        const uint32_t rt = (uint32_t)((rre >> 1) - rl0);
        const uint32_t q = (uint32_t)((rro >> 1) - rh0 + lh);
        const uint32_t dtr = rt * (uint32_t)det_pitch, dbr = q * (uint32_t)det_pitch;
#pragma unroll
        for (int h = 0; h < NS; ++h)
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                r[h][c][0] = ld(detp[c], dtr + (uint32_t)hi_off[h]);
                r[h][c][1] = ld(detp[c], dbr + (uint32_t)lo_idx[h]);
                r[h][c][2] = ld(detp[c], dbr + (uint32_t)hi_off[h]);
            }
    };

    // ---- inner level: geometry of process_item for the ring's columns
    // [mbase, mbase + RW) (band-relative), made by lanes 2..61
    const int ia0 = J.ia0, ia1 = J.ia1, ib0 = J.ib0, ib1 = J.ib1;
    const int ilw = J.ilw, ilh = J.ilh, irl0 = J.irl0, irh0 = J.irh0;
    const int ill_pitch = J.ill_pitch;
    const int ill_float = J.ill_kind == SRC_FLOAT;
    const __attribute__((address_space(1))) uint32_t *illp[NC];
    float ign[NC][4];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        illp[c] = gptr<uint32_t>(J.ill[c]);
#pragma unroll
        for (int j = 0; j < 4; ++j) ign[c][j] = J.igain[c][j];
    }
    const int inh = ib1 - ib0;            // inner output rows (the final level's LL rows)
    const int W0 = ia0 + mbase;           // absolute inner column of slot 0
    const int ipair0 = (W0 >> 1) - HP;    // floor: W0 may be negative
    const int ice = 2 * (ipair0 + lane), ico = ice + 1;
    const int ilo_idx = (reflect(ice, ia0, ia1) >> 1) - J.icl0;
    const int ihi_off = ilw + (reflect(ico, ia0, ia1) >> 1) - J.ich0;
    const int xe = ice - W0, xo = ico - W0;
    const bool ilane = lane >= HP && lane < 64 - HP;
    const bool we = ilane && xe >= 0 && xe < RW;
    const bool wo = ilane && xo >= 0 && xo < RW;
    const int inrefl = (ib1 - ib0) >= 8 ? 1 : 3;
    auto ireflect_row = [&](int r) {
#pragma unroll 1
        for (int i = 0; i < inrefl; ++i) {
            r = r < ib0 ? 2 * ib0 - r : r;
            r = r >= ib1 ? 2 * (ib1 - 1) - r : r;
        }
        return r;
    };
    auto ifetch = [&](int q, uint32_t (&r)[NC][4]) {
        const int rre = ireflect_row(2 * q), rro = ireflect_row(2 * q + 1);
        const uint32_t rt = (uint32_t)((rre >> 1) - irl0);
        const uint32_t qq = (uint32_t)((rro >> 1) - irh0 + ilh);
        const uint32_t llr = rt * (uint32_t)ill_pitch, dtr = rt * (uint32_t)det_pitch;
        const uint32_t dbr = qq * (uint32_t)det_pitch;
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            r[c][0] = ld(illp[c], llr + (uint32_t)ilo_idx);
            r[c][1] = ld(detp[c], dtr + (uint32_t)ihi_off);
            r[c][2] = ld(detp[c], dbr + (uint32_t)ilo_idx);
            r[c][3] = ld(detp[c], dbr + (uint32_t)ihi_off);
        }
    };

    // inner rows this segment reads: [rmin, rmax], relative to ib0
    int rmin = 0x7fffffff, rmax = -1;
#pragma unroll 1
    for (int k = ks; k <= ke; ++k) {
        const int r = ll_row(k);
        rmin = min(rmin, r);
        rmax = max(rmax, r);
    }
    const int qo0 = (ib0 + rmin) >> 1, qo1 = (ib0 + rmax) >> 1;  // inner pairs to emit
    // inner pipeline range, warm-up included
    const int iqs = qo0 - ((FILT == 97) ? 2 : (J.iquirk ? 2 : 1)), iqe = qo1 + LAG;
    const int ikq = (FILT == 53 && J.iquirk) ? (((ib0 & ~1) >> 1) + (inh >> 1)) : -0x7fffffff;
    const bool inh_even = !(inh & 1);

    Emitter<FILT, NC, OUTK, PL, SIMPLE> em;
    em.init(J);
    bool single = false;  // some lane has one pixel of its pair in the tile
#pragma unroll
    for (int h = 0; h < NS; ++h) single = single || (ce_ok[h] != co_ok[h]);
    em.set_lanes(single, false);
    const bool full = em.full_ok();

    // One lifting step of a pipeline (the final level's or the inner
    // level's), as process_item: top-down pair k in, pair k - 2 out;
    // bottom-up pair k in, pair k + 2 out.
    // top-down:  p0 = H_{k-1}, p1 = e3_{k-1}, p2 = o4_{k-2}, p3 = e5_{k-2}
    // bottom-up: p0 = L_{k+1}, p1 = H_{k+1}, p2 = e3_{k+2}, p3 = o4_{k+2},
    //            p4 = e5_{k+3}
    // 5/3 (process_item's 5/3 pipeline), top-down: pair k in, pair k - 1
    // out; p0 = H_{k-1}, p1 = e1_{k-1}, p2 = L_{k-1}, p3 = H_{k-2}; at the
    // bottom-edge pair (at_q) the reference's VDWT_R53_LINES tail.
    // Bottom-up: pair k in, pair k + 1 out; p0 = L_{k+1}, p1 = H_{k+1},
    // p2 = e1_{k+2} (no bottom-edge pair in its range).
    auto lift53 = [](auto dir, int Lk, int Hk, int &p0, int &p1, int &p2, int &p3, int &ev, int &od,
                     bool at_q, bool h_even) {
        if constexpr (decltype(dir)::value) {
            const int e1 = p0 - ((Hk + p1 + 2) >> 2);
            ev = e1;
            od = p1 + ((e1 + p2) >> 1);
            p0 = Lk;
            p1 = Hk;
            p2 = e1;
            return;
        }
        int e1 = Lk - ((p0 + Hk + 2) >> 2);
        int o2 = p0 + ((p1 + e1) >> 1);
        if (at_q) {
            if (h_even) {
                o2 = Hk + p1;
            } else {
                e1 = p2 - ((p3 + p3 + 2) >> 2);
                o2 = p3 + ((p1 + e1) >> 1);
            }
        }
        ev = p1;
        od = o2;
        p3 = p0;
        p0 = Hk;
        p2 = Lk;
        p1 = e1;
    };
    auto lift_step = [](auto dir, float Lk, float Hk, float &p0, float &p1, float &p2, float &p3,
                        float &p4, float &ev, float &od) {
        constexpr bool REV = decltype(dir)::value;
        if constexpr (!REV) {
            const float e3 = lift97(Lk, kDelta, p0, Hk);
            const float o4 = lift97(p0, kGamma, p1, e3);
            const float e5 = lift97(p1, kBeta, p2, o4);
            const float o6 = lift97(p2, kAlpha, p3, e5);
            ev = p3;
            od = o6;
            p0 = Hk;
            p1 = e3;
            p2 = o4;
            p3 = e5;
        } else {
            const float e3 = lift97(p0, kDelta, Hk, p1);
            const float o4 = lift97(p1, kGamma, e3, p2);
            const float e5 = lift97(p2, kBeta, o4, p3);
            const float o6 = lift97(p3, kAlpha, e5, p4);
            ev = e5;
            od = o6;
            p0 = Lk;
            p1 = Hk;
            p2 = e3;
            p3 = o4;
            p4 = e5;
        }
    };

    // Odd segments run bottom-up, both levels, as in process_item: the rows
    // two vertically neighbouring segments both read (their halos) are then
    // read by both at about the same time and come from L2 the second time.
    auto run = [&](auto dir) {
        constexpr bool REV = decltype(dir)::value;
        // inner level state; two steps' loads in flight (one buffer each)
        T t0[NC][2], t1[NC][2], t2[NC][2], t3[NC][2], t4[NC][2];
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
            for (int e = 0; e < 2; ++e) t0[c][e] = t1[c][e] = t2[c][e] = t3[c][e] = t4[c][e] = 0.0f;
        int q = REV ? iqe : iqs;  // next inner pair to consume
        int have = rmin;          // top-down: rows [rmin, have) are in the ring
        int low = rmax + 1;       // bottom-up: rows [low, rmax] are in the ring
        uint32_t islot[2][NC][4];
        ifetch(q, islot[0]);
        ifetch(REV ? max(q - 1, iqs) : min(q + 1, iqe), islot[1]);

        // one inner step with pair q, whose loads are in sl; refills sl
        // with pair q + 2 (the step after next, which uses the same buffer)
        auto inner_step = [&](uint32_t (&sl)[NC][4]) {
            T smp[NC][4];
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                smp[c][0] = to_sample<FILT, SK_ILL>(sl[c][0], ill_float, ign[c][0]);
                smp[c][1] = to_sample<FILT, SK_DET>(sl[c][1], det_float, ign[c][1]);
                smp[c][2] = to_sample<FILT, SK_DET>(sl[c][2], det_float, ign[c][2]);
                smp[c][3] = to_sample<FILT, SK_DET>(sl[c][3], det_float, ign[c][3]);
            }
            pin_samples(smp);
            ifetch(REV ? max(q - 2, iqs) : min(q + 2, iqe), sl);
            const int po = REV ? q + LAG : q - LAG;  // pair emitted by this step
            const int r0 = 2 * po - ib0;         // its first row, relative
            const bool emit = po >= qo0 && po <= qo1;
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                T tl = smp[c][0], th = smp[c][1], bl = smp[c][2], bh = smp[c][3];
                hlift<FILT>(tl, th);
                hlift<FILT>(bl, bh);
                const T Lk[2] = {tl, th}, Hk[2] = {bl, bh};
                T ev[2], od[2];
#pragma unroll
                for (int e = 0; e < 2; ++e) {
                    if constexpr (FILT == 97)
                        lift_step(dir, Lk[e], Hk[e], t0[c][e], t1[c][e], t2[c][e], t3[c][e], t4[c][e],
                                  ev[e], od[e]);
                    else
                        lift53(dir, Lk[e], Hk[e], t0[c][e], t1[c][e], t2[c][e], t3[c][e], ev[e], od[e],
                               q == ikq, inh_even);
                }
                if (emit) {
                    // rows r0 (even) and r0 + 1 (odd) of the inner output
                    if (r0 >= 0 && r0 < inh) {
                        T *row = ring + ((r0 & (kRingRows - 1)) * NC + c) * RW;
                        if (we) row[xe] = ev[0];
                        if (wo) row[xo] = ev[1];
                    }
                    if (r0 + 1 >= 0 && r0 + 1 < inh) {
                        T *row = ring + (((r0 + 1) & (kRingRows - 1)) * NC + c) * RW;
                        if (we) row[xe] = od[0];
                        if (wo) row[xo] = od[1];
                    }
                }
            }
            if (emit) {
                have = max(have, r0 + 2);
                low = min(low, r0);
            }
            q += REV ? -1 : 1;
        };

        // final-level pipeline state
        T s0[NS][NC][2], s1[NS][NC][2], s2[NS][NC][2], s3[NS][NC][2], s4[NS][NC][2];
#pragma unroll
        for (int h = 0; h < NS; ++h)
#pragma unroll
            for (int c = 0; c < NC; ++c)
#pragma unroll
                for (int e = 0; e < 2; ++e)
                    s0[h][c][e] = s1[h][c][e] = s2[h][c][e] = s3[h][c][e] = s4[h][c][e] = 0.0f;

        uint32_t slot[D][NS][NC][3];
#pragma unroll
        for (int d = 0; d < D; ++d) fetch(REV ? max(ke - d, ks) : min(ks + d, ke), slot[d]);
        // the LL samples of the next pair, read from the ring one pair ahead
        // (the inner level runs as far as that pair needs, two steps at a
        // time; a step past the pipeline's range emits nothing)
        T llnext[NS][NC];
        auto read_ll = [&](int k) {
            const int rk = ll_row(min(max(k, ks), ke));
            if constexpr (REV) {
#pragma unroll 1
                while (low > rk && q >= iqs) {
                    inner_step(islot[0]);
                    inner_step(islot[1]);
                }
            } else {
#pragma unroll 1
                while (have <= rk && q <= iqe) {
                    inner_step(islot[0]);
                    inner_step(islot[1]);
                }
            }
            const T *row = ring + (rk & (kRingRows - 1)) * NC * RW;
#pragma unroll
            for (int h = 0; h < NS; ++h)
#pragma unroll
                for (int c = 0; c < NC; ++c) llnext[h][c] = row[c * RW + xl[h]];
        };
        read_ll(REV ? ke : ks);
        const int rounds = (ke - ks + D) / D;
        for (int it = 0; it < rounds; ++it) {
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const int k = REV ? ke - (it * D + d) : ks + it * D + d;
                T smp[NS][NC][4];
#pragma unroll
                for (int h = 0; h < NS; ++h)
#pragma unroll
                    for (int c = 0; c < NC; ++c) {
                        if constexpr (FILT == 97) smp[h][c][0] = __fmul_rn(llnext[h][c], gn[c][0]);
                        else smp[h][c][0] = llnext[h][c];
                        smp[h][c][1] = to_sample<FILT, SK_DET>(slot[d][h][c][0], det_float, gn[c][1]);
                        smp[h][c][2] = to_sample<FILT, SK_DET>(slot[d][h][c][1], det_float, gn[c][2]);
                        smp[h][c][3] = to_sample<FILT, SK_DET>(slot[d][h][c][2], det_float, gn[c][3]);
                    }
#pragma unroll
                for (int h = 0; h < NS; ++h) pin_samples(smp[h]);
                fetch(REV ? max(k - D, ks) : min(k + D, ke), slot[d]);
                read_ll(REV ? k - 1 : k + 1);

                T evn[NS][2][NC], odd[NS][2][NC];
#pragma unroll
                for (int h = 0; h < NS; ++h)
#pragma unroll
                    for (int c = 0; c < NC; ++c) {
                        T tl = smp[h][c][0], th = smp[h][c][1], bl = smp[h][c][2], bh = smp[h][c][3];
                        hlift<FILT>(tl, th);
                        hlift<FILT>(bl, bh);
                        const T Lk[2] = {tl, th}, Hk[2] = {bl, bh};
#pragma unroll
                        for (int e = 0; e < 2; ++e) {
                            if constexpr (FILT == 97)
                                lift_step(dir, Lk[e], Hk[e], s0[h][c][e], s1[h][c][e], s2[h][c][e],
                                          s3[h][c][e], s4[h][c][e], evn[h][e][c], odd[h][e][c]);
                            else
                                lift53(dir, Lk[e], Hk[e], s0[h][c][e], s1[h][c][e], s2[h][c][e],
                                       s3[h][c][e], evn[h][e][c], odd[h][e][c], k == kq, nh_even);
                        }
                    }
                const int ko = REV ? k + LAG : k - LAG;
                if (ko >= kout0 && ko <= kout1) {
                    if (full && 2 * ko >= b0 && 2 * ko + 1 < b1) {
#pragma unroll
                        for (int h = 0; h < NS; ++h) {
                            em.put2_full(2 * ko, ce[h], evn[h][0], evn[h][1], ce_ok[h], co_ok[h]);
                            em.put2_full(2 * ko + 1, ce[h], odd[h][0], odd[h][1], ce_ok[h], co_ok[h]);
                        }
                    } else {
#pragma unroll
                        for (int h = 0; h < NS; ++h) {
                            if (2 * ko >= b0 && 2 * ko < b1)
                                em.put2(2 * ko, ce[h], evn[h][0], evn[h][1], ce_ok[h], co_ok[h]);
                            if (2 * ko + 1 >= b0 && 2 * ko + 1 < b1)
                                em.put2(2 * ko + 1, ce[h], odd[h][0], odd[h][1], ce_ok[h], co_ok[h]);
                        }
                    }
                }
                __builtin_amdgcn_sched_barrier(0);
            }
        }
    };
    // odd segments run bottom-up, except 5/3 ones whose final or inner
    // pipeline range holds the pair of the reference's bottom-edge rule
    const bool rev = (seg & 1) && (FILT == 97 || ((kq < ks || kq > ke) && (ikq < iqs || ikq > iqe)));
    if (rev)
        run(std::true_type{});
    else
        run(std::false_type{});
}

// Kernel variants: V = (filt == 97) * 32 + sub * 4 + kd.
// sub: 0 plane (1 comp), 1 merge (1 comp), 2..5 RGB of layout
// PL_BGR8..PL_BGRA16 (3 comps), 6 merge (2-3 comps), 7 merge (4 comps).
// kd: coefficient encodings (9/7 planes and RGB only; KD_ANY otherwise).
// Thumbnail jobs (lev < 0): V = 64 + (filt == 97) * 8 + sub.
// Two-level jobs: V = 96 + 64 * (filt == 53) + 32 * (strips == 2) + sub * 4
// + kd, kd from the inner level's LL encoding and the detail bands' (9/7).
__host__ __device__ inline int variant_of(const LevelJob &J)
{
    int sub;
    if (J.out_kind == OUT_SCRATCH || J.out_kind == OUT_PLANE) sub = 0;
    else if (J.out_kind == OUT_RGB) sub = 2 + pixel_layout(J.fmt);
    else sub = J.nc == 1 ? 1 : (J.nc <= 3 ? 6 : 7);
    if (J.lev < 0) return 64 + (J.filt == 97 ? 8 : 0) + sub;
    int kd = KD_ANY;
    if (J.filt == 97 && (sub == 0 || (sub >= 2 && sub <= 5))) {
        const bool llf = (J.fuse2 ? J.ill_kind : J.ll_kind) == SRC_FLOAT, detf = J.det_kind == SRC_FLOAT;
        kd = (llf && detf) ? KD_FF : (llf ? KD_FS : (detf ? KD_ANY : KD_SS));
    }
    if (J.fuse2) return 96 + (J.filt == 53 ? 64 : 0) + (J.fuse2 == 2 ? 32 : 0) + sub * 4 + kd;
    return (J.filt == 97 ? 32 : 0) + sub * 4 + kd;
}

// the two-level variants instantiated, as X(V); a final level whose
// variant is not listed keeps the two-launch form.  There are none for one
// component: its inner level is a third of the work of three components'
// but runs on the same half-idle lanes, and on the 256 mixed tiles the
// two launches took 55 us against 72-84 us for the two-level form.
// Two-strip kernels exist for 24-bit and 48-bit BGR; 5/3 two-level
// kernels for those two outputs (168, 176; two strips 200, 208).
#define BMI_FOR_EACH_FUSED_VARIANT(X)                                                       \
    X(104) X(105) X(106) X(107) X(109) X(113) X(120) X(137) X(138) X(139) X(145)             \
    X(168) X(176) X(200) X(208)

__host__ __device__ inline bool fused_variant_available(int v)
{
    switch (v) {
#define BMI_FUSED_CASE(V) case V:
        BMI_FOR_EACH_FUSED_VARIANT(BMI_FUSED_CASE)
#undef BMI_FUSED_CASE
        return true;
    default:
        return false;
    }
}

// lockstep components and strips of a two-level variant (its LDS ring
// per wave holds kRingRows rows of 64 x strips floats per component)
__host__ __device__ constexpr int fused_variant_nc(int v)
{
    return (((v - 96) >> 2) & 7) <= 1 ? 1 : ((((v - 96) >> 2) & 7) == 7 ? 4 : 3);
}
__host__ __device__ constexpr int fused_variant_strips(int v) { return ((v - 96) & 32) ? 2 : 1; }
__host__ __device__ constexpr int fused_variant_filt(int v) { return ((v - 96) & 64) ? 53 : 97; }

template <int V>
__device__ __attribute__((always_inline)) inline void run_variant(const LevelJob &__restrict__ J,
                                                                 int strip, int seg, int lane)
{
    if constexpr (V >= 64) {
        constexpr int TF = (V & 8) ? 97 : 53;
        constexpr int TS = V & 7;
        if constexpr (TS == 0) thumb_item<TF, 1, K_PLANE, PL_ANY>(J, strip, seg, lane);
        else if constexpr (TS == 1) thumb_item<TF, 1, K_MERGE, PL_ANY>(J, strip, seg, lane);
        else if constexpr (TS >= 2 && TS <= 5) thumb_item<TF, 3, K_RGB, TS - 2>(J, strip, seg, lane);
        else if constexpr (TS == 6) thumb_item<TF, 3, K_MERGE, PL_ANY>(J, strip, seg, lane);
        else thumb_item<TF, 4, K_MERGE, PL_ANY>(J, strip, seg, lane);
        return;
    }
    constexpr int FILT = (V >= 32) ? 97 : 53;
    constexpr int SUB = (V >> 2) & 7;
    constexpr int KD = V & 3;
    if constexpr (SUB == 0) process_item<FILT, 1, K_PLANE, PL_ANY, KD>(J, strip, seg, lane);
    else if constexpr (SUB == 1) process_item<FILT, 1, K_MERGE, PL_ANY, KD>(J, strip, seg, lane);
    else if constexpr (SUB >= 2 && SUB <= 5) process_item<FILT, 3, K_RGB, SUB - 2, KD>(J, strip, seg, lane);
    else if constexpr (SUB == 6) process_item<FILT, 3, K_MERGE, PL_ANY, KD>(J, strip, seg, lane);
    else process_item<FILT, 4, K_MERGE, PL_ANY, KD>(J, strip, seg, lane);
}

// every instantiated variant, as X(V)
#define BMI_FOR_EACH_VARIANT(X)                                                             \
    X(0) X(4) X(8) X(12) X(16) X(20) X(24) X(28)                                            \
    X(32) X(33) X(34) X(35) X(36) X(40) X(41) X(42) X(43) X(44) X(45) X(46) X(47) X(48)     \
    X(49) X(50) X(51) X(52) X(53) X(54) X(55) X(56) X(60)                               \
    X(64) X(65) X(66) X(67) X(68) X(69) X(70) X(71)                                         \
    X(72) X(73) X(74) X(75) X(76) X(77) X(78) X(79)

__device__ __attribute__((always_inline)) inline void run_item(
    const LevelJob *__restrict__ jobs, const WorkItem &it, int lane)
{
    const LevelJob &__restrict__ J = jobs[it.job];
    // consecutive items are neighbouring strips of one row segment: waves
    // running together read whole image rows, sharing the cache lines that
    // straddle strip boundaries
    const int strip = it.local % J.nstrips;
    const int seg = it.local / J.nstrips;
    switch (variant_of(J)) {
#define BMI_RUN_CASE(V) \
    case V: run_variant<V>(J, strip, seg, lane); break;
        BMI_FOR_EACH_VARIANT(BMI_RUN_CASE)
#undef BMI_RUN_CASE
    default: break;
    }
}

// spin limit: a dependency that never completes (a bug) ends the wait and
// raises the error flag instead of hanging the device
constexpr int kSpinLimit = 1 << 22;

// Counter block layout: [0, njobs) per-job completion counts, then the
// queue head, the count of exited waves and the sticky error flag.  The
// last wave to leave zeroes the counts, head and exit count, so the next
// launch (e.g. a graph replay) starts from a clean queue without a
// separate reset kernel.
__global__ void __launch_bounds__(256) fused_persistent_kernel(const LevelJob *jobs,
                                                                const WorkItem *items, int n_items,
                                                                int *counters, int njobs,
                                                                int total_waves)
{
    // All control flow below depends only on wave-uniform values
    // (readfirstlane results and kernel arguments).  Every lane executes the
    // atomics, lane 0 contributing 1 and the others 0, so no branch is
    // lane-divergent and the compiler cannot split the wave across loop
    // iterations around the queue operations.
    int *head = counters + njobs;
    int *exited = counters + njobs + 1;
    int *error_flag = counters + njobs + 2;
    const int lane = threadIdx.x & 63;
    const int one = (lane == 0) ? 1 : 0;
    for (;;) {
        const int i = __builtin_amdgcn_readfirstlane(atomicAdd(head, one));
        if (i >= n_items) break;
        const int job = __builtin_amdgcn_readfirstlane(items[i].job);
        const int local = __builtin_amdgcn_readfirstlane(items[i].local);
        const LevelJob &J = jobs[job];
        const int ndeps = J.ndeps;
        for (int d = 0; d < ndeps; ++d) {
            const int dj = J.deps[d];
            const int need = jobs[dj].n_items;
            for (int spins = 0;; ++spins) {
                const int done = __builtin_amdgcn_readfirstlane(__hip_atomic_load(
                    &counters[dj], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
                if (done >= need) break;
                const int err = __builtin_amdgcn_readfirstlane(
                    __hip_atomic_load(error_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
                if (err || spins > kSpinLimit) {
                    atomicOr(error_flag, one);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        run_item(jobs, WorkItem{job, local}, lane);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_fetch_add(&counters[job], one, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
    const int prev = __builtin_amdgcn_readfirstlane(
        __hip_atomic_fetch_add(exited, one, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT));
    if (prev == total_waves - 1) {
        for (int j = lane; j < njobs; j += 64)
            __hip_atomic_store(&counters[j], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(head, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(exited, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    }
}

// level-step mode: one launch per dependency step and item variant (no
// waiting); each variant gets its own register budget
template <int V>
__global__ void __launch_bounds__(256) fused_step_kernel(const LevelJob *__restrict__ jobs,
                                                         const WorkItem *__restrict__ items,
                                                         int first, int count, int remap)
{
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    // Workgroups are dispatched to the 8 XCDs round robin (workgroup g on
    // XCD g % 8).  Give each XCD a contiguous range of items instead, so
    // the rows and columns that neighbouring items both read (the lifting
    // halo) are fetched through one XCD's L2.
    int g = blockIdx.x;
    if (remap) {
        const int G = gridDim.x, x = g & 7, j = g >> 3, q = G >> 3, r = G & 7;
        g = x * q + min(x, r) + j;
    }
    const int i = g * (blockDim.x >> 6) + wave;
    if (i >= count) return;
    const int job = __builtin_amdgcn_readfirstlane(items[first + i].job);
    const int local = __builtin_amdgcn_readfirstlane(items[first + i].local);
    const LevelJob &__restrict__ J = jobs[job];
    run_variant<V>(J, local % J.nstrips, local / J.nstrips, lane);
}

// level-step mode, two-level jobs: the inner level's rows live in the
// dynamic LDS of each wave (kRingRows x components x 64 x strips floats)
// (Capping it at 128 VGPRs for four waves per SIMD spills and runs about
// 1.5x slower.)
template <int V>
__global__ void __launch_bounds__(256) fused2_step_kernel(const LevelJob *__restrict__ jobs,
                                                          const WorkItem *__restrict__ items,
                                                          int first, int count, int remap)
{
    extern __shared__ float bmi_ring_lds[];
    constexpr int SUB = ((V - 96) >> 2) & 7;
    constexpr int KD = (V - 96) & 3;
    constexpr int NC = fused_variant_nc(V);
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    int g = blockIdx.x;
    if (remap) {
        const int G = gridDim.x, x = g & 7, j = g >> 3, q = G >> 3, r = G & 7;
        g = x * q + min(x, r) + j;
    }
    const int i = g * (blockDim.x >> 6) + wave;
    if (i >= count) return;
    const int job = __builtin_amdgcn_readfirstlane(items[first + i].job);
    const int local = __builtin_amdgcn_readfirstlane(items[first + i].local);
    const LevelJob &__restrict__ J = jobs[job];
    constexpr int NS = fused_variant_strips(V);
    constexpr int F = fused_variant_filt(V);
    float *ring = bmi_ring_lds + wave * (kRingRows * NC * 64 * NS);
    const int strip = local % J.nstrips, seg = local / J.nstrips;
    // a tile whose columns start and end even and whose packed output has
    // an even pitch (the common case) takes the stores without the
    // one-pixel-lane and odd-row cases
    constexpr int OUTK = SUB == 0 ? K_PLANE : ((SUB >= 2 && SUB <= 5) ? K_RGB : K_MERGE);
    const bool simple = !(J.a0 & 1) && !(J.a1 & 1) && (OUTK == K_PLANE || !(J.out_pitch_px & 1));
    auto item = [&](auto simple_t) {
        constexpr bool S = decltype(simple_t)::value;
        if constexpr (SUB == 0) process_item2<F, 1, K_PLANE, PL_ANY, KD, NS, S>(J, strip, seg, lane, ring);
        else if constexpr (SUB == 1) process_item2<F, 1, K_MERGE, PL_ANY, KD, NS, S>(J, strip, seg, lane, ring);
        else if constexpr (SUB >= 2 && SUB <= 5)
            process_item2<F, 3, K_RGB, SUB - 2, KD, NS, S>(J, strip, seg, lane, ring);
        else if constexpr (SUB == 6) process_item2<F, 3, K_MERGE, PL_ANY, KD, NS, S>(J, strip, seg, lane, ring);
        else process_item2<F, 4, K_MERGE, PL_ANY, KD, NS, S>(J, strip, seg, lane, ring);
    };
    if (simple)
        item(std::true_type{});
    else
        item(std::false_type{});
}

// level-step mode with all kernel variants of a step in one launch (each
// wave dispatches on its item's variant): one launch per step instead of
// one per step and variant, at the register budget of the largest variant
__global__ void __launch_bounds__(256) fused_step_any_kernel(const LevelJob *__restrict__ jobs,
                                                             const WorkItem *__restrict__ items,
                                                             int first, int count, int remap)
{
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    int g = blockIdx.x;
    if (remap) {
        const int G = gridDim.x, x = g & 7, j = g >> 3, q = G >> 3, r = G & 7;
        g = x * q + min(x, r) + j;
    }
    const int i = g * (blockDim.x >> 6) + wave;
    if (i >= count) return;
    const int job = __builtin_amdgcn_readfirstlane(items[first + i].job);
    const int local = __builtin_amdgcn_readfirstlane(items[first + i].local);
    run_item(jobs, WorkItem{job, local}, lane);
}

}  // namespace bmi

// ---------------------------------------------------------------------------
// host side: plans
// ---------------------------------------------------------------------------

using namespace bmi;

struct BmiBatchPlan {
    std::vector<BmiTileDesc> tiles;
    std::vector<LevelJob> jobs;
    std::vector<WorkItem> items;
    std::vector<int> step_first, step_count;  // level-step mode
    struct Launch {
        int first, count, variant, step;
        int wpg = 0;  // waves per workgroup of this launch
    };
    std::vector<Launch> launches;  // level-step mode: per step and variant
    std::vector<int> job_tile;                // tile of each job
    struct {
        int seg_rows, waves_per_group, groups_per_cu, persistent, min_waves_per_cu, fixed_seg, fuse2,
            from_coarse;
    } tuning;
    std::vector<char> tile_fused;  // per tile: its last two levels run as one two-level step
    LevelJob *d_jobs = nullptr;
    WorkItem *d_items = nullptr;
    int *d_counters = nullptr;  // [njobs] completion counters; [njobs] queue head,
                                // [njobs+1] exited waves, [njobs+2] error flag
                                // (persistent mode)
    void *scratch = nullptr;     // device base of the plan's memory (scratch first)
    size_t scratch_bytes = 0;
    size_t device_bytes = 0;     // scratch + job table + items + counters
    bool owns_device = false;    // memory came from hipMalloc (else a caller's arena)
    std::vector<size_t> tile_scratch_off;  // per tile: base offset
    std::vector<int> tile_scratch_pitch;
    std::vector<size_t> tile_scratch_plane;  // bytes per plane
    int device = 0;
    int num_cus = 256;
    hipGraphExec_t graph_exec = nullptr;
    hipGraph_t graph = nullptr;
};

namespace {

int out_kind_for(const BmiTileDesc &t)
{
    switch (t.out_format) {
    case IMAGE_FOTMAT_NONE:
        return OUT_PLANE;
    case IMAGE_FOTMAT_32BITS_ARGB:
    case IMAGE_FOTMAT_48BITS_RGB:
    case IMAGE_FOTMAT_24BITS_RGB:
    case IMAGE_FOTMAT_64BITS_ARGB:
        return OUT_RGB;
    case IMAGE_FOTMAT_32BITS_4COMPONENTS:
    case IMAGE_FOTMAT_64BITS_4COMPONENTS:
    case IMAGE_FOTMAT_48BITS_3COMPONENTS:
    case IMAGE_FOTMAT_24BITS_3COMPONENTS:
        return OUT_MERGE;
    default:
        return -1;
    }
}

int validate_tile(const BmiTileDesc &t)
{
    if (t.width < 1 || t.height < 1 || t.width > 32767 || t.height > 32767) return BMI_E_INVALID;
    if (t.x0 < 0 || t.y0 < 0) return BMI_E_INVALID;
    if (t.levels < 1 || t.levels > MAX_DWT_LEVEL) return BMI_E_INVALID;
    if (t.top_level < -1 || t.top_level > t.levels) return BMI_E_INVALID;
    if (t.ncomp < 1 || t.ncomp > MAX_COMPONENT) return BMI_E_INVALID;
    if (t.method != Ckernels_W5X3 && t.method != Ckernels_W9X7) return BMI_E_INVALID;
    if (t.method == Ckernels_W9X7 && t.coef_kind != BMI_COEF_SIGNMAG &&
        t.coef_kind != BMI_COEF_FLOAT)
        return BMI_E_INVALID;
    if (t.bit_depth < 1 || t.bit_depth > 16) return BMI_E_INVALID;
    if (t.coef_pitch < t.width) return BMI_E_INVALID;
    // planes addressed with 32-bit offsets: at most 2^31 bytes each
    if ((long long)t.coef_pitch * t.height * 4 > (1ll << 31)) return BMI_E_UNSUPPORTED;
    if (t.out_format == IMAGE_FOTMAT_NONE && (long long)t.planes_pitch * t.height * 4 > (1ll << 31))
        return BMI_E_UNSUPPORTED;
    const int ok = out_kind_for(t);
    if (ok < 0) return BMI_E_UNSUPPORTED;
    if (ok == OUT_RGB && t.ncomp < 3) return BMI_E_UNSUPPORTED;
    if (ok == OUT_RGB && (t.check_out < -1 || t.check_out > 2)) return BMI_E_INVALID;
    if (ok == OUT_MERGE && (t.check_out < -1 || t.check_out >= t.ncomp)) return BMI_E_INVALID;
    if (ok == OUT_PLANE) {
        for (int c = 0; c < t.ncomp; ++c)
            if (!t.planes_out[c]) return BMI_E_INVALID;
    } else if (!t.out || t.out_pitch_px < 1 || t.out_x < 0 || t.out_y < 0) {
        return BMI_E_INVALID;
    } else if ((long long)t.out_pitch_px * (t.out_y + t.height) * 8 > (1ll << 31)) {
        return BMI_E_UNSUPPORTED;  // packed output addressed with 32-bit offsets
    }
    if (t.r53_mode != BMI_R53_REFERENCE && t.r53_mode != BMI_R53_TEXTBOOK) return BMI_E_INVALID;
    return BMI_OK;
}

void level_span(int t0, int n, int L, int lev, int *a0, int *a1)
{
    const int s = L - 1 - lev;
    *a0 = ceil_shift(t0, s);
    *a1 = ceil_shift(t0 + n, s);
}

// components whose planes the final level reads
void needed_comps(const BmiTileDesc &t, int ok, int *list, int *n)
{
    *n = 0;
    if (ok == OUT_RGB) {
        for (int c = 0; c < 3; ++c) list[(*n)++] = c;
    } else if (ok == OUT_MERGE && t.check_out >= 0) {
        list[(*n)++] = t.check_out;
    } else {
        for (int c = 0; c < t.ncomp; ++c) list[(*n)++] = c;
    }
}

}  // namespace

int bmi_apply_ignore(const BmiTileDesc &in, BmiTileDesc *out)
{
    BmiTileDesc t = in;
    const unsigned mask = in.ignore_mask & ((1u << MAX_COMPONENT) - 1);
    if (in.ncomp < 1 || in.ncomp > MAX_COMPONENT) return BMI_E_INVALID;
    if (mask) {
        int n = 0, co = -1;
        for (int c = 0; c < in.ncomp; ++c) {
            if ((mask >> c) & 1) continue;
            t.coef[n] = in.coef[c];
            t.planes_out[n] = in.planes_out[c];
            memcpy(t.steps[n], in.steps[c], sizeof(t.steps[n]));
            if (in.check_out == c) co = n;
            ++n;
        }
        if (n == 0) return BMI_E_INVALID;
        for (int c = n; c < MAX_COMPONENT; ++c) {
            t.coef[c] = nullptr;
            t.planes_out[c] = nullptr;
        }
        t.ncomp = n;
        // an ignored (or missing) check-out component means no check-out
        // (bmi_cpu_dwt_decoder.cpp:1222-1225); RGB check-outs name output
        // channels, which keep their meaning
        const int ok = out_kind_for(in);
        if (ok != OUT_RGB) t.check_out = in.check_out >= 0 ? co : -1;
        if (ok == OUT_RGB && n < 3) return BMI_E_UNSUPPORTED;
    }
    t.ignore_mask = 0;
    *out = t;
    return BMI_OK;
}

// A tile runs its last two synthesis levels as one two-level step when
// both levels are 9/7, at least two samples in each direction, its final
// level spans at least `min_side` samples each way, and a two-level kernel
// exists for its final level's variant.
static bool tile_fuses(const BmiTileDesc &t, int min_side)
{
    const int L = t.levels;
    const int Ls = (t.top_level < 0) ? L : t.top_level;
    if (Ls < 2) return false;
    for (int lev = Ls - 2; lev < Ls; ++lev) {
        int a0, a1, b0, b1;
        level_span(t.x0, t.width, L, lev, &a0, &a1);
        level_span(t.y0, t.height, L, lev, &b0, &b1);
        if (a1 - a0 < 2 || b1 - b0 < 2) return false;
        if (lev == Ls - 1 && (a1 - a0 < min_side || b1 - b0 < min_side)) return false;
    }
    const int ok = out_kind_for(t);
    int comps[4], n = 0;
    needed_comps(t, ok, comps, &n);
    LevelJob J;
    memset(&J, 0, sizeof(J));
    J.filt = (t.method == Ckernels_W9X7) ? 97 : 53;
    J.lev = Ls - 1;
    J.out_kind = ok;
    J.fmt = t.out_format;
    J.nc = (ok == OUT_PLANE) ? 1 : n;
    J.fuse2 = 1;
    J.det_kind = J.filt == 53 ? SRC_INT : (t.coef_kind == BMI_COEF_FLOAT ? SRC_FLOAT : SRC_SIGNMAG);
    J.ill_kind = (Ls - 2 == 0) ? J.det_kind : (J.filt == 97 ? SRC_FLOAT : SRC_INT);
    return fused_variant_available(variant_of(J));
}

static int default_seg_rows(const BmiTuning *t) { return (t && t->seg_rows > 0) ? t->seg_rows : 32; }

static size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

// Host half of plan construction: validates the tiles and builds the job
// table and work items.  Device pointers into the plan's own memory
// (scratch planes) are stored as offsets from address 0 and relocated by
// bmi_plan_bind once the memory exists.
int bmi_plan_build(BmiBatchPlan **out_plan, const BmiTileDesc *tiles, int ntiles,
                   const BmiTuning *tuning)
{
    if (!out_plan || !tiles || ntiles < 1) return bmi_fail(BMI_E_INVALID, "batch_plan_create: bad args");
    *out_plan = nullptr;
    // ignored components are dropped here: everything below sees the tile
    // with its remaining components in the lower slots
    std::vector<BmiTileDesc> eff(tiles, tiles + ntiles);
    for (int i = 0; i < ntiles; ++i) {
        int rc = bmi_apply_ignore(tiles[i], &eff[i]);
        if (!rc) rc = validate_tile(eff[i]);
        if (rc) return bmi_fail(rc, "batch_plan_create: tile %d rejected", i);
    }
    tiles = eff.data();
    BmiBatchPlan *P = new BmiBatchPlan();
    P->tiles.assign(tiles, tiles + ntiles);
    P->tuning.seg_rows = default_seg_rows(tuning);
    if (P->tuning.seg_rows & 1) P->tuning.seg_rows += 1;
    const bool wpg_auto = !(tuning && tuning->waves_per_group > 0);
    P->tuning.waves_per_group = wpg_auto ? 4 : tuning->waves_per_group;
    P->tuning.groups_per_cu = (tuning && tuning->groups_per_cu > 0) ? tuning->groups_per_cu : 8;
    P->tuning.persistent = tuning ? tuning->persistent : 0;
    P->tuning.min_waves_per_cu = (tuning && tuning->min_waves_per_cu > 0) ? tuning->min_waves_per_cu : 4;
    P->tuning.fixed_seg = (tuning && tuning->fixed_seg) ? 1 : 0;
    P->tuning.fuse2 = tuning ? (tuning->fuse2 < 0 ? -1 : tuning->fuse2) : 0;
    P->tuning.from_coarse = (tuning && tuning->from_coarse) ? 1 : 0;
    if (P->tuning.waves_per_group > 4) P->tuning.waves_per_group = 4;
    if (hipGetDevice(&P->device) != hipSuccess) P->device = 0;
    {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, P->device) ==
                hipSuccess &&
            cus > 0)
            P->num_cus = cus;
    }

    // which tiles run their last two levels as one two-level step
    P->tile_fused.assign(ntiles, 0);
    const bool fuse_mode = P->tuning.fuse2 >= 0 && P->tuning.persistent != 1 && P->tuning.persistent != 2;
    for (int ti = 0; ti < ntiles && fuse_mode; ++ti)
        P->tile_fused[ti] = tile_fuses(tiles[ti], P->tuning.fuse2) ? 1 : 0;

    // scratch layout: per tile, per needed component, two planes sized for
    // the largest intermediate level written to HBM (a two-level step keeps
    // its inner level in LDS)
    size_t total = 0;
    P->tile_scratch_off.resize(ntiles);
    P->tile_scratch_pitch.resize(ntiles);
    P->tile_scratch_plane.resize(ntiles);
    for (int ti = 0; ti < ntiles; ++ti) {
        const BmiTileDesc &t = tiles[ti];
        const int L = t.levels;
        const int Ls = (t.top_level < 0) ? L : t.top_level;
        const int Lw = Ls - (P->tile_fused[ti] ? 1 : 0);
        int maxw = 0, maxh = 0;
        for (int lev = 0; lev + 1 < Lw; ++lev) {
            int a0, a1, b0, b1;
            level_span(t.x0, t.width, L, lev, &a0, &a1);
            level_span(t.y0, t.height, L, lev, &b0, &b1);
            maxw = std::max(maxw, a1 - a0);
            maxh = std::max(maxh, b1 - b0);
        }
        const int pitch = (maxw + 63) & ~63;
        const size_t plane = (size_t)pitch * maxh * 4;
        P->tile_scratch_off[ti] = total;
        P->tile_scratch_pitch[ti] = pitch;
        P->tile_scratch_plane[ti] = plane;
        total += plane * 2 * (size_t)t.ncomp;
        total = (total + 255) & ~(size_t)255;
    }
    P->scratch_bytes = total;
    P->scratch = nullptr;  // offsets until bmi_plan_bind

    // jobs, ordered by level step then tile; items follow job order
    int max_steps = 0;
    for (int ti = 0; ti < ntiles; ++ti)
        max_steps = std::max(max_steps, tiles[ti].top_level < 0 ? tiles[ti].levels : tiles[ti].top_level);
    for (int ti = 0; ti < ntiles; ++ti)
        if (tiles[ti].top_level == 0) max_steps = std::max(max_steps, 1);  // thumbnails run in step 0
    std::vector<std::vector<int>> job_of(ntiles);  // [tile][lev*4+comp] -> job
    for (int ti = 0; ti < ntiles; ++ti) job_of[ti].assign(MAX_DWT_LEVEL * 4, -1);
    const int seg_rows = P->tuning.seg_rows;
    P->step_first.assign(max_steps, 0);
    P->step_count.assign(max_steps, 0);
    for (int step = 0; step < max_steps; ++step) {
        P->step_first[step] = (int)P->items.size();
        const int step_job0 = (int)P->jobs.size();
        for (int ti = 0; ti < ntiles; ++ti) {
            const BmiTileDesc &t = tiles[ti];
            const int L = t.levels;
            const int Ls = (t.top_level < 0) ? L : t.top_level;
            // the tile's level in this step: steps aligned at every tile's
            // finest level (by default: the final levels, most of the work,
            // share the last step; 256 mixed tiles 0.49 -> 0.43 ms) or at
            // the coarsest
            const int lev = P->tuning.from_coarse ? step : step - (max_steps - Ls);
            if (Ls == 0 && step == 0) {
                // thumbnail: the LL band of the coarsest level, no synthesis
                const int filt = (t.method == Ckernels_W9X7) ? 97 : 53;
                const int ok = out_kind_for(t);
                LevelJob J;
                memset(&J, 0, sizeof(J));
                J.a0 = ceil_shift(t.x0, L);
                J.a1 = ceil_shift(t.x0 + t.width, L);
                J.b0 = ceil_shift(t.y0, L);
                J.b1 = ceil_shift(t.y0 + t.height, L);
                J.lw = J.a1 - J.a0;
                J.lh = J.b1 - J.b0;
                J.filt = filt;
                J.lev = -1;
                J.out_kind = ok;
                J.ll_kind = (filt == 97) ? (t.coef_kind == BMI_COEF_FLOAT ? SRC_FLOAT : SRC_SIGNMAG)
                                         : SRC_INT;
                J.det_kind = J.ll_kind;
                J.ll_pitch = J.det_pitch = t.coef_pitch;
                J.nstrips = (J.lw + 63) / 64;
                J.nsegs = J.lh;
                J.seg_pairs = 1;
                J.n_items = J.nstrips * J.nsegs;
                J.fmt = t.out_format;
                J.bit_depth = t.bit_depth;
                J.check_out = t.check_out;
                J.out = t.out;
                J.out_pitch_px = t.out_pitch_px;
                J.out_x = t.out_x;
                J.out_y = t.out_y;
                int comps[4], ncomps = 0;
                needed_comps(t, ok, comps, &ncomps);
                auto fill = [&](LevelJob &K, int slot, int c) {
                    K.comp[slot] = c;
                    K.ll[slot] = K.det[slot] = t.coef[c];
                    K.gain[slot][0] = (filt == 97) ? t.steps[c][0] : 1.0f;
                };
                if (ok == OUT_PLANE) {
                    for (int i = 0; i < ncomps; ++i) {
                        LevelJob K = J;
                        K.nc = 1;
                        fill(K, 0, comps[i]);
                        K.dst[0] = t.planes_out[comps[i]];
                        K.dst_pitch = t.planes_pitch;
                        P->jobs.push_back(K);
                        P->job_tile.push_back(ti);
                    }
                } else {
                    J.nc = ncomps;
                    for (int i = 0; i < 4; ++i) fill(J, i, comps[i < ncomps ? i : 0]);
                    if (ok == OUT_MERGE && t.check_out >= 0) J.check_out = 0;
                    P->jobs.push_back(J);
                    P->job_tile.push_back(ti);
                }
                continue;
            }
            if (lev >= Ls) continue;
            if (lev < 0) continue;
            const bool fused = P->tile_fused[ti] != 0;
            if (fused && lev == Ls - 2) continue;  // synthesised inside the final level's items
            const bool final_level = (lev == Ls - 1);
            const int ok = final_level ? out_kind_for(t) : (int)OUT_SCRATCH;
            const int filt = (t.method == Ckernels_W9X7) ? 97 : 53;
            int a0, a1, b0, b1;
            level_span(t.x0, t.width, L, lev, &a0, &a1);
            level_span(t.y0, t.height, L, lev, &b0, &b1);
            LevelJob base;
            memset(&base, 0, sizeof(base));
            base.a0 = a0;
            base.a1 = a1;
            base.b0 = b0;
            base.b1 = b1;
            base.lw = ceil_div2(a1) - ceil_div2(a0);
            base.lh = ceil_div2(b1) - ceil_div2(b0);
            base.cl0 = ceil_div2(a0);
            base.ch0 = floor_div2(a0);
            base.rl0 = ceil_div2(b0);
            base.rh0 = floor_div2(b0);
            const int owp = (filt == 97) ? 60 : 62;
            const int ncols = a1 - (a0 & ~1);
            base.nstrips = (ncols + 2 * owp - 1) / (2 * owp);
            const int nrows = b1 - (b0 & ~1);
            base.seg_pairs = seg_rows / 2;
            base.nsegs = (nrows + seg_rows - 1) / seg_rows;
            // An empty level (a tile of a few samples at an odd offset can
            // have coarse levels of no rows or no columns) gets a job with
            // no items: the kernels' row and column reflections assume at
            // least one sample.  The level above then has at most one row
            // (or column), a high-band one, and reads none of its output.
            if (a1 <= a0 || b1 <= b0) base.nstrips = base.nsegs = 0;
            base.filt = filt;
            base.lev = lev;
            base.out_kind = ok;
            base.quirk = (filt == 53 && t.r53_mode == BMI_R53_REFERENCE && !(b0 & 1) &&
                          (b1 - b0) >= 2)
                             ? 1
                             : 0;
            base.det_kind = (filt == 97) ? (t.coef_kind == BMI_COEF_FLOAT ? SRC_FLOAT : SRC_SIGNMAG)
                                         : SRC_INT;
            base.ll_kind = (lev == 0) ? base.det_kind : (filt == 97 ? SRC_FLOAT : SRC_INT);
            base.det_pitch = t.coef_pitch;
            base.ll_pitch = (lev == 0) ? t.coef_pitch : P->tile_scratch_pitch[ti];
            base.n_items = base.nstrips * base.nsegs;
            base.fmt = t.out_format;
            base.bit_depth = t.bit_depth;
            base.check_out = t.check_out;
            base.out = t.out;
            base.out_pitch_px = t.out_pitch_px;
            base.out_x = t.out_x;
            base.out_y = t.out_y;

            auto scratch_plane = [&](int c, int which) -> void * {
                return (char *)P->scratch + P->tile_scratch_off[ti] +
                       ((size_t)c * 2 + which) * P->tile_scratch_plane[ti];
            };
            auto fill_comp = [&](LevelJob &J, int slot, int c) {
                J.comp[slot] = c;
                J.det[slot] = t.coef[c];
                J.ll[slot] = (lev == 0) ? t.coef[c] : scratch_plane(c, (lev - 1) & 1);
                const float *s = t.steps[c];
                if (filt == 97) {
                    J.gain[slot][0] = lev == 0 ? (s[0] * kLoGain) * kLoGain : kLoGain * kLoGain;
                    J.gain[slot][1] = (s[3 * lev + 1] * kHiGain) * kLoGain;
                    J.gain[slot][2] = (s[3 * lev + 2] * kLoGain) * kHiGain;
                    J.gain[slot][3] = (s[3 * lev + 3] * kHiGain) * kHiGain;
                } else {
                    J.gain[slot][0] = J.gain[slot][1] = J.gain[slot][2] = J.gain[slot][3] = 1.0f;
                }
            };
            auto add_deps = [&](LevelJob &J, int c) {
                const int src = lev - (fused && final_level ? 2 : 1);
                if (src < 0) return;
                const int dj = job_of[ti][src * 4 + c];
                if (dj >= 0 && J.ndeps < kMaxDeps) J.deps[J.ndeps++] = dj;
            };
            // a two-level final job: geometry, LL source and gains of the
            // level below (level lev - 1), which never gets a job of its own
            auto set_fuse_geom = [&](LevelJob &J) {
                if (!(fused && final_level)) return;
                int ia0, ia1, ib0, ib1;
                level_span(t.x0, t.width, L, lev - 1, &ia0, &ia1);
                level_span(t.y0, t.height, L, lev - 1, &ib0, &ib1);
                J.fuse2 = 1;
                J.ia0 = ia0;
                J.ia1 = ia1;
                J.ib0 = ib0;
                J.ib1 = ib1;
                J.ilw = ceil_div2(ia1) - ceil_div2(ia0);
                J.ilh = ceil_div2(ib1) - ceil_div2(ib0);
                J.icl0 = ceil_div2(ia0);
                J.ich0 = floor_div2(ia0);
                J.irl0 = ceil_div2(ib0);
                J.irh0 = floor_div2(ib0);
                J.ill_kind = (lev - 1 == 0) ? J.det_kind : (filt == 97 ? SRC_FLOAT : SRC_INT);
                J.ill_pitch = (lev - 1 == 0) ? t.coef_pitch : P->tile_scratch_pitch[ti];
                // the inner level's own 5/3 bottom-edge rule (as base.quirk)
                J.iquirk = (filt == 53 && t.r53_mode == BMI_R53_REFERENCE && !(ib0 & 1) &&
                            (ib1 - ib0) >= 2)
                               ? 1
                               : 0;
                // two strips per item where that kernel exists (faster for
                // one 4096x2160x3 image, 46.5 against 47.8 us, and for the
                // 68 three-component 9/7 tiles of the 256-tile batch, 111
                // against 126 us, although 230-column items leave more
                // lanes past the tile edges of narrow tiles)
                J.fuse2 = 2;
                if (!fused_variant_available(variant_of(J))) J.fuse2 = 1;
                const int w2 = 2 * fused2_out_pairs(J.fuse2);
                J.nstrips = (a1 - (a0 & ~1) + w2 - 1) / w2;
                J.n_items = J.nstrips * J.nsegs;
            };
            auto fill_fuse = [&](LevelJob &J, int slot, int c) {
                if (!(fused && final_level)) return;
                const int il = lev - 1;
                J.ll[slot] = nullptr;  // the final level's LL comes from LDS
                J.ill[slot] = (il == 0) ? t.coef[c] : scratch_plane(c, (il - 1) & 1);
                const float *st = t.steps[c];
                if (filt == 97) {
                    J.igain[slot][0] = il == 0 ? (st[0] * kLoGain) * kLoGain : kLoGain * kLoGain;
                    J.igain[slot][1] = (st[3 * il + 1] * kHiGain) * kLoGain;
                    J.igain[slot][2] = (st[3 * il + 2] * kLoGain) * kHiGain;
                    J.igain[slot][3] = (st[3 * il + 3] * kHiGain) * kHiGain;
                } else {
                    for (int g = 0; g < 4; ++g) J.igain[slot][g] = 1.0f;
                }
            };
            auto push_job = [&](const LevelJob &J, int key_comp) {
                const int jid = (int)P->jobs.size();
                P->jobs.push_back(J);
                P->job_tile.push_back(ti);
                if (key_comp >= 0) job_of[ti][lev * 4 + key_comp] = jid;
                return jid;
            };

            int comps[4], ncomps = 0;
            needed_comps(t, final_level ? ok : (int)OUT_PLANE, comps, &ncomps);
            if (!final_level) {
                // one job per component the final level will need
                int fcomps[4], nf = 0;
                needed_comps(t, out_kind_for(t), fcomps, &nf);
                for (int i = 0; i < nf; ++i) {
                    const int c = fcomps[i];
                    LevelJob J = base;
                    J.nc = 1;
                    fill_comp(J, 0, c);
                    J.dst[0] = scratch_plane(c, lev & 1);
                    J.dst_pitch = P->tile_scratch_pitch[ti];
                    add_deps(J, c);
                    push_job(J, c);
                }
            } else if (ok == OUT_PLANE) {
                for (int i = 0; i < ncomps; ++i) {
                    const int c = comps[i];
                    LevelJob J = base;
                    J.nc = 1;
                    fill_comp(J, 0, c);
                    set_fuse_geom(J);
                    fill_fuse(J, 0, c);
                    J.dst[0] = t.planes_out[c];
                    J.dst_pitch = t.planes_pitch;
                    add_deps(J, c);
                    push_job(J, c);
                }
            } else {
                LevelJob J = base;
                J.nc = ncomps;
                set_fuse_geom(J);
                for (int i = 0; i < ncomps; ++i) {
                    fill_comp(J, i, comps[i]);
                    fill_fuse(J, i, comps[i]);
                    add_deps(J, comps[i]);
                }
                // a merge that checks out one component shows it as grey
                if (ok == OUT_MERGE && t.check_out >= 0) J.check_out = 0;
                // the kernel loads every lockstep slot: pad unused ones with
                // component slot 0 (never emitted: slots >= nc are ignored)
                for (int i = ncomps; i < 4; ++i) {
                    J.comp[i] = J.comp[0];
                    J.ll[i] = J.ll[0];
                    J.det[i] = J.det[0];
                    J.ill[i] = J.ill[0];
                    for (int g = 0; g < 4; ++g) J.gain[i][g] = J.gain[0][g];
                    for (int g = 0; g < 4; ++g) J.igain[i][g] = J.igain[0][g];
                }
                push_job(J, -1);
            }
        }
        // Segment height of a launch: the same number of row pairs for every
        // job in it, chosen by a cost model of the launch.  When the items
        // at the tuned height fit on the GPU at once (at most four waves per
        // SIMD), they all run side by side and the step lasts about as long
        // as its busiest SIMD: rounds = ceil(items / SIMDs) waves (at least
        // two), each covering its rows plus the vertical halo it re-reads (8
        // rows for 9/7, 4 for 5/3).  The model then picks, up to twice the tuned
        // height, the one with the fewest rounds x (rows + halo) among those
        // giving the step at least min_waves_per_cu items per CU (small
        // coarse levels are latency-bound, not bandwidth-bound); ties go to
        // the taller segment.  4096x2160 final level: 38-row segments (1995
        // items, two per SIMD) run in 31.0 us, 32-row ones (2380 items,
        // some SIMDs three) in 34.4 us.  Larger steps, which the dispatcher
        // refills as waves finish, search only up to the tuned height (256
        // mixed tiles: 0.49 ms at 32 rows, 0.50 ms with 64 allowed).
        // Thumbnail jobs keep their one-pair items.  The launches of a step
        // are the jobs the dispatching kernel (or one specialised kernel)
        // takes, and each two-level variant, which launches on its own.
        {
            const long target = (long)P->num_cus * P->tuning.min_waves_per_cu;
            const long simds = 4L * P->num_cus;
            const int mode = P->tuning.persistent;
            auto group_of = [&](const LevelJob &J) {
                return (J.fuse2 || mode == 3) ? variant_of(J) : -1;
            };
            std::vector<int> groups;
            for (int j = step_job0; j < (int)P->jobs.size(); ++j) {
                const int g = group_of(P->jobs[j]);
                if (std::find(groups.begin(), groups.end(), g) == groups.end()) groups.push_back(g);
            }
            for (const int g : groups) {
                long fixed = 0;
                bool any97 = false, any = false, any2 = false;
                int max_pairs = 1;  // row pairs of the tallest job: no taller segment helps
                for (int j = step_job0; j < (int)P->jobs.size(); ++j) {
                    const LevelJob &J = P->jobs[j];
                    if (group_of(J) != g) continue;
                    if (J.seg_pairs <= 1) {
                        fixed += J.n_items;
                    } else {
                        any = true;
                        any97 = any97 || J.filt == 97;
                        any2 = any2 || J.fuse2;
                        max_pairs = std::max(max_pairs, (J.b1 - (J.b0 & ~1) + 1) / 2);
                    }
                }
                if (!any || P->tuning.fixed_seg) continue;
                // a two-level item also re-synthesises about 8 rows' worth
                // of the level below at each end of its segment
                const int halo = any2 ? 16 : (any97 ? 8 : 4);
                auto items_at = [&](int sp) {
                    long total = fixed;
                    for (int j = step_job0; j < (int)P->jobs.size(); ++j) {
                        const LevelJob &J = P->jobs[j];
                        if (J.seg_pairs <= 1 || group_of(J) != g) continue;
                        const int nrows = J.b1 - (J.b0 & ~1);
                        total += (long)J.nstrips * ((nrows + 2 * sp - 1) / (2 * sp));
                    }
                    return total;
                };
                const int tuned = std::min(P->tuning.seg_rows / 2, max_pairs);
                // two-level steps may always go to twice the tuned height:
                // their halo costs twice as much (four 4096x2160 images per
                // plan: 64-row segments 38.8 us per image, 32-row 55.5 us)
                const int top = std::min(max_pairs, (items_at(tuned) <= 4 * simds || any2) ? 2 * tuned : tuned);
                int best = 1;
                long best_cost = -1;
                // a two-level launch is never a small coarse level: one item
                // per SIMD is enough (the 68 three-component 9/7 tiles of
                // the 256 mixed tiles: 112 us against 129 us at eight waves
                // per CU)
                const long need = any2 ? std::min(target, simds) : target;
                for (int sp = top; sp >= 1; --sp) {
                    const long total = items_at(sp);
                    if (total < need && sp > 1) continue;
                    // one wave per SIMD leaves memory latency unhidden: it
                    // costs about as much as two (final level above in
                    // 76-row segments, 1015 items: 44 us against 31 us)
                    const long rounds = std::max(2L, (total + simds - 1) / simds);
                    const long cost = rounds * (2 * sp + halo);
                    if (best_cost < 0 || cost < best_cost) {
                        best_cost = cost;
                        best = sp;
                    }
                }
                for (int j = step_job0; j < (int)P->jobs.size(); ++j) {
                    LevelJob &J = P->jobs[j];
                    if (J.seg_pairs <= 1 || group_of(J) != g) continue;
                    J.seg_pairs = best;
                    const int nrows = J.b1 - (J.b0 & ~1);
                    J.nsegs = (nrows + 2 * J.seg_pairs - 1) / (2 * J.seg_pairs);
                    J.n_items = J.nstrips * J.nsegs;
                }
            }
            // the step's items grouped by kernel variant (jobs stable-sorted
            // by variant, tile order kept), one launch per variant.  In a
            // two-level launch the tiles whose items take the simple stores
            // (even first and last column, even output pitch) come first:
            // the two kinds run different code, and kept apart in the grid
            // (each XCD takes a contiguous range) they do not share an
            // instruction cache as often (68 three-component 9/7 tiles of
            // mixed widths: 122.6 us in tile order)
            auto store_kind = [&](const LevelJob &J) {
                return J.fuse2 && ((J.a0 & 1) || (J.a1 & 1) || (J.out_pitch_px & 1)) ? 1 : 0;
            };
            std::vector<std::pair<int, int>> order;  // (variant, job)
            for (int j = step_job0; j < (int)P->jobs.size(); ++j)
                order.emplace_back(variant_of(P->jobs[j]), j);
            std::stable_sort(order.begin(), order.end(),
                             [&](const std::pair<int, int> &a, const std::pair<int, int> &b) {
                                 if (a.first != b.first) return a.first < b.first;
                                 return store_kind(P->jobs[a.second]) < store_kind(P->jobs[b.second]);
                             });
            long step_items = 0;
            for (const auto &o : order) step_items += P->jobs[o.second].n_items;
            P->items.reserve(P->items.size() + step_items);
            for (size_t i = 0; i < order.size();) {
                size_t k = i;
                const int first = (int)P->items.size();
                for (; k < order.size() && order[k].first == order[i].first; ++k) {
                    const int j = order[k].second;
                    const int n = P->jobs[j].n_items;
                    for (int li = 0; li < n; ++li) P->items.push_back(WorkItem{j, li});
                }
                P->launches.push_back({first, (int)P->items.size() - first, order[i].first, step});
                i = k;
            }
            // A step that needs several kernel variants (a batch of mixed
            // filters / components / outputs) becomes one launch of the
            // dispatching kernel (variant -1) unless the tuning asks for one
            // launch per variant (3); mode 2 merges every step.  One launch
            // per step beats per-variant launches for mixed batches (256
            // mixed tiles: 0.49 against 0.72 ms).  With two variants doing
            // large levels the specialised kernels can win (256
            // three-component 9/7 tiles: 0.55 against 0.59 ms), since the
            // dispatching kernel carries every variant's code and the
            // largest register budget; the autotuner tries both.
            {
                // two-level launches (variants 96 and up, sorted last) keep
                // their own kernels: the dispatching kernel does not carry them
                int n0 = -1, nv = 0;
                std::vector<BmiBatchPlan::Launch> two;
                for (size_t q = 0; q < P->launches.size(); ++q)
                    if (P->launches[q].step == step) {
                        if (P->launches[q].variant >= 96) {
                            two.push_back(P->launches[q]);
                            continue;
                        }
                        if (n0 < 0) n0 = (int)q;
                        ++nv;
                    }
                const int mode = P->tuning.persistent;
                if (n0 >= 0 && ((mode == 0 && nv > 1) || mode == 2)) {
                    const int f = P->launches[n0].first;
                    const int end = two.empty() ? (int)P->items.size() : two.front().first;
                    P->launches.resize(n0);
                    P->launches.push_back({f, end - f, -1, step});
                    for (const auto &l : two) P->launches.push_back(l);
                }
            }
        }
        P->step_count[step] = (int)P->items.size() - P->step_first[step];
    }

    // a step whose jobs are all empty levels has nothing to launch (a grid
    // of no workgroups is an invalid launch)
    P->launches.erase(std::remove_if(P->launches.begin(), P->launches.end(),
                                     [](const BmiBatchPlan::Launch &l) { return l.count <= 0; }),
                      P->launches.end());

    // Waves per workgroup of each step launch.  A launch of a specialised
    // kernel with fewer than four waves' worth of items per SIMD runs one
    // wave per workgroup: the dispatcher then spreads the items over the
    // CUs in single waves, and no CU holds a whole extra workgroup while
    // others idle at the end (4096x2160x3 final level: 35.4 -> 34.6 us).
    // Larger launches and the dispatching kernel keep the tuning's value.
    for (auto &l : P->launches)
        l.wpg = (wpg_auto && l.variant >= 0 && l.count < 4 * 4 * P->num_cus) ? 1 : P->tuning.waves_per_group;

    const size_t jb = sizeof(LevelJob) * P->jobs.size();
    const size_t ib = sizeof(WorkItem) * P->items.size();
    const size_t cb = sizeof(int) * (P->jobs.size() + 3);
    P->device_bytes = align256(P->scratch_bytes) + align256(jb) + align256(ib) + align256(cb);
    *out_plan = P;
    return BMI_OK;
}

size_t bmi_plan_device_bytes(const BmiBatchPlan *P) { return P ? P->device_bytes : 0; }
size_t bmi_plan_table_bytes(const BmiBatchPlan *P)
{
    return P ? align256(sizeof(LevelJob) * P->jobs.size()) + sizeof(WorkItem) * P->items.size() : 0;
}

// Device half: place scratch, job table, items and counters in the
// device_bytes at `base`, relocate the scratch pointers of the job table
// and upload it on `stream` (asynchronous; the plan's host vectors stay
// alive with the plan).  The counters are cleared with a memset, so a
// caller that captures graphs must bind outside the capture.
size_t bmi_plan_estimate_bytes(const BmiTileDesc *tin, const BmiTuning *tuning)
{
    (void)tuning;
    BmiTileDesc eff;
    if (!tin || bmi_apply_ignore(*tin, &eff) != BMI_OK) return 0;
    const BmiTileDesc *t = &eff;
    if (validate_tile(*t) != BMI_OK) return 0;
    const int L = t->levels;
    const int Ls = (t->top_level < 0) ? L : t->top_level;
    int maxw = 0, maxh = 0;
    size_t jobs = 0, items = 0;
    for (int lev = 0; lev < Ls; ++lev) {
        int a0, a1, b0, b1;
        level_span(t->x0, t->width, L, lev, &a0, &a1);
        level_span(t->y0, t->height, L, lev, &b0, &b1);
        if (lev + 1 < Ls) {
            maxw = std::max(maxw, a1 - a0);
            maxh = std::max(maxh, b1 - b0);
        }
        const int owp = (t->method == Ckernels_W9X7) ? 60 : 62;
        const size_t nstrips = (size_t)(a1 - (a0 & ~1) + 2 * owp - 1) / (2 * owp);
        const size_t nrows = (size_t)(b1 - (b0 & ~1));
        // at most one job per component; segments of one row pair at least
        jobs += (size_t)t->ncomp;
        items += (size_t)t->ncomp * nstrips * ((nrows + 1) / 2);
    }
    const int pitch = (maxw + 63) & ~63;
    const size_t scratch = align256((size_t)pitch * maxh * 4 * 2 * (size_t)t->ncomp);
    return scratch + align256(sizeof(LevelJob) * jobs) + align256(sizeof(WorkItem) * items) +
           align256(sizeof(int) * (jobs + 3));
}

// Place scratch, job table, items and counters in the device_bytes at
// `base` and relocate the scratch pointers of the job table (host only).
static void plan_place(BmiBatchPlan *P, void *base)
{
    char *b = (char *)base;
    P->scratch = b;
    const size_t jb = sizeof(LevelJob) * P->jobs.size();
    const size_t ib = sizeof(WorkItem) * P->items.size();
    size_t off = align256(P->scratch_bytes);
    P->d_jobs = (LevelJob *)(b + off);
    off += align256(jb);
    P->d_items = (WorkItem *)(b + off);
    off += align256(ib);
    P->d_counters = (int *)(b + off);
    for (LevelJob &J : P->jobs) {
        const bool packed = (J.out_kind == OUT_RGB || J.out_kind == OUT_MERGE);
        const int nslots = packed ? 4 : 1;
        if (J.lev > 0 && !J.fuse2)
            for (int s = 0; s < nslots; ++s) J.ll[s] = (const char *)J.ll[s] + (uintptr_t)b;
        if (J.fuse2 && J.lev > 1)  // the inner level reads a scratch plane
            for (int s = 0; s < nslots; ++s) J.ill[s] = (const char *)J.ill[s] + (uintptr_t)b;
        if (J.out_kind == OUT_SCRATCH) J.dst[0] = (char *)J.dst[0] + (uintptr_t)b;
    }
}

int bmi_plan_bind(BmiBatchPlan *P, void *base, hipStream_t stream, void *staging,
                  size_t staging_bytes)
{
    if (!P || !base || P->scratch) return bmi_fail(BMI_E_INVALID, "plan_bind: bad args");
    plan_place(P, base);
    const size_t jb = sizeof(LevelJob) * P->jobs.size();
    const size_t ib = sizeof(WorkItem) * P->items.size();
    const size_t cb = sizeof(int) * (P->jobs.size() + 3);
    if (staging && staging_bytes >= align256(jb) + ib) {
        // one copy of both tables from pinned memory (same layout as on the device)
        memcpy(staging, P->jobs.data(), jb);
        memcpy((char *)staging + align256(jb), P->items.data(), ib);
        if (hipMemcpyAsync(P->d_jobs, staging, align256(jb) + ib, hipMemcpyHostToDevice, stream) !=
            hipSuccess)
            return bmi_fail(BMI_E_HIP, "plan_bind: descriptor upload");
    } else if (hipMemcpyAsync(P->d_jobs, P->jobs.data(), jb, hipMemcpyHostToDevice, stream) !=
                   hipSuccess ||
               hipMemcpyAsync(P->d_items, P->items.data(), ib, hipMemcpyHostToDevice, stream) !=
                   hipSuccess) {
        return bmi_fail(BMI_E_HIP, "plan_bind: descriptor upload");
    }
    if (hipMemsetAsync(P->d_counters, 0, cb, stream) != hipSuccess)
        return bmi_fail(BMI_E_HIP, "plan_bind: counters");
    return BMI_OK;
}

extern "C" int BMI_batch_plan_create(BmiBatchPlan **out_plan, const BmiTileDesc *tiles, int ntiles,
                                     const BmiTuning *tuning)
{
    BmiBatchPlan *P = nullptr;
    int rc = bmi_plan_build(&P, tiles, ntiles, tuning);
    if (rc) return rc;
    void *mem = nullptr;
    if (hipMalloc(&mem, P->device_bytes) != hipSuccess) {
        delete P;
        return bmi_fail(BMI_E_NOMEM, "batch_plan_create: %zu device bytes", P->device_bytes);
    }
    P->owns_device = true;
    rc = bmi_plan_bind(P, mem, nullptr, nullptr, 0);
    if (!rc && hipStreamSynchronize(nullptr) != hipSuccess)
        rc = bmi_fail(BMI_E_HIP, "batch_plan_create: descriptor upload");
    if (rc) {
        BMI_batch_plan_destroy(P);
        return rc;
    }
    *out_plan = P;
    return BMI_OK;
}

static int launch_one(const BmiBatchPlan *P, const BmiBatchPlan::Launch &l, hipStream_t s)
{
    const int wpg = l.wpg > 0 ? l.wpg : P->tuning.waves_per_group;
    const int xcd_remap = 1;
    const dim3 grid((l.count + wpg - 1) / wpg), block(64 * wpg);
    if (l.variant < 0) {
        hipLaunchKernelGGL(fused_step_any_kernel, grid, block, 0, s, P->d_jobs, P->d_items, l.first,
                           l.count, xcd_remap);
        return BMI_OK;
    }
    if (l.variant >= 96) {
        const size_t lds = (size_t)wpg * kRingRows * fused_variant_nc(l.variant) * 64 *
                           fused_variant_strips(l.variant) * sizeof(float);
        switch (l.variant) {
#define BMI_STEP2_CASE(V)                                                                        \
    case V:                                                                                     \
        hipLaunchKernelGGL((fused2_step_kernel<V>), grid, block, lds, s, P->d_jobs, P->d_items, \
                           l.first, l.count, xcd_remap);                                        \
        break;
            BMI_FOR_EACH_FUSED_VARIANT(BMI_STEP2_CASE)
#undef BMI_STEP2_CASE
        default:
            return bmi_fail(BMI_E_INVALID, "fused launch: no two-level kernel for variant %d", l.variant);
        }
        return BMI_OK;
    }
    switch (l.variant) {
#define BMI_STEP_CASE(V)                                                                        \
    case V:                                                                                    \
        hipLaunchKernelGGL((fused_step_kernel<V>), grid, block, 0, s, P->d_jobs, P->d_items, \
                           l.first, l.count, xcd_remap);                                       \
        break;
        BMI_FOR_EACH_VARIANT(BMI_STEP_CASE)
#undef BMI_STEP_CASE
    default:
        return bmi_fail(BMI_E_INVALID, "fused launch: no kernel for variant %d", l.variant);
    }
    return BMI_OK;
}

// Enqueue the plan on stream s.  With `branches` (graph capture only), the
// launches of one level step -- different kernel variants over disjoint
// tiles, independent of each other -- are forked onto those streams and
// joined before the next step, so the captured graph runs them side by
// side instead of one after another.
static int plan_enqueue(BmiBatchPlan *P, hipStream_t s, const std::vector<hipStream_t> *branches = nullptr,
                        std::vector<hipEvent_t> *events = nullptr)
{
    const int nj = (int)P->jobs.size();
    const int n = (int)P->items.size();
    if (n == 0) return BMI_OK;
    const int wpg = P->tuning.waves_per_group;
    (void)hipGetLastError();  // report this plan's launches, not an earlier error
    if (P->tuning.persistent == 1) {
        const int max_groups = P->num_cus * P->tuning.groups_per_cu;
        const int need_groups = (n + wpg - 1) / wpg;
        const int groups = std::max(1, std::min(max_groups, need_groups));
        hipLaunchKernelGGL(fused_persistent_kernel, dim3(groups), dim3(64 * wpg), 0, s, P->d_jobs,
                           P->d_items, n, P->d_counters, nj, groups * wpg);
    } else {
        const size_t nl = P->launches.size();
        for (size_t i = 0; i < nl;) {
            size_t j = i;
            while (j < nl && P->launches[j].step == P->launches[i].step) ++j;
            const int nb = branches ? (int)std::min(branches->size(), j - i) : 0;
            if (nb <= 1) {
                for (size_t k = i; k < j; ++k) {
                    const int rc = launch_one(P, P->launches[k], s);
                    if (rc) return rc;
                }
            } else {
                // fork: every branch waits for what s has enqueued so far
                hipEvent_t fork;
                if (hipEventCreateWithFlags(&fork, hipEventDisableTiming) != hipSuccess)
                    return bmi_fail(BMI_E_HIP, "fused launch: event");
                events->push_back(fork);
                if (hipEventRecord(fork, s) != hipSuccess)
                    return bmi_fail(BMI_E_HIP, "fused launch: fork");
                for (int b = 0; b < nb; ++b)
                    if (hipStreamWaitEvent((*branches)[b], fork, 0) != hipSuccess)
                        return bmi_fail(BMI_E_HIP, "fused launch: fork wait");
                for (size_t k = i; k < j; ++k) {
                    const int rc = launch_one(P, P->launches[k], (*branches)[(k - i) % nb]);
                    if (rc) return rc;
                }
                // join: s waits for every branch
                for (int b = 0; b < nb; ++b) {
                    hipEvent_t join;
                    if (hipEventCreateWithFlags(&join, hipEventDisableTiming) != hipSuccess)
                        return bmi_fail(BMI_E_HIP, "fused launch: event");
                    events->push_back(join);
                    if (hipEventRecord(join, (*branches)[b]) != hipSuccess ||
                        hipStreamWaitEvent(s, join, 0) != hipSuccess)
                        return bmi_fail(BMI_E_HIP, "fused launch: join");
                }
            }
            i = j;
        }
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return bmi_fail(BMI_E_HIP, "fused launch: %s", hipGetErrorString(e));
    return BMI_OK;
}

extern "C" int BMI_batch_plan_num_launches(const BmiBatchPlan *P)
{
    if (!P) return 0;
    return P->tuning.persistent == 1 ? 1 : (int)P->launches.size();
}

extern "C" int BMI_batch_plan_execute_launches(BmiBatchPlan *P, int first, int count, StreamId stream)
{
    if (!P || P->tuning.persistent == 1 || first < 0 || count < 0 ||
        first + count > (int)P->launches.size())
        return bmi_fail(BMI_E_INVALID, "batch_plan_execute_launches: bad range");
    auto saved = P->launches;
    std::vector<BmiBatchPlan::Launch> sub(saved.begin() + first, saved.begin() + first + count);
    P->launches.swap(sub);
    int rc = plan_enqueue(P, (hipStream_t)stream);
    P->launches.swap(sub);
    return rc;
}

extern "C" int BMI_batch_plan_execute(BmiBatchPlan *P, StreamId stream)
{
    if (!P) return bmi_fail(BMI_E_INVALID, "batch_plan_execute: null plan");
    return plan_enqueue(P, (hipStream_t)stream);
}

extern "C" int BMI_batch_plan_execute_graph(BmiBatchPlan *P, StreamId stream)
{
    if (!P) return bmi_fail(BMI_E_INVALID, "batch_plan_execute_graph: null plan");
    hipStream_t s = (hipStream_t)stream;
    if (!P->graph_exec) {
        hipStream_t cap;
        if (hipStreamCreateWithFlags(&cap, hipStreamNonBlocking) != hipSuccess)
            return bmi_fail(BMI_E_HIP, "graph: stream");
        if (hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            (void)hipStreamDestroy(cap);
            return bmi_fail(BMI_E_HIP, "graph: begin capture");
        }
        // independent launches of a level step become parallel branches
        std::vector<hipStream_t> branches;
        std::vector<hipEvent_t> events;
        for (int b = 0; b < 4; ++b) {
            hipStream_t bs;
            if (hipStreamCreateWithFlags(&bs, hipStreamNonBlocking) != hipSuccess) break;
            branches.push_back(bs);
        }
        int rc = plan_enqueue(P, cap, &branches, &events);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(cap, &g);
        for (hipStream_t bs : branches) (void)hipStreamDestroy(bs);
        for (hipEvent_t ev : events) (void)hipEventDestroy(ev);
        (void)hipStreamDestroy(cap);
        if (rc || e != hipSuccess || !g) return bmi_fail(BMI_E_HIP, "graph: capture failed");
        if (hipGraphInstantiate(&P->graph_exec, g, nullptr, nullptr, 0) != hipSuccess) {
            (void)hipGraphDestroy(g);
            P->graph_exec = nullptr;
            return bmi_fail(BMI_E_HIP, "graph: instantiate");
        }
        P->graph = g;
    }
    if (hipGraphLaunch(P->graph_exec, s) != hipSuccess) return bmi_fail(BMI_E_HIP, "graph: launch");
    return BMI_OK;
}

extern "C" int BMI_batch_plan_rebind(BmiBatchPlan *P, int i, const BmiTileDesc *tin)
{
    if (!P || !tin || i < 0 || i >= (int)P->tiles.size())
        return bmi_fail(BMI_E_INVALID, "batch_plan_rebind: bad args");
    BmiTileDesc eff;
    if (int rc0 = bmi_apply_ignore(*tin, &eff)) return bmi_fail(rc0, "batch_plan_rebind: tile rejected");
    const BmiTileDesc *t = &eff;
    const BmiTileDesc &o = P->tiles[i];
    if (t->width != o.width || t->height != o.height || t->levels != o.levels ||
        t->ncomp != o.ncomp || t->method != o.method || t->out_format != o.out_format ||
        t->top_level != o.top_level || t->x0 != o.x0 || t->y0 != o.y0)
        return bmi_fail(BMI_E_INVALID, "batch_plan_rebind: geometry differs");
    int rc = validate_tile(*t);
    if (rc) return bmi_fail(rc, "batch_plan_rebind: tile rejected");
    for (size_t j = 0; j < P->jobs.size(); ++j) {
        if (P->job_tile[j] != i) continue;
        LevelJob &J = P->jobs[j];
        const bool packed = (J.out_kind == OUT_RGB || J.out_kind == OUT_MERGE);
        for (int s = 0; s < (packed ? 4 : 1); ++s) {
            // padded lockstep slots (s >= nc) repeat slot 0's component
            const int c = J.comp[s];
            J.det[s] = t->coef[c];
            if (J.lev <= 0) J.ll[s] = t->coef[c];
            if (J.fuse2 && J.lev == 1) J.ill[s] = t->coef[c];
            if (J.out_kind == OUT_PLANE && s < J.nc) J.dst[s] = t->planes_out[c];
        }
        if (J.out_kind == OUT_RGB || J.out_kind == OUT_MERGE) {
            J.out = t->out;
            J.out_pitch_px = t->out_pitch_px;
            J.out_x = t->out_x;
            J.out_y = t->out_y;
        }
        if (J.out_kind == OUT_PLANE) J.dst_pitch = t->planes_pitch;
    }
    P->tiles[i] = *t;
    if (hipMemcpy(P->d_jobs, P->jobs.data(), sizeof(LevelJob) * P->jobs.size(),
                  hipMemcpyHostToDevice) != hipSuccess)
        return bmi_fail(BMI_E_HIP, "batch_plan_rebind: upload");
    return BMI_OK;
}

extern "C" size_t BMI_batch_plan_scratch_bytes(const BmiBatchPlan *P)
{
    return P ? P->scratch_bytes : 0;
}

extern "C" int BMI_batch_plan_num_items(const BmiBatchPlan *P) { return P ? (int)P->items.size() : 0; }

// Host-side description of the plan BMI_batch_plan_create would build for
// these tiles, without device memory: its job table as raw LevelJob
// records (pointers into the plan's own memory are offsets from 0, below
// *device_bytes) and the item count of each launch.  For tests and tools
// that check, on the CPU, what the kernels will address.
extern "C" int BMI_batch_plan_describe(const BmiTileDesc *tiles, int ntiles, const BmiTuning *tuning,
                                       void *jobs, int max_jobs, int *njobs, int *job_bytes,
                                       int *launch_counts, int max_launches, int *nlaunches,
                                       size_t *device_bytes)
{
    if (!njobs || !job_bytes || !nlaunches || !device_bytes || max_jobs < 0 || max_launches < 0 ||
        (max_jobs > 0 && !jobs) || (max_launches > 0 && !launch_counts))
        return bmi_fail(BMI_E_INVALID, "batch_plan_describe: bad args");
    BmiBatchPlan *P = nullptr;
    const int rc = bmi_plan_build(&P, tiles, ntiles, tuning);
    if (rc) return rc;
    *njobs = (int)P->jobs.size();
    *job_bytes = (int)sizeof(LevelJob);
    *nlaunches = (int)P->launches.size();
    *device_bytes = P->device_bytes;
    const int nj = std::min(max_jobs, (int)P->jobs.size());
    if (nj > 0) memcpy(jobs, P->jobs.data(), sizeof(LevelJob) * (size_t)nj);
    const int nl = std::min(max_launches, (int)P->launches.size());
    for (int i = 0; i < nl; ++i) launch_counts[i] = P->launches[i].count;
    delete P;
    return BMI_OK;
}

extern "C" int BMI_batch_plan_destroy(BmiBatchPlan *P)
{
    if (!P) return BMI_OK;
    if (P->graph_exec) (void)hipGraphExecDestroy(P->graph_exec);
    if (P->graph) (void)hipGraphDestroy(P->graph);
    if (P->owns_device && P->scratch) (void)hipFree(P->scratch);
    delete P;
    return BMI_OK;
}

extern "C" int BMI_batch_plan_error_flag(BmiBatchPlan *P)
{
    if (!P || !P->d_counters) return 0;
    int v = 0;
    (void)hipMemcpy(&v, P->d_counters + P->jobs.size() + 2, sizeof(int), hipMemcpyDeviceToHost);
    return v;
}

// ---------------------------------------------------------------------------
// Autotuner: time candidate launch shapes on the caller's own tiles.

extern "C" int BMI_autotune(const BmiTileDesc *tiles, int ntiles, const BmiTuning *candidates,
                            int ncand, int reps, StreamId stream, BmiTuning *best, float *best_us)
{
    if (!tiles || ntiles < 1 || !best || reps < 1)
        return bmi_fail(BMI_E_INVALID, "autotune: bad args");
    std::vector<BmiTuning> cand;
    if (candidates && ncand > 0) {
        cand.assign(candidates, candidates + ncand);
    } else {
        // default grid: segment height x items per CU x waves per group x
        // (merged or per-variant launches of steps that mix variants) x
        // (the last two 9/7 levels of a tile in one pass, for tiles of any
        // size or of final levels of 256 samples and more, or in two) x
        // (level steps aligned at the finest or the coarsest level)
        for (int seg : {16, 32, 64, 128})
            for (int mw : {4, 8, 16})
                for (int wpg : {0, 4})   // 0: chosen per launch
                    for (int mode : {0, 3})
                        for (int fuse : {0, 256, -1})
                            for (int coarse : {0, 1})
                                cand.push_back(BmiTuning{seg, wpg, 0, mode, mw, 0, fuse, coarse});
    }
    hipStream_t s = (hipStream_t)stream;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return bmi_fail(BMI_E_HIP, "autotune: event");
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return bmi_fail(BMI_E_HIP, "autotune: event");
    }
    // one candidate timed: warm-up (captures the plan's graph and
    // first-touches the scratch), then `n` timed graph replays and `n`
    // timed direct launches, the two ways plans run; it scores the faster
    auto time_candidate = [&](const BmiTuning &c, int n, float *us_out) -> int {
        BmiBatchPlan *P = nullptr;
        int r = BMI_batch_plan_create(&P, tiles, ntiles, &c);
        if (r) return r;
        r = BMI_batch_plan_execute_graph(P, stream);
        float us = 0.0f;
        for (int way = 0; way < 2 && !r; ++way) {
            if (hipEventRecord(e0, s) != hipSuccess) r = BMI_E_HIP;
            for (int k = 0; k < n && !r; ++k)
                r = way ? BMI_batch_plan_execute(P, stream) : BMI_batch_plan_execute_graph(P, stream);
            if (!r && hipEventRecord(e1, s) != hipSuccess) r = BMI_E_HIP;
            float ms = 0.0f;
            if (!r && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
                r = BMI_E_HIP;
            if (!r && (way == 0 || ms * 1000.0f / n < us)) us = ms * 1000.0f / n;
        }
        if (!r && BMI_batch_plan_error_flag(P) != 0) r = BMI_E_HIP;
        BMI_batch_plan_destroy(P);
        *us_out = us;
        return r;
    };
    // every candidate with `reps` repetitions, then the six fastest again
    // with four times as many: single short timings differ by a few
    // percent from run to run, enough to rank near-equal plans wrongly
    std::vector<std::pair<float, int>> t;
    int rc = BMI_OK;
    for (size_t i = 0; i < cand.size() && !rc; ++i) {
        float us = 0.0f;
        rc = time_candidate(cand[i], reps, &us);
        if (!rc) t.emplace_back(us, (int)i);
    }
    float best_t = 0.0f;
    int best_i = -1;
    if (!rc) {
        std::sort(t.begin(), t.end());
        const size_t top = std::min<size_t>(6, t.size());
        for (size_t j = 0; j < top && !rc; ++j) {
            float us = 0.0f;
            rc = time_candidate(cand[t[j].second], 4 * reps, &us);
            if (!rc && (best_i < 0 || us < best_t)) {
                best_t = us;
                best_i = t[j].second;
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return bmi_fail(rc, "autotune: candidate failed");
    *best = cand[best_i];
    if (best_us) *best_us = best_t;
    return BMI_OK;
}

// ---------------------------------------------------------------------------
// Decode without allocating: the plan's device memory (ping-pong scratch
// planes, job and item tables, persistent-mode counters) lives in one
// file-scope device array instead of a hipMalloc'd block.  For callers that
// cannot call hipMalloc (code under graph capture, tight serving loops,
// harnesses that forbid it).  The array exists once per GPU, from module
// load to process end; every call on a GPU reuses it, so calls that overlap
// on one GPU (other host threads, or other streams whose work is still
// running) race.  Host-side tables are staged in a per-device host buffer
// that the next call reuses only after the previous upload has been read.

constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

__global__ void zero_ints_kernel(int *p, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

namespace {

constexpr int kNoAllocMaxDevices = 64;

struct NoAllocDevice {
    bool ready = false;
    void *ptr = nullptr;               // this device's address of `scratch`
    hipEvent_t uploaded = nullptr;     // recorded after the last table upload
    bool recorded = false;             // `uploaded` has been recorded
    std::vector<unsigned char> staging;  // host copy of the last tables
};

std::mutex g_noalloc_mu;
NoAllocDevice g_noalloc[kNoAllocMaxDevices];

}  // namespace

extern "C" size_t BMI_decode_tiles_noalloc_capacity(void) { return kNoAllocScratchBytes; }

extern "C" int BMI_decode_tiles_noalloc(const BmiTileDesc *tiles, int ntiles,
                                        const BmiTuning *tuning, StreamId stream)
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kNoAllocMaxDevices)
        return bmi_fail(BMI_E_NODEVICE, "decode_tiles_noalloc: no current device");
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    NoAllocDevice &D = g_noalloc[dev];
    if (!D.ready) {
        // once per device, with that device current: the array's device
        // address (the symbol's host-side name is not a device address)
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return bmi_fail(BMI_E_HIP, "decode_tiles_noalloc: scratch address: %s",
                            hipGetErrorString(e));
        if (hipEventCreateWithFlags(&D.uploaded, hipEventDisableTiming) != hipSuccess)
            return bmi_fail(BMI_E_HIP, "decode_tiles_noalloc: event");
        D.ptr = p;
        D.ready = true;
    }
    BmiBatchPlan *P = nullptr;
    int rc = bmi_plan_build(&P, tiles, ntiles, tuning);
    if (rc) return rc;
    if (P->device_bytes > kNoAllocScratchBytes) {
        const size_t need = P->device_bytes;
        BMI_batch_plan_destroy(P);
        return bmi_fail(BMI_E_NOMEM, "decode_tiles_noalloc: %zu bytes of scratch needed, %zu available",
                        need, kNoAllocScratchBytes);
    }
    plan_place(P, D.ptr);
    const hipStream_t s = (hipStream_t)stream;
    const size_t jb = sizeof(LevelJob) * P->jobs.size();
    const size_t ib = sizeof(WorkItem) * P->items.size();
    const int nc = (int)P->jobs.size() + 3;
    // the previous call's upload must have read the staging copy
    if (D.recorded && hipEventSynchronize(D.uploaded) != hipSuccess) {
        BMI_batch_plan_destroy(P);
        return bmi_fail(BMI_E_HIP, "decode_tiles_noalloc: previous upload");
    }
    D.staging.resize(align256(jb) + ib);
    memcpy(D.staging.data(), P->jobs.data(), jb);
    memcpy(D.staging.data() + align256(jb), P->items.data(), ib);
    (void)hipGetLastError();
    bool ok = hipMemcpyAsync(P->d_jobs, D.staging.data(), align256(jb) + ib, hipMemcpyHostToDevice,
                             s) == hipSuccess;
    ok = ok && hipEventRecord(D.uploaded, s) == hipSuccess;
    D.recorded = D.recorded || ok;
    // the array keeps the previous call's contents: the persistent queue's
    // counters must start at zero (a fill kernel, not a memset node, so the
    // call can be captured)
    if (ok) {
        hipLaunchKernelGGL(zero_ints_kernel, dim3((nc + 255) / 256), dim3(256), 0, s,
                           P->d_counters, nc);
        ok = hipGetLastError() == hipSuccess;
    }
    if (!ok) {
        BMI_batch_plan_destroy(P);
        return bmi_fail(BMI_E_HIP, "decode_tiles_noalloc: table upload");
    }
    rc = plan_enqueue(P, s);
    BMI_batch_plan_destroy(P);  // host tables only: the memory is the array's
    return rc;
}
