// bmi_device.h -- device helpers shared by the gfx950 kernels.
//
// Every kernel here is written for CDNA4 wave64.  Floating-point code is
// compiled with -ffp-contract=off so each lifting update is rounded after
// the add, the multiply and the subtract, the evaluation the reference's
// C paths get (vdwt_analyst_line, decoder/codec/include/bmi_cpu_dwt.h:179-185).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bmi_gpulib.h"

namespace bmi {

// decoder/codec/include/bmi_cpu_dwt.h:31-37 (same values in
// decoder/cudainterface/include/gpu_setting.h:16-21)
constexpr float kAlpha = -1.58613434205992f;
constexpr float kBeta = -0.05298011857296f;
constexpr float kGamma = 0.88291107553093f;
constexpr float kDelta = 0.44350685204397f;
constexpr float kLoGain = 1.23017410558578f;
constexpr float kHiGain = 1.62578613134411f;

// ICT constants, float, bmi_cpu_dwt_decoder.cpp:3573-3576
constexpr float kCrR = (float)(1.402);
constexpr float kCbG = (float)(-0.34413);
constexpr float kCrG = (float)(-0.71414);
constexpr float kCbB = (float)(1.772);

__host__ __device__ inline int ceil_div2(int v) { return (v + 1) >> 1; }
__host__ __device__ inline int floor_div2(int v) { return v >> 1; }
__host__ __device__ inline int ceil_shift(int v, int s) { return (v + (1 << s) - 1) >> s; }

// Reflect absolute coordinate c into [a0, a1) (whole-sample symmetric
// extension, T.800 F.3.7 PSE); n = a1 - a0 >= 2.  Reflection keeps the
// parity of c, so even coordinates stay low-band samples.
__host__ __device__ inline int reflect(int c, int a0, int a1)
{
    int n = a1 - a0;
    int period = 2 * (n - 1);
    int i = (c - a0) % period;
    if (i < 0) i += period;
    return a0 + (i < n ? i : period - i);
}

// Global-address-space views of pointers that reach a kernel through
// descriptors in memory.  Without the cast the compiler emits FLAT
// accesses, which share a wait counter with LDS and scalar loads and
// serialise every load behind an s_waitcnt vmcnt(0) lgkmcnt(0).
template <typename T>
__device__ inline const __attribute__((address_space(1))) T *gptr(const void *p)
{
    return (const __attribute__((address_space(1))) T *)p;
}
template <typename T>
__device__ inline __attribute__((address_space(1))) T *gptr_w(void *p)
{
    return (__attribute__((address_space(1))) T *)p;
}

// two's complement of a T1 sign-magnitude word (FETCHPIXEL_INT,
// bmi_cpu_dwt_decoder.cpp:2287)
__device__ inline int signmag_to_int(uint32_t w)
{
    int m = (int)(w & 0x7fffffffu);
    return (w & 0x80000000u) ? -m : m;
}

// x - c*(l + r), rounded per operation
__device__ inline float lift97(float x, float c, float l, float r)
{
    float s = __fadd_rn(l, r);
    float p = __fmul_rn(c, s);
    return __fsub_rn(x, p);
}

// wave64 neighbour exchange with DPP wave shifts (a VALU modifier, no
// LDS round trip): wave_shr:1 gives each lane the value of lane-1,
// wave_shl:1 the value of lane+1.  With bound_ctrl the edge lanes read 0,
// which lets the compiler fold the shift into the consuming add; edge
// lanes are halo lanes and never stored.
__device__ inline int wave_from_left(int v) { return __builtin_amdgcn_mov_dpp(v, 0x138, 0xf, 0xf, true); }
__device__ inline int wave_from_right(int v) { return __builtin_amdgcn_mov_dpp(v, 0x130, 0xf, 0xf, true); }
__device__ inline float wave_from_left(float v) { return __int_as_float(wave_from_left(__float_as_int(v))); }
__device__ inline float wave_from_right(float v) { return __int_as_float(wave_from_right(__float_as_int(v))); }

// LIFT, bmi_cpu_dwt.h:22
__device__ inline int lift_depth(int v, int in_depth, int out_depth)
{
    return in_depth >= out_depth ? (v >> (in_depth - out_depth)) : (v << (out_depth - in_depth));
}


// ---- pixel packing (element layouts of codec_base.h:51-143: B,G,R[,A]) ----

__device__ inline void put_pixel(uint8_t *pg, int fmt, int R, int G, int B, int A)
{
    auto p = gptr_w<uint8_t>(pg);
    switch (fmt) {
    case IMAGE_FOTMAT_32BITS_ARGB:
    case IMAGE_FOTMAT_32BITS_4COMPONENTS:
        *gptr_w<uint32_t>(pg) = (uint32_t)(B & 0xff) | ((uint32_t)(G & 0xff) << 8) |
                                ((uint32_t)(R & 0xff) << 16) | ((uint32_t)(A & 0xff) << 24);
        break;
    case IMAGE_FOTMAT_24BITS_RGB:
    case IMAGE_FOTMAT_24BITS_3COMPONENTS:
        p[0] = (uint8_t)B;
        p[1] = (uint8_t)G;
        p[2] = (uint8_t)R;
        break;
    case IMAGE_FOTMAT_48BITS_RGB:
    case IMAGE_FOTMAT_48BITS_3COMPONENTS: {
        auto q = gptr_w<uint16_t>(pg);
        q[0] = (uint16_t)B;
        q[1] = (uint16_t)G;
        q[2] = (uint16_t)R;
    } break;
    default: {  // 64-bit ARGB / 4 components
        auto q = gptr_w<uint16_t>(pg);
        q[0] = (uint16_t)B;
        q[1] = (uint16_t)G;
        q[2] = (uint16_t)R;
        q[3] = (uint16_t)A;
    } break;
    }
}

__host__ __device__ inline int fmt_depth(int fmt)
{
    switch (fmt) {
    case IMAGE_FOTMAT_48BITS_RGB:
    case IMAGE_FOTMAT_64BITS_4COMPONENTS:
    case IMAGE_FOTMAT_48BITS_3COMPONENTS:
    case IMAGE_FOTMAT_64BITS_ARGB:
        return 16;
    default:
        return 8;
    }
}

__host__ __device__ inline int fmt_bytes(int fmt)
{
    switch (fmt) {
    case IMAGE_FOTMAT_32BITS_ARGB:
    case IMAGE_FOTMAT_32BITS_4COMPONENTS:
        return 4;
    case IMAGE_FOTMAT_48BITS_RGB:
    case IMAGE_FOTMAT_48BITS_3COMPONENTS:
        return 6;
    case IMAGE_FOTMAT_24BITS_RGB:
    case IMAGE_FOTMAT_24BITS_3COMPONENTS:
        return 3;
    case IMAGE_FOTMAT_64BITS_4COMPONENTS:
    case IMAGE_FOTMAT_64BITS_ARGB:
        return 8;
    default:
        return 0;
    }
}

__device__ inline int clamp_i(int v, int mx) { return v < 0 ? 0 : (v > mx ? mx : v); }

// (int)(SET_BOUNDARY(0, v, maxV)) with the comparison in float; med3 gives
// the same value for every non-NaN v (including -0.0, which truncates to 0)
__device__ inline int clamp_trunc_f(float v, int mx)
{
    return (int)__builtin_amdgcn_fmed3f(v, 0.0f, (float)mx);
}

// inverse ICT + level shift + clamp + truncate + LIFT, float throughout
// (RICT_97_FLOAT, bmi_cpu_dwt_decoder.cpp:3570-3841)
__device__ inline void ict_pixel(float y, float cb, float cr, int bd, int od, int check_out,
                                 int &R, int &G, int &B)
{
    const int mx = (1 << bd) - 1;
    const float foff = (float)(1 << (bd - 1));
    const float r = __fadd_rn(__fadd_rn(y, __fmul_rn(kCrR, cr)), foff);
    const float g =
        __fadd_rn(__fadd_rn(__fadd_rn(y, __fmul_rn(kCbG, cb)), __fmul_rn(kCrG, cr)), foff);
    const float b = __fadd_rn(__fadd_rn(y, __fmul_rn(kCbB, cb)), foff);
    R = lift_depth(clamp_trunc_f(r, mx), bd, od);
    G = lift_depth(clamp_trunc_f(g, mx), bd, od);
    B = lift_depth(clamp_trunc_f(b, mx), bd, od);
    if (check_out == 0) G = B = R;
    else if (check_out == 1) R = B = G;
    else if (check_out == 2) R = G = B;
}

// inverse RCT with floor, level shift, clamp, LIFT (RRct_53,
// bmi_cpu_dwt_decoder.cpp:3936-4130)
__device__ inline void rct_pixel(int y0, int y1, int y2, int bd, int od, int check_out, int &R,
                                 int &G, int &B)
{
    const int mx = (1 << bd) - 1;
    const int g = y0 - ((y1 + y2) >> 2) + (1 << (bd - 1));
    const int r = y2 + g, b = y1 + g;
    R = lift_depth(clamp_i(r, mx), bd, od);
    G = lift_depth(clamp_i(g, mx), bd, od);
    B = lift_depth(clamp_i(b, mx), bd, od);
    if (check_out == 0) G = B = R;
    else if (check_out == 1) R = B = G;
    else if (check_out == 2) R = G = B;
}

// merge: truncate, level shift, clamp, LIFT (MergeComponents<T>,
// bmi_cpu_dwt_decoder.cpp:4282-4596)
__device__ inline int merge_value(int s, int bd, int od)
{
    const int mx = (1 << bd) - 1;
    return lift_depth(clamp_i(s + (1 << (bd - 1)), mx), bd, od);
}

// place merged components: 4-component slots Comp0=A, Comp1=R, Comp2=G,
// Comp3=B; 3-component Comp0=R, Comp1=G, Comp2=B; check_out 0 = grey
__device__ inline void merge_slots(const int *t, int fmt, int grey, int &R, int &G, int &B, int &A)
{
    const bool four =
        (fmt == IMAGE_FOTMAT_32BITS_4COMPONENTS || fmt == IMAGE_FOTMAT_64BITS_4COMPONENTS ||
         fmt == IMAGE_FOTMAT_32BITS_ARGB || fmt == IMAGE_FOTMAT_64BITS_ARGB);
    if (grey) {
        R = G = B = t[0];
        A = four ? t[0] : 0;
    } else if (four) {
        A = t[0];
        R = t[1];
        G = t[2];
        B = t[3];
    } else {
        R = t[0];
        G = t[1];
        B = t[2];
        A = 0;
    }
}

}  // namespace bmi
