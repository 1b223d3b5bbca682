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
#include <cstdio>
#include <cstring>
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

// write one output position (all lockstep components of one pixel)
template <int FILT, int NC>
__device__ inline void emit_pixel(const LevelJob &J, int row, int col,
                                  const typename SampleT<FILT>::T *v)
{
    using T = typename SampleT<FILT>::T;
    const int ok = J.out_kind;
    if (ok == OUT_SCRATCH || ok == OUT_PLANE) {
        size_t idx = (size_t)(row - J.b0) * J.dst_pitch + (col - J.a0);
#pragma unroll
        for (int c = 0; c < NC; ++c)
            if (c < J.nc) gptr_w<T>(J.dst[c])[idx] = v[c];
        return;
    }
    const int bd = J.bit_depth;
    const int od = fmt_depth(J.fmt);
    int R = 0, G = 0, B = 0, A = 0;
    if (ok == OUT_RGB) {
        if constexpr (NC >= 3) {
            if constexpr (FILT == 97)
                ict_pixel(v[0], v[1], v[2], bd, od, J.check_out, R, G, B);
            else
                rct_pixel(v[0], v[1], v[2], bd, od, J.check_out, R, G, B);
        }
    } else {  // OUT_MERGE
        int t[4] = {0, 0, 0, 0};
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            if (c < J.nc) {
                int s;
                if constexpr (FILT == 97) s = (int)v[c];
                else s = v[c];
                t[c] = merge_value(s, bd, od);
            }
        }
        merge_slots(t, J.fmt, J.check_out == 0, R, G, B, A);
    }
    uint8_t *p = (uint8_t *)J.out +
                 ((size_t)(J.out_y + row - J.b0) * J.out_pitch_px + (J.out_x + col - J.a0)) *
                     fmt_bytes(J.fmt);
    put_pixel(p, J.fmt, R, G, B, A);
}

// Prefetch depth of the row-pair pipeline: loads for pair k+D are issued
// while pair k is lifted, so each wave keeps D pairs of loads in flight
// (deeper for single-component items, which need fewer registers).
template <int NC>
struct Prefetch {
    static constexpr int D = (NC == 1) ? 4 : 2;
};

// convert a raw 32-bit word to a sample: sign-magnitude or float words
// times the folded gain (9/7), two's complement (5/3); select, no branch
template <int FILT>
__device__ inline typename SampleT<FILT>::T to_sample(uint32_t w, int is_float, float gain)
{
    if constexpr (FILT == 97) {
        const float v = is_float ? __uint_as_float(w) : (float)signmag_to_int(w);
        return __fmul_rn(v, gain);
    } else {
        return (int)w;
    }
}

// Degenerate levels (a single row or a single column): no prefetching.
template <int FILT, int NC>
__device__ __attribute__((always_inline)) inline void process_item_small(const LevelJob &__restrict__ J,
                                                             int strip, int seg, int lane);

// One work item: strip `strip`, row segment `seg` of level job J, run by
// one wave64 (lane = 0..63).
template <int FILT, int NC>
__device__ __attribute__((always_inline)) inline void process_item(
    const LevelJob &__restrict__ J, int strip, int seg, int lane)
{
    using T = typename SampleT<FILT>::T;
    constexpr int HP = (FILT == 97) ? 2 : 1;  // halo pairs on each side
    constexpr int OWP = 64 - 2 * HP;          // output pairs per strip
    constexpr int D = Prefetch<NC>::D;

    const int a0 = J.a0, a1 = J.a1, b0 = J.b0, b1 = J.b1;
    if ((a1 - a0) == 1 || (b1 - b0) == 1) {
        process_item_small<FILT, NC>(J, strip, seg, lane);
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

    // raw words of pair k: [comp][top-low, top-high, bottom-low, bottom-high]
    auto fetch = [&](int k, uint32_t (&r)[NC][4]) {
        const int rre = reflect(2 * k, b0, b1), rro = reflect(2 * k + 1, b0, b1);
        const size_t rt = (size_t)((rre >> 1) - rl0);
        const size_t q = (size_t)((rro >> 1) - rh0 + lh);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            r[c][0] = llp[c][rt * ll_pitch + lo_idx];
            r[c][1] = detp[c][rt * det_pitch + hi_off];
            r[c][2] = detp[c][q * det_pitch + lo_idx];
            r[c][3] = detp[c][q * det_pitch + hi_off];
        }
    };

    auto emit_row = [&](int row, const T *ve, const T *vo) {
        if (!lane_out || row < b0 || row >= b1) return;
        if (ce >= a0 && ce < a1) emit_pixel<FILT, NC>(J, row, ce, ve);
        if (co >= a0 && co < a1) emit_pixel<FILT, NC>(J, row, co, vo);
    };

    // vertical pipeline state per component and column (0 = ce, 1 = co)
    // 9/7: s0 = H_{k-1}, s1 = e3_{k-1}, s2 = o4_{k-2}, s3 = e5_{k-2}
    // 5/3: s0 = H_{k-1}, s1 = e1_{k-1}, s2 = L_{k-1}, s3 = H_{k-2}
    T s0[NC][2], s1[NC][2], s2[NC][2], s3[NC][2];
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int e = 0; e < 2; ++e) s0[c][e] = s1[c][e] = s2[c][e] = s3[c][e] = (T)0;

    const int nh = b1 - b0;
    // iteration at which the reference's VDWT_R53_LINES tail applies:
    // local pair nh/2 (bmi_cpu_dwt_decoder.cpp:3359-3378)
    const int kq = (FILT == 53 && quirk) ? (K0 + (nh >> 1)) : -0x7fffffff;
    const bool nh_even = !(nh & 1);
    constexpr int LAG = (FILT == 97) ? 2 : 1;  // pair k in, pair k-LAG out
    const int ks = kout0 - ((FILT == 97) ? 2 : (quirk ? 2 : 1));
    const int ke = kout1 + LAG;

    // one pipeline step with pair k
    auto consume = [&](int k, const uint32_t (&raw)[NC][4]) {
        T evn[2][NC], odd[2][NC];  // [column][component]
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            T tl = to_sample<FILT>(raw[c][0], ll_float, gn[c][0]);
            T th = to_sample<FILT>(raw[c][1], det_float, gn[c][1]);
            T bl = to_sample<FILT>(raw[c][2], det_float, gn[c][2]);
            T bh = to_sample<FILT>(raw[c][3], det_float, gn[c][3]);
            hlift<FILT>(tl, th);
            hlift<FILT>(bl, bh);
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
        const int ko = k - LAG;
        if (ko >= kout0 && ko <= kout1) {
            emit_row(2 * ko, evn[0], evn[1]);
            emit_row(2 * ko + 1, odd[0], odd[1]);
        }
    };

    // software pipeline, D pairs of loads in flight; loads are issued
    // unconditionally (rows past the segment end are clamped to it and the
    // pipeline steps that use them emit nothing)
    uint32_t ring[D][NC][4];
#pragma unroll
    for (int d = 0; d < D; ++d) fetch(min(ks + d, ke), ring[d]);
    const int rounds = (ke - ks + D) / D;
    for (int it = 0; it < rounds; ++it) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int k = ks + it * D + d;
            uint32_t cur[NC][4];
#pragma unroll
            for (int c = 0; c < NC; ++c)
#pragma unroll
                for (int j = 0; j < 4; ++j) cur[c][j] = ring[d][c][j];
            fetch(min(k + D, ke), ring[d]);
            consume(k, cur);
        }
    }
}

template <int FILT, int NC>
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
    auto emit_row = [&](int row, const T *ve, const T *vo) {
        if (!lane_out || row < b0 || row >= b1) return;
        if (ce >= a0 && ce < a1) emit_pixel<FILT, NC>(J, row, ce, ve);
        if (co >= a0 && co < a1) emit_pixel<FILT, NC>(J, row, co, vo);
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

template <int FILT>
__device__ __attribute__((always_inline)) inline void dispatch_nc(const LevelJob &__restrict__ J, int strip, int seg, int lane)
{
    if (J.nc == 1)
        process_item<FILT, 1>(J, strip, seg, lane);
    else if (J.nc <= 3)
        process_item<FILT, 3>(J, strip, seg, lane);
    else
        process_item<FILT, 4>(J, strip, seg, lane);
}

__device__ __attribute__((always_inline)) inline void run_item(const LevelJob *__restrict__ jobs, const WorkItem &it, int lane)
{
    const LevelJob &__restrict__ J = jobs[it.job];
    const int strip = it.local % J.nstrips;
    const int seg = it.local / J.nstrips;
    if (J.filt == 97)
        dispatch_nc<97>(J, strip, seg, lane);
    else
        dispatch_nc<53>(J, strip, seg, lane);
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
template <int FILT, int NC>
__global__ void __launch_bounds__(256) fused_step_kernel(const LevelJob *__restrict__ jobs,
                                                         const WorkItem *__restrict__ items,
                                                         int first, int count)
{
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int i = blockIdx.x * (blockDim.x >> 6) + wave;
    if (i >= count) return;
    const int job = __builtin_amdgcn_readfirstlane(items[first + i].job);
    const int local = __builtin_amdgcn_readfirstlane(items[first + i].local);
    const LevelJob &__restrict__ J = jobs[job];
    process_item<FILT, NC>(J, local % J.nstrips, local / J.nstrips, lane);
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
        int first, count, variant;
    };
    std::vector<Launch> launches;  // level-step mode: per step and variant
    std::vector<int> job_tile;                // tile of each job
    BmiTuning tuning;
    LevelJob *d_jobs = nullptr;
    WorkItem *d_items = nullptr;
    int *d_counters = nullptr;  // [njobs] completion counters, [njobs] = head,
                                // [njobs+1] = error flag
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
    std::vector<size_t> tile_scratch_off;  // per tile: base offset
    std::vector<int> tile_scratch_pitch;
    std::vector<size_t> tile_scratch_plane;  // bytes per plane
    int device = 0;
    int num_cus = 256;
    hipGraphExec_t graph_exec = nullptr;
    hipGraph_t graph = nullptr;
};

namespace {

// kernel variant of a job: filter x lockstep width class
int variant_of(const LevelJob &J)
{
    const int ncc = J.nc == 1 ? 0 : (J.nc <= 3 ? 1 : 2);
    return (J.filt == 97 ? 3 : 0) + ncc;
}

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
    if (t.top_level == 0) return BMI_E_UNSUPPORTED;
    if (t.top_level < -1 || t.top_level > t.levels) return BMI_E_INVALID;
    if (t.ncomp < 1 || t.ncomp > MAX_COMPONENT) return BMI_E_INVALID;
    if (t.method != Ckernels_W5X3 && t.method != Ckernels_W9X7) return BMI_E_INVALID;
    if (t.method == Ckernels_W9X7 && t.coef_kind != BMI_COEF_SIGNMAG &&
        t.coef_kind != BMI_COEF_FLOAT)
        return BMI_E_INVALID;
    if (t.bit_depth < 1 || t.bit_depth > 16) return BMI_E_INVALID;
    if (t.coef_pitch < t.width) return BMI_E_INVALID;
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

static int default_seg_rows(const BmiTuning *t) { return (t && t->seg_rows > 0) ? t->seg_rows : 32; }

extern "C" int BMI_batch_plan_create(BmiBatchPlan **out_plan, const BmiTileDesc *tiles, int ntiles,
                                     const BmiTuning *tuning)
{
    if (!out_plan || !tiles || ntiles < 1) return bmi_fail(BMI_E_INVALID, "batch_plan_create: bad args");
    *out_plan = nullptr;
    for (int i = 0; i < ntiles; ++i) {
        int rc = validate_tile(tiles[i]);
        if (rc) return bmi_fail(rc, "batch_plan_create: tile %d rejected", i);
    }
    BmiBatchPlan *P = new BmiBatchPlan();
    P->tiles.assign(tiles, tiles + ntiles);
    P->tuning.seg_rows = default_seg_rows(tuning);
    if (P->tuning.seg_rows & 1) P->tuning.seg_rows += 1;
    P->tuning.waves_per_group = (tuning && tuning->waves_per_group > 0) ? tuning->waves_per_group : 4;
    P->tuning.groups_per_cu = (tuning && tuning->groups_per_cu > 0) ? tuning->groups_per_cu : 8;
    P->tuning.persistent = tuning ? tuning->persistent : 1;
    if (P->tuning.waves_per_group > 4) P->tuning.waves_per_group = 4;
    if (hipGetDevice(&P->device) != hipSuccess) P->device = 0;
    {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, P->device) ==
                hipSuccess &&
            cus > 0)
            P->num_cus = cus;
    }

    // scratch layout: per tile, per needed component, two planes sized for
    // the largest intermediate level
    size_t total = 0;
    P->tile_scratch_off.resize(ntiles);
    P->tile_scratch_pitch.resize(ntiles);
    P->tile_scratch_plane.resize(ntiles);
    for (int ti = 0; ti < ntiles; ++ti) {
        const BmiTileDesc &t = tiles[ti];
        const int L = t.levels;
        const int Ls = (t.top_level < 0) ? L : t.top_level;
        int maxw = 0, maxh = 0;
        for (int lev = 0; lev + 1 < Ls; ++lev) {
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
    if (total) {
        if (hipMalloc(&P->scratch, total) != hipSuccess) {
            delete P;
            return bmi_fail(BMI_E_NOMEM, "batch_plan_create: scratch of %zu bytes", total);
        }
    }

    // jobs, ordered by level step then tile; items follow job order
    int max_steps = 0;
    for (int ti = 0; ti < ntiles; ++ti)
        max_steps = std::max(max_steps, tiles[ti].top_level < 0 ? tiles[ti].levels : tiles[ti].top_level);
    std::vector<std::vector<int>> job_of(ntiles);  // [tile][lev*4+comp] -> job
    for (int ti = 0; ti < ntiles; ++ti) job_of[ti].assign(MAX_DWT_LEVEL * 4, -1);
    const int seg_rows = P->tuning.seg_rows;
    P->step_first.assign(max_steps, 0);
    P->step_count.assign(max_steps, 0);
    for (int step = 0; step < max_steps; ++step) {
        P->step_first[step] = (int)P->items.size();
        for (int ti = 0; ti < ntiles; ++ti) {
            const BmiTileDesc &t = tiles[ti];
            const int L = t.levels;
            const int Ls = (t.top_level < 0) ? L : t.top_level;
            const int lev = step;
            if (lev >= Ls) continue;
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
                if (lev == 0) return;
                const int dj = job_of[ti][(lev - 1) * 4 + c];
                if (dj >= 0 && J.ndeps < kMaxDeps) J.deps[J.ndeps++] = dj;
            };
            auto push_job = [&](const LevelJob &J, int key_comp) {
                const int jid = (int)P->jobs.size();
                P->jobs.push_back(J);
                P->job_tile.push_back(ti);
                if (key_comp >= 0) job_of[ti][lev * 4 + key_comp] = jid;
                for (int li = 0; li < J.n_items; ++li) P->items.push_back(WorkItem{jid, li});
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
                    J.dst[0] = t.planes_out[c];
                    J.dst_pitch = t.planes_pitch;
                    add_deps(J, c);
                    push_job(J, c);
                }
            } else {
                LevelJob J = base;
                J.nc = ncomps;
                for (int i = 0; i < ncomps; ++i) {
                    fill_comp(J, i, comps[i]);
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
                    for (int g = 0; g < 4; ++g) J.gain[i][g] = J.gain[0][g];
                }
                push_job(J, -1);
            }
        }
        P->step_count[step] = (int)P->items.size() - P->step_first[step];
        // group the step's items by kernel variant (stable: tile order kept)
        auto vid = [&](const WorkItem &w) { return variant_of(P->jobs[w.job]); };
        std::stable_sort(P->items.begin() + P->step_first[step], P->items.end(),
                         [&](const WorkItem &x, const WorkItem &y) { return vid(x) < vid(y); });
        for (int i = P->step_first[step]; i < (int)P->items.size();) {
            int j = i;
            while (j < (int)P->items.size() && vid(P->items[j]) == vid(P->items[i])) ++j;
            P->launches.push_back({i, j - i, vid(P->items[i])});
            i = j;
        }
    }

    const size_t jb = sizeof(LevelJob) * P->jobs.size();
    const size_t ib = sizeof(WorkItem) * P->items.size();
    const size_t cb = sizeof(int) * (P->jobs.size() + 3);
    if (hipMalloc(&P->d_jobs, jb) != hipSuccess || hipMalloc(&P->d_items, ib) != hipSuccess ||
        hipMalloc(&P->d_counters, cb) != hipSuccess) {
        BMI_batch_plan_destroy(P);
        return bmi_fail(BMI_E_NOMEM, "batch_plan_create: descriptor allocation");
    }
    if (hipMemcpy(P->d_jobs, P->jobs.data(), jb, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(P->d_items, P->items.data(), ib, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(P->d_counters, 0, cb) != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
        BMI_batch_plan_destroy(P);
        return bmi_fail(BMI_E_HIP, "batch_plan_create: descriptor upload");
    }
    *out_plan = P;
    return BMI_OK;
}

static int plan_enqueue(BmiBatchPlan *P, hipStream_t s)
{
    const int nj = (int)P->jobs.size();
    const int n = (int)P->items.size();
    if (n == 0) return BMI_OK;
    const int wpg = P->tuning.waves_per_group;
    if (P->tuning.persistent) {
        const int max_groups = P->num_cus * P->tuning.groups_per_cu;
        const int need_groups = (n + wpg - 1) / wpg;
        const int groups = std::max(1, std::min(max_groups, need_groups));
        hipLaunchKernelGGL(fused_persistent_kernel, dim3(groups), dim3(64 * wpg), 0, s, P->d_jobs,
                           P->d_items, n, P->d_counters, nj, groups * wpg);
    } else {
        for (const auto &l : P->launches) {
            const dim3 grid((l.count + wpg - 1) / wpg), block(64 * wpg);
            switch (l.variant) {
            case 0: hipLaunchKernelGGL((fused_step_kernel<53, 1>), grid, block, 0, s, P->d_jobs, P->d_items, l.first, l.count); break;
            case 1: hipLaunchKernelGGL((fused_step_kernel<53, 3>), grid, block, 0, s, P->d_jobs, P->d_items, l.first, l.count); break;
            case 2: hipLaunchKernelGGL((fused_step_kernel<53, 4>), grid, block, 0, s, P->d_jobs, P->d_items, l.first, l.count); break;
            case 3: hipLaunchKernelGGL((fused_step_kernel<97, 1>), grid, block, 0, s, P->d_jobs, P->d_items, l.first, l.count); break;
            case 4: hipLaunchKernelGGL((fused_step_kernel<97, 3>), grid, block, 0, s, P->d_jobs, P->d_items, l.first, l.count); break;
            default: hipLaunchKernelGGL((fused_step_kernel<97, 4>), grid, block, 0, s, P->d_jobs, P->d_items, l.first, l.count); break;
            }
        }
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return bmi_fail(BMI_E_HIP, "fused launch: %s", hipGetErrorString(e));
    return BMI_OK;
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
        int rc = plan_enqueue(P, cap);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(cap, &g);
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

extern "C" int BMI_batch_plan_rebind(BmiBatchPlan *P, int i, const BmiTileDesc *t)
{
    if (!P || !t || i < 0 || i >= (int)P->tiles.size())
        return bmi_fail(BMI_E_INVALID, "batch_plan_rebind: bad args");
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
        for (int s = 0; s < 4; ++s) {
            if (s >= J.nc && J.nc == 1) break;
            const int c = J.comp[s];
            J.det[s] = t->coef[c];
            if (J.lev == 0) J.ll[s] = t->coef[c];
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

extern "C" int BMI_batch_plan_destroy(BmiBatchPlan *P)
{
    if (!P) return BMI_OK;
    if (P->graph_exec) (void)hipGraphExecDestroy(P->graph_exec);
    if (P->graph) (void)hipGraphDestroy(P->graph);
    if (P->d_jobs) (void)hipFree(P->d_jobs);
    if (P->d_items) (void)hipFree(P->d_items);
    if (P->d_counters) (void)hipFree(P->d_counters);
    if (P->scratch) (void)hipFree(P->scratch);
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
