// bmi_plain.hip -- the reference-shaped ("plain") GPU path behind the
// drop-in entry points BMI_comp_idwt / BMI_resolution_idwt /
// BMI_tile_ICT / BMI_tile_MCT / BMI_tile_Merge / BMI_Comp_DC_SHIFT /
// BMI_tile_One_Component.
//
// Same structure as the reference's CUDA back end (gpufunc.cu:419-546,
// 953-1089): per synthesis level one launch over the rows (orig -> temp,
// transposed store) and one over the columns (temp -> orig, transposed
// store), then a separate colour kernel.  Written for gfx950: a block of
// 256 threads owns 64 lines x 64 samples (+ lifting halo) in LDS, lifts
// stepwise, and stores transposed with 64 consecutive lines per row so the
// store is a 256-byte run.  Arithmetic is the reference CPU's (gains folded
// into the row pass, x - c*(l + r) rounded per operation), so results are
// bit-identical to the fused path and the CPU oracle.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "bmi_device.h"
#include "bmi_fused.h"
#include "bmi_internal.h"

namespace bmi {

constexpr int kPlLines = 64;  // lines per block
constexpr int kPlSeg = 64;    // output samples per block and line

__device__ inline int mirror_local(int i, int n)
{
    if (n == 1) return 0;
    return reflect(i, 0, n);
}

template <int FILT>
struct PlainT;
template <>
struct PlainT<97> {
    using T = float;
};
template <>
struct PlainT<53> {
    using T = int;
};

// Load one sample of a line stored as [low part | high part]; kind and
// gain chosen per part.
template <int FILT>
__device__ inline typename PlainT<FILT>::T plain_load(const void *base, size_t idx, int kind,
                                                      float gain)
{
    if constexpr (FILT == 97) {
        const uint32_t w = gptr<uint32_t>(base)[idx];
        const float v = (kind == SRC_SIGNMAG) ? (float)signmag_to_int(w) : __uint_as_float(w);
        return __fmul_rn(v, gain);
    } else {
        return gptr<int>(base)[idx];
    }
}

struct LineArgs {
    const void *in;
    int in_pitch;  // elements
    void *out;
    int out_pitch;  // elements; out[pos * out_pitch + line]
    int nlines, n, lo_n, parity;
    int split;      // lines < split are "top" lines (first gain pair / LL kind)
    float g[4];     // top-lo, top-hi, bottom-lo, bottom-hi
    int kind_top_lo, kind_rest;
    int quirk;      // 5/3 reference bottom-edge quirk (column pass)
};

// One 1-D synthesis over 64 lines x 64 output samples, stored transposed.
template <int FILT>
__global__ void __launch_bounds__(256) plain_line_kernel(LineArgs A)
{
    using T = typename PlainT<FILT>::T;
    constexpr int H = (FILT == 97) ? 4 : 2;  // lifting halo
    constexpr int W = kPlSeg + 2 * H;
    __shared__ T x[kPlLines][W + 1];
    const int seg0 = blockIdx.x * kPlSeg;
    const int line0 = blockIdx.y * kPlLines;
    const int tid = threadIdx.x;
    const int n = A.n;

    constexpr int NLOAD = (kPlLines * W + 255) / 256;
#pragma unroll
    for (int it = 0; it < NLOAD; ++it) {
        const int e = tid + 256 * it;
        if (e >= kPlLines * W) break;
        const int l = e / W, p = e - l * W;
        const int line = line0 + l;
        T v = (T)0;
        if (line < A.nlines) {
            const int i = mirror_local(seg0 - H + p, n);
            const bool low = ((i + A.parity) & 1) == 0;
            const bool top = line < A.split;
            const int idx = low ? ((i - A.parity) >> 1) : (A.lo_n + ((i - 1 + A.parity) >> 1));
            const float g = low ? (top ? A.g[0] : A.g[2]) : (top ? A.g[1] : A.g[3]);
            const int kind = (low && top) ? A.kind_top_lo : A.kind_rest;
            v = plain_load<FILT>(A.in, (size_t)line * A.in_pitch + idx, kind, g);
            if (n == 1 && !low) {
                if constexpr (FILT == 97) v = __fmul_rn(v, 0.5f);
                else v = v >> 1;
            }
        }
        x[l][p] = v;
    }
    __syncthreads();

    if (n > 1) {
        // one update class per step (low samples on even steps); a step reads
        // only the other class, so all of a thread's reads can be in flight
        // together: fixed trip counts, fully unrolled
        constexpr int NSTEPS = (FILT == 97) ? 4 : 2;
        constexpr int NE = kPlLines * W;
        constexpr int NIT = (NE + 255) / 256;
#pragma unroll
        for (int st = 0; st < NSTEPS; ++st) {
            const int want_low = (st & 1) == 0;
            T nv[NIT];
#pragma unroll
            for (int it = 0; it < NIT; ++it) {
                const int e = tid + 256 * it;
                const int l = e / W, p = e - l * W;
                const int i = seg0 - H + p;
                const bool low = (((i + A.parity) & 1) == 0);
                const bool upd = e < NE && p > 0 && p < W - 1 && low == (bool)want_low;
                nv[it] = (T)0;
                if (upd) {
                    if constexpr (FILT == 97) {
                        const float c = st == 0 ? kDelta : st == 1 ? kGamma : st == 2 ? kBeta : kAlpha;
                        nv[it] = lift97(x[l][p], c, x[l][p - 1], x[l][p + 1]);
                    } else {
                        if (st == 0)
                            nv[it] = x[l][p] - ((x[l][p - 1] + x[l][p + 1] + 2) >> 2);
                        else
                            nv[it] = x[l][p] + ((x[l][p - 1] + x[l][p + 1]) >> 1);
                    }
                }
            }
#pragma unroll
            for (int it = 0; it < NIT; ++it) {
                const int e = tid + 256 * it;
                const int l = e / W, p = e - l * W;
                const int i = seg0 - H + p;
                const bool low = (((i + A.parity) & 1) == 0);
                if (e < NE && p > 0 && p < W - 1 && low == (bool)want_low) x[l][p] = nv[it];
            }
            __syncthreads();
        }
    }

    if constexpr (FILT == 53) {
        // VDWT_R53_LINES tail (bmi_cpu_dwt_decoder.cpp:3359-3378): with
        // m = n/2 and P = H_{max(m-2,0)}, even n: X[n-1] = P + X[n-2];
        // odd n: X[n-1] = L_{m-1} - ((P+1)>>1), X[n-2] = P + ((X[n-1]+X[n-3])>>1)
        if (A.quirk && n >= 2 && (seg0 <= n - 1) && (n - 2 < seg0 + kPlSeg)) {
            int v1 = 0, v2 = 0;
            const int l = tid;
            const int line = line0 + l;
            const bool act = tid < kPlLines && line < A.nlines;
            const int m = n >> 1;
            const int pn1 = (n - 1) - seg0 + H, pn2 = (n - 2) - seg0 + H, pn3 = (n - 3) - seg0 + H;
            if (act) {
                const int *row = (const int *)A.in + (size_t)line * A.in_pitch;
                const int P = row[A.lo_n + (m >= 2 ? m - 2 : 0)];
                if (!(n & 1)) {
                    v1 = P + x[l][pn2];
                } else {
                    v1 = row[m - 1] - ((P + 1) >> 1);
                    v2 = P + ((v1 + x[l][pn3]) >> 1);
                }
            }
            __syncthreads();
            if (act) {
                x[l][pn1] = v1;
                if (n & 1) x[l][pn2] = v2;
            }
        }
        __syncthreads();
    }

#pragma unroll
    for (int it = 0; it < kPlSeg * kPlLines / 256; ++it) {
        const int e = tid + 256 * it;
        const int p = e / kPlLines, l = e - p * kPlLines;
        const int i = seg0 + p, line = line0 + l;
        if (i < n && line < A.nlines)
            gptr_w<T>(A.out)[(size_t)i * A.out_pitch + line] = x[l][H + p];
    }
}

template <int FILT>
static hipError_t launch_lines(const LineArgs &A, hipStream_t s)
{
    dim3 grid((A.n + kPlSeg - 1) / kPlSeg, (A.nlines + kPlLines - 1) / kPlLines);
    hipLaunchKernelGGL(plain_line_kernel<FILT>, grid, dim3(256), 0, s, A);
    return hipGetLastError();
}

// ---------------------------------------------------------------- colour --

struct ColourArgs {
    const void *c[4];
    int in_pitch;  // elements
    int w, h;
    uint8_t *out;
    int out_pitch_px;
    int fmt, bd, check_out, nc, is_float;
};

template <int FILT>
__global__ void plain_rgb_kernel(ColourArgs A)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= A.w) return;
    const size_t i = (size_t)y * A.in_pitch + x;
    const int od = fmt_depth(A.fmt);
    int R, G, B;
    if constexpr (FILT == 97)
        ict_pixel(((const float *)A.c[0])[i], ((const float *)A.c[1])[i], ((const float *)A.c[2])[i],
                  A.bd, od, A.check_out, R, G, B);
    else
        rct_pixel(((const int *)A.c[0])[i], ((const int *)A.c[1])[i], ((const int *)A.c[2])[i], A.bd,
                  od, A.check_out, R, G, B);
    put_pixel(A.out + ((size_t)y * A.out_pitch_px + x) * fmt_bytes(A.fmt), A.fmt, R, G, B, 0);
}

__global__ void plain_merge_kernel(ColourArgs A)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= A.w) return;
    const size_t i = (size_t)y * A.in_pitch + x;
    const int od = fmt_depth(A.fmt);
    int t[4] = {0, 0, 0, 0};
    for (int c = 0; c < A.nc && c < 4; ++c) {
        if (!A.c[c]) continue;
        const int s = A.is_float ? (int)((const float *)A.c[c])[i] : ((const int *)A.c[c])[i];
        t[c] = merge_value(s, A.bd, od);
    }
    int R, G, B, Al;
    merge_slots(t, A.fmt, A.check_out == 0, R, G, B, Al);
    put_pixel(A.out + ((size_t)y * A.out_pitch_px + x) * fmt_bytes(A.fmt), A.fmt, R, G, B, Al);
}

// bmi_dc_shift_kernel (bmi_dwt_gpu_kernel.h:709-751)
__global__ void plain_dc_shift_kernel(const void *src, int *dst, int sx, int sy, int dx, int dy,
                                      int pitch, int w, int is_float, int bd)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= w) return;
    const size_t si = (size_t)(sy + y) * pitch + sx + x;
    int t = is_float ? (int)((const float *)src)[si] : ((const int *)src)[si];
    t += 1 << (bd - 1);
    dst[(size_t)(dy + y) * pitch + dx + x] = clamp_i(t, (1 << bd) - 1);
}

// Grey preview of one component (bmi_get_one_component, gpufunc.cu:1404-1480;
// kernels bmi_dwt_gpu_kernel.h:1644-1836).  Follows the reference kernels
// literally: the output index uses the component pitch, 5/3 chroma uses
// (v + 2^B - 1) >> 1, and the 9/7 kernel reads the plane words as the
// fixed-point integers of the reference's older integer 9/7 path.
__global__ void plain_one_comp_kernel(int mode, int comp_id, const int *in, int *out, int xo, int yo,
                                      int in_pitch, int w, int bd)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= w) return;
    const size_t idx = (size_t)(yo + y) * in_pitch + xo + x;
    const int maxV = (1 << bd) - 1;
    int o;
    if (mode == 0) {  // from a packed 32-bit result
        o = (out[idx] >> ((2 - comp_id) * 8)) & 0xff;
    } else if (mode == 1) {  // 5/3 integer plane
        o = in[idx];
        if (comp_id == 0) o += 1 << (bd - 1);
        else o = (o + (1 << bd) - 1) >> 1;
        o = clamp_i(o, maxV);
        o = bd >= 8 ? (o >> (bd - 8)) : (o << (8 - bd));
        o &= 0xff;
    } else {  // 9/7: fixed-point rounding of the raw words
        int t = in[idx];
        const int flag = (t < 0) ? -1 : 1;
        t = flag * ((flag * t + 0x1000) & 0xfffe000);
        t >>= 13;
        t += 1 << (bd - 1);
        o = clamp_i(t, maxV) >> (bd - 8);
    }
    out[idx] = o | (o << 8) | (o << 16);
}

}  // namespace bmi

using namespace bmi;

// ------------------------------------------------------------- IDWT entry --

static int level_count(const CompInfo_c *comp) { return (comp->mNumOfSubband - 1) / 3; }

// True when the region of `h` rows of `wbytes` bytes starting at byte x0 of
// row y0 lies inside `b` addressed with `pitch` bytes per row.  Buffers
// whose mSize is 0 (described by the caller, not allocated by
// MallocGpuBuf) are not checked.
static bool region_fits(const CudaBuf *b, size_t pitch, long long x0, long long y0,
                        long long wbytes, long long h)
{
    if (!b->mSize || wbytes <= 0 || h <= 0) return true;
    if (x0 < 0 || y0 < 0) return false;
    return (unsigned long long)(y0 + h - 1) * pitch + x0 + wbytes <= b->mSize;
}

// One synthesis level `level` of a tile-component, in place in orig.
static int plain_level(TileInfo_c *tile, CompInfo_c *comp, CudaBuf *orig, CudaBuf *temp,
                       hipStream_t s, EncodeMathod method, int level, int coef_kind, int r53_mode)
{
    const SubbandInfo_c *sb = comp->mSubbandInfo;
    const int hl = 3 * level + 1, lh = 3 * level + 2, hh = 3 * level + 3;
    // geometry from the subband table, as gpufunc.cu:458-468
    const int lw = sb[lh].mSize.x, nw = sb[lh].mSize.x + sb[hh].mSize.x;
    const int lhh = sb[hl].mSize.y, nh = sb[hl].mSize.y + sb[hh].mSize.y;
    if (nw <= 0 || nh <= 0) return BMI_OK;
    const int px = tile->mXParity[level] & 1, py = tile->mYParity[level] & 1;
    const int opitch = (int)(orig->mPitch / 4), tpitch = (int)(temp->mPitch / 4);
    if (tpitch < nh || opitch < nw) return bmi_fail(BMI_E_INVALID, "comp_idwt: pitch too small");
    if (!region_fits(orig, orig->mPitch, 4LL * tile->mOff.x, tile->mOff.y, 4LL * nw, nh) ||
        !region_fits(temp, temp->mPitch, 4LL * tile->mOff.y, tile->mOff.x, 4LL * nh, nw))
        return bmi_fail(BMI_E_INVALID, "comp_idwt: tile outside its buffers");
    char *obase = (char *)orig->mBuf + (size_t)orig->mPitch * tile->mOff.y + 4 * (size_t)tile->mOff.x;
    char *tbase = (char *)temp->mBuf + (size_t)temp->mPitch * tile->mOff.x + 4 * (size_t)tile->mOff.y;

    LineArgs R;
    memset(&R, 0, sizeof(R));
    R.in = obase;
    R.in_pitch = opitch;
    R.out = tbase;
    R.out_pitch = tpitch;
    R.nlines = nh;
    R.n = nw;
    R.lo_n = lw;
    R.parity = px;
    R.split = lhh;
    LineArgs C = R;
    C.in = tbase;
    C.in_pitch = tpitch;
    C.out = obase;
    C.out_pitch = opitch;
    C.nlines = nw;
    C.n = nh;
    C.lo_n = lhh;
    C.parity = py;
    C.split = nw;
    hipError_t e;
    if (method == Ckernels_W9X7) {
        const int ck = (coef_kind == BMI_COEF_FLOAT) ? SRC_FLOAT : SRC_SIGNMAG;
        // CPU gain folding, bmi_cpu_dwt_decoder.cpp:1433-1473, 1599, 1626
        R.g[0] = level == 0 ? (sb[0].mfAbsStepNorm * kLoGain) * kLoGain : kLoGain * kLoGain;
        R.g[1] = (sb[hl].mfAbsStepNorm * kHiGain) * kLoGain;
        R.g[2] = (sb[lh].mfAbsStepNorm * kLoGain) * kHiGain;
        R.g[3] = (sb[hh].mfAbsStepNorm * kHiGain) * kHiGain;
        R.kind_top_lo = level == 0 ? ck : SRC_FLOAT;
        R.kind_rest = ck;
        C.g[0] = C.g[1] = C.g[2] = C.g[3] = 1.0f;
        C.kind_top_lo = C.kind_rest = SRC_FLOAT;
        if ((e = launch_lines<97>(R, s)) != hipSuccess) return bmi_check_hip(e, "comp_idwt rows");
        if ((e = launch_lines<97>(C, s)) != hipSuccess) return bmi_check_hip(e, "comp_idwt cols");
    } else if (method == Ckernels_W5X3) {
        R.kind_top_lo = R.kind_rest = SRC_INT;
        C.kind_top_lo = C.kind_rest = SRC_INT;
        C.quirk = (r53_mode == BMI_R53_REFERENCE && !py) ? 1 : 0;
        if ((e = launch_lines<53>(R, s)) != hipSuccess) return bmi_check_hip(e, "comp_idwt rows");
        if ((e = launch_lines<53>(C, s)) != hipSuccess) return bmi_check_hip(e, "comp_idwt cols");
    } else {
        return bmi_fail(BMI_E_INVALID, "comp_idwt: method %d", (int)method);
    }
    return BMI_OK;
}

static int check_idwt_args(TileInfo_c *tile, CompInfo_c *comp, CudaBuf *orig, CudaBuf *temp,
                           short wordShift)
{
    if (!tile || !comp || !orig || !temp || !orig->mBuf || !temp->mBuf || !comp->mSubbandInfo)
        return bmi_fail(BMI_E_INVALID, "comp_idwt: null argument");
    if (wordShift != 2) return bmi_fail(BMI_E_UNSUPPORTED, "comp_idwt: wordShift %d", wordShift);
    const int L = level_count(comp);
    if (L < 0 || L > MAX_DWT_LEVEL) return bmi_fail(BMI_E_INVALID, "comp_idwt: %d levels", L);
    return BMI_OK;
}

extern "C" int bmi_comp_idwt_ex(TileInfo_c *tile, CompInfo_c *comp, CudaBuf *orig, CudaBuf *temp,
                                StreamId asyncId, short wordShift, EncodeMathod mathod,
                                int toplevel, int coef_kind, int r53_mode)
{
    int rc = check_idwt_args(tile, comp, orig, temp, wordShift);
    if (rc) return rc;
    int L = level_count(comp);
    if (toplevel != -1) L = std::min(L, toplevel);
    for (int level = 0; level < L; ++level) {
        rc = plain_level(tile, comp, orig, temp, bmi_stream(asyncId), mathod, level, coef_kind,
                         r53_mode);
        if (rc) return rc;
    }
    return BMI_OK;
}

extern "C" int BMI_comp_idwt(TileInfo_c *tile, CompInfo_c *comp, CudaBuf *orig, CudaBuf *temp,
                             StreamId asyncId, short wordShift, EncodeMathod mathod, int toplevel)
{
    return bmi_comp_idwt_ex(tile, comp, orig, temp, asyncId, wordShift, mathod, toplevel,
                            BMI_COEF_SIGNMAG, BMI_get_r53_mode());
}

extern "C" int BMI_resolution_idwt(TileInfo_c *tile, CompInfo_c *comp, CudaBuf *orig,
                                   CudaBuf *temp, StreamId asyncId, short wordShift,
                                   EncodeMathod mathod, int resolutionLevel)
{
    int rc = check_idwt_args(tile, comp, orig, temp, wordShift);
    if (rc) return rc;
    if (resolutionLevel < 0 || resolutionLevel >= level_count(comp))
        return bmi_fail(BMI_E_INVALID, "resolution_idwt: level %d", resolutionLevel);
    return plain_level(tile, comp, orig, temp, bmi_stream(asyncId), mathod, resolutionLevel,
                       BMI_COEF_SIGNMAG, BMI_get_r53_mode());
}

// ----------------------------------------------------------- colour entry --

// decoded size and result offset of a tile at decodeLevel (gpufunc.cu:1100-1119)
static bool colour_geometry(TileInfo_c *t, int dwtLevel, int decodeLevel, int *w, int *h, int *rx,
                            int *ry)
{
    *w = t->mSize.x;
    *h = t->mSize.y;
    *rx = t->mOff.x;
    *ry = t->mOff.y;
    if (decodeLevel < 0 || decodeLevel > dwtLevel) decodeLevel = -1;
    if (decodeLevel == 0) return false;  // the reference indexes subband -2 here
    if (decodeLevel != -1) {
        const SubbandInfo_c *sb = t->mCompInfo[0].mSubbandInfo;
        *h = sb[decodeLevel * 3 - 2].mSize.y + sb[decodeLevel * 3 - 1].mSize.y;
        *w = sb[decodeLevel * 3 - 2].mSize.x + sb[decodeLevel * 3 - 1].mSize.x;
        while (dwtLevel > decodeLevel) {
            *rx = (*rx + 1) / 2;
            *ry = (*ry + 1) / 2;
            --dwtLevel;
        }
    }
    return true;
}

static bool rgb_format(int f)
{
    return f == IMAGE_FOTMAT_32BITS_ARGB || f == IMAGE_FOTMAT_48BITS_RGB ||
           f == IMAGE_FOTMAT_24BITS_RGB || f == IMAGE_FOTMAT_64BITS_ARGB;
}

static void tile_rgb(TileInfo_c *tile, short bitDepth, short wordShift, int dwtLevel,
                     int decodeLevel, OutputFormat fmt, StreamId asyncId, bool ict, int check_out)
{
    if (!tile || !tile->mCompInfo || !tile->mResultBuf || !tile->mResultBuf->mBuf) {
        bmi_fail(BMI_E_INVALID, "tile colour: null argument");
        return;
    }
    if (!rgb_format(fmt)) {
        // *_COMPONENTS / NONE: nothing to do (gpulib.c:277-283); the 12-bit
        // formats are rejected as in the CPU back end
        if (fmt == IMAGE_FOTMAT_36BITS_RGB || fmt == IMAGE_FOTMAT_48BITS_ARGB)
            bmi_fail(BMI_E_UNSUPPORTED, "tile colour: format %d", (int)fmt);
        return;
    }
    if (wordShift != 2) {
        bmi_fail(BMI_E_UNSUPPORTED, "tile colour: wordShift %d", wordShift);
        return;
    }
    int w, h, rx, ry;
    if (!colour_geometry(tile, dwtLevel, decodeLevel, &w, &h, &rx, &ry)) {
        bmi_fail(BMI_E_UNSUPPORTED, "tile colour: decodeLevel 0");
        return;
    }
    ColourArgs A;
    memset(&A, 0, sizeof(A));
    const int pb = fmt_bytes(fmt);
    for (int c = 0; c < 3; ++c) {
        CudaBuf *b = tile->mCompInfo[c].mCompBuf;
        if (!b || !b->mBuf) {
            bmi_fail(BMI_E_INVALID, "tile colour: component %d buffer", c);
            return;
        }
        A.c[c] = (const char *)b->mBuf + (size_t)b->mPitch * tile->mOff.y + 4 * (size_t)tile->mOff.x;
    }
    A.in_pitch = (int)(tile->mCompInfo[0].mCompBuf->mPitch / 4);
    A.w = w;
    A.h = h;
    if (!region_fits(tile->mResultBuf, tile->mResultBuf->mPitch, (long long)rx * pb, ry,
                     (long long)w * pb, h)) {
        bmi_fail(BMI_E_INVALID, "tile colour: result region outside the result buffer");
        return;
    }
    A.out_pitch_px = (int)(tile->mResultBuf->mPitch / pb);
    A.out = (uint8_t *)tile->mResultBuf->mBuf + ((size_t)ry * A.out_pitch_px + rx) * pb;
    A.fmt = fmt;
    A.bd = bitDepth;
    A.check_out = check_out;
    if (w <= 0 || h <= 0) return;
    dim3 grid((w + 255) / 256, h);
    if (ict)
        hipLaunchKernelGGL(plain_rgb_kernel<97>, grid, dim3(256), 0, bmi_stream(asyncId), A);
    else
        hipLaunchKernelGGL(plain_rgb_kernel<53>, grid, dim3(256), 0, bmi_stream(asyncId), A);
    bmi_check_hip(hipGetLastError(), "tile colour launch");
}

extern "C" void BMI_tile_MCT(TileInfo_c *tileInfo, short bitDepth, short wordShift, int dwtlevel,
                             int decodeLevel, OutputFormat outFormat, StreamId asyncId)
{
    bmi_clear_error();
    tile_rgb(tileInfo, bitDepth, wordShift, dwtlevel, decodeLevel, outFormat, asyncId, false, -1);
}

extern "C" void BMI_tile_ICT(TileInfo_c *tileInfo, short bitDepth, short wordShift, int dwtlevel,
                             int decodeLevel, OutputFormat outFormat, StreamId asyncId)
{
    bmi_clear_error();
    tile_rgb(tileInfo, bitDepth, wordShift, dwtlevel, decodeLevel, outFormat, asyncId, true, -1);
}

static void tile_merge(TileInfo_c *tile, int componentNum, int dwtLevel, int decodeLevel,
                       short bitDepth, short wordShift, int compBufPitchBytes, OutputFormat fmt,
                       EncodeMathod method, StreamId asyncId, int check_out)
{
    if (fmt == IMAGE_FOTMAT_NONE) return;  // gpulib.c:333-335
    if (!tile || !tile->mCompInfo || !tile->mResultBuf || !tile->mResultBuf->mBuf) {
        bmi_fail(BMI_E_INVALID, "tile merge: null argument");
        return;
    }
    if (fmt == IMAGE_FOTMAT_36BITS_RGB || fmt == IMAGE_FOTMAT_48BITS_ARGB || fmt > IMAGE_FOTMAT_48BITS_ARGB) {
        bmi_fail(BMI_E_UNSUPPORTED, "tile merge: format %d", (int)fmt);
        return;
    }
    if (wordShift != 2 || componentNum < 1 || componentNum > 4) {
        bmi_fail(BMI_E_INVALID, "tile merge: wordShift %d components %d", wordShift, componentNum);
        return;
    }
    int w, h, rx, ry;
    if (!colour_geometry(tile, dwtLevel, decodeLevel, &w, &h, &rx, &ry)) {
        bmi_fail(BMI_E_UNSUPPORTED, "tile merge: decodeLevel 0");
        return;
    }
    ColourArgs A;
    memset(&A, 0, sizeof(A));
    const int pb = fmt_bytes(fmt);
    const int pitch = compBufPitchBytes > 0 ? compBufPitchBytes : (int)tile->mCompInfo[0].mCompBuf->mPitch;
    const int first = check_out >= 0 ? check_out : 0;
    const int n = check_out >= 0 ? 1 : componentNum;
    for (int c = 0; c < n; ++c) {
        CudaBuf *b = tile->mCompInfo[first + c].mCompBuf;
        if (!b || !b->mBuf) {
            bmi_fail(BMI_E_INVALID, "tile merge: component %d buffer", first + c);
            return;
        }
        A.c[c] = (const char *)b->mBuf + (size_t)pitch * tile->mOff.y + 4 * (size_t)tile->mOff.x;
    }
    A.nc = n;
    A.in_pitch = pitch / 4;
    A.w = w;
    A.h = h;
    if (!region_fits(tile->mResultBuf, tile->mResultBuf->mPitch, (long long)rx * pb, ry,
                     (long long)w * pb, h)) {
        bmi_fail(BMI_E_INVALID, "tile merge: result region outside the result buffer");
        return;
    }
    A.out_pitch_px = (int)(tile->mResultBuf->mPitch / pb);
    A.out = (uint8_t *)tile->mResultBuf->mBuf + ((size_t)ry * A.out_pitch_px + rx) * pb;
    A.fmt = fmt;
    A.bd = bitDepth;
    A.check_out = check_out >= 0 ? 0 : -1;
    A.is_float = method == Ckernels_W9X7;
    if (w <= 0 || h <= 0) return;
    hipLaunchKernelGGL(plain_merge_kernel, dim3((w + 255) / 256, h), dim3(256), 0,
                       bmi_stream(asyncId), A);
    bmi_check_hip(hipGetLastError(), "tile merge launch");
}

extern "C" void BMI_tile_Merge(TileInfo_c *tileInfo, int componentNum, int dwtLevel,
                               int decodeLevel, short bitDepth, short wordShift,
                               int compBufPitchBytes, OutputFormat outFormat, EncodeMathod method,
                               StreamId asyncId)
{
    bmi_clear_error();
    tile_merge(tileInfo, componentNum, dwtLevel, decodeLevel, bitDepth, wordShift,
               compBufPitchBytes, outFormat, method, asyncId, -1);
}

extern "C" void BMI_Comp_DC_SHIFT(CudaBuf *compBuf, CudaBuf *compBufTemp, EncodeMathod method,
                                  int src_x_off, int src_y_off, int dst_x_off, int dst_y_off,
                                  int width, int height, int bitDepth, int pixelWordLength,
                                  StreamId asyncId)
{
    bmi_clear_error();
    if (!compBuf || !compBufTemp || !compBuf->mBuf || !compBufTemp->mBuf || pixelWordLength != 4 ||
        width < 0 || height < 0 || bitDepth < 1 || bitDepth > 30) {
        bmi_fail(BMI_E_INVALID, "Comp_DC_SHIFT: bad args");
        return;
    }
    if (!width || !height) return;
    // both buffers addressed with compBuf's pitch, as gpufunc.cu:397-413
    if (!region_fits(compBuf, compBuf->mPitch, 4LL * src_x_off, src_y_off, 4LL * width, height) ||
        !region_fits(compBufTemp, compBuf->mPitch, 4LL * dst_x_off, dst_y_off, 4LL * width, height)) {
        bmi_fail(BMI_E_INVALID, "Comp_DC_SHIFT: region outside a buffer");
        return;
    }
    hipLaunchKernelGGL(plain_dc_shift_kernel, dim3((width + 255) / 256, height), dim3(256), 0,
                       bmi_stream(asyncId), compBuf->mBuf, (int *)compBufTemp->mBuf, src_x_off,
                       src_y_off, dst_x_off, dst_y_off, (int)(compBuf->mPitch / pixelWordLength),
                       width, method == Ckernels_W9X7 ? 1 : 0, bitDepth);
    bmi_check_hip(hipGetLastError(), "Comp_DC_SHIFT launch");
}

extern "C" void BMI_tile_One_Component(TileInfo_c *tile, int component_id, EncodeMathod method,
                                       int hasMCT, int bitDepth, int wordShift, int dwtLevel,
                                       int decodeLevel, StreamId asyncId)
{
    bmi_clear_error();
    (void)dwtLevel;
    if (!tile || !tile->mCompInfo || !tile->mResultBuf || !tile->mResultBuf->mBuf ||
        component_id < 0 || component_id > 3 || wordShift != 2) {
        bmi_fail(BMI_E_INVALID, "tile_One_Component: bad args");
        return;
    }
    int w = tile->mSize.x, h = tile->mSize.y;
    if (decodeLevel == 0) {
        bmi_fail(BMI_E_UNSUPPORTED, "tile_One_Component: decodeLevel 0");
        return;
    }
    if (decodeLevel > 0) {
        const SubbandInfo_c *sb = tile->mCompInfo[0].mSubbandInfo;
        h = sb[decodeLevel * 3 - 2].mSize.y + sb[decodeLevel * 3 - 1].mSize.y;
        w = sb[decodeLevel * 3 - 2].mSize.x + sb[decodeLevel * 3 - 1].mSize.x;
    }
    if (w <= 0 || h <= 0) return;
    int mode;
    const int *in = nullptr;
    int pitch;
    if (hasMCT) {
        mode = 0;
        pitch = (int)(tile->mResultBuf->mPitch / 4);
    } else {
        CudaBuf *b = tile->mCompInfo[component_id].mCompBuf;
        if (!b || !b->mBuf) {
            bmi_fail(BMI_E_INVALID, "tile_One_Component: component buffer");
            return;
        }
        if (method == Ckernels_W9X7 && bitDepth < 8) {
            bmi_fail(BMI_E_UNSUPPORTED, "tile_One_Component: 9/7 below 8 bits");
            return;
        }
        mode = method == Ckernels_W5X3 ? 1 : 2;
        in = (const int *)b->mBuf;
        pitch = (int)(b->mPitch / 4);
    }
    // the kernels index the result with the same pitch as their input
    // (gpufunc.cu:1443-1455): the whole region must fit the result buffer
    if (!region_fits(tile->mResultBuf, 4LL * pitch, 4LL * tile->mOff.x, tile->mOff.y, 4LL * w, h) ||
        (in && !region_fits(tile->mCompInfo[component_id].mCompBuf, 4LL * pitch,
                            4LL * tile->mOff.x, tile->mOff.y, 4LL * w, h))) {
        bmi_fail(BMI_E_INVALID, "tile_One_Component: region outside a buffer");
        return;
    }
    hipLaunchKernelGGL(plain_one_comp_kernel, dim3((w + 255) / 256, h), dim3(256), 0,
                       bmi_stream(asyncId), mode, component_id, in, (int *)tile->mResultBuf->mBuf,
                       (int)tile->mOff.x, (int)tile->mOff.y, pitch, w, bitDepth);
    bmi_check_hip(hipGetLastError(), "tile_One_Component launch");
}

// ------------------------------------------------------------ plain driver --

// The CudaManager sequence for one tile (cudainterface.cpp:1241-1630):
// BMI_comp_idwt per component, then ICT / MCT / Merge into the result.
extern "C" int BMI_plain_decode(const BmiTileDesc *tin, void *temp, size_t temp_bytes, StreamId stream)
{
    if (!tin || !temp) return bmi_fail(BMI_E_INVALID, "plain_decode: null argument");
    // ignored components are not decoded; the rest fill the lower slots
    BmiTileDesc eff;
    if (int rc = bmi_apply_ignore(*tin, &eff)) return bmi_fail(rc, "plain_decode: ignore mask");
    const BmiTileDesc *t = &eff;
    if (t->levels < 0 || t->levels > MAX_DWT_LEVEL || t->ncomp < 1 || t->ncomp > 4 ||
        t->width < 1 || t->height < 1)
        return bmi_fail(BMI_E_INVALID, "plain_decode: bad tile");
    if (t->top_level == 0)
        return bmi_fail(BMI_E_UNSUPPORTED, "plain_decode: thumbnails (top_level 0) are batch-plan only");
    const int L = t->levels;
    // subband table (sizes as cudainterface.cpp:674-703 from the standard geometry)
    std::vector<SubbandInfo_c> sbs((size_t)t->ncomp * (3 * L + 1));
    for (int c = 0; c < t->ncomp; ++c) {
        SubbandInfo_c *sb = &sbs[(size_t)c * (3 * L + 1)];
        memset(sb, 0, sizeof(SubbandInfo_c) * (3 * L + 1));
        for (int lev = 0; lev < L; ++lev) {
            const int s = L - 1 - lev;
            const int a0 = ceil_shift(t->x0, s), a1 = ceil_shift(t->x0 + t->width, s);
            const int b0 = ceil_shift(t->y0, s), b1 = ceil_shift(t->y0 + t->height, s);
            const int lw = ceil_div2(a1) - ceil_div2(a0), hw = (a1 - a0) - lw;
            const int lh = ceil_div2(b1) - ceil_div2(b0), hh = (b1 - b0) - lh;
            if (lev == 0) {
                sb[0].mSize.x = (short)lw;
                sb[0].mSize.y = (short)lh;
                sb[0].mfAbsStepNorm = t->steps[c][0];
            }
            SubbandInfo_c *b = &sb[3 * lev + 1];
            b[0].mSize.x = (short)hw; b[0].mSize.y = (short)lh; b[0].mOff.x = (short)lw;
            b[1].mSize.x = (short)lw; b[1].mSize.y = (short)hh; b[1].mOff.y = (short)lh;
            b[2].mSize.x = (short)hw; b[2].mSize.y = (short)hh; b[2].mOff.x = (short)lw; b[2].mOff.y = (short)lh;
            for (int k = 0; k < 3; ++k) {
                b[k].mId = 3 * lev + 1 + k;
                b[k].mfAbsStepNorm = t->steps[c][3 * lev + 1 + k];
            }
        }
    }
    std::vector<CudaBuf> bufs(t->ncomp);
    std::vector<CompInfo_c> comps(t->ncomp);
    for (int c = 0; c < t->ncomp; ++c) {
        memset(&bufs[c], 0, sizeof(CudaBuf));
        bufs[c].mType = GPU_ALLIGN_BUF;
        bufs[c].mBuf = (void *)t->coef[c];
        bufs[c].mPitch = (unsigned)t->coef_pitch * 4;
        bufs[c].mWidth = (unsigned)t->width * 4;
        bufs[c].mHeight = (unsigned)t->height;
        memset(&comps[c], 0, sizeof(CompInfo_c));
        comps[c].mSubbandInfo = &sbs[(size_t)c * (3 * L + 1)];
        comps[c].mNumOfSubband = 3 * L + 1;
        comps[c].mCompBuf = &bufs[c];
    }
    CudaBuf tempbuf;
    memset(&tempbuf, 0, sizeof(tempbuf));
    tempbuf.mType = GPU_ALLIGN_BUF;
    tempbuf.mBuf = temp;
    tempbuf.mPitch = (unsigned)t->height * 4;
    tempbuf.mHeight = (unsigned)t->width;
    if ((size_t)tempbuf.mPitch * tempbuf.mHeight > temp_bytes)
        return bmi_fail(BMI_E_INVALID, "plain_decode: temp needs %zu bytes",
                        (size_t)tempbuf.mPitch * tempbuf.mHeight);
    TileInfo_c tile;
    memset(&tile, 0, sizeof(tile));
    tile.mSize.x = (short)t->width;
    tile.mSize.y = (short)t->height;
    tile.mCompInfo = comps.data();
    // parities as cudainterface.cpp:607-615
    for (int d = L - 1, xo = t->x0, yo = t->y0; d >= 0; --d) {
        tile.mXParity[d] = xo & 1;
        tile.mYParity[d] = yo & 1;
        xo = ceil_div2(xo);
        yo = ceil_div2(yo);
    }
    const int top = t->top_level;
    const bool need_all = t->out_format == IMAGE_FOTMAT_NONE || !rgb_format(t->out_format);
    for (int c = 0; c < t->ncomp; ++c) {
        if (!need_all && c > 2) continue;
        int rc = bmi_comp_idwt_ex(&tile, &comps[c], &bufs[c], &tempbuf, stream, 2,
                                  (EncodeMathod)t->method, top, t->coef_kind, t->r53_mode);
        if (rc) return rc;
    }
    if (t->out_format == IMAGE_FOTMAT_NONE) return BMI_OK;
    const int pb = fmt_bytes(t->out_format);
    CudaBuf result;
    memset(&result, 0, sizeof(result));
    result.mType = GPU_BUF;
    result.mPitch = (unsigned)(t->out_pitch_px * pb);
    result.mBuf = (char *)t->out + ((size_t)t->out_y * t->out_pitch_px + t->out_x) * pb;
    tile.mResultBuf = &result;
    const int decode_level = (top < 0 || top >= L) ? -1 : top;
    bmi_clear_error();
    if (rgb_format(t->out_format)) {
        if (t->ncomp < 3) return bmi_fail(BMI_E_UNSUPPORTED, "plain_decode: RGB needs 3 components");
        tile_rgb(&tile, (short)t->bit_depth, 2, L, decode_level, (OutputFormat)t->out_format, stream,
                 t->method == Ckernels_W9X7, t->check_out);
    } else {
        tile_merge(&tile, t->ncomp, L, decode_level, (short)t->bit_depth, 2, 0,
                   (OutputFormat)t->out_format, (EncodeMathod)t->method, stream, t->check_out);
    }
    return BMI_last_error();
}
