/*
 * j2k_oracle.c -- CPU restatement of the reference decoder's back end.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the thing measured or
 * shipped).  See j2k_oracle.h for the plane conventions.
 *
 * Built with -O2 -ffp-contract=off -fno-fast-math on x86-64 (SSE scalar
 * float, no FMA), which is the float evaluation the reference's C paths get
 * on its x64 and Mac/Linux SSE builds.
 *
 * Parity status of each function (what pins it, besides the GPU):
 *  - 5/3 textbook: forward/inverse identity (or_fdwt53), impulse and
 *    constant cases, brute force on all sizes 1..33 (tests/test_oracle.py).
 *  - 5/3 reference mode: differs from textbook only on the rows the
 *    closed form of the reference's VDWT_R53_LINES bottom edge names
 *    (tests/test_oracle.py::test_r53_quirk_closed_form).  That the
 *    reference binary behaves this way is a hand trace of the source; the
 *    reference cannot be built here (needs cuda_runtime.h and 32-bit asm).
 *  - 9/7: FP64 oracle and FP64 forward/inverse round trip; float vs FP64
 *    max-abs / RMS bounded; symmetric-extension identities.
 *  - RCT: exhaustive forward/inverse identity on 8-bit and sampled 12-bit.
 *  - ICT / merge / DC shift: FP64 formula, exhaustive grid over the clamp
 *    and truncation order.
 *  - One-component grey preview: byte extraction against numpy bit
 *    operations on packed words; the 9/7 fixed-point reading against
 *    round-half-away-from-zero of w / 2^13 on a dense sample and at the
 *    half-way points; the 5/3 shift against the 8-bit slot rule at 4..16
 *    bits (tests/test_oracle.py).
 */
#include "j2k_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* decoder/codec/include/bmi_cpu_dwt.h:31-37 */
static const float F_ALPHA = -1.58613434205992f;
static const float F_BETA = -0.05298011857296f;
static const float F_GAMMA = 0.88291107553093f;
static const float F_DELTA = 0.44350685204397f;
static const float F_LO = 1.23017410558578f;
static const float F_HI = 1.62578613134411f;

/* ITU-T T.800 Table F.4 values (FP64 oracle) */
static const double D_ALPHA = -1.586134342059924;
static const double D_BETA = -0.052980118572961;
static const double D_GAMMA = 0.882911075530934;
static const double D_DELTA = 0.443506852043971;
static const double D_K = 1.230174104914001;

static int ceil_shift(int v, int s) { return (v + (1 << s) - 1) >> s; }
static int ceil2(int v) { return (v + 1) >> 1; }

void or_level_span(int t0, int n, int L, int lev, int *a0, int *a1)
{
    /* resolution produced by synthesis level lev; parities as in
       cudainterface.cpp:607-615 (ceil-halving the tile offset) */
    int s = L - 1 - lev;
    *a0 = ceil_shift(t0, s);
    *a1 = ceil_shift(t0 + n, s);
}

/* two's complement of a T1 sign-magnitude word: FETCHPIXEL_INT,
   bmi_cpu_dwt_decoder.cpp:2287 ((w-1)^0x7fffffff for negatives) */
static int32_t signmag(uint32_t w)
{
    return (w & 0x80000000u) ? -(int32_t)(w & 0x7fffffffu) : (int32_t)w;
}

static int mirror(int i, int n)
{
    if (n == 1) return 0;
    int period = 2 * (n - 1);
    i %= period;
    if (i < 0) i += period;
    return i < n ? i : period - i;
}

/* ---------------------------------------------------------------- 5/3 -- */

void or_h53_line_ref(int32_t *t, const int32_t *lo, const int32_t *hi, int lw, int hw)
{
    /* HDWT_R53_INT, bmi_cpu_dwt_decoder.cpp:2609-2659, one line.  The
       reference treats the first sample as low whatever the parity (H2).
       Length 1: the reference copies lo[0] without advancing its pointers
       (correct only for a single line); the single-line result is kept. */
    int width = lw + hw;
    if (width < 2) {
        t[0] = lo[0];
        return;
    }
    width -= 2;
    t[0] = *lo++ - ((*hi + 1) >> 1);
    while (width > 0) {
        if (width >= 2)
            t[2] = *lo++ - ((hi[0] + hi[1] + 2) >> 2);
        else
            t[2] = *lo++ - ((hi[0] + 1) >> 1);
        t[1] = *hi++ + ((t[0] + t[2]) >> 1);
        t += 2;
        width -= 2;
    }
    if (!width) t[1] = *hi + t[0];
}

void or_53_line_textbook(int32_t *x, int n, int p)
{
    /* ITU-T T.800 F.3.7/F.3.8.1 (1D_SR + 1D_FILTR_5-3R), whole-sample
       symmetric extension; position i is a low sample when (i+p) is even */
    if (n == 1) {
        if (p) x[0] >>= 1;
        return;
    }
    for (int i = p ? 1 : 0; i < n; i += 2)
        x[i] -= (x[mirror(i - 1, n)] + x[mirror(i + 1, n)] + 2) >> 2;
    for (int i = p ? 0 : 1; i < n; i += 2)
        x[i] += (x[mirror(i - 1, n)] + x[mirror(i + 1, n)]) >> 1;
}

void or_v53_ref(int32_t *out, const int32_t *hres, int lh, int hh, int width)
{
    /* VDWT_R53 -> VDWT_R53_LINES with isStart && isEnd
       (bmi_cpu_dwt_decoder.cpp:3311-3389, 3391-3425): one pass down the
       rows.  `pre` is set from the current high row before the input
       pointers advance and is never re-synchronised, so the last odd row
       (even heights >= 4) and the last two rows (odd heights) use stale
       rows (H1).  Height 1 reads past the high rows in the reference; the
       textbook identity is used there. */
    int n = lh + hh;
    if (n == 1) {
        memcpy(out, hres, sizeof(int32_t) * width);
        return;
    }
    const int32_t *L = hres;              /* inputLine[0] */
    const int32_t *H = hres + (size_t)lh * width; /* inputLine[1] */
    const int32_t *pre = H;
    int32_t *d0 = out, *d1 = out + width, *d2 = out + 2 * width;
    int left = n - 2;
    for (int x = 0; x < width; ++x) d0[x] = L[x] - ((pre[x] + H[x] + 2) >> 2);
    while (left >= 2) {
        left -= 2;
        pre = H;
        L += width;
        H += width;
        for (int x = 0; x < width; ++x) {
            d2[x] = L[x] - ((pre[x] + H[x] + 2) >> 2);
            d1[x] = pre[x] + ((d2[x] + d0[x]) >> 1);
        }
        d0 = d2;
        d1 = d2 + width;
        d2 = d1 + width;
    }
    if (left) {
        for (int x = 0; x < width; ++x) {
            d2[x] = L[x] - ((pre[x] + 1) >> 1);
            d1[x] = pre[x] + ((d2[x] + d0[x]) >> 1);
        }
    } else {
        for (int x = 0; x < width; ++x) d1[x] = pre[x] + d0[x];
    }
}

static void interleave_i32(int32_t *x, const int32_t *lo, const int32_t *hi, int n, int p)
{
    for (int i = 0; i < n; ++i)
        x[i] = ((i + p) & 1) ? hi[(i - 1 + p) >> 1] : lo[(i - p) >> 1];
}

int or_idwt53(int32_t *plane, int pitch, int w, int h, int x0, int y0, int levels,
              int toplevel, int mode)
{
    /* bmi_cpu_dwt_decoder.cpp:1509-1532 (HDWT_R53_INT per stripe) and
       :1561-1580 (VDWT_R53 per resolution), level by level from coarse to
       fine, on the Mallat plane instead of the CPU's packed buffers. */
    if (levels < 0 || levels > 32 || w <= 0 || h <= 0) return -1;
    int L = levels;
    int top = (toplevel >= 0 && toplevel < L) ? toplevel : L;
    int32_t *tmp = (int32_t *)malloc(sizeof(int32_t) * (size_t)w * h);
    int32_t *line = (int32_t *)malloc(sizeof(int32_t) * (size_t)(w > h ? w : h));
    int32_t *col = (int32_t *)malloc(sizeof(int32_t) * (size_t)h * 2);
    if (!tmp || !line || !col) { free(tmp); free(line); free(col); return -3; }
    for (int lev = 0; lev < top; ++lev) {
        int a0, a1, b0, b1;
        or_level_span(x0, w, L, lev, &a0, &a1);
        or_level_span(y0, h, L, lev, &b0, &b1);
        int nw = a1 - a0, nh = b1 - b0;
        /* an empty level (a tile of a few samples at an odd offset) has no
           samples to transform; the reference's line loops are undefined
           for it (its 5/3 column tail would index row -1) */
        if (nw <= 0 || nh <= 0) continue;
        int lw = ceil2(a1) - ceil2(a0), lh = ceil2(b1) - ceil2(b0);
        int px = a0 & 1, py = b0 & 1;
        int refx = (mode == OR_R53_REFERENCE) && !px;
        int refy = (mode == OR_R53_REFERENCE) && !py;
        /* horizontal: rows [0,lh) from LL|HL, rows [lh,nh) from LH|HH */
        for (int y = 0; y < nh; ++y) {
            const int32_t *row = plane + (size_t)y * pitch;
            int32_t *dst = tmp + (size_t)y * nw;
            if (refx) {
                or_h53_line_ref(dst, row, row + lw, lw, nw - lw);
            } else {
                interleave_i32(dst, row, row + lw, nw, px);
                or_53_line_textbook(dst, nw, px);
            }
        }
        /* vertical */
        if (refy) {
            int32_t *out = (int32_t *)malloc(sizeof(int32_t) * (size_t)nw * nh);
            if (!out) { free(tmp); free(line); free(col); return -3; }
            or_v53_ref(out, tmp, lh, nh - lh, nw);
            for (int y = 0; y < nh; ++y)
                memcpy(plane + (size_t)y * pitch, out + (size_t)y * nw, sizeof(int32_t) * nw);
            free(out);
        } else {
            for (int x = 0; x < nw; ++x) {
                for (int i = 0; i < nh; ++i) {
                    int r = ((i + py) & 1) ? lh + ((i - 1 + py) >> 1) : ((i - py) >> 1);
                    col[i] = tmp[(size_t)r * nw + x];
                }
                or_53_line_textbook(col, nh, py);
                for (int i = 0; i < nh; ++i) plane[(size_t)i * pitch + x] = col[i];
            }
        }
    }
    free(tmp);
    free(line);
    free(col);
    return 0;
}

static void fwd53_line(int32_t *x, int n, int p, int32_t *lo, int32_t *hi)
{
    /* T.800 F.4.8.1 (1D_FILTD_5-3R) with symmetric extension, then split */
    if (n == 1) {
        if (p) hi[0] = x[0] * 2; else lo[0] = x[0];
        return;
    }
    for (int i = p ? 0 : 1; i < n; i += 2)
        x[i] -= (x[mirror(i - 1, n)] + x[mirror(i + 1, n)]) >> 1;
    for (int i = p ? 1 : 0; i < n; i += 2)
        x[i] += (x[mirror(i - 1, n)] + x[mirror(i + 1, n)] + 2) >> 2;
    int nl = 0, nhh = 0;
    for (int i = 0; i < n; ++i) {
        if ((i + p) & 1) hi[nhh++] = x[i]; else lo[nl++] = x[i];
    }
}

int or_fdwt53(const int32_t *in, int w, int h, int x0, int y0, int levels, int32_t *out)
{
    /* forward 2D_SD: vertical then horizontal at each level, finest first */
    int L = levels;
    memcpy(out, in, sizeof(int32_t) * (size_t)w * h);
    int32_t *x = (int32_t *)malloc(sizeof(int32_t) * (size_t)(w > h ? w : h));
    int32_t *lo = (int32_t *)malloc(sizeof(int32_t) * (size_t)(w > h ? w : h));
    int32_t *hi = (int32_t *)malloc(sizeof(int32_t) * (size_t)(w > h ? w : h));
    for (int lev = L - 1; lev >= 0; --lev) {
        int a0, a1, b0, b1;
        or_level_span(x0, w, L, lev, &a0, &a1);
        or_level_span(y0, h, L, lev, &b0, &b1);
        int nw = a1 - a0, nh = b1 - b0;
        /* an empty level (a tile of a few samples at an odd offset) has no
           samples to transform; the reference's line loops are undefined
           for it (its 5/3 column tail would index row -1) */
        if (nw <= 0 || nh <= 0) continue;
        int lw = ceil2(a1) - ceil2(a0), lh = ceil2(b1) - ceil2(b0);
        int px = a0 & 1, py = b0 & 1;
        for (int c = 0; c < nw; ++c) {
            for (int i = 0; i < nh; ++i) x[i] = out[(size_t)i * w + c];
            fwd53_line(x, nh, py, lo, hi);
            for (int i = 0; i < lh; ++i) out[(size_t)i * w + c] = lo[i];
            for (int i = 0; i < nh - lh; ++i) out[(size_t)(lh + i) * w + c] = hi[i];
        }
        for (int r = 0; r < nh; ++r) {
            memcpy(x, out + (size_t)r * w, sizeof(int32_t) * nw);
            fwd53_line(x, nw, px, lo, hi);
            memcpy(out + (size_t)r * w, lo, sizeof(int32_t) * lw);
            memcpy(out + (size_t)r * w + lw, hi, sizeof(int32_t) * (nw - lw));
        }
    }
    free(x);
    free(lo);
    free(hi);
    return 0;
}

/* ---------------------------------------------------------------- 9/7 -- */

void or_97_line(float *x, int n, int p)
{
    /* The four lifting sweeps of HDWT_R97 (bmi_cpu_dwt_decoder.cpp:2283-2485)
       and VDWT_R97(_LINES) (:2766-3309): each update is x - c*(l + r) with
       float rounding after the add, the multiply and the subtract
       (vdwt_analyst_line, bmi_cpu_dwt.h:179-185).  The reference runs the
       sweeps as a rolling pipeline; every sample gets the same expression
       on the same operands, and a mirrored neighbour is the same value as
       its source, so the stepwise form is the same arithmetic.
       Length 1 with parity 0 keeps the gain-scaled sample (HDWT_R97 :2300,
       VDWT_R97 :3095).  Parity 1 with length 1 is a TODO in the reference
       (:2307-2310); T.800 F.3.7 (X = Y/2) is used. */
    if (n == 1) {
        if (p) x[0] = x[0] * 0.5f;
        return;
    }
    const float c[4] = {F_DELTA, F_GAMMA, F_BETA, F_ALPHA};
    for (int s = 0; s < 4; ++s) {
        int start = ((s & 1) == 0) ? (p ? 1 : 0) : (p ? 0 : 1);
        float cs = c[s];
        for (int i = start; i < n; i += 2) {
            float l = x[mirror(i - 1, n)], r = x[mirror(i + 1, n)];
            float sum = l + r;
            float prod = cs * sum;
            x[i] = x[i] - prod;
        }
    }
}

void or_97_line_f64(double *x, int n, int p)
{
    if (n == 1) {
        if (p) x[0] *= 0.5;
        return;
    }
    const double c[4] = {D_DELTA, D_GAMMA, D_BETA, D_ALPHA};
    for (int s = 0; s < 4; ++s) {
        int start = ((s & 1) == 0) ? (p ? 1 : 0) : (p ? 0 : 1);
        for (int i = start; i < n; i += 2)
            x[i] -= c[s] * (x[mirror(i - 1, n)] + x[mirror(i + 1, n)]);
    }
}

static float load97(uint32_t wd, int as_float)
{
    if (as_float) {
        float f;
        memcpy(&f, &wd, 4);
        return f;
    }
    return (float)signmag(wd);
}

int or_idwt97(uint32_t *plane, int pitch, int w, int h, int x0, int y0, int levels,
              int toplevel, const float *steps, int coef_kind)
{
    /* gains as folded by the CPU back end into the row pass
       (bmi_cpu_dwt_decoder.cpp:1433-1473, 1599, 1626): horizontal gain x
       vertical gain, each product rounded to float left to right; level-0
       LL and all detail bands are coefficients, later LL is the float
       output of the previous level. */
    if (levels < 0 || levels > 32 || w <= 0 || h <= 0) return -1;
    int L = levels;
    int top = (toplevel >= 0 && toplevel < L) ? toplevel : L;
    int cf = (coef_kind == OR_COEF_FLOAT);
    float *tmp = (float *)malloc(sizeof(float) * (size_t)w * h);
    float *col = (float *)malloc(sizeof(float) * (size_t)h);
    if (!tmp || !col) { free(tmp); free(col); return -3; }
    if (top == 0) {
        /* thumbnail (no synthesis): the LL band dequantised with its own
           step only, (float)twos(w) * step.  With the reference's gains
           folded into the synthesis passes each level has DC gain 1, so
           this is the flat-region scale of the full decode.  The
           reference's GPU thumbnail (cudainterface.cpp:1278-1293,
           1496-1510) colour-converts the LL band; it defines no 9/7
           dequantisation for it. */
        int lw = ((x0 + w + (1 << L) - 1) >> L) - ((x0 + (1 << L) - 1) >> L);
        int lh = ((y0 + h + (1 << L) - 1) >> L) - ((y0 + (1 << L) - 1) >> L);
        for (int y = 0; y < lh; ++y)
            for (int x = 0; x < lw; ++x) {
                float v = load97(plane[(size_t)y * pitch + x], cf) * steps[0];
                memcpy(&plane[(size_t)y * pitch + x], &v, 4);
            }
        free(tmp);
        free(col);
        return 0;
    }
    for (int lev = 0; lev < top; ++lev) {
        int a0, a1, b0, b1;
        or_level_span(x0, w, L, lev, &a0, &a1);
        or_level_span(y0, h, L, lev, &b0, &b1);
        int nw = a1 - a0, nh = b1 - b0;
        /* an empty level (a tile of a few samples at an odd offset) has no
           samples to transform; the reference's line loops are undefined
           for it (its 5/3 column tail would index row -1) */
        if (nw <= 0 || nh <= 0) continue;
        int lw = ceil2(a1) - ceil2(a0), lh = ceil2(b1) - ceil2(b0);
        int px = a0 & 1, py = b0 & 1;
        float gTopLo = lev == 0 ? (steps[0] * F_LO) * F_LO : F_LO * F_LO;
        float gTopHi = (steps[3 * lev + 1] * F_HI) * F_LO;
        float gBotLo = (steps[3 * lev + 2] * F_LO) * F_HI;
        float gBotHi = (steps[3 * lev + 3] * F_HI) * F_HI;
        for (int y = 0; y < nh; ++y) {
            const uint32_t *row = plane + (size_t)y * pitch;
            float *dst = tmp + (size_t)y * nw;
            int topRow = y < lh;
            float gl = topRow ? gTopLo : gBotLo, gh = topRow ? gTopHi : gBotHi;
            int loFloat = topRow ? (lev > 0 || cf) : cf;
            for (int i = 0; i < nw; ++i) {
                if ((i + px) & 1)
                    dst[i] = load97(row[lw + ((i - 1 + px) >> 1)], cf) * gh;
                else
                    dst[i] = load97(row[(i - px) >> 1], loFloat) * gl;
            }
            or_97_line(dst, nw, px);
        }
        for (int x = 0; x < nw; ++x) {
            for (int i = 0; i < nh; ++i) {
                int r = ((i + py) & 1) ? lh + ((i - 1 + py) >> 1) : ((i - py) >> 1);
                col[i] = tmp[(size_t)r * nw + x];
            }
            or_97_line(col, nh, py);
            for (int i = 0; i < nh; ++i) memcpy(&plane[(size_t)i * pitch + x], &col[i], 4);
        }
    }
    free(tmp);
    free(col);
    return 0;
}

int or_idwt97_f64(const uint32_t *plane, int pitch, int w, int h, int x0, int y0,
                  int levels, const float *steps, int coef_kind, double *out)
{
    if (levels < 0 || levels > 32 || w <= 0 || h <= 0) return -1;
    int L = levels;
    int cf = (coef_kind == OR_COEF_FLOAT);
    double *cur = out; /* Mallat double plane, pitch w */
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            uint32_t wd = plane[(size_t)y * pitch + x];
            double v;
            if (cf) { float f; memcpy(&f, &wd, 4); v = f; } else v = (double)signmag(wd);
            cur[(size_t)y * w + x] = v;
        }
    if (L == 0) return 0;
    double *tmp = (double *)malloc(sizeof(double) * (size_t)w * h);
    double *col = (double *)malloc(sizeof(double) * (size_t)h);
    const double lo = D_K, hi = 2.0 / D_K;
    for (int lev = 0; lev < L; ++lev) {
        int a0, a1, b0, b1;
        or_level_span(x0, w, L, lev, &a0, &a1);
        or_level_span(y0, h, L, lev, &b0, &b1);
        int nw = a1 - a0, nh = b1 - b0;
        /* an empty level (a tile of a few samples at an odd offset) has no
           samples to transform; the reference's line loops are undefined
           for it (its 5/3 column tail would index row -1) */
        if (nw <= 0 || nh <= 0) continue;
        int lw = ceil2(a1) - ceil2(a0), lh = ceil2(b1) - ceil2(b0);
        int px = a0 & 1, py = b0 & 1;
        double gTopLo = (lev == 0 ? (double)steps[0] : 1.0) * lo * lo;
        double gTopHi = (double)steps[3 * lev + 1] * hi * lo;
        double gBotLo = (double)steps[3 * lev + 2] * lo * hi;
        double gBotHi = (double)steps[3 * lev + 3] * hi * hi;
        for (int y = 0; y < nh; ++y) {
            const double *row = cur + (size_t)y * w;
            double *dst = tmp + (size_t)y * nw;
            int topRow = y < lh;
            double gl = topRow ? gTopLo : gBotLo, gh = topRow ? gTopHi : gBotHi;
            for (int i = 0; i < nw; ++i) {
                if ((i + px) & 1) dst[i] = row[lw + ((i - 1 + px) >> 1)] * gh;
                else dst[i] = row[(i - px) >> 1] * gl;
            }
            or_97_line_f64(dst, nw, px);
        }
        for (int x = 0; x < nw; ++x) {
            for (int i = 0; i < nh; ++i) {
                int r = ((i + py) & 1) ? lh + ((i - 1 + py) >> 1) : ((i - py) >> 1);
                col[i] = tmp[(size_t)r * nw + x];
            }
            or_97_line_f64(col, nh, py);
            for (int i = 0; i < nh; ++i) cur[(size_t)i * w + x] = col[i];
        }
    }
    free(tmp);
    free(col);
    return 0;
}

static void fwd97_line_f64(double *x, int n, int p, double *lo, double *hi)
{
    /* inverse of or_97_line_f64 plus the reference gain convention
       (synthesis scales low by K and high by 2/K) */
    if (n == 1) {
        if (p) hi[0] = x[0] * 2.0 / (2.0 / D_K);
        else lo[0] = x[0] / D_K;
        return;
    }
    const double c[4] = {D_ALPHA, D_BETA, D_GAMMA, D_DELTA};
    for (int s = 0; s < 4; ++s) {
        int start = ((s & 1) == 0) ? (p ? 0 : 1) : (p ? 1 : 0);
        for (int i = start; i < n; i += 2)
            x[i] += c[s] * (x[mirror(i - 1, n)] + x[mirror(i + 1, n)]);
    }
    int nl = 0, nhh = 0;
    for (int i = 0; i < n; ++i) {
        if ((i + p) & 1) hi[nhh++] = x[i] / (2.0 / D_K);
        else lo[nl++] = x[i] / D_K;
    }
}

int or_fdwt97_f64(const double *in, int w, int h, int x0, int y0, int levels, double *out)
{
    int L = levels;
    memcpy(out, in, sizeof(double) * (size_t)w * h);
    int m = w > h ? w : h;
    double *x = (double *)malloc(sizeof(double) * m);
    double *lo = (double *)malloc(sizeof(double) * m);
    double *hi = (double *)malloc(sizeof(double) * m);
    for (int lev = L - 1; lev >= 0; --lev) {
        int a0, a1, b0, b1;
        or_level_span(x0, w, L, lev, &a0, &a1);
        or_level_span(y0, h, L, lev, &b0, &b1);
        int nw = a1 - a0, nh = b1 - b0;
        /* an empty level (a tile of a few samples at an odd offset) has no
           samples to transform; the reference's line loops are undefined
           for it (its 5/3 column tail would index row -1) */
        if (nw <= 0 || nh <= 0) continue;
        int lw = ceil2(a1) - ceil2(a0), lh = ceil2(b1) - ceil2(b0);
        int px = a0 & 1, py = b0 & 1;
        for (int c = 0; c < nw; ++c) {
            for (int i = 0; i < nh; ++i) x[i] = out[(size_t)i * w + c];
            fwd97_line_f64(x, nh, py, lo, hi);
            for (int i = 0; i < lh; ++i) out[(size_t)i * w + c] = lo[i];
            for (int i = 0; i < nh - lh; ++i) out[(size_t)(lh + i) * w + c] = hi[i];
        }
        for (int r = 0; r < nh; ++r) {
            memcpy(x, out + (size_t)r * w, sizeof(double) * nw);
            fwd97_line_f64(x, nw, px, lo, hi);
            memcpy(out + (size_t)r * w, lo, sizeof(double) * lw);
            memcpy(out + (size_t)r * w + lw, hi, sizeof(double) * (nw - lw));
        }
    }
    free(x);
    free(lo);
    free(hi);
    return 0;
}

/* ------------------------------------------------------ colour / pack -- */

enum {
    FMT_32_ARGB = 0,
    FMT_48_RGB = 1,
    FMT_24_RGB = 2,
    FMT_32_4C = 3,
    FMT_64_4C = 4,
    FMT_48_3C = 5,
    FMT_24_3C = 6,
    FMT_NONE = 7,
    FMT_64_ARGB = 8,
    FMT_36_RGB = 9,
    FMT_48_ARGB = 10
};

int or_pixel_bytes(int fmt)
{
    switch (fmt) {
    case FMT_32_ARGB: case FMT_32_4C: return 4;
    case FMT_48_RGB: case FMT_48_3C: return 6;
    case FMT_24_RGB: case FMT_24_3C: return 3;
    case FMT_64_4C: case FMT_64_ARGB: return 8;
    default: return 0;
    }
}

/* LIFT, bmi_cpu_dwt.h:22 */
static int lift(int v, int in_depth, int out_depth)
{
    return in_depth >= out_depth ? (v >> (in_depth - out_depth)) : (v << (out_depth - in_depth));
}

/* write one channel value at byte slot `slot` (0 = first element) */
static void put_elem(uint8_t *px, int slot, int wide, int v)
{
    if (wide) { uint16_t u = (uint16_t)v; memcpy(px + 2 * slot, &u, 2); }
    else px[slot] = (uint8_t)v;
}

/* element layout of the RGB/ARGB formats (codec_base.h:51-143): B,G,R[,A] */
static int fmt_rgb_info(int fmt, int *wide, int *alpha)
{
    switch (fmt) {
    case FMT_32_ARGB: *wide = 0; *alpha = 1; return 8;
    case FMT_24_RGB: *wide = 0; *alpha = 0; return 8;
    case FMT_48_RGB: *wide = 1; *alpha = 0; return 16;
    case FMT_64_ARGB: *wide = 1; *alpha = 1; return 16;
    default: return 0;
    }
}

int or_rct_pack(uint8_t *out, int out_pitch_px, const int32_t *c0, const int32_t *c1,
                const int32_t *c2, int in_pitch, int w, int h, int bd, int fmt, int check_out)
{
    /* RRct_53: G = Y0 - ((Y1+Y2)>>2) + 2^(B-1); R = clamp(Y2+G), B =
       clamp(Y1+G), G = clamp(G) to [0,2^B-1], then LIFT to the container */
    int wide, alpha;
    int od = fmt_rgb_info(fmt, &wide, &alpha);
    if (!od) return -2;
    int pb = or_pixel_bytes(fmt);
    int32_t off = 1 << (bd - 1), mx = (off << 1) - 1;
    for (int y = 0; y < h; ++y) {
        uint8_t *px = out + (size_t)y * out_pitch_px * pb;
        for (int x = 0; x < w; ++x, px += pb) {
            size_t i = (size_t)y * in_pitch + x;
            int32_t g = c0[i] - ((c1[i] + c2[i]) >> 2) + off;
            int32_t r = c2[i] + g, b = c1[i] + g;
            int R = lift(r < 0 ? 0 : (r > mx ? mx : r), bd, od);
            int G = lift(g < 0 ? 0 : (g > mx ? mx : g), bd, od);
            int B = lift(b < 0 ? 0 : (b > mx ? mx : b), bd, od);
            if (check_out == 0) G = B = R;
            else if (check_out == 1) R = B = G;
            else if (check_out == 2) R = G = B;
            put_elem(px, 0, wide, B);
            put_elem(px, 1, wide, G);
            put_elem(px, 2, wide, R);
            if (alpha) put_elem(px, 3, wide, 0);
        }
    }
    return 0;
}

static int ict_clamp_trunc(float v, int mx)
{
    /* (int)(SET_BOUNDARY(0, tmpf, maxV)): compare in float, truncate */
    float c = v < 0.0f ? 0.0f : (v > (float)mx ? (float)mx : v);
    return (int)c;
}

void or_ict_float(float *r, float *g, float *b, const float *c0, const float *c1,
                  const float *c2, int n, int bd)
{
    const float C20 = (float)(1.402), C11 = (float)(-0.34413), C21 = (float)(-0.71414),
                C12 = (float)(1.772);
    float off = (float)(1 << (bd - 1));
    for (int i = 0; i < n; ++i) {
        float t;
        t = c0[i] + C20 * c2[i]; t += off; r[i] = t;
        t = c0[i] + C11 * c1[i] + C21 * c2[i]; t += off; g[i] = t;
        t = c0[i] + C12 * c1[i]; t += off; b[i] = t;
    }
}

int or_ict_pack(uint8_t *out, int out_pitch_px, const float *c0, const float *c1,
                const float *c2, int in_pitch, int w, int h, int bd, int fmt, int check_out)
{
    /* RICT_97_FLOAT (bmi_cpu_dwt_decoder.cpp:3570-3841): float constants,
       R = Y + 1.402 Cr, G = Y - 0.34413 Cb - 0.71414 Cr, B = Y + 1.772 Cb,
       + 2^(B-1), clamp in float, truncate, LIFT */
    int wide, alpha;
    int od = fmt_rgb_info(fmt, &wide, &alpha);
    if (!od) return -2;
    int pb = or_pixel_bytes(fmt);
    int mx = (1 << bd) - 1;
    for (int y = 0; y < h; ++y) {
        uint8_t *px = out + (size_t)y * out_pitch_px * pb;
        for (int x = 0; x < w; ++x, px += pb) {
            size_t i = (size_t)y * in_pitch + x;
            float r, g, b;
            or_ict_float(&r, &g, &b, c0 + i, c1 + i, c2 + i, 1, bd);
            int R = lift(ict_clamp_trunc(r, mx), bd, od);
            int G = lift(ict_clamp_trunc(g, mx), bd, od);
            int B = lift(ict_clamp_trunc(b, mx), bd, od);
            if (check_out == 0) G = B = R;
            else if (check_out == 1) R = B = G;
            else if (check_out == 2) R = G = B;
            put_elem(px, 0, wide, B);
            put_elem(px, 1, wide, G);
            put_elem(px, 2, wide, R);
            if (alpha) put_elem(px, 3, wide, 0);
        }
    }
    return 0;
}

int or_merge_pack(uint8_t *out, int out_pitch_px, const void *const *comps, int ncomp,
                  int in_pitch, int w, int h, int bd, int fmt, int check_out, int is_float)
{
    /* MergeComponents<T> (bmi_cpu_dwt_decoder.cpp:4282-4596): per sample
       static_cast<int>(v) + 2^(B-1) (truncation before the shift, H5),
       clamp, LIFT.  4-component slots: Comp0 = A, Comp1 = R, Comp2 = G,
       Comp3 = B; 3-component: Comp0 = R, Comp1 = G, Comp2 = B. */
    int wide, nslots;
    switch (fmt) {
    case FMT_32_4C: wide = 0; nslots = 4; break;
    case FMT_64_4C: wide = 1; nslots = 4; break;
    case FMT_24_3C: wide = 0; nslots = 3; break;
    case FMT_48_3C: wide = 1; nslots = 3; break;
    default: return -2;
    }
    int od = wide ? 16 : 8;
    int pb = or_pixel_bytes(fmt);
    int lf = 1 << (bd - 1), mx = (1 << bd) - 1;
    for (int y = 0; y < h; ++y) {
        uint8_t *px = out + (size_t)y * out_pitch_px * pb;
        for (int x = 0; x < w; ++x, px += pb) {
            size_t i = (size_t)y * in_pitch + x;
            int v[4] = {0, 0, 0, 0};
            int have[4] = {0, 0, 0, 0};
            int nc = check_out == 0 ? 1 : ncomp;
            for (int c = 0; c < nc && c < 4; ++c) {
                if (!comps[c]) continue;
                int t = is_float ? (int)((const float *)comps[c])[i] : ((const int32_t *)comps[c])[i];
                t += lf;
                t = t < 0 ? 0 : (t > mx ? mx : t);
                v[c] = lift(t, bd, od);
                have[c] = 1;
            }
            /* byte slot of Comp k: 4-comp: A=3,R=2,G=1,B=0; 3-comp: R=2,G=1,B=0 */
            if (check_out == 0) {
                int t = v[0];
                put_elem(px, 0, wide, t);
                put_elem(px, 1, wide, t);
                put_elem(px, 2, wide, t);
                if (nslots == 4) put_elem(px, 3, wide, t);
            } else {
                for (int k = 0; k < nslots; ++k) {
                    int slot = nslots - 1 - k;
                    put_elem(px, slot, wide, have[k] ? v[k] : 0);
                }
            }
        }
    }
    return 0;
}

void or_dc_shift(int32_t *dst, int dst_pitch, const uint32_t *src, int src_pitch, int w, int h,
                 int bd, int is_float)
{
    /* bmi_dc_shift_kernel (bmi_dwt_gpu_kernel.h:709-751): (int)v for float
       input, + 2^(B-1), clamp to [0, 2^B-1] */
    int mx = (1 << bd) - 1;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            uint32_t wd = src[(size_t)y * src_pitch + x];
            int t;
            if (is_float) { float f; memcpy(&f, &wd, 4); t = (int)f; } else t = (int32_t)wd;
            t += 1 << (bd - 1);
            t = t < 0 ? 0 : (t > mx ? mx : t);
            dst[(size_t)y * dst_pitch + x] = t;
        }
}

void or_one_component(int32_t *out, int out_pitch, const int32_t *in, int in_pitch, int w, int h,
                      int mode, int comp_id, int bd)
{
    /* BMI_tile_One_Component (gpufunc.cu:1404-1480), kernels
       bmi_tile_one_component_int / _float and bmi_get_one_component_from_result
       (bmi_dwt_gpu_kernel.h:1644-1836).  Every mode writes the grey word
       o | o<<8 | o<<16.
       mode 0: o = byte (2 - comp_id) of the packed 32-bit result word `in`.
       mode 1 (5/3 plane): comp 0 adds 2^(B-1), other comps add 2^B - 1 and
         halve; clamp to [0, 2^B-1]; to 8 bits (>> B-8, or << 8-B); & 0xff.
       mode 2 (9/7 kernel): the word is read as a fixed-point value with 13
         fraction bits, rounded half away from zero to an integer, + 2^(B-1),
         clamped to [0, 2^B-1] and shifted right by B-8 (B >= 8). */
    const int mx = (1 << bd) - 1;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const int32_t v = in[(size_t)y * in_pitch + x];
            int o;
            if (mode == 0) {
                o = (int)(((uint32_t)v >> ((2 - comp_id) * 8)) & 0xffu);
            } else if (mode == 1) {
                int t = v;
                if (comp_id == 0) t += 1 << (bd - 1);
                else t = (t + (1 << bd) - 1) >> 1;
                t = t < 0 ? 0 : (t > mx ? mx : t);
                t = bd >= 8 ? (t >> (bd - 8)) : (t << (8 - bd));
                o = t & 0xff;
            } else {
                /* flag * ((flag * w + 0x1000) & 0xfffe000) >> 13, with the
                   products in 32-bit two's complement */
                const int neg = v < 0;
                uint32_t m = neg ? (uint32_t)0 - (uint32_t)v : (uint32_t)v;
                m = (m + 0x1000u) & 0xfffe000u;
                int32_t t = neg ? -(int32_t)m : (int32_t)m;
                t >>= 13;
                t += 1 << (bd - 1);
                t = t < 0 ? 0 : (t > mx ? mx : t);
                o = t >> (bd - 8);
            }
            out[(size_t)y * out_pitch + x] = o | (o << 8) | (o << 16);
        }
}
