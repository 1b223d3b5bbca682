/*
 * cpu_pipeline.c -- the multithreaded CPU baseline: the reference CPU back
 * end's 9/7 + ICT pipeline for one tile, run on all host cores.
 *
 * TEST AND BASELINE INFRASTRUCTURE ONLY: bench.py times it as the CPU
 * baseline beside the GPU; tests check it equals the scalar oracle bit for
 * bit.  Nothing in the product links or loads it.
 *
 * Structure of the reference's multithreaded CPU decode
 * (decoderconcrete.cpp:1442-1812, bmi_cpu_dwt_decoder.cpp:1180-1330):
 * per component and level the horizontal pass is spread over the worker
 * threads by rows, the vertical pass by bands (the reference splits the top
 * level into per-thread row bands with overlapping lifting windows; here a
 * thread owns a band of columns and lifts it row by row, which computes the
 * same values and keeps each row segment contiguous), and the colour
 * transform and packing by row bands.  Each sample gets exactly the
 * oracle's float operations in the oracle's order (x - c*(l + r) rounded
 * per operation, gains folded into the row pass, float ICT, clamp then
 * truncate), so the output is bit-identical to or_idwt97 + or_ict_pack;
 * only the loop order differs, which lets the compiler vectorise across
 * columns.  Built -O3 for x86-64-v3 (AVX2/FMA hosts) with
 * -ffp-contract=off, so no FMA is formed.
 */
#define _POSIX_C_SOURCE 200809L
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* decoder/codec/include/bmi_cpu_dwt.h:31-37 */
static const float P_ALPHA = -1.58613434205992f;
static const float P_BETA = -0.05298011857296f;
static const float P_GAMMA = 0.88291107553093f;
static const float P_DELTA = 0.44350685204397f;
static const float P_LO = 1.23017410558578f;
static const float P_HI = 1.62578613134411f;

static int pmirror(int i, int n)
{
    if (n == 1) return 0;
    int period = 2 * (n - 1);
    i %= period;
    if (i < 0) i += period;
    return i < n ? i : period - i;
}

static int cdiv2(int v) { return (v + 1) >> 1; }

/* one 9/7 synthesis line in place (HDWT_R97, bmi_cpu_dwt_decoder.cpp:2283-2485) */
static void line97(float *x, int n, int p)
{
    if (n == 1) {
        if (p) x[0] = x[0] * 0.5f;
        return;
    }
    const float c[4] = {P_DELTA, P_GAMMA, P_BETA, P_ALPHA};
    for (int s = 0; s < 4; ++s) {
        const int start = ((s & 1) == 0) ? (p ? 1 : 0) : (p ? 0 : 1);
        const float cs = c[s];
        int i = start;
        if (i == 0) {  /* left edge: mirrored neighbour */
            const float sum = x[pmirror(-1, n)] + x[pmirror(1, n)];
            x[0] = x[0] - cs * sum;
            i += 2;
        }
        for (; i < n - 1; i += 2) {  /* interior: both neighbours exist */
            const float sum = x[i - 1] + x[i + 1];
            const float prod = cs * sum;
            x[i] = x[i] - prod;
        }
        if (i == n - 1) {  /* right edge */
            const float sum = x[pmirror(n - 2, n)] + x[pmirror(n, n)];
            x[i] = x[i] - cs * sum;
        }
    }
}

typedef struct {
    /* shared */
    const float *const *coef;  /* ncomp coefficient planes (pitch `pitch`) */
    int pitch, w, h, levels, ncomp;
    const float *steps;        /* [ncomp][3L+1] */
    float **plane;             /* ncomp working planes, pitch w */
    float *tmp;                /* row-pass result, w*h */
    uint8_t *out;              /* 24-bit BGR, pitch out_pitch bytes */
    int out_pitch;
    int nthreads;
    pthread_barrier_t *bar;
    /* per thread */
    int id;
} Work;

static void band(int n, int parts, int id, int *b0, int *b1)
{
    *b0 = (int)((long long)n * id / parts);
    *b1 = (int)((long long)n * (id + 1) / parts);
}

static void *worker(void *arg)
{
    Work *W = (Work *)arg;
    const int L = W->levels, w = W->w, h = W->h, T = W->nthreads, id = W->id;
    for (int c = 0; c < W->ncomp; ++c) {
        float *plane = W->plane[c];
        const float *coef = W->coef[c];
        const float *st = W->steps + (size_t)c * (3 * L + 1);
        for (int lev = 0; lev < L; ++lev) {
            const int s = L - 1 - lev;
            const int nw = (w + (1 << s) - 1) >> s, nh = (h + (1 << s) - 1) >> s;
            const int lw = cdiv2(nw), lh = cdiv2(nh);
            const float gTopLo = lev == 0 ? (st[0] * P_LO) * P_LO : P_LO * P_LO;
            const float gTopHi = (st[3 * lev + 1] * P_HI) * P_LO;
            const float gBotLo = (st[3 * lev + 2] * P_LO) * P_HI;
            const float gBotHi = (st[3 * lev + 3] * P_HI) * P_HI;
            /* horizontal pass, rows spread over the threads: low part from
               LL (coefficients at level 0, the previous output after),
               high part from the detail bands, gains folded in */
            int r0, r1;
            band(nh, T, id, &r0, &r1);
            for (int y = r0; y < r1; ++y) {
                const int top = y < lh;
                const float gl = top ? gTopLo : gBotLo, gh = top ? gTopHi : gBotHi;
                const float *lo = (top && lev > 0) ? plane + (size_t)y * w : coef + (size_t)y * W->pitch;
                const float *hi = coef + (size_t)y * W->pitch + lw;
                float *dst = W->tmp + (size_t)y * nw;
                for (int i = 0; i < lw; ++i) dst[2 * i] = lo[i] * gl;
                for (int i = 0; i < nw - lw; ++i) dst[2 * i + 1] = hi[i] * gh;
                line97(dst, nw, 0);
            }
            pthread_barrier_wait(W->bar);
            /* vertical pass, a band of columns per thread, lifted row by row */
            int c0, c1;
            band(nw, T, id, &c0, &c1);
            if (c1 > c0) {
                const int n = c1 - c0;
                for (int i = 0; i < nh; ++i) {
                    const int r = (i & 1) ? lh + (i >> 1) : (i >> 1);
                    memcpy(plane + (size_t)i * w + c0, W->tmp + (size_t)r * nw + c0, sizeof(float) * n);
                }
                if (nh == 1) {
                    /* even parity: a one-sample column keeps its value */
                } else {
                    const float cc[4] = {P_DELTA, P_GAMMA, P_BETA, P_ALPHA};
                    for (int k = 0; k < 4; ++k) {
                        const float cs = cc[k];
                        for (int i = (k & 1); i < nh; i += 2) {
                            float *x = plane + (size_t)i * w + c0;
                            const float *a = plane + (size_t)pmirror(i - 1, nh) * w + c0;
                            const float *b = plane + (size_t)pmirror(i + 1, nh) * w + c0;
                            for (int j = 0; j < n; ++j) {
                                const float sum = a[j] + b[j];
                                const float prod = cs * sum;
                                x[j] = x[j] - prod;
                            }
                        }
                    }
                }
            }
            pthread_barrier_wait(W->bar);
        }
    }
    /* inverse ICT, + 2^(B-1) (B = 8), clamp in float, truncate, pack B,G,R
       (RICT_97_FLOAT, bmi_cpu_dwt_decoder.cpp:3570-3841), row bands */
    if (W->ncomp >= 3) {
        const float C20 = (float)(1.402), C11 = (float)(-0.34413), C21 = (float)(-0.71414),
                    C12 = (float)(1.772);
        const float off = 128.0f;
        int y0, y1;
        band(h, T, id, &y0, &y1);
        for (int y = y0; y < y1; ++y) {
            const float *Y = W->plane[0] + (size_t)y * w, *Cb = W->plane[1] + (size_t)y * w,
                        *Cr = W->plane[2] + (size_t)y * w;
            uint8_t *px = W->out + (size_t)y * W->out_pitch;
            for (int x = 0; x < w; ++x) {
                float r = Y[x] + C20 * Cr[x];
                r += off;
                float g = Y[x] + C11 * Cb[x] + C21 * Cr[x];
                g += off;
                float b = Y[x] + C12 * Cb[x];
                b += off;
                r = r < 0.0f ? 0.0f : (r > 255.0f ? 255.0f : r);
                g = g < 0.0f ? 0.0f : (g > 255.0f ? 255.0f : g);
                b = b < 0.0f ? 0.0f : (b > 255.0f ? 255.0f : b);
                px[3 * x + 0] = (uint8_t)(int)b;
                px[3 * x + 1] = (uint8_t)(int)g;
                px[3 * x + 2] = (uint8_t)(int)r;
            }
        }
    }
    return NULL;
}

/* Decode one tile at offset (0, 0): `ncomp` float coefficient planes
   (Mallat layout, pitch `pitch` floats, left untouched), `levels` 9/7
   levels with per-band steps [ncomp][3L+1] (NULL: 1.0), then (ncomp >= 3)
   the inverse ICT of components 0..2 to 8-bit B,G,R at `out` (pitch
   out_pitch bytes).  `planes_out` (ncomp planes of w*h floats, pitch w)
   receives the synthesised planes.  Returns 0, or < 0 on bad arguments or
   failed allocation. */
int cpu_decode97_ict(const float *const *coef, int pitch, int w, int h, int levels, int ncomp,
                     const float *steps, float *const *planes_out, uint8_t *out, int out_pitch,
                     int nthreads)
{
    if (!coef || !planes_out || w < 1 || h < 1 || levels < 1 || levels > 6 || ncomp < 1 ||
        ncomp > 4 || pitch < w || nthreads < 1 || (ncomp >= 3 && !out))
        return -1;
    if (nthreads > 256) nthreads = 256;
    float *ones = NULL;
    if (!steps) {
        ones = (float *)malloc(sizeof(float) * (size_t)ncomp * (3 * levels + 1));
        if (!ones) return -3;
        for (size_t i = 0; i < (size_t)ncomp * (3 * levels + 1); ++i) ones[i] = 1.0f;
        steps = ones;
    }
    float *tmp = (float *)malloc(sizeof(float) * (size_t)w * h);
    if (!tmp) {
        free(ones);
        return -3;
    }
    pthread_barrier_t bar;
    pthread_barrier_init(&bar, NULL, (unsigned)nthreads);
    Work base;
    memset(&base, 0, sizeof base);
    base.coef = coef;
    base.pitch = pitch;
    base.w = w;
    base.h = h;
    base.levels = levels;
    base.ncomp = ncomp;
    base.steps = steps;
    base.plane = (float **)planes_out;
    base.tmp = tmp;
    base.out = out;
    base.out_pitch = out_pitch;
    base.nthreads = nthreads;
    base.bar = &bar;
    Work *ws = (Work *)malloc(sizeof(Work) * (size_t)nthreads);
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
    int rc = 0;
    if (!ws || !th) {
        rc = -3;
    } else {
        for (int i = 0; i < nthreads; ++i) {
            ws[i] = base;
            ws[i].id = i;
        }
        for (int i = 1; i < nthreads; ++i) pthread_create(&th[i], NULL, worker, &ws[i]);
        worker(&ws[0]);
        for (int i = 1; i < nthreads; ++i) pthread_join(th[i], NULL);
    }
    pthread_barrier_destroy(&bar);
    free(ws);
    free(th);
    free(tmp);
    free(ones);
    return rc;
}
