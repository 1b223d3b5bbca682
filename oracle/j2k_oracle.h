/*
 * j2k_oracle.h -- CPU restatement of the reference decoder's back end.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * back end: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  Nothing in the product links it.
 *
 * Every function restates a function of the reference CPU back end
 * (decoder/codec/src/bmi_cpu_dwt_decoder.cpp) in this project's own code;
 * each definition in j2k_oracle.c cites the reference lines it follows.
 *
 * Plane convention (same as the reference GPU "orig" buffer,
 * decoder/cudainterface/src/cudainterface.cpp:674-703): one tile-component
 * is a pitched 2-D array in Mallat layout.  At synthesis level `lev`
 * (0 = coarsest) the resolution being produced occupies [0,nw)x[0,nh)
 * relative to the tile origin; LL sits in [0,lw)x[0,lh), HL in
 * [lw,nw)x[0,lh), LH in [0,lw)x[lh,nh), HH in [lw,nw)x[lh,nh).
 */
#ifndef J2K_ORACLE_H
#define J2K_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 5/3 modes */
#define OR_R53_REFERENCE 0 /* CPU decoder behaviour: parity ignored, bottom-edge quirk */
#define OR_R53_TEXTBOOK 1  /* ITU-T T.800 F.3 with symmetric extension and parity */

/* coefficient kinds for the 9/7 path */
#define OR_COEF_SIGNMAG 0 /* T1 sign-magnitude words (FETCHPIXEL_INT) */
#define OR_COEF_FLOAT 1   /* float32 coefficients */

/* geometry of one synthesis level of a tile at absolute offset t0 with full
 * size n, L levels, lev in [0,L): output span [a0,a1) in level coordinates */
void or_level_span(int t0, int n, int L, int lev, int *a0, int *a1);

/* ---- 1-D lines -------------------------------------------------------- */
/* HDWT_R53_INT for one line (bmi_cpu_dwt_decoder.cpp:2609-2659) */
void or_h53_line_ref(int32_t *out, const int32_t *lo, const int32_t *hi, int lw, int hw);
/* textbook 5/3 synthesis of an interleaved line in place, parity p */
void or_53_line_textbook(int32_t *x, int n, int p);
/* 9/7 synthesis of an interleaved, already gain-scaled float line (stepwise,
 * symmetric extension, op order x - c*(l+r)) */
void or_97_line(float *x, int n, int p);
void or_97_line_f64(double *x, int n, int p);

/* VDWT_R53 over whole resolution (bmi_cpu_dwt_decoder.cpp:3311-3425):
 * hres = lh low rows then hh high rows, each `width` ints. */
void or_v53_ref(int32_t *out, const int32_t *hres, int lh, int hh, int width);

/* ---- multi-level tile-component IDWT, in place on a Mallat plane ------- */
/* 5/3: plane holds two's-complement int32 coefficients; result int32. */
int or_idwt53(int32_t *plane, int pitch, int w, int h, int x0, int y0, int levels,
              int toplevel, int mode);
/* 9/7: plane holds 32-bit words (sign-magnitude ints or floats per coef_kind
 * for every band; bands above level 0 LL are always coefficients); result is
 * float in place.  steps[3L+1] = mfAbsStepNorm per subband id. */
int or_idwt97(uint32_t *plane, int pitch, int w, int h, int x0, int y0, int levels,
              int toplevel, const float *steps, int coef_kind);
/* same transform in FP64 with exact T.800 constants; out is a w*h double
 * array (pitch w) */
int or_idwt97_f64(const uint32_t *plane, int pitch, int w, int h, int x0, int y0,
                  int levels, const float *steps, int coef_kind, double *out);
/* FP64 forward 9/7 (analysis) matching the reference gain convention, used
 * to build round-trip checks: in is w*h doubles, out Mallat doubles */
int or_fdwt97_f64(const double *in, int w, int h, int x0, int y0, int levels, double *out);
/* integer forward 5/3 (T.800 F.4), for reversibility checks */
int or_fdwt53(const int32_t *in, int w, int h, int x0, int y0, int levels, int32_t *out);

/* ---- colour / level shift / pack ----------------------------------------- */
/* OutputFormat values of decoder/common/include/codec_datatype.h:38-51 */
/* RRct_53<int_32> (bmi_cpu_dwt_decoder.cpp:3936-4130) */
int or_rct_pack(uint8_t *out, int out_pitch_px, const int32_t *c0, const int32_t *c1,
                const int32_t *c2, int in_pitch, int w, int h, int bit_depth, int fmt,
                int check_out);
/* RICT_97_FLOAT (bmi_cpu_dwt_decoder.cpp:3570-3841) */
int or_ict_pack(uint8_t *out, int out_pitch_px, const float *c0, const float *c1,
                const float *c2, int in_pitch, int w, int h, int bit_depth, int fmt,
                int check_out);
/* MergeComponents<T> (bmi_cpu_dwt_decoder.cpp:4282-4596); is_float selects T */
int or_merge_pack(uint8_t *out, int out_pitch_px, const void *const *comps, int ncomp,
                  int in_pitch, int w, int h, int bit_depth, int fmt, int check_out,
                  int is_float);
/* bmi_dc_shift_kernel semantics (decoder/cudainterface/src/bmi_dwt_gpu_kernel.h:709-751) */
void or_dc_shift(int32_t *dst, int dst_pitch, const uint32_t *src, int src_pitch, int w, int h,
                 int bit_depth, int is_float);
/* grey preview of one component (BMI_tile_One_Component, gpufunc.cu:1404-1480,
   bmi_dwt_gpu_kernel.h:1644-1836); mode 0 from packed 32-bit result words,
   1 from a 5/3 integer plane, 2 the 9/7 kernel's fixed-point reading */
void or_one_component(int32_t *out, int out_pitch, const int32_t *in, int in_pitch, int w, int h,
                      int mode, int comp_id, int bit_depth);
/* the unclamped ICT value before truncation, for pre-clamp parity checks */
void or_ict_float(float *r, float *g, float *b, const float *c0, const float *c1,
                  const float *c2, int n, int bit_depth);

int or_pixel_bytes(int fmt);

/* ---- EBCOT tier 1 (t1_oracle.c) --------------------------------------- */
/* decode_code_block (bmi_block_decoder.cpp:589-853) of one w x h code-block
 * of band orientation `orient` (0 LL, 1 HL, 2 LH, 3 HH): `numbps` as the
 * reference stores it (9/7: exponent<<11 | mantissa, plus guard bits - 1;
 * 5/3: exponent + guard bits - 1), `missing` all-zero leading bit-planes,
 * `num_pass` coding passes in one segment of `len` bytes.  Writes the T1
 * words (sign bit 31, magnitude below) to out (pitch out_pitch). */
int or_t1_decode_block(const uint8_t *data, long len, int w, int h, int orient, int numbps,
                       int missing, int num_pass, int32_t *out, int out_pitch);
/* the same with code-block style `style` (T.800 Table A.19: 0x01 bypass,
 * 0x02 context reset, 0x04 termination on every pass, 0x08 vertically
 * causal contexts, 0x10 predictable termination, 0x20 segmentation
 * symbols).  With bypass or termination on every pass the `len` bytes
 * hold `nseg` codeword segments of seg_len[k] bytes each, nseg being
 * or_t1_num_segments(num_pass, style); otherwise one segment.  Returns -1
 * for a segment table that does not fit. */
int or_t1_decode_block_ex(const uint8_t *data, long len, int w, int h, int orient, int numbps,
                          int missing, int num_pass, int style, const int32_t *seg_len, int nseg,
                          int32_t *out, int out_pitch);
/* codeword segments that the first num_pass passes of a code-block open */
int or_t1_num_segments(int num_pass, int style);
/* dec_cb_stripe's 5/3 conversion (dec_t1_comp.cpp:343-367) */
void or_t1_words_to_int(const int32_t *words, int w, int h, int pitch, int numbps, int32_t *out,
                        int out_pitch);
/* the decoder's MQ table (for the test against ITU-T T.800 Table C.2) and
 * its zero-coding / sign contexts of a 3x3 neighbourhood (nb9 row-major,
 * centre 4; bit 0 significant, bit 1 negative); sign returns 2*cx + xor */
void or_t1_tables(uint16_t *qe, uint8_t *nmps, uint8_t *nlps, uint8_t *sw);
int or_t1_zc(const uint8_t *nb9, int orient);
int or_t1_sc(const uint8_t *nb9);

#ifdef __cplusplus
}
#endif
#endif
