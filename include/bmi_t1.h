/*
 * bmi_t1.h -- C-ABI of the GPU EBCOT tier-1 (MQ) code-block decoder.
 *
 * Additive: the reference decodes code-blocks on its host worker threads
 * (decode_code_block, decoder/codec/src/bmi_block_decoder.cpp:589-853,
 * driven by dec_cb_stripe, decoder/codec/src/dec_t1_comp.cpp:283-426) and
 * uploads the resulting words to the GPU (cudainterface.cpp:996-1115).
 * This entry point decodes the code-blocks on the GPU, one code-block per
 * wavefront, and writes the words straight into the device coefficient
 * planes the IDWT reads, with the same results bit for bit.
 */
#ifndef BMI_T1_H
#define BMI_T1_H

#include "bmi_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* output conversion of a code-block */
#define BMI_T1_WORDS 0 /* T1 sign-magnitude words, as dec_cb_stripe copies them
                          for 9/7 (BMI_COEF_SIGNMAG input of the IDWT) */
#define BMI_T1_INT 1   /* two's complement integers, dec_cb_stripe's 5/3
                          conversion: -( (w & 0x7fffffff) >> (31-numbps) ) */

/* Band orientations (zero-coding context table of the band) */
#define BMI_T1_LL 0
#define BMI_T1_HL 1
#define BMI_T1_LH 2
#define BMI_T1_HH 3

typedef struct BmiT1Block {
    unsigned long long data_offset; /* first byte of the block's codeword segment in `data` */
    int len;       /* bytes of the segment (all coding passes of the block) */
    int width;     /* code-block width, 1..64 */
    int height;    /* code-block height, 1..64 */
    int orient;    /* BMI_T1_LL / HL / LH / HH */
    int numbps;    /* cur_band->numbps as the reference stores it: 9/7
                      (exponent<<11 | mantissa) + guard bits - 1, 5/3
                      exponent + guard bits - 1 */
    int missing;   /* missing_bit_plan: all-zero leading bit-planes */
    int num_pass;  /* coding passes present (0: the block decodes to zeros) */
    int kind;      /* BMI_T1_WORDS / BMI_T1_INT */
    void *out;     /* device address of the block's top-left output word */
    int out_pitch; /* elements between output rows */
    /* Code-block style (COD/COC SPcod, T.800 Table A.19; 0 in every stream
       the reference decodes).  With BMI_T1_BYPASS or BMI_T1_TERMALL the
       `len` bytes hold `nseg` codeword segments back to back, segment k
       being L[k] bytes, where L is an int32 array at byte offset
       `seg_offset` (a multiple of 4) of `data`; nseg must be the number of
       segments the block's num_pass passes open (BMI_t1_num_segments).
       Otherwise the data is one segment and nseg / seg_offset are unused.
       The reference asserts on these styles (bmi_block_decoder.cpp:792-797,
       818-822); they are decoded here as T.800 Annex D defines them. */
    int style;
    int nseg;
    unsigned long long seg_offset;
} BmiT1Block;

/* code-block style bits (T.800 Table A.19) */
#define BMI_T1_BYPASS 0x01   /* raw significance / refinement passes from pass 10 on (D.6) */
#define BMI_T1_RESET 0x02    /* context states reset on every pass boundary */
#define BMI_T1_TERMALL 0x04  /* every pass is a terminated codeword segment */
#define BMI_T1_VSC 0x08      /* vertically stripe-causal contexts (D.7) */
#define BMI_T1_PTERM 0x10    /* predictable termination (no effect on decoding) */
#define BMI_T1_SEGSYM 0x20   /* segmentation symbol after every cleanup pass (D.5) */
#define BMI_T1_STYLES 0x3F

/* Decode `nblocks` code-blocks described by the DEVICE array `blocks`, whose
   segments are in the DEVICE buffer `data`; `data` must stay readable for
   16 bytes past the end of every segment (the decoder fetches 16-byte
   aligned windows; bytes past a segment's end are read as 0xFF, as the
   reference's decoder appends 0xFF 0xFF, bmi_mq_decoder.cpp:195).
   Asynchronous on `stream`.  Blocks larger than 64x64 are rejected with
   BMI_E_UNSUPPORTED by BMI_t1_check_blocks (host copy of the table). */
int BMI_t1_decode_blocks(const BmiT1Block *blocks, int nblocks, const void *data,
                         StreamId stream);
/* Host-side validation of a block table before upload: returns BMI_OK or
   the first error (index in *bad when not NULL).  Segment lengths live in
   device memory and are not read here: the decoder clamps each segment to
   the block's `len` bytes, so a wrong table decodes wrong words but reads
   nothing outside the block's data. */
int BMI_t1_check_blocks(const BmiT1Block *blocks, int nblocks, int *bad);
/* Reorder a host block table in place, longest code-block data first
   (stable for equal lengths).  The decoder runs blocks in table order, one
   per wavefront; starting the longest ones first shortens the tail of the
   launch (4096x2160x3 image: 24.9 -> 23.8 ms).  Outputs do not depend on
   the order. */
int BMI_t1_order_blocks(BmiT1Block *blocks, int nblocks);
/* Codeword segments opened by the first num_pass coding passes of a block
   of code-block style `style` (1 for num_pass > 0 without BMI_T1_BYPASS /
   BMI_T1_TERMALL). */
int BMI_t1_num_segments(int num_pass, int style);

#ifdef __cplusplus
}
#endif
#endif
