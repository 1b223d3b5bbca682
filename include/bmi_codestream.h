/*
 * bmi_codestream.h -- C-ABI of the JPEG 2000 codestream front end (host).
 *
 * Parses a JPEG 2000 Part 1 codestream (raw .j2k/.j2c, or a JP2 file whose
 * jp2c box holds one) and decodes its tier-2 packet headers on the host,
 * producing for every tile the code-block table and code-block segments
 * that BMI_t1_decode_blocks consumes and a BmiTileDesc for the fused
 * IDWT + colour plan.  Replaces the reference's T2 stage (BMI_dec_info /
 * dec_t2: decoder/codec/src/dec_jp2.cpp:420-560, dec_tsk.cpp:107-185,
 * dec_syntax.cpp, dec_packet.cpp:37-730, dec_tagtree.cpp:40-142,
 * dec_io.cpp:99-177, dec_init.cpp:645-1056) for the streams the
 * reference's block decoder handles:
 *   - any number of tiles and tile-parts; 1..4 components of one bit depth,
 *     unsigned, not sub-sampled (the reference's GPU path, which has no
 *     sub-sampled or signed formats);
 *   - COD/COC: LRCP, RLCP, RPCL, PCRL, CPRL; any number of layers; default
 *     or explicit precincts; SOP / EPH markers; every code-block style of
 *     T.800 Table A.19 (bypass, reset, termination on each pass, vertically
 *     causal contexts, predictable termination, segmentation symbols: the
 *     reference decodes style 0 only and asserts on bypass,
 *     bmi_block_decoder.cpp:792-797), with one segment length per codeword
 *     segment in the packet headers (B.10.7.2); code-blocks up to 64x64;
 *   - QCD/QCC: no quantisation (5/3), scalar derived or expounded (9/7);
 *   - COM, TLM, PLM, PLT, CRG are skipped; RGN, POC, PPM, PPT are rejected
 *     with BMI_E_UNSUPPORTED.
 */
#ifndef BMI_CODESTREAM_H
#define BMI_CODESTREAM_H

#include <stddef.h>

#include "bmi_gpulib.h"
#include "bmi_t1.h"

#ifdef __cplusplus
extern "C" {
#endif

/* BMI_codestream_open flags */
#define BMI_CS_REFERENCE 0    /* numbps exactly as the reference stores it */
#define BMI_CS_GUARD_T800 1   /* 9/7: place the most significant bit-plane at
                                 exponent + guard bits - 1 (ITU-T T.800 E.1).
                                 The reference adds guard bits - 1 to the
                                 mantissa field of numbps (dec_jp2.cpp:177-178),
                                 so its 9/7 decode of a stream with more than
                                 one guard bit is scaled by 2^-(guard bits - 1);
                                 identical results for one guard bit and 5/3. */

typedef struct BmiCodestreamInfo {
    int width, height;          /* image size: Xsiz - XOsiz, Ysiz - YOsiz */
    int x0, y0;                 /* image origin on the reference grid */
    int tile_w, tile_h;         /* nominal tile size */
    int tile_x0, tile_y0;       /* tile grid origin */
    int ntiles_x, ntiles_y;
    int ncomp;
    int bit_depth;              /* of every component */
    int levels;                 /* decomposition levels (COD) */
    int reversible;             /* 1: 5/3, 0: 9/7 */
    int mct;                    /* COD multiple component transform */
    int layers;
    int progression;            /* 0 LRCP, 1 RLCP, 2 RPCL, 3 PCRL, 4 CPRL */
    int cb_w, cb_h;             /* nominal code-block size (COD) */
    int guard_bits;             /* QCD */
    int colour_space;           /* JP2 colr enumerated space (16 sRGB, 17 grey,
                                   18 sYCC), 0 for a raw codestream */
    int total_blocks;           /* code-blocks over all tiles */
} BmiCodestreamInfo;

typedef struct BmiCodestream BmiCodestream;

/* Parse the main header, all tile-parts and every packet header.  `data`
   is copied; it need not outlive the call.  Returns BMI_OK, BMI_E_INVALID
   for a malformed or truncated stream, BMI_E_UNSUPPORTED for a feature
   listed above as rejected. */
int BMI_codestream_open(const void *data, size_t len, int flags, BmiCodestream **cs,
                        BmiCodestreamInfo *info);
/* Tile t (raster order): its rectangle on the reference grid, number of
   code-blocks and bytes of code-block segments. */
int BMI_codestream_tile_info(const BmiCodestream *cs, int t, int *x0, int *y0, int *w, int *h,
                             int *nblocks, size_t *seg_bytes);
/* Fill, for tile t:
     blocks[0..nblocks)  the T1 code-block table; `out` addresses point into
                         the device coefficient planes planes[c] (pitch
                         `plane_pitch` 32-bit elements, tile-sized, Mallat
                         layout: the layout the IDWT reads);
     segs                the code-block data (host buffer of
                         seg_capacity >= seg_bytes + 16 bytes: the GPU
                         decoder reads up to 16 bytes past a segment): first
                         the segment-length tables of blocks whose code-block
                         style splits their passes into several codeword
                         segments (int32, at blocks[i].seg_offset - seg_base,
                         16-byte aligned in segs), then each block's codeword
                         segments, every one gathered over the layers, at
                         blocks[i].data_offset - seg_base.  seg_base is the
                         offset of segs[0] in the `data` buffer later given to
                         BMI_t1_decode_blocks (a multiple of 16);
     desc                the tile descriptor: coefficient planes, geometry,
                         levels, filter, bit depth, steps, colour transform.
                         Output fields (out, out_format, out_pitch_px, out_x,
                         out_y, planes_out) are left for the caller.
   Blocks never included in a packet get num_pass 0 (decoded to zeros). */
int BMI_codestream_tile_blocks(const BmiCodestream *cs, int t, void *const *planes, int plane_pitch,
                               unsigned long long seg_base, BmiT1Block *blocks, void *segs,
                               size_t seg_capacity, BmiTileDesc *desc);
int BMI_codestream_close(BmiCodestream *cs);

#ifdef __cplusplus
}
#endif
#endif
