// bmi_fused.h -- work descriptors of the fused / batched IDWT engine.
#pragma once

#include <stdint.h>

namespace bmi {

// what a level job writes
enum OutKind : int {
    OUT_SCRATCH = 0,  // intermediate level: float / int32 plane for the next level
    OUT_PLANE = 1,    // final level, IMAGE_FOTMAT_NONE: float (9/7) / int32 (5/3) plane
    OUT_RGB = 2,      // final level: inverse RCT / ICT + level shift + pack
    OUT_MERGE = 3,    // final level: level shift + pack of 1..4 components
};

// how a source plane stores its samples
enum SrcKind : int { SRC_SIGNMAG = 0, SRC_FLOAT = 1, SRC_INT = 2 };

constexpr int kMaxDeps = 4;

// One synthesis level of one tile, for `nc` components processed in
// lockstep (nc > 1 only on a final level that mixes components).
struct LevelJob {
    int a0, a1, b0, b1;  // absolute column / row span of the output resolution
    int lw, lh;          // low-band width / height
    int cl0, ch0;        // ceil(a0/2), floor(a0/2): band column origins
    int rl0, rh0;        // ceil(b0/2), floor(b0/2): band row origins
    int nstrips, nsegs;  // work-item grid of this job
    int seg_pairs;       // output row pairs per item
    int filt;            // 53 or 97
    int nc;              // components in lockstep (1..4)
    int ll_kind, det_kind;
    int ll_pitch, det_pitch, dst_pitch;  // elements
    int out_kind;
    int quirk;  // 5/3 reference bottom-edge quirk applies to this level
    int lev;    // synthesis level (0 = coarsest)
    int comp[4];         // component of each lockstep slot
    const void *ll[4];   // LL source per component (level-0: coefficient plane)
    const void *det[4];  // coefficient plane per component (detail bands)
    void *dst[4];        // OUT_SCRATCH / OUT_PLANE destination planes
    float gain[4][4];    // 9/7: top-lo, top-hi, bottom-lo, bottom-hi per component
    // pixel output (OUT_RGB / OUT_MERGE)
    void *out;
    int out_pitch_px, out_x, out_y;
    int fmt, bit_depth, check_out;  // check_out: -1 colour, 0..2 grey of R/G/B
    int ndeps;
    int deps[kMaxDeps];  // jobs that must be complete first
    int n_items;
    // two-level jobs (fuse2 = 1): the level below this final level is
    // synthesised inside the same items, into a ring of rows in LDS, and
    // never goes to HBM.  Its output span and band origins (absolute):
    int fuse2;
    int ia0, ia1, ib0, ib1;
    int ilw, ilh, icl0, ich0, irl0, irh0;
    int ill_kind, ill_pitch;
    int iquirk;          // 5/3 reference bottom-edge rule applies to the inner level
    const void *ill[4];  // LL source of the inner level per slot
    float igain[4][4];   // 9/7 gains of the inner level per slot
};

struct WorkItem {
    int job;
    int local;  // strip + nstrips * seg
};

}  // namespace bmi
