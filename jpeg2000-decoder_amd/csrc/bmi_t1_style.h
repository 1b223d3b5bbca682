// bmi_t1_style.h -- coding passes and codeword segments of a code-block
// under the code-block style switches of ITU-T T.800 Table A.19, shared by
// the tier-1 kernel (device) and the tier-2 parser (host).
#pragma once

#include "bmi_t1.h"

#if defined(__HIPCC__)
#define BMI_HD __host__ __device__
#else
#define BMI_HD
#endif

// Pass i of a code-block: 0 is the first cleanup pass, then significance
// propagation, magnitude refinement and cleanup per bit-plane, so its type
// is (i + 2) % 3 (0 significance, 1 refinement, 2 cleanup).

// With bypass (D.6) the significance and refinement passes from the fifth
// coded bit-plane on (pass 10 on) are raw.
BMI_HD inline bool bmi_pass_raw(int i, int style)
{
    return (style & BMI_T1_BYPASS) && i >= 10 && i % 3 != 0;
}

// Pass i opens a codeword segment: the first pass; every pass when each is
// terminated; with bypass, from pass 10 on, the raw significance pass of
// each bit-plane (its raw segment also holds the refinement pass) and each
// cleanup pass.
BMI_HD inline bool bmi_pass_new_segment(int i, int style)
{
    if (i == 0 || (style & BMI_T1_TERMALL)) return true;
    return (style & BMI_T1_BYPASS) && i >= 10 && i % 3 != 2;
}

// the data of the block is more than one segment
BMI_HD inline bool bmi_style_segmented(int style) { return (style & (BMI_T1_BYPASS | BMI_T1_TERMALL)) != 0; }
