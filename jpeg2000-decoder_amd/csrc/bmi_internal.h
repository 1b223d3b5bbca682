// bmi_internal.h -- host-side helpers shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include "bmi_gpulib.h"

// record `code` (and a message on stderr when BMI_VERBOSE is set) as the
// calling thread's last error; returns `code`
int bmi_fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
// reset the calling thread's last error to BMI_OK
void bmi_clear_error();
// map a HIP status to BMI_OK / BMI_E_HIP, recording it
int bmi_check_hip(hipError_t e, const char *what);

// reference stream convention: a null / non-positive StreamId is the
// legacy default stream
inline hipStream_t bmi_stream(StreamId s) { return (hipStream_t)s; }


// Plan construction in two halves, for callers that own the device memory
// (the request server carves it from its HBM arena):
//   bmi_plan_build      host only: validate tiles, build jobs and items
//   bmi_plan_device_bytes  bytes the plan needs on the device
//   bmi_plan_bind       place the plan in [base, base+bytes) and upload
//                       its tables on `stream` (asynchronous); with a
//                       pinned `staging` buffer of at least
//                       bmi_plan_table_bytes, one copy from it (the staging
//                       must stay untouched until the stream reaches it)
//   bmi_plan_estimate_bytes  upper bound of device bytes of a one-tile
//                       plan, without building it (0 for an invalid tile)
// A plan bound this way does not free its memory on destroy.
int bmi_plan_build(BmiBatchPlan **plan, const BmiTileDesc *tiles, int ntiles,
                   const BmiTuning *tuning);
size_t bmi_plan_device_bytes(const BmiBatchPlan *plan);
size_t bmi_plan_table_bytes(const BmiBatchPlan *plan);
size_t bmi_plan_estimate_bytes(const BmiTileDesc *tile, const BmiTuning *tuning);
int bmi_plan_bind(BmiBatchPlan *plan, void *base, hipStream_t stream, void *staging,
                  size_t staging_bytes);

// The reference's ignore-component mask (mIgnoreComponents): the tile with
// its ignored components removed and the rest moved down to the lower
// component slots, in order (coefficient planes, steps, result planes), the
// check-out component renumbered (-1 when it is ignored).  Returns BMI_OK,
// BMI_E_INVALID when every component is ignored, or BMI_E_UNSUPPORTED for
// an RGB format with fewer than three components left (the reference's
// MergeComponents asserts on RGB formats).  `out` may alias `in`.
int bmi_apply_ignore(const BmiTileDesc &in, BmiTileDesc *out);
