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

extern "C" int BMI_batch_plan_error_flag(BmiBatchPlan *plan);
