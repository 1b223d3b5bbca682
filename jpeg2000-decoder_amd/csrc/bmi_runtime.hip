// bmi_runtime.hip -- device bring-up, buffers, streams, copies and error
// reporting behind the reference's GPU library entry points
// (decoder/cudainterface/src/gpulib.c:12-78, gpufunc.cu:55-259).

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "bmi_internal.h"

static thread_local int g_last_error = BMI_OK;
static int g_r53_mode = BMI_R53_REFERENCE;

int bmi_fail(int code, const char *fmt, ...)
{
    g_last_error = code;
    static const bool verbose = getenv("BMI_VERBOSE") != nullptr;
    if (verbose && code != BMI_OK) {
        va_list ap;
        va_start(ap, fmt);
        fprintf(stderr, "[bmi] error %d: ", code);
        vfprintf(stderr, fmt, ap);
        fprintf(stderr, "\n");
        va_end(ap);
    }
    return code;
}

void bmi_clear_error() { g_last_error = BMI_OK; }

int bmi_check_hip(hipError_t e, const char *what)
{
    if (e == hipSuccess) return BMI_OK;
    return bmi_fail(BMI_E_HIP, "%s: %s", what, hipGetErrorString(e));
}

extern "C" int BMI_last_error(void) { return g_last_error; }

extern "C" const char *BMI_error_string(int code)
{
    switch (code) {
    case BMI_OK: return "ok";
    case BMI_E_INVALID: return "invalid argument";
    case BMI_E_NODEVICE: return "no HIP device";
    case BMI_E_HIP: return "HIP runtime error";
    case BMI_E_NOMEM: return "out of memory or over budget";
    case BMI_E_UNSUPPORTED: return "unsupported mode or format";
    default: return "unknown error";
    }
}

extern "C" void BMI_set_r53_mode(int mode)
{
    if (mode == BMI_R53_REFERENCE || mode == BMI_R53_TEXTBOOK)
        g_r53_mode = mode;
    else
        bmi_fail(BMI_E_INVALID, "BMI_set_r53_mode(%d)", mode);
}

extern "C" int BMI_get_r53_mode(void) { return g_r53_mode; }

// gpufunc.cu:55-123: device 0 only, CU count, clock, memory
extern "C" int GpuInit(GpuInfo *gpu)
{
    if (!gpu) return bmi_fail(BMI_E_INVALID, "GpuInit(NULL)");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        gpu->device = -1;
        gpu->active = 0;
        bmi_fail(BMI_E_NODEVICE, "GpuInit: no device");
        return -1;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) {
        gpu->device = -1;
        bmi_fail(BMI_E_HIP, "GpuInit: properties");
        return -1;
    }
    gpu->device = 0;
    gpu->major = prop.major;
    gpu->minor = prop.minor;
    strncpy(gpu->name, prop.name, sizeof(gpu->name) - 1);
    gpu->name[sizeof(gpu->name) - 1] = 0;
    gpu->multiProcessorCount = prop.multiProcessorCount;
    gpu->clockRate = prop.clockRate;
    // the reference stores the byte count in an unsigned int; saturate
    // instead of wrapping for devices above 4 GiB
    gpu->memory = prop.totalGlobalMem > 0xffffffffull ? 0xffffffffu : (unsigned)prop.totalGlobalMem;
    if (hipSetDevice(0) != hipSuccess) {
        bmi_fail(BMI_E_HIP, "GpuInit: set device");
        return -1;
    }
    gpu->active = 1;
    return 0;
}

extern "C" int GpuRemove(GpuInfo *gpu)
{
    if (!gpu) return bmi_fail(BMI_E_INVALID, "GpuRemove(NULL)");
    (void)hipDeviceSynchronize();
    gpu->active = 0;
    return 0;
}

// gpufunc.cu:134-187
extern "C" void MallocGpuBuf(CudaBuf *buf)
{
    bmi_clear_error();
    if (!buf || buf->mBuf) {
        bmi_fail(BMI_E_INVALID, "MallocGpuBuf: null or already allocated");
        return;
    }
    hipError_t e = hipSuccess;
    switch (buf->mType) {
    case GPU_BUF:
        if (buf->mWidth != 0 && buf->mHeight != 0) buf->mSize = buf->mWidth * buf->mHeight;
        buf->mPitch = buf->mWidth;
        e = hipMalloc(&buf->mBuf, buf->mSize ? buf->mSize : 1);
        break;
    case GPU_ALLIGN_BUF: {
        size_t pitch = 0;
        e = hipMallocPitch(&buf->mBuf, &pitch, buf->mWidth ? buf->mWidth : 1,
                           buf->mHeight ? buf->mHeight : 1);
        buf->mPitch = (unsigned)pitch;
        buf->mSize = buf->mPitch * buf->mHeight;
    } break;
    case HOST_ALLLIGN_BUF:
    case HOST_BUF:
        buf->mSize = buf->mWidth * buf->mHeight;
        buf->mPitch = buf->mWidth;
        e = hipHostMalloc(&buf->mBuf, buf->mSize ? buf->mSize : 1, hipHostMallocDefault);
        break;
    default:
        bmi_fail(BMI_E_INVALID, "MallocGpuBuf: type %d", (int)buf->mType);
        return;
    }
    if (e != hipSuccess) {
        buf->mBuf = nullptr;
        bmi_check_hip(e, "MallocGpuBuf");
    }
}

extern "C" void FreeGpuBuf(CudaBuf *buf)
{
    bmi_clear_error();
    if (!buf || !buf->mBuf) return;
    hipError_t e = hipSuccess;
    switch (buf->mType) {
    case GPU_BUF:
    case GPU_ALLIGN_BUF:
        e = hipFree(buf->mBuf);
        break;
    case HOST_ALLLIGN_BUF:
    case HOST_BUF:
        e = hipHostFree(buf->mBuf);
        break;
    default:
        bmi_fail(BMI_E_INVALID, "FreeGpuBuf: type %d", (int)buf->mType);
        return;
    }
    buf->mBuf = nullptr;
    bmi_check_hip(e, "FreeGpuBuf");
}

extern "C" StreamId CreateANewStream(void)
{
    hipStream_t s = nullptr;
    if (bmi_check_hip(hipStreamCreate(&s), "CreateANewStream")) return nullptr;
    return (StreamId)s;
}

extern "C" void DestroyStream(StreamId *s)
{
    bmi_clear_error();
    if (s && *s) bmi_check_hip(hipStreamDestroy((hipStream_t)*s), "DestroyStream");
}

extern "C" void GpuSync(StreamId *s)
{
    bmi_clear_error();
    if (s)
        bmi_check_hip(hipStreamSynchronize((hipStream_t)*s), "GpuSync(stream)");
    else
        bmi_check_hip(hipDeviceSynchronize(), "GpuSync(device)");
}

static hipMemcpyKind copy_kind(CudaCopyDirection d, bool *ok)
{
    *ok = true;
    switch (d) {
    case CUDA_COPY_HOST_TO_DEVICE: return hipMemcpyHostToDevice;
    case CUDA_COPY_DEVICE_TO_HOST: return hipMemcpyDeviceToHost;
    case CUDA_COPY_HOST_TO_HOST: return hipMemcpyHostToHost;
    case CUDA_COPY_DEVICE_TO_DEVICE: return hipMemcpyDeviceToDevice;
    default: *ok = false; return hipMemcpyDefault;
    }
}

static void copy_2d(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width,
                    size_t height, StreamId *async_id, CudaCopyDirection dir)
{
    bool ok;
    hipMemcpyKind k = copy_kind(dir, &ok);
    if (!ok) {
        bmi_fail(BMI_E_INVALID, "2-D copy: direction %d", (int)dir);
        return;
    }
    if (width == 0 || height == 0) return;
    if (async_id)
        bmi_check_hip(hipMemcpy2DAsync(dst, dpitch, src, spitch, width, height, k,
                                       (hipStream_t)*async_id),
                      "Gpu2DCopy(async)");
    else
        bmi_check_hip(hipMemcpy2D(dst, dpitch, src, spitch, width, height, k), "Gpu2DCopy");
}

// gpulib.c:52-61
extern "C" void Gpu2DCopy(CudaBuf *srcBuf, CudaBuf *destBuf, int src_x_off, int src_y_off,
                          int dest_x_off, int dest_y_off, int width, int height,
                          StreamId *P_asyncId, CudaCopyDirection direction, int src_no_pitch,
                          int dest_no_pitch)
{
    bmi_clear_error();
    if (!srcBuf || !destBuf || width < 0 || height < 0) {
        bmi_fail(BMI_E_INVALID, "Gpu2DCopy: bad args");
        return;
    }
    const unsigned char *src = (const unsigned char *)srcBuf->mBuf + src_x_off +
                               (size_t)src_y_off * srcBuf->mPitch;
    unsigned char *dst = (unsigned char *)destBuf->mBuf + dest_x_off + (size_t)dest_y_off * destBuf->mPitch;
    copy_2d(dst, dest_no_pitch ? width : destBuf->mPitch, src, src_no_pitch ? width : srcBuf->mPitch,
            width, height, P_asyncId, direction);
}

// gpulib.c:63-72
extern "C" void Gpu2DCopyWithDiffPitch(CudaBuf *srcBuf, CudaBuf *destBuf, int src_x_off,
                                       int src_y_off, int dest_x_off, int dest_y_off, int src_pitch,
                                       int dest_pitch, int width, int height, StreamId *P_asyncId,
                                       CudaCopyDirection direction, int src_no_pitch,
                                       int dest_no_pitch)
{
    bmi_clear_error();
    if (!srcBuf || !destBuf || width < 0 || height < 0) {
        bmi_fail(BMI_E_INVALID, "Gpu2DCopyWithDiffPitch: bad args");
        return;
    }
    const unsigned char *src =
        (const unsigned char *)srcBuf->mBuf + src_x_off + (size_t)src_y_off * src_pitch;
    unsigned char *dst = (unsigned char *)destBuf->mBuf + dest_x_off + (size_t)dest_y_off * dest_pitch;
    copy_2d(dst, dest_no_pitch ? width : dest_pitch, src, src_no_pitch ? width : src_pitch, width,
            height, P_asyncId, direction);
}

// gpulib.c:74-78 / gpufunc.cu:219-238: asynchronous when a stream is given
extern "C" void GpuLinearMemCopy(CudaBuf *srcBuf, CudaBuf *destBuf, int copy_size_in_byte,
                                 StreamId async_id, CudaCopyDirection direction)
{
    bmi_clear_error();
    if (!srcBuf || !destBuf || copy_size_in_byte < 0) {
        bmi_fail(BMI_E_INVALID, "GpuLinearMemCopy: bad args");
        return;
    }
    bool ok;
    hipMemcpyKind k = copy_kind(direction, &ok);
    if (!ok) {
        bmi_fail(BMI_E_INVALID, "GpuLinearMemCopy: direction");
        return;
    }
    if (async_id)
        bmi_check_hip(hipMemcpyAsync(destBuf->mBuf, srcBuf->mBuf, copy_size_in_byte, k,
                                     (hipStream_t)async_id),
                      "GpuLinearMemCopy(async)");
    else
        bmi_check_hip(hipMemcpy(destBuf->mBuf, srcBuf->mBuf, copy_size_in_byte, k),
                      "GpuLinearMemCopy");
}
