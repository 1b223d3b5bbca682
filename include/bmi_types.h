/*
 * bmi_types.h -- C types that cross the back-end ABI.
 *
 * Field order and sizes follow the reference so that host code written
 * against it compiles and links unchanged:
 *   EncodeMathod, OutputFormat, s2d      decoder/common/include/codec_datatype.h:16-84
 *   CudaBufType, CudaBufStatus,
 *   CudaCopyDirection, CudaBuf, GpuInfo,
 *   SubbandInfo_c, CompInfo_c, TileInfo_c decoder/cudainterface/include/cudainterfacedatatype.h:13-152
 *   MAX_DWT_LEVEL, MAX_COMPONENT          decoder/codec/include/codec_macro.h:19-20
 * StreamId is a HIP stream (the reference's cudaStream_t).
 */
#ifndef BMI_TYPES_H
#define BMI_TYPES_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t *hipStream_t;
typedef hipStream_t StreamId;

#ifndef MAX_DWT_LEVEL
#define MAX_DWT_LEVEL 6
#endif
#ifndef MAX_COMPONENT
#define MAX_COMPONENT 4
#endif

typedef unsigned char uint_8;

typedef enum EncodeMathod_e { Ckernels_W5X3, Ckernels_W9X7, UNKNOWN_MATHOD } EncodeMathod;

typedef enum OutputFormat_e {
    IMAGE_FOTMAT_32BITS_ARGB = 0,
    IMAGE_FOTMAT_48BITS_RGB,
    IMAGE_FOTMAT_24BITS_RGB,
    IMAGE_FOTMAT_32BITS_4COMPONENTS,
    IMAGE_FOTMAT_64BITS_4COMPONENTS,
    IMAGE_FOTMAT_48BITS_3COMPONENTS,
    IMAGE_FOTMAT_24BITS_3COMPONENTS,
    IMAGE_FOTMAT_NONE,
    IMAGE_FOTMAT_64BITS_ARGB,
    IMAGE_FOTMAT_36BITS_RGB,
    IMAGE_FOTMAT_48BITS_ARGB
} OutputFormat;

typedef struct ShortSize_s {
    short x;
    short y;
} sSize, sPos, s2d;

typedef enum CudaBufType_e { HOST_BUF, HOST_ALLLIGN_BUF, GPU_BUF, GPU_ALLIGN_BUF, UNKNOWN_BUF_TYPE } CudaBufType;

typedef enum CudaBufStatus_e {
    BUF_STATUS_UNINITIALIZED,
    BUF_STATUS_IDLE,
    BUF_STATUS_DECODING,
    BUF_STATUS_UPLOADING,
    BUF_STATUS_SYNCING
} CudaBufStatus;

typedef enum CudaCopyDirection_e {
    CUDA_COPY_HOST_TO_DEVICE,
    CUDA_COPY_DEVICE_TO_HOST,
    CUDA_COPY_HOST_TO_HOST,
    CUDA_COPY_DEVICE_TO_DEVICE,
    INVALID_COPY_DIRECTION
} CudaCopyDirection;

typedef struct SpcValue_s {
    unsigned char mRes_Subband;
    unsigned char mComponentId;
    unsigned short mTileId;
} SpcValue;

typedef union BufContainInfo_s {
    unsigned int mContainInfo;
    SpcValue mSpcValue;
} BufContainInfo;

typedef struct CudaBuf_s {
    CudaBufType mType;
    unsigned int mSize;
    unsigned int mPitch;
    unsigned int mWidth;
    unsigned int mHeight;
    CudaBufStatus mStatus;
    StreamId mAsyncId;
    int mOwnerThreadIndex;
    BufContainInfo mBufContain;
    void *mBuf;
} CudaBuf, *P_CudaBuf;

typedef struct bmi_gpu_info_s {
    int active;
    int device;
    char name[256];
    int major;
    int minor;
    unsigned int memory;
    int multiProcessorCount;
    int clockRate;
} GpuInfo;

typedef struct SubbandInfo_s_c {
    int mId;
    s2d mOff;
    s2d mSize;
    union {
        int miAbsStep;
        float mfAbsStepNorm;
    };
    int mQquanShift;
    int mLinesInserted;
} SubbandInfo_c;

typedef struct CompInfo_s_c {
    SubbandInfo_c *mSubbandInfo;
    int mNumOfSubband;
    int mSubbandFinishedBitmap;
    int mResolutionFinishedId;
    int mSubbandFinishedCounter;
    int mAsyncOperationStreamIndex;
    union {
        CudaBuf *mCompBuf;
        void *mDwtBuf;
    };
    uint_8 mDeShifted;
    uint_8 mCompBufMoved;
} CompInfo_c;

typedef struct TileInfo_s_c {
    s2d mOff;
    s2d mSize;
    CompInfo_c *mCompInfo;
    int mComFinishedCounter;
    int mLastCompId;
    union {
        CudaBuf *mResultBuf;
        void *mCpuResultBuf;
    };
    uint_8 mXParity[MAX_DWT_LEVEL];
    uint_8 mYParity[MAX_DWT_LEVEL];
} TileInfo_c;

#ifdef __cplusplus
}
#endif
#endif
