/*
 * bmi_gpulib.h -- C-ABI of the MI355X (gfx950) back end.
 *
 * Part 1 keeps, name for name and argument for argument, the entry points
 * the reference host code calls into its GPU library
 * (decoder/cudainterface/include/gpulib.h:24-57, implemented in
 * decoder/cudainterface/src/gpulib.c:12-377), so CudaManager-style host
 * code links against this library unchanged.  StreamId is hipStream_t.
 *
 * Compatible strengthening of the error convention: where the reference
 * printed and called exit(-1) on any CUDA error (gpufunc.cu:9-22) or
 * asserted on an unsupported mode, these functions never exit: int
 * functions return a negative BMI_E* code, void functions record it, and
 * BMI_last_error() returns the most recent code of the calling thread
 * (each void entry point of Part 1 resets it to BMI_OK on entry, so after
 * such a call it tells whether that call failed).
 *
 * Part 2 is additive (no counterpart in the reference): the fused /
 * batched persistent decode, the plain-path driver used as its baseline,
 * the request server and GPU T1.
 */
#ifndef BMI_GPULIB_H
#define BMI_GPULIB_H

#include "bmi_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* error codes */
#define BMI_OK 0
#define BMI_E_INVALID (-1)     /* bad argument / unsupported mode */
#define BMI_E_NODEVICE (-2)    /* no HIP device */
#define BMI_E_HIP (-3)         /* a HIP runtime call failed */
#define BMI_E_NOMEM (-4)       /* allocation failed or over budget */
#define BMI_E_UNSUPPORTED (-5) /* format/mode the reference also rejects */

/* ------------------------------------------------------------------------
 * Part 1: drop-in entry points
 * --------------------------------------------------------------------- */

/* gpulib.h:24 / gpufunc.cu:55-123: select device 0, fill name, CU count,
   clock and memory; 0 or -1 when no device */
int GpuInit(GpuInfo *gpfInfo);
/* gpulib.h:25 / gpufunc.cu:125-131 */
int GpuRemove(GpuInfo *gpfInfo);

/* gpulib.h:27-28 / gpufunc.cu:134-187: by mType -- GPU_BUF linear
   (mSize = mWidth*mHeight bytes when both are set), GPU_ALLIGN_BUF pitched
   (mPitch filled, mSize = mPitch*mHeight), HOST_* pinned host memory */
void MallocGpuBuf(CudaBuf *buf);
void FreeGpuBuf(CudaBuf *buf);

/* gpulib.h:30-33 / gpufunc.cu:189-217; GpuSync(NULL) syncs the device */
StreamId CreateANewStream(void);
void DestroyStream(StreamId *async_id);
void GpuSync(StreamId *P_asyncId);

/* gpulib.h:35-37 / gpulib.c:52-78: x offsets and widths in bytes, y in
   rows; P_asyncId == NULL -> synchronous; *_no_pitch -> pitch = width.
   GpuLinearMemCopy is asynchronous when async_id is not the null stream. */
void Gpu2DCopy(CudaBuf *stripeBuf, CudaBuf *deviceBuf, int src_x_off, int src_y_off,
               int dest_x_off, int dest_y_off, int width, int height, StreamId *P_asyncId,
               CudaCopyDirection direction, int src_no_pitch, int dest_no_pitch);
void Gpu2DCopyWithDiffPitch(CudaBuf *srcBuf, CudaBuf *destBuf, int src_x_off, int src_y_off,
                            int dest_x_off, int dest_y_off, int src_pitch, int dest_pitch,
                            int width, int height, StreamId *async_id,
                            CudaCopyDirection direction, int src_no_pitch, int dest_no_pitch);
void GpuLinearMemCopy(CudaBuf *srcBuf, CudaBuf *destBuf, int copy_size_in_byte,
                      StreamId async_id, CudaCopyDirection direction);

/* gpulib.h:44 / gpulib.c:165-185: multi-level IDWT of one tile-component,
   in place in `orig` (Mallat layout at the tile offset, pitch
   orig->mPitch); `temp` is transposed scratch (rows = tile columns, pitch
   >= 4*tile height, starting at row tile->mOff.x, column tile->mOff.y).
   toplevel -1 = all levels.  wordShift must be 2 (32-bit words).
   5/3: two's-complement int32 in and out.  9/7: level-0 LL and every
   detail band hold T1 sign-magnitude words, dequantised with
   mfAbsStepNorm and the lifting gains; the output is float32. */
int BMI_comp_idwt(TileInfo_c *tile, CompInfo_c *comp, CudaBuf *orig, CudaBuf *temp,
                  StreamId asyncId, short wordShift, EncodeMathod mathod, int toplevel);
/* gpulib.h:45 / gpulib.c:101-162: one synthesis level.  The reference
   implements only 9/7 here (its 5/3 branch asserts, gpufunc.cu:602-612);
   both filters are implemented. */
int BMI_resolution_idwt(TileInfo_c *tile, CompInfo_c *comp, CudaBuf *orig, CudaBuf *temp,
                        StreamId asyncId, short wordShift, EncodeMathod mathod,
                        int resolutionLevel);

/* gpulib.h:47-48 / gpulib.c:187-290: inverse RCT (5/3) / ICT (9/7) of
   components 0..2 of the tile into tile->mResultBuf at the tile offset
   (reduced to decodeLevel), packed per OutputFormat.  *_COMPONENTS and
   NONE formats: MCT asserts in the reference, ICT returns; both return. */
void BMI_tile_MCT(TileInfo_c *tileInfo, short bitDepth, short wordShift, int dwtlevel,
                  int decodeLevel, OutputFormat outFormat, StreamId asyncId);
void BMI_tile_ICT(TileInfo_c *tileInfo, short bitDepth, short wordShift, int dwtlevel,
                  int decodeLevel, OutputFormat outFormat, StreamId asyncId);
/* gpulib.h:49 / gpulib.c:292-341: N-component pack without colour
   transform (CPU MergeComponents semantics) */
void BMI_tile_Merge(TileInfo_c *tileInfo, int componentNum, int dwtLevel, int decodeLevel,
                    short bitDepth, short wordShift, int compBufPitchBytes,
                    OutputFormat outFormat, EncodeMathod method, StreamId asyncId);
/* gpulib.h:50 / gpulib.c:343-350: grey preview of one component into the
   32-bit result buffer (0x00RRGGBB words) */
void BMI_tile_One_Component(TileInfo_c *tile, int component_id, EncodeMathod method, int hasMCT,
                            int bitDepth, int wordShift, int dwtLevel, int decodeLevel,
                            StreamId asyncId);
/* gpulib.h:53-54 / gpulib.c:364-369 -> bmi_dc_shift_kernel: (int)v (9/7)
   + 2^(B-1), clamp to [0, 2^B-1], int32 out */
void BMI_Comp_DC_SHIFT(CudaBuf *compBuf, CudaBuf *compBufTemp, EncodeMathod methos,
                       int src_x_off, int src_y_off, int dst_x_off, int dst_y_off, int width,
                       int height, int bitDepth, int pixelWordLength, StreamId asyncId);

/* ------------------------------------------------------------------------
 * Part 2: additive API (no counterpart in the reference)
 * --------------------------------------------------------------------- */

int BMI_last_error(void);
const char *BMI_error_string(int code);

/* 5/3 synthesis flavour.  BMI_R53_REFERENCE (default) reproduces the
   reference CPU decoder bit for bit, including its bottom-edge quirk in
   the vertical pass (bmi_cpu_dwt_decoder.cpp:3311-3389); where that
   decoder is undefined (odd parity, length-1 lines) the T.800 result is
   produced.  BMI_R53_TEXTBOOK is ITU-T T.800 F.3 throughout. */
#define BMI_R53_REFERENCE 0
#define BMI_R53_TEXTBOOK 1
void BMI_set_r53_mode(int mode);
int BMI_get_r53_mode(void);

/* 9/7 coefficient encodings */
#define BMI_COEF_SIGNMAG 0 /* T1 sign-magnitude words */
#define BMI_COEF_FLOAT 1   /* float32 coefficients */

/* One tile (all its components) for the fused / batched decode. */
typedef struct BmiTileDesc {
    const void *coef[MAX_COMPONENT]; /* device Mallat planes (32-bit words) */
    int coef_pitch;                  /* elements */
    int width, height;               /* tile size at full resolution */
    int x0, y0;                      /* tile offset in the image (parities) */
    int levels;                      /* decomposition levels, 0..MAX_DWT_LEVEL */
    int top_level;                   /* -1 = all; else synthesise this many levels */
    int ncomp;                       /* 1..4 */
    int method;                      /* EncodeMathod */
    int coef_kind;                   /* 9/7: BMI_COEF_SIGNMAG or BMI_COEF_FLOAT */
    int bit_depth;                   /* 1..16 */
    int out_format;                  /* OutputFormat; NONE -> float/int planes */
    int check_out;                   /* -1 = all components, else one as grey */
    int mct;                         /* 1 = RCT/ICT of components 0..2 */
    int r53_mode;                    /* BMI_R53_REFERENCE / BMI_R53_TEXTBOOK */
    float steps[MAX_COMPONENT][3 * MAX_DWT_LEVEL + 1]; /* 9/7 mfAbsStepNorm */
    void *out;                       /* device packed pixels */
    int out_pitch_px;                /* pixels */
    int out_x, out_y;                /* tile position in `out`, pixels */
    void *planes_out[MAX_COMPONENT]; /* NONE: per-component result planes */
    int planes_pitch;                /* elements */
    unsigned int ignore_mask;        /* bit c set: component c is not decoded
                                        (the reference's mIgnoreComponents,
                                        decoder/common/include/codec_base.h:215-228);
                                        the others fill the output slots in
                                        order (bmi_cpu_dwt_decoder.cpp:1180-1290),
                                        an ignored check-out component means
                                        none; RGB formats need three left */
} BmiTileDesc;

/* Launch shape of the fused kernel (0 = library default / autotuned). */
typedef struct BmiTuning {
    int seg_rows;        /* tuned output rows per work item (even, default
                            32).  Each level step takes, from a cost model
                            of its waves per SIMD, a height up to this one,
                            or up to twice it when all the step's items fit
                            on the GPU at once */
    int waves_per_group; /* 1..4 wave64s per workgroup (0: 1 for launches of
                            under four items per SIMD, else 4) */
    int groups_per_cu;   /* persistent workgroups per CU */
    int persistent;      /* 0 (default) = one launch per level step (a step
                            mixing kernel variants runs them in one
                            dispatching launch); 1 = one persistent launch
                            with a device work queue; 2 = every step as one
                            dispatching launch; 3 = one launch per step and
                            variant */
    int min_waves_per_cu; /* segment heights are chosen among those giving
                             each level step this many items per CU, or
                             the smallest height (default 4) */
    int fixed_seg;       /* 1: every level step uses segments of exactly
                            seg_rows output rows (the cost model is off) */
    int fuse2;           /* the last two synthesis levels of a tile in
                            one pass (the lower one kept in LDS, never in
                            HBM): 0 default (on), -1 off, N > 0 on for tiles
                            whose final level spans at least N samples each
                            way.  Step modes 0 and 3 only; persistent and
                            all-merged modes do not fuse */
    int from_coarse;     /* 0 (default): level steps count from each
                            tile's finest level (the last step holds every
                            tile's final level, so all two-level launches
                            fall in one step); 1: from the coarsest (step k
                            holds every tile's level k) */
} BmiTuning;

typedef struct BmiBatchPlan BmiBatchPlan;

/* Build a plan for decoding `ntiles` tiles (of any size, level count,
   filter and component count) in one persistent launch.  Validates the
   descriptors, builds the device work queue and allocates the scratch it
   needs.  tuning may be NULL. */
int BMI_batch_plan_create(BmiBatchPlan **plan, const BmiTileDesc *tiles, int ntiles,
                          const BmiTuning *tuning);
/* Enqueue the decode on `stream` (asynchronous). */
int BMI_batch_plan_execute(BmiBatchPlan *plan, StreamId stream);
/* Capture execute() in a hipGraph on first use, replay it afterwards. */
int BMI_batch_plan_execute_graph(BmiBatchPlan *plan, StreamId stream);
/* Replace device pointers (coefficients / outputs) of tile i keeping its
   geometry, so one plan (and its graph) serves a stream of same-shaped
   frames.  Synchronous upload: no execute of this plan may be in flight. */
int BMI_batch_plan_rebind(BmiBatchPlan *plan, int i, const BmiTileDesc *tile);
size_t BMI_batch_plan_scratch_bytes(const BmiBatchPlan *plan);
/* Level-step plans: number of kernel launches one execute() enqueues, and
   enqueue only launches [first, first+count) -- for timing a single level
   (launch order is the dependency order, so a range starting at 0 is a
   valid partial decode). */
int BMI_batch_plan_num_launches(const BmiBatchPlan *plan);
int BMI_batch_plan_execute_launches(BmiBatchPlan *plan, int first, int count, StreamId stream);
int BMI_batch_plan_num_items(const BmiBatchPlan *plan);
int BMI_batch_plan_destroy(BmiBatchPlan *plan);
/* Host-side description of the plan BMI_batch_plan_create would build,
   made without device memory or a GPU: up to max_jobs job records (raw
   internal LevelJob structs of *job_bytes bytes each; pointers into the
   plan's own memory are offsets below *device_bytes) and the item count of
   up to max_launches launches.  No counterpart in the reference: it lets
   tests and tools check on the CPU every address the kernels will touch. */
int BMI_batch_plan_describe(const BmiTileDesc *tiles, int ntiles, const BmiTuning *tuning,
                            void *jobs, int max_jobs, int *njobs, int *job_bytes,
                            int *launch_counts, int max_launches, int *nlaunches,
                            size_t *device_bytes);
/* Non-zero when a persistent launch gave up waiting for a dependency
   (bounded spin) instead of hanging; synchronous. */
int BMI_batch_plan_error_flag(BmiBatchPlan *plan);

/* Autotuner: build a plan for `tiles` with each candidate tuning, replay
   its graph once untimed, then time `reps` graph replays and `reps` direct
   executions with HIP events on `stream`; a candidate scores the faster
   mean.  The six fastest are timed again with 4 x reps each, and the best
   of that second round is returned in *best (its mean time per execute in
   *best_us).  candidates == NULL selects the built-in grid (segment rows
   16..128 x items per CU 4..16 x per-launch or 4 waves per workgroup x
   merged or per-variant launches of mixed steps x two-level final steps
   on, from 256 samples, or off x steps aligned at the finest or the
   coarsest level).  Runs the real decode: outputs of the tiles are
   written. */
int BMI_autotune(const BmiTileDesc *tiles, int ntiles, const BmiTuning *candidates, int ncand,
                 int reps, StreamId stream, BmiTuning *best, float *best_us);

/* Plain path: the reference's own sequence of drop-in calls for one tile
   (BMI_comp_idwt per component, then BMI_tile_ICT / BMI_tile_MCT /
   BMI_tile_Merge), on device buffers described by `tile`.  The
   coefficient planes are overwritten (in-place IDWT).  `temp` must hold
   width*height*4 bytes (transposed scratch). */
int BMI_plain_decode(const BmiTileDesc *tile, void *temp, size_t temp_bytes, StreamId stream);

/* Decode `tiles` once, as BMI_batch_plan_create + BMI_batch_plan_execute +
   BMI_batch_plan_destroy would, without allocating device memory.  No
   counterpart in the reference: for callers that cannot allocate (code
   under hipGraph capture, tight serving loops, harnesses that forbid
   hipMalloc).  The plan's scratch planes, tables and counters live in one
   file-scope device array of BMI_decode_tiles_noalloc_capacity() bytes
   (64 MiB) per GPU, held until the process ends; a batch needing more
   returns BMI_E_NOMEM before any device work.  Asynchronous on `stream`
   after a host-side table build.  Every call on a GPU shares that GPU's
   array: calls that overlap on one GPU (from other threads, or on other
   streams while earlier work still runs) race. */
int BMI_decode_tiles_noalloc(const BmiTileDesc *tiles, int ntiles, const BmiTuning *tuning,
                             StreamId stream);
size_t BMI_decode_tiles_noalloc_capacity(void);

/* ------------------------------------------------------------------------
 * Request server (continuous batching over a fixed HBM budget)
 * --------------------------------------------------------------------- */

#define BMI_SERVE_FUSED 0 /* continuous batching: waiting requests that fit
                             are decoded together in one batch plan, several
                             batches in flight on separate streams */
#define BMI_SERVE_PLAIN 1 /* baseline: one request at a time through the
                             drop-in call sequence (BMI_plain_decode) */

typedef struct BmiServerConfig {
    size_t hbm_budget;  /* bytes of HBM the server holds for requests
                           (one allocation at create) */
    int num_streams;    /* batches in flight at once, 1..8 (plain: 1) */
    int max_batch;      /* requests per batch, >= 1 (plain: 1) */
    int mode;           /* BMI_SERVE_FUSED / BMI_SERVE_PLAIN */
    BmiTuning tuning;   /* launch shape of the batch plans */
} BmiServerConfig;

typedef struct BmiRequest {
    /* Geometry, filter and output format of the image.  host_io = 1: coef[]
       are host planes (pitch coef_pitch elements, pinned memory for
       asynchronous copies) and `out` a host pixel buffer (out_pitch_px,
       out_x, out_y), both copied by the server.  host_io = 0: they are
       device buffers of the caller, read and written in place.
       out_format NONE is not served. */
    BmiTileDesc tile;
    int host_io;
} BmiRequest;

typedef struct BmiRequestResult {
    int status;         /* BMI_OK or the error that failed the request */
    double submit_us;   /* server clock (us since create) at submit */
    double start_us;    /* admitted: its batch was enqueued */
    double done_us;     /* the server saw its batch complete */
    int batch_id;       /* batch the request ran in */
    int batch_size;     /* requests in that batch */
} BmiRequestResult;

typedef struct BmiServerStats {
    long long submitted, completed, failed, batches;
    size_t peak_bytes;   /* high-water mark of the arena */
    int peak_inflight;   /* most batches on the device at once */
} BmiServerStats;

typedef struct BmiServer BmiServer;

int BMI_server_create(BmiServer **server, const BmiServerConfig *cfg);
/* Validate and queue one request (returns at once).  Rejects, with no
   effect on other requests, a request that is malformed (BMI_E_INVALID /
   BMI_E_UNSUPPORTED) or whose footprint exceeds the whole budget
   (BMI_E_NOMEM).  *ticket identifies it for BMI_server_wait. */
int BMI_server_submit(BmiServer *server, const BmiRequest *req, unsigned long long *ticket);
/* Block until the request completes (timeout_ms < 0: no limit); fills
   *result and forgets the ticket.  Returns the request's status, or
   BMI_E_INVALID for an unknown ticket, or 1 on timeout. */
int BMI_server_wait(BmiServer *server, unsigned long long ticket, int timeout_ms,
                    BmiRequestResult *result);
int BMI_server_stats(BmiServer *server, BmiServerStats *stats);
/* Finish every queued request, then release the arena and streams. */
int BMI_server_destroy(BmiServer *server);

#ifdef __cplusplus
}
#endif
#endif
