"""Python mirror of the gfx950 back end's C-ABI (include/bmi_gpulib.h).

The product is the C-ABI shared library ``libbmi_j2k.so`` built from
``csrc/``; this module only binds it with ctypes so tests, the benchmark and
Python callers can drive it on device buffers they own (torch tensors are
used for device memory and streams).  Loading fails loudly when the library
is missing: there is no CPU fallback anywhere in the product path.

Constants and the meaning of every argument follow the reference's own
interface (decoder/cudainterface/include/gpulib.h:24-57,
decoder/common/include/codec_datatype.h:16-51).
"""
from __future__ import annotations

import ctypes
import os
import pathlib

PKG_DIR = pathlib.Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "libbmi_j2k.so"

# EncodeMathod (codec_datatype.h:16-21)
W5X3 = 0
W9X7 = 1

# OutputFormat (codec_datatype.h:38-51)
FMT_32BITS_ARGB = 0
FMT_48BITS_RGB = 1
FMT_24BITS_RGB = 2
FMT_32BITS_4COMPONENTS = 3
FMT_64BITS_4COMPONENTS = 4
FMT_48BITS_3COMPONENTS = 5
FMT_24BITS_3COMPONENTS = 6
FMT_NONE = 7
FMT_64BITS_ARGB = 8
FMT_36BITS_RGB = 9
FMT_48BITS_ARGB = 10

PIXEL_BYTES = {0: 4, 1: 6, 2: 3, 3: 4, 4: 8, 5: 6, 6: 3, 8: 8}
RGB_FORMATS = (0, 1, 2, 8)
MERGE_FORMATS = (3, 4, 5, 6)

COEF_SIGNMAG = 0
COEF_FLOAT = 1
R53_REFERENCE = 0
R53_TEXTBOOK = 1

BMI_OK = 0
BMI_E_INVALID = -1
BMI_E_NODEVICE = -2
BMI_E_HIP = -3
BMI_E_NOMEM = -4
BMI_E_UNSUPPORTED = -5

MAX_COMPONENT = 4
MAX_DWT_LEVEL = 6


class BmiTileDesc(ctypes.Structure):
    """struct BmiTileDesc (include/bmi_gpulib.h)."""

    _fields_ = [
        ("coef", ctypes.c_void_p * MAX_COMPONENT),
        ("coef_pitch", ctypes.c_int),
        ("width", ctypes.c_int),
        ("height", ctypes.c_int),
        ("x0", ctypes.c_int),
        ("y0", ctypes.c_int),
        ("levels", ctypes.c_int),
        ("top_level", ctypes.c_int),
        ("ncomp", ctypes.c_int),
        ("method", ctypes.c_int),
        ("coef_kind", ctypes.c_int),
        ("bit_depth", ctypes.c_int),
        ("out_format", ctypes.c_int),
        ("check_out", ctypes.c_int),
        ("mct", ctypes.c_int),
        ("r53_mode", ctypes.c_int),
        ("steps", (ctypes.c_float * (3 * MAX_DWT_LEVEL + 1)) * MAX_COMPONENT),
        ("out", ctypes.c_void_p),
        ("out_pitch_px", ctypes.c_int),
        ("out_x", ctypes.c_int),
        ("out_y", ctypes.c_int),
        ("planes_out", ctypes.c_void_p * MAX_COMPONENT),
        ("planes_pitch", ctypes.c_int),
        ("ignore_mask", ctypes.c_uint),
    ]


class BmiTuning(ctypes.Structure):
    _fields_ = [
        ("seg_rows", ctypes.c_int),
        ("waves_per_group", ctypes.c_int),
        ("groups_per_cu", ctypes.c_int),
        ("persistent", ctypes.c_int),
        ("min_waves_per_cu", ctypes.c_int),
        ("fixed_seg", ctypes.c_int),
        ("fuse2", ctypes.c_int),
        ("from_coarse", ctypes.c_int),
    ]


SERVE_FUSED = 0
SERVE_PLAIN = 1


class BmiServerConfig(ctypes.Structure):
    _fields_ = [
        ("hbm_budget", ctypes.c_size_t),
        ("num_streams", ctypes.c_int),
        ("max_batch", ctypes.c_int),
        ("mode", ctypes.c_int),
        ("tuning", BmiTuning),
    ]


class BmiRequest(ctypes.Structure):
    _fields_ = [("tile", BmiTileDesc), ("host_io", ctypes.c_int)]


class BmiRequestResult(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int),
        ("submit_us", ctypes.c_double),
        ("start_us", ctypes.c_double),
        ("done_us", ctypes.c_double),
        ("batch_id", ctypes.c_int),
        ("batch_size", ctypes.c_int),
    ]


class BmiServerStats(ctypes.Structure):
    _fields_ = [
        ("submitted", ctypes.c_longlong),
        ("completed", ctypes.c_longlong),
        ("failed", ctypes.c_longlong),
        ("batches", ctypes.c_longlong),
        ("peak_bytes", ctypes.c_size_t),
        ("peak_inflight", ctypes.c_int),
    ]


T1_WORDS = 0
T1_INT = 1


class BmiT1Block(ctypes.Structure):
    """struct BmiT1Block (include/bmi_t1.h)."""

    _fields_ = [
        ("data_offset", ctypes.c_ulonglong),
        ("len", ctypes.c_int),
        ("width", ctypes.c_int),
        ("height", ctypes.c_int),
        ("orient", ctypes.c_int),
        ("numbps", ctypes.c_int),
        ("missing", ctypes.c_int),
        ("num_pass", ctypes.c_int),
        ("kind", ctypes.c_int),
        ("out", ctypes.c_void_p),
        ("out_pitch", ctypes.c_int),
        ("style", ctypes.c_int),
        ("nseg", ctypes.c_int),
        ("seg_offset", ctypes.c_ulonglong),
    ]


# code-block style bits (include/bmi_t1.h, ITU-T T.800 Table A.19)
T1_BYPASS, T1_RESET, T1_TERMALL, T1_VSC, T1_PTERM, T1_SEGSYM = 0x01, 0x02, 0x04, 0x08, 0x10, 0x20


# BMI_codestream_open flags (include/bmi_codestream.h)
CS_REFERENCE = 0
CS_GUARD_T800 = 1


class BmiCodestreamInfo(ctypes.Structure):
    """struct BmiCodestreamInfo (include/bmi_codestream.h)."""

    _fields_ = [(n, ctypes.c_int) for n in (
        "width", "height", "x0", "y0", "tile_w", "tile_h", "tile_x0", "tile_y0", "ntiles_x",
        "ntiles_y", "ncomp", "bit_depth", "levels", "reversible", "mct", "layers", "progression",
        "cb_w", "cb_h", "guard_bits", "colour_space", "total_blocks")]


class GpuInfo(ctypes.Structure):
    _fields_ = [
        ("active", ctypes.c_int),
        ("device", ctypes.c_int),
        ("name", ctypes.c_char * 256),
        ("major", ctypes.c_int),
        ("minor", ctypes.c_int),
        ("memory", ctypes.c_uint),
        ("multiProcessorCount", ctypes.c_int),
        ("clockRate", ctypes.c_int),
    ]


# ---------------------------------------------------------------------------
# Types of the drop-in boundary (include/bmi_types.h; the reference's
# decoder/cudainterface/include/cudainterfacedatatype.h:13-152)

# CudaBufType
HOST_BUF, HOST_ALLLIGN_BUF, GPU_BUF, GPU_ALLIGN_BUF = 0, 1, 2, 3
# CudaCopyDirection
COPY_H2D, COPY_D2H, COPY_H2H, COPY_D2D = 0, 1, 2, 3


class S2d(ctypes.Structure):
    _fields_ = [("x", ctypes.c_short), ("y", ctypes.c_short)]


class CudaBuf(ctypes.Structure):
    _fields_ = [
        ("mType", ctypes.c_int),
        ("mSize", ctypes.c_uint),
        ("mPitch", ctypes.c_uint),
        ("mWidth", ctypes.c_uint),
        ("mHeight", ctypes.c_uint),
        ("mStatus", ctypes.c_int),
        ("mAsyncId", ctypes.c_void_p),
        ("mOwnerThreadIndex", ctypes.c_int),
        ("mBufContain", ctypes.c_uint),
        ("mBuf", ctypes.c_void_p),
    ]


class _StepUnion(ctypes.Union):
    _fields_ = [("miAbsStep", ctypes.c_int), ("mfAbsStepNorm", ctypes.c_float)]


class SubbandInfo_c(ctypes.Structure):
    _anonymous_ = ("_u",)
    _fields_ = [
        ("mId", ctypes.c_int),
        ("mOff", S2d),
        ("mSize", S2d),
        ("_u", _StepUnion),
        ("mQquanShift", ctypes.c_int),
        ("mLinesInserted", ctypes.c_int),
    ]


class _CompBufUnion(ctypes.Union):
    _fields_ = [("mCompBuf", ctypes.POINTER(CudaBuf)), ("mDwtBuf", ctypes.c_void_p)]


class CompInfo_c(ctypes.Structure):
    _anonymous_ = ("_u",)
    _fields_ = [
        ("mSubbandInfo", ctypes.POINTER(SubbandInfo_c)),
        ("mNumOfSubband", ctypes.c_int),
        ("mSubbandFinishedBitmap", ctypes.c_int),
        ("mResolutionFinishedId", ctypes.c_int),
        ("mSubbandFinishedCounter", ctypes.c_int),
        ("mAsyncOperationStreamIndex", ctypes.c_int),
        ("_u", _CompBufUnion),
        ("mDeShifted", ctypes.c_ubyte),
        ("mCompBufMoved", ctypes.c_ubyte),
    ]


class _ResultBufUnion(ctypes.Union):
    _fields_ = [("mResultBuf", ctypes.POINTER(CudaBuf)), ("mCpuResultBuf", ctypes.c_void_p)]


class TileInfo_c(ctypes.Structure):
    _anonymous_ = ("_u",)
    _fields_ = [
        ("mOff", S2d),
        ("mSize", S2d),
        ("mCompInfo", ctypes.POINTER(CompInfo_c)),
        ("mComFinishedCounter", ctypes.c_int),
        ("mLastCompId", ctypes.c_int),
        ("_u", _ResultBufUnion),
        ("mXParity", ctypes.c_ubyte * MAX_DWT_LEVEL),
        ("mYParity", ctypes.c_ubyte * MAX_DWT_LEVEL),
    ]


_LIB = None


def lib() -> ctypes.CDLL:
    """Load libbmi_j2k.so (raises if it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `make -C {PKG_DIR}` "
            "(or __graft_entry__.build()); there is no CPU fallback")
    # The library needs the HIP runtime (NEEDED libamdhip64.so.7).  torch
    # bundles its own copy of that runtime; loading torch first makes the
    # library bind to the copy torch already uses, so one process never runs
    # two HIP runtimes side by side (device buffers and streams are shared
    # with torch).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(str(LIB_PATH))
    vp, i, f = ctypes.c_void_p, ctypes.c_int, ctypes.c_float
    L.BMI_last_error.restype = i
    L.BMI_error_string.restype = ctypes.c_char_p
    L.BMI_error_string.argtypes = [i]
    L.BMI_set_r53_mode.argtypes = [i]
    L.BMI_get_r53_mode.restype = i
    L.GpuInit.argtypes = [ctypes.POINTER(GpuInfo)]
    L.GpuInit.restype = i
    L.GpuRemove.argtypes = [ctypes.POINTER(GpuInfo)]
    L.GpuRemove.restype = i
    bp, tp, cp = ctypes.POINTER(CudaBuf), ctypes.POINTER(TileInfo_c), ctypes.POINTER(CompInfo_c)
    sp = ctypes.POINTER(vp)
    sh = ctypes.c_short
    L.MallocGpuBuf.argtypes = [bp]
    L.MallocGpuBuf.restype = None
    L.FreeGpuBuf.argtypes = [bp]
    L.FreeGpuBuf.restype = None
    L.CreateANewStream.argtypes = []
    L.CreateANewStream.restype = vp
    L.DestroyStream.argtypes = [sp]
    L.DestroyStream.restype = None
    L.GpuSync.argtypes = [sp]
    L.GpuSync.restype = None
    L.Gpu2DCopy.argtypes = [bp, bp, i, i, i, i, i, i, sp, i, i, i]
    L.Gpu2DCopy.restype = None
    L.Gpu2DCopyWithDiffPitch.argtypes = [bp, bp, i, i, i, i, i, i, i, i, sp, i, i, i]
    L.Gpu2DCopyWithDiffPitch.restype = None
    L.GpuLinearMemCopy.argtypes = [bp, bp, i, vp, i]
    L.GpuLinearMemCopy.restype = None
    L.BMI_comp_idwt.argtypes = [tp, cp, bp, bp, vp, sh, i, i]
    L.BMI_comp_idwt.restype = i
    L.BMI_resolution_idwt.argtypes = [tp, cp, bp, bp, vp, sh, i, i]
    L.BMI_resolution_idwt.restype = i
    L.BMI_tile_MCT.argtypes = [tp, sh, sh, i, i, i, vp]
    L.BMI_tile_MCT.restype = None
    L.BMI_tile_ICT.argtypes = [tp, sh, sh, i, i, i, vp]
    L.BMI_tile_ICT.restype = None
    L.BMI_tile_Merge.argtypes = [tp, i, i, i, sh, sh, i, i, i, vp]
    L.BMI_tile_Merge.restype = None
    L.BMI_tile_One_Component.argtypes = [tp, i, i, i, i, i, i, i, vp]
    L.BMI_tile_One_Component.restype = None
    L.BMI_Comp_DC_SHIFT.argtypes = [bp, bp, i, i, i, i, i, i, i, i, i, vp]
    L.BMI_Comp_DC_SHIFT.restype = None
    L.BMI_batch_plan_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(BmiTileDesc), i,
                                        ctypes.POINTER(BmiTuning)]
    L.BMI_batch_plan_create.restype = i
    L.BMI_batch_plan_execute.argtypes = [vp, vp]
    L.BMI_batch_plan_execute.restype = i
    L.BMI_batch_plan_execute_graph.argtypes = [vp, vp]
    L.BMI_batch_plan_execute_graph.restype = i
    L.BMI_batch_plan_rebind.argtypes = [vp, i, ctypes.POINTER(BmiTileDesc)]
    L.BMI_batch_plan_rebind.restype = i
    L.BMI_batch_plan_scratch_bytes.argtypes = [vp]
    L.BMI_batch_plan_scratch_bytes.restype = ctypes.c_size_t
    L.BMI_batch_plan_num_items.argtypes = [vp]
    L.BMI_batch_plan_num_items.restype = i
    L.BMI_batch_plan_destroy.argtypes = [vp]
    L.BMI_batch_plan_destroy.restype = i
    L.BMI_batch_plan_error_flag.argtypes = [vp]
    L.BMI_batch_plan_error_flag.restype = i
    L.BMI_batch_plan_num_launches.argtypes = [vp]
    L.BMI_batch_plan_num_launches.restype = i
    L.BMI_batch_plan_execute_launches.argtypes = [vp, i, i, vp]
    L.BMI_batch_plan_execute_launches.restype = i
    L.BMI_autotune.argtypes = [ctypes.POINTER(BmiTileDesc), i, ctypes.POINTER(BmiTuning), i, i, vp,
                               ctypes.POINTER(BmiTuning), ctypes.POINTER(f)]
    L.BMI_autotune.restype = i
    L.BMI_server_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(BmiServerConfig)]
    L.BMI_server_create.restype = i
    L.BMI_server_submit.argtypes = [vp, ctypes.POINTER(BmiRequest), ctypes.POINTER(ctypes.c_ulonglong)]
    L.BMI_server_submit.restype = i
    L.BMI_server_wait.argtypes = [vp, ctypes.c_ulonglong, i, ctypes.POINTER(BmiRequestResult)]
    L.BMI_server_wait.restype = i
    L.BMI_server_stats.argtypes = [vp, ctypes.POINTER(BmiServerStats)]
    L.BMI_server_stats.restype = i
    L.BMI_server_destroy.argtypes = [vp]
    L.BMI_server_destroy.restype = i
    L.BMI_t1_decode_blocks.argtypes = [vp, i, vp, vp]
    L.BMI_t1_decode_blocks.restype = i
    L.BMI_t1_check_blocks.argtypes = [ctypes.POINTER(BmiT1Block), i, ctypes.POINTER(i)]
    L.BMI_t1_num_segments.argtypes = [i, i]
    L.BMI_t1_order_blocks.argtypes = [ctypes.POINTER(BmiT1Block), i]
    L.BMI_t1_check_blocks.restype = i
    L.BMI_decode_tiles_noalloc.argtypes = [ctypes.POINTER(BmiTileDesc), i, ctypes.POINTER(BmiTuning), vp]
    L.BMI_decode_tiles_noalloc.restype = i
    L.BMI_decode_tiles_noalloc_capacity.argtypes = []
    L.BMI_decode_tiles_noalloc_capacity.restype = ctypes.c_size_t
    L.BMI_plain_decode.argtypes = [ctypes.POINTER(BmiTileDesc), vp, ctypes.c_size_t, vp]
    L.BMI_plain_decode.restype = i
    L.BMI_codestream_open.argtypes = [vp, ctypes.c_size_t, i, ctypes.POINTER(vp),
                                      ctypes.POINTER(BmiCodestreamInfo)]
    L.BMI_codestream_open.restype = i
    ip = ctypes.POINTER(i)
    L.BMI_codestream_tile_info.argtypes = [vp, i, ip, ip, ip, ip, ip, ctypes.POINTER(ctypes.c_size_t)]
    L.BMI_codestream_tile_info.restype = i
    L.BMI_codestream_tile_blocks.argtypes = [vp, i, ctypes.POINTER(vp), i, ctypes.c_ulonglong,
                                             ctypes.POINTER(BmiT1Block), vp, ctypes.c_size_t,
                                             ctypes.POINTER(BmiTileDesc)]
    L.BMI_codestream_tile_blocks.restype = i
    L.BMI_codestream_close.argtypes = [vp]
    L.BMI_codestream_close.restype = i
    _LIB = L
    return L


class BmiError(RuntimeError):
    pass


def check(rc: int, what: str) -> None:
    if rc != BMI_OK:
        msg = lib().BMI_error_string(rc).decode()
        raise BmiError(f"{what}: {msg} ({rc})")


def tile_desc(*, coef_ptrs, coef_pitch, width, height, x0=0, y0=0, levels, method,
              ncomp, coef_kind=COEF_FLOAT, bit_depth=8, out_format=FMT_NONE, check_out=-1,
              mct=None, r53_mode=R53_REFERENCE, steps=None, out_ptr=0, out_pitch_px=0,
              out_x=0, out_y=0, planes_ptrs=None, planes_pitch=0, top_level=-1, ignore_mask=0) -> BmiTileDesc:
    """Fill a BmiTileDesc; steps is [ncomp][3L+1] (defaults to 1.0)."""
    d = BmiTileDesc()
    for c, p in enumerate(coef_ptrs):
        d.coef[c] = p
    d.coef_pitch = coef_pitch
    d.width, d.height, d.x0, d.y0 = width, height, x0, y0
    d.levels, d.top_level, d.ncomp, d.method = levels, top_level, ncomp, method
    d.coef_kind, d.bit_depth, d.out_format, d.check_out = coef_kind, bit_depth, out_format, check_out
    d.mct = int(out_format in RGB_FORMATS) if mct is None else int(mct)
    d.r53_mode = r53_mode
    for c in range(MAX_COMPONENT):
        for b in range(3 * MAX_DWT_LEVEL + 1):
            d.steps[c][b] = 1.0
    if steps is not None:
        for c, row in enumerate(steps):
            for b, v in enumerate(row):
                d.steps[c][b] = float(v)
    d.out, d.out_pitch_px, d.out_x, d.out_y = out_ptr or None, out_pitch_px, out_x, out_y
    if planes_ptrs is not None:
        for c, p in enumerate(planes_ptrs):
            d.planes_out[c] = p
    d.planes_pitch = planes_pitch
    d.ignore_mask = ignore_mask
    return d


class BatchPlan:
    """BMI_batch_plan_* wrapper: one persistent launch for many tiles."""

    def __init__(self, descs, tuning: BmiTuning | None = None):
        L = lib()
        arr = (BmiTileDesc * len(descs))(*descs)
        self._descs = arr
        self.handle = ctypes.c_void_p()
        rc = L.BMI_batch_plan_create(ctypes.byref(self.handle), arr, len(descs),
                                     ctypes.byref(tuning) if tuning is not None else None)
        check(rc, "BMI_batch_plan_create")

    def execute(self, stream: int = 0) -> None:
        check(lib().BMI_batch_plan_execute(self.handle, stream or None), "BMI_batch_plan_execute")

    def execute_graph(self, stream: int = 0) -> None:
        check(lib().BMI_batch_plan_execute_graph(self.handle, stream or None),
              "BMI_batch_plan_execute_graph")

    @property
    def num_launches(self) -> int:
        return lib().BMI_batch_plan_num_launches(self.handle)

    def execute_launches(self, first: int, count: int, stream: int = 0) -> None:
        check(lib().BMI_batch_plan_execute_launches(self.handle, first, count, stream or None),
              "BMI_batch_plan_execute_launches")

    def rebind(self, i: int, desc: BmiTileDesc) -> None:
        check(lib().BMI_batch_plan_rebind(self.handle, i, ctypes.byref(desc)), "rebind")

    @property
    def num_items(self) -> int:
        return lib().BMI_batch_plan_num_items(self.handle)

    @property
    def scratch_bytes(self) -> int:
        return lib().BMI_batch_plan_scratch_bytes(self.handle)

    def error_flag(self) -> int:
        return lib().BMI_batch_plan_error_flag(self.handle)

    def close(self) -> None:
        if self.handle:
            lib().BMI_batch_plan_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune(descs, candidates=None, reps: int = 5, stream: int = 0):
    """BMI_autotune: fastest tuning for these tiles -> (BmiTuning, us per execute)."""
    arr = (BmiTileDesc * len(descs))(*descs)
    cand = None
    n = 0
    if candidates:
        cand = (BmiTuning * len(candidates))(*candidates)
        n = len(candidates)
    best = BmiTuning()
    us = ctypes.c_float()
    check(lib().BMI_autotune(arr, len(descs), cand, n, reps, stream or None, ctypes.byref(best),
                             ctypes.byref(us)), "BMI_autotune")
    return best, float(us.value)


def decode_tiles_noalloc(descs, tuning: BmiTuning | None = None, stream: int = 0) -> None:
    """BMI_decode_tiles_noalloc: decode the tiles once with no device
    allocation (the library's fixed per-GPU scratch array)."""
    arr = (BmiTileDesc * len(descs))(*descs)
    check(lib().BMI_decode_tiles_noalloc(arr, len(descs), ctypes.byref(tuning) if tuning else None,
                                         stream or None), "BMI_decode_tiles_noalloc")


def plain_decode(desc: BmiTileDesc, temp_ptr: int, temp_bytes: int, stream: int = 0) -> None:
    """The reference's drop-in call sequence for one tile (in place)."""
    check(lib().BMI_plain_decode(ctypes.byref(desc), temp_ptr, temp_bytes, stream or None),
          "BMI_plain_decode")


def level_span(t0: int, n: int, levels: int, lev: int):
    s = levels - 1 - lev
    return -(-t0 // (1 << s)), -(-(t0 + n) // (1 << s))


def tile_geometry(w: int, h: int, x0: int, y0: int, levels: int):
    """Per synthesis level (0 = coarsest): dict of spans and band sizes."""
    out = []
    for lev in range(levels):
        a0, a1 = level_span(x0, w, levels, lev)
        b0, b1 = level_span(y0, h, levels, lev)
        lw = (a1 + 1) // 2 - (a0 + 1) // 2
        lh = (b1 + 1) // 2 - (b0 + 1) // 2
        out.append(dict(a0=a0, a1=a1, b0=b0, b1=b1, nw=a1 - a0, nh=b1 - b0, lw=lw, lh=lh))
    return out


def output_size(w: int, h: int, x0: int, y0: int, levels: int, top_level: int = -1):
    """Output size of a tile decoded to `top_level` synthesis levels (-1:
    all; 0: the thumbnail, i.e. the LL band)."""
    if top_level < 0 or top_level >= levels:
        return w, h
    if top_level == 0:
        s = 1 << levels

        def cd(a):
            return -(-a // s)
        return cd(x0 + w) - cd(x0), cd(y0 + h) - cd(y0)
    g = tile_geometry(w, h, x0, y0, levels)[top_level - 1]
    return g["nw"], g["nh"]


class Server:
    """BMI_server_*: continuous-batching request server over an HBM budget."""

    def __init__(self, hbm_budget: int, num_streams: int = 4, max_batch: int = 64,
                 mode: int = SERVE_FUSED, tuning: BmiTuning | None = None):
        cfg = BmiServerConfig()
        cfg.hbm_budget, cfg.num_streams, cfg.max_batch, cfg.mode = (hbm_budget, num_streams,
                                                                    max_batch, mode)
        if tuning is not None:
            cfg.tuning = tuning
        self.handle = ctypes.c_void_p()
        check(lib().BMI_server_create(ctypes.byref(self.handle), ctypes.byref(cfg)),
              "BMI_server_create")

    def submit(self, desc: BmiTileDesc, host_io: bool = False) -> int:
        req = BmiRequest()
        req.tile = desc
        req.host_io = int(host_io)
        t = ctypes.c_ulonglong()
        check(lib().BMI_server_submit(self.handle, ctypes.byref(req), ctypes.byref(t)),
              "BMI_server_submit")
        return t.value

    def wait(self, ticket: int, timeout_ms: int = -1) -> BmiRequestResult:
        res = BmiRequestResult()
        rc = lib().BMI_server_wait(self.handle, ticket, timeout_ms, ctypes.byref(res))
        if rc == 1:
            raise TimeoutError(f"request {ticket} not done after {timeout_ms} ms")
        check(rc, "BMI_server_wait")
        return res

    def stats(self) -> BmiServerStats:
        st = BmiServerStats()
        check(lib().BMI_server_stats(self.handle, ctypes.byref(st)), "BMI_server_stats")
        return st

    def close(self) -> None:
        if self.handle:
            lib().BMI_server_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def t1_check_blocks(blocks) -> None:
    """BMI_t1_check_blocks on a host array of BmiT1Block."""
    arr = (BmiT1Block * len(blocks))(*blocks)
    bad = ctypes.c_int(-1)
    rc = lib().BMI_t1_check_blocks(arr, len(blocks), ctypes.byref(bad))
    if rc != BMI_OK:
        raise BmiError(f"BMI_t1_check_blocks: block {bad.value}: {lib().BMI_error_string(rc).decode()}")


def t1_order_blocks(blocks) -> None:
    """BMI_t1_order_blocks on a ctypes BmiT1Block array (in place): longest
    code-block data first."""
    check(lib().BMI_t1_order_blocks(blocks, len(blocks)), "BMI_t1_order_blocks")


def t1_decode_blocks(blocks_dev_ptr: int, nblocks: int, data_dev_ptr: int, stream: int = 0) -> None:
    """BMI_t1_decode_blocks on device block table and code-stream buffer."""
    check(lib().BMI_t1_decode_blocks(blocks_dev_ptr, nblocks, data_dev_ptr, stream or None),
          "BMI_t1_decode_blocks")


# ------------------------------------------------------------ codestream --

class Codestream:
    """BMI_codestream_* wrapper: host-side parse and tier-2 decode of a JPEG
    2000 codestream (.j2k) or JP2 file, feeding the GPU tier-1 decoder and
    the batched IDWT / colour plan (the reference's BMI_dec_info + dec_t2,
    decoder/codec/src/dec_jp2.cpp:420-560, dec_tsk.cpp:107-185)."""

    def __init__(self, data: bytes, flags: int = CS_REFERENCE):
        L = lib()
        self._buf = bytes(data)
        self.handle = ctypes.c_void_p()
        self.info = BmiCodestreamInfo()
        check(L.BMI_codestream_open(self._buf, len(self._buf), flags, ctypes.byref(self.handle),
                                    ctypes.byref(self.info)), "BMI_codestream_open")

    @property
    def ntiles(self) -> int:
        return self.info.ntiles_x * self.info.ntiles_y

    def tile_info(self, t: int):
        """(x0, y0, w, h, nblocks, seg_bytes) of tile t."""
        v = [ctypes.c_int() for _ in range(5)]
        sb = ctypes.c_size_t()
        check(lib().BMI_codestream_tile_info(self.handle, t, *[ctypes.byref(x) for x in v],
                                             ctypes.byref(sb)), "BMI_codestream_tile_info")
        return tuple(x.value for x in v) + (sb.value,)

    def tile_blocks(self, t: int, plane_ptrs, plane_pitch: int, seg_base: int = 0, segs=None):
        """Block table, segments and tile descriptor of tile t.  `segs` may
        be a writable uint8 numpy view of at least seg_bytes + 16 bytes;
        otherwise one is allocated.  Returns (BmiT1Block array, segs, desc)."""
        import numpy as np
        _, _, _, _, nb, sb = self.tile_info(t)
        blocks = (BmiT1Block * max(nb, 1))()
        if segs is None:
            segs = np.zeros(sb + 16, dtype=np.uint8)
        planes = (ctypes.c_void_p * MAX_COMPONENT)(*[p or None for p in plane_ptrs])
        desc = BmiTileDesc()
        check(lib().BMI_codestream_tile_blocks(self.handle, t, planes, plane_pitch, seg_base, blocks,
                                               segs.ctypes.data_as(ctypes.c_void_p), segs.nbytes,
                                               ctypes.byref(desc)), "BMI_codestream_tile_blocks")
        return blocks, segs, desc

    def close(self) -> None:
        if self.handle:
            lib().BMI_codestream_close(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def default_format(info: BmiCodestreamInfo):
    """Packed output format and check_out for a codestream's components:
    RCT / ICT to BGR when the stream signals a component transform, the
    merge formats otherwise (a single component as grey in every slot)."""
    wide = info.bit_depth > 8
    if info.ncomp >= 3 and info.mct and info.ncomp == 3:
        return (FMT_48BITS_RGB if wide else FMT_24BITS_RGB), -1
    if info.ncomp == 4:
        return (FMT_64BITS_4COMPONENTS if wide else FMT_32BITS_4COMPONENTS), -1
    if info.ncomp == 1:
        return (FMT_48BITS_3COMPONENTS if wide else FMT_24BITS_3COMPONENTS), 0
    return (FMT_48BITS_3COMPONENTS if wide else FMT_24BITS_3COMPONENTS), -1


def _check_block_table(blocks, nb: int, plane_ptrs, pitch: int, rows: int, seg_base: int,
                       seg_bytes: int, segs=None) -> None:
    """Host-side bounds check of a tile's block table before it reaches
    the GPU: every block's output rectangle lies inside its tile plane,
    every segment inside the tile's share of the segment buffer, and
    (given the host copy `segs` of that share) every segment-length table
    inside it, aligned, and adding up to the block's bytes."""
    if nb:
        t1_check_blocks([blocks[i] for i in range(nb)])
    for i in range(nb):
        b = blocks[i]
        c = next((k for k, p in enumerate(plane_ptrs) if p <= b.out < p + 4 * pitch * rows), -1)
        if c < 0:
            raise BmiError(f"block {i}: output outside the tile planes")
        y, x = divmod((b.out - plane_ptrs[c]) // 4, pitch)
        if x + b.width > pitch or y + b.height > rows or b.out_pitch != pitch:
            raise BmiError(f"block {i}: output rectangle outside the tile plane")
        if b.data_offset < seg_base or b.data_offset + b.len > seg_base + seg_bytes:
            raise BmiError(f"block {i}: segment outside the tile's data")
        if b.style & (T1_BYPASS | T1_TERMALL) and b.nseg:
            t0 = b.seg_offset - seg_base
            if t0 < 0 or t0 % 4 or t0 + 4 * b.nseg > seg_bytes:
                raise BmiError(f"block {i}: segment-length table outside the tile's data")
            if segs is not None:
                import numpy as np
                lens = np.frombuffer(bytes(segs[t0:t0 + 4 * b.nseg]), dtype=np.int32)
                if (lens < 0).any() or int(lens.sum()) != b.len:
                    raise BmiError(f"block {i}: segment lengths do not add up to the block's bytes")


class CodestreamDecoder:
    """Decode a whole codestream on the GPU: host tier-2, one tier-1 launch
    over every code-block of every tile, then one batched IDWT / colour plan
    over all tiles writing into one output image.  Device buffers are torch
    tensors kept for the life of the object, so run() can be repeated
    (e.g. timed) on the resident inputs."""

    def __init__(self, data: bytes, *, flags: int = CS_REFERENCE, out_format=None, check_out=None,
                 r53_mode: int = R53_REFERENCE, top_level: int = -1, ignore_mask: int = 0,
                 device: str = "cuda"):
        """top_level: synthesis levels to run (-1 all; k < levels gives the
        image at 2^-(levels-k) of full size, 0 the LL thumbnail, the
        reference's mDwtDecodeTopLevel / GetThumbnail).  Tier 1 then decodes
        only the code-blocks of the bands that resolution needs.
        ignore_mask: components not to decode (bit c = component c, the
        reference's mIgnoreComponents); their code-blocks are not decoded
        either, and the remaining components fill the output slots in
        order."""
        import numpy as np
        import torch
        self.cs = cs = Codestream(data, flags)
        info = cs.info
        fmt, co = default_format(info)
        self.out_format = fmt if out_format is None else out_format
        self.check_out = co if check_out is None else check_out
        reduce = 0 <= top_level < info.levels
        shift = info.levels - top_level if reduce else 0

        def cd(a):
            return -(-a // (1 << shift))
        W, H = cd(info.x0 + info.width) - cd(info.x0), cd(info.y0 + info.height) - cd(info.y0)
        self.width, self.height = W, H
        self.planes, descs, tables, seg_parts = [], [], [], []
        self.out_planes = []
        if self.out_format == FMT_NONE:
            for c in range(info.ncomp):
                self.out_planes.append(torch.zeros((H, W), dtype=torch.int32, device=device))
            self.out = None
        else:
            self.out = torch.zeros((H, W * PIXEL_BYTES[self.out_format]), dtype=torch.uint8,
                                   device=device)
        seg_base = 0
        nkept = 0
        for t in range(cs.ntiles):
            x0, y0, w, h, nb, sb = cs.tile_info(t)
            pitch = (w + 63) // 64 * 64
            pl = torch.zeros((info.ncomp, h, pitch), dtype=torch.int32, device=device)
            self.planes.append(pl)
            ptrs = [pl[c].data_ptr() for c in range(info.ncomp)]
            blocks, segs, desc = cs.tile_blocks(t, ptrs, pitch, seg_base)
            if reduce and desc.levels != info.levels:
                raise BmiError("reduced-resolution decode of tiles with different level counts (-5)")
            desc.out_format = self.out_format
            desc.check_out = self.check_out
            desc.r53_mode = r53_mode
            desc.top_level = top_level if reduce else -1
            desc.ignore_mask = ignore_mask
            ox, oy = cd(x0) - cd(info.x0), cd(y0) - cd(info.y0)
            if self.out is not None:
                desc.out = self.out.data_ptr()
                desc.out_pitch_px = W
                desc.out_x, desc.out_y = ox, oy
            else:
                for c in range(info.ncomp):
                    desc.planes_out[c] = self.out_planes[c][oy:, ox:].data_ptr()
                desc.planes_pitch = W
            descs.append(desc)
            _check_block_table(blocks, nb, ptrs, pitch, h, seg_base, sb, segs)
            size = ctypes.sizeof(BmiT1Block)
            if (reduce and desc.levels > 0) or ignore_mask:
                # keep the blocks of the components decoded, inside the
                # Mallat region of the target resolution: the LL band and
                # the detail bands of the levels that are synthesised
                if reduce and desc.levels > 0:
                    rw, rh = cd(x0 + w) - cd(x0), cd(y0 + h) - cd(y0)
                else:
                    rw, rh = w, h
                raw = bytes(blocks)
                keep = []
                for i in range(nb):
                    b = blocks[i]
                    c = next(k for k, p in enumerate(ptrs) if p <= b.out < p + 4 * pitch * h)
                    if (ignore_mask >> c) & 1:
                        continue
                    y, x = divmod((b.out - ptrs[c]) // 4, pitch)
                    if x < rw and y < rh:
                        keep.append(raw[i * size:(i + 1) * size])
                tables.append(b"".join(keep))
                nkept += len(keep)
            else:
                tables.append(bytes(blocks)[:nb * size])
                nkept += nb
            # each tile's share starts 16-byte aligned (its segment-length
            # tables are read as aligned words)
            pad = -sb % 16
            seg_parts.append(segs[:sb].tobytes() + b"\xff" * pad)
            seg_base += sb + pad
        self.nblocks = nkept
        seg_all = b"".join(seg_parts) + b"\xff" * 16
        self.seg_bytes = seg_base
        self.d_segs = torch.from_numpy(np.frombuffer(seg_all, np.uint8).copy()).to(device)
        tab = b"".join(tables)
        if tab:
            # longest code-block data first: a shorter tail of the launch
            arr = (BmiT1Block * (len(tab) // ctypes.sizeof(BmiT1Block))).from_buffer_copy(tab)
            t1_order_blocks(arr)
            tab = bytes(arr)
        self.d_blocks = torch.from_numpy(np.frombuffer(tab or b"\0", np.uint8).copy()).to(device)
        self.descs = descs
        # tiles without decomposition levels have no synthesis: the plain
        # sequence (colour / merge of the coefficient planes) serves them;
        # for 9/7 the dequantisation happens in the synthesis row pass, so
        # such tiles are refused (as the reference cannot decode them)
        fused = [d for d in descs if d.levels > 0]
        self.plain = [(t, d) for t, d in enumerate(descs) if d.levels == 0]
        if any(d.method == W9X7 for _, d in self.plain):
            raise BmiError("9/7 tile without decomposition levels: unsupported (-5)")
        self.plan = BatchPlan(fused) if fused else None
        self.temp = None
        if self.plain:
            n = max(d.width * d.height for _, d in self.plain) + 64
            self.temp = torch.zeros(n, dtype=torch.int32, device=device)

    def run_t1(self, stream: int = 0) -> None:
        if self.nblocks:
            t1_decode_blocks(self.d_blocks.data_ptr(), self.nblocks, self.d_segs.data_ptr(), stream)

    def run(self, stream: int = 0) -> None:
        """Tier 1 then IDWT / colour, asynchronous on `stream`."""
        self.run_t1(stream)
        if self.plan is not None:
            self.plan.execute(stream)
        for t, d in self.plain:
            plain_decode(d, self.temp.data_ptr(), self.temp.numel() * 4, stream)
            if self.out is None:
                # the plain sequence leaves planes in the coefficient buffers
                ox, oy = d.x0 - self.cs.info.x0, d.y0 - self.cs.info.y0
                for c in range(d.ncomp):
                    self.out_planes[c][oy:oy + d.height, ox:ox + d.width].copy_(
                        self.planes[t][c, :d.height, :d.width])

    def close(self) -> None:
        if self.plan is not None:
            self.plan.close()
        self.cs.close()


def decode_codestream(data: bytes, **kw):
    """Decode a .j2k / .jp2 on the GPU: returns the packed image (uint8
    tensor, rows of width * pixel bytes) or, with out_format=FMT_NONE, the
    list of component planes (int32 bits; float32 for 9/7)."""
    import torch
    dec = CodestreamDecoder(data, **kw)
    dec.run(0)
    torch.cuda.synchronize()
    out = dec.out if dec.out is not None else dec.out_planes
    dec.close()
    return out
