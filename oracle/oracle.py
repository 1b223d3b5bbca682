"""ctypes wrapper of the CPU oracle (oracle/libj2k_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product.  Each function restates
the reference CPU back end (see j2k_oracle.c for the file:line of every
function it follows).
"""
from __future__ import annotations

import ctypes
import pathlib
import subprocess

import numpy as np

ORACLE_DIR = pathlib.Path(__file__).resolve().parent
LIB_PATH = ORACLE_DIR / "libj2k_oracle.so"

R53_REFERENCE = 0
R53_TEXTBOOK = 1
COEF_SIGNMAG = 0
COEF_FLOAT = 1

_LIB = None


def build() -> None:
    subprocess.run(["make", "-B", "-C", str(ORACLE_DIR)], check=True,
                   stdout=subprocess.DEVNULL)


def lib() -> ctypes.CDLL:
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(LIB_PATH))
        P = ctypes.c_void_p
        i = ctypes.c_int
        L.or_idwt53.argtypes = [P, i, i, i, i, i, i, i, i]
        L.or_idwt97.argtypes = [P, i, i, i, i, i, i, i, P, i]
        L.or_idwt97_f64.argtypes = [P, i, i, i, i, i, i, P, i, P]
        L.or_fdwt97_f64.argtypes = [P, i, i, i, i, i, P]
        L.or_fdwt53.argtypes = [P, i, i, i, i, i, P]
        L.or_53_line_textbook.argtypes = [P, i, i]
        L.or_97_line.argtypes = [P, i, i]
        L.or_h53_line_ref.argtypes = [P, P, P, i, i]
        L.or_v53_ref.argtypes = [P, P, i, i, i]
        L.or_rct_pack.argtypes = [P, i, P, P, P, i, i, i, i, i, i]
        L.or_ict_pack.argtypes = [P, i, P, P, P, i, i, i, i, i, i]
        L.or_merge_pack.argtypes = [P, i, P, i, i, i, i, i, i, i, i]
        L.or_dc_shift.argtypes = [P, i, P, i, i, i, i, i]
        L.or_one_component.argtypes = [P, i, P, i, i, i, i, i, i]
        L.or_ict_float.argtypes = [P, P, P, P, P, P, i, i]
        L.or_pixel_bytes.argtypes = [i]
        L.or_t1_decode_block.argtypes = [P, ctypes.c_long, i, i, i, i, i, i, P, i]
        L.or_t1_decode_block_ex.argtypes = [P, ctypes.c_long, i, i, i, i, i, i, i, P, i, P, i]
        L.or_t1_num_segments.argtypes = [i, i]
        L.or_t1_words_to_int.argtypes = [P, i, i, i, i, P, i]
        L.or_t1_tables.argtypes = [P, P, P, P]
        L.or_t1_zc.argtypes = [P, i]
        L.or_t1_sc.argtypes = [P]
        _LIB = L
    return _LIB


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def idwt53(plane: np.ndarray, levels: int, *, x0=0, y0=0, toplevel=-1, mode=R53_REFERENCE):
    """In-place-style 5/3 IDWT of one tile-component; returns a new array."""
    a = np.ascontiguousarray(plane, dtype=np.int32).copy()
    h, w = a.shape
    rc = lib().or_idwt53(_p(a), w, w, h, x0, y0, levels, toplevel, mode)
    assert rc == 0, rc
    return a


def idwt97(words: np.ndarray, levels: int, steps=None, *, x0=0, y0=0, toplevel=-1,
           coef_kind=COEF_FLOAT):
    """9/7 IDWT; words is a float32 or uint32 (sign-magnitude) plane."""
    a = np.ascontiguousarray(words).view(np.uint32).copy()
    h, w = a.shape
    st = np.ones(3 * max(levels, 1) + 1, np.float32) if steps is None else np.asarray(steps, np.float32)
    rc = lib().or_idwt97(_p(a), w, w, h, x0, y0, levels, toplevel, _p(st), coef_kind)
    assert rc == 0, rc
    return a.view(np.float32)


def idwt97_f64(words: np.ndarray, levels: int, steps=None, *, x0=0, y0=0, coef_kind=COEF_FLOAT):
    a = np.ascontiguousarray(words).view(np.uint32)
    h, w = a.shape
    st = np.ones(3 * max(levels, 1) + 1, np.float32) if steps is None else np.asarray(steps, np.float32)
    out = np.zeros((h, w), np.float64)
    rc = lib().or_idwt97_f64(_p(a), w, w, h, x0, y0, levels, _p(st), coef_kind, _p(out))
    assert rc == 0, rc
    return out


def fdwt53(img: np.ndarray, levels: int, *, x0=0, y0=0):
    a = np.ascontiguousarray(img, dtype=np.int32)
    h, w = a.shape
    out = np.zeros_like(a)
    assert lib().or_fdwt53(_p(a), w, h, x0, y0, levels, _p(out)) == 0
    return out


def fdwt97_f64(img: np.ndarray, levels: int, *, x0=0, y0=0):
    a = np.ascontiguousarray(img, dtype=np.float64)
    h, w = a.shape
    out = np.zeros_like(a)
    assert lib().or_fdwt97_f64(_p(a), w, h, x0, y0, levels, _p(out)) == 0
    return out


def line53_textbook(x: np.ndarray, parity: int):
    a = np.ascontiguousarray(x, dtype=np.int32).copy()
    lib().or_53_line_textbook(_p(a), a.size, parity)
    return a


def line97(x: np.ndarray, parity: int):
    a = np.ascontiguousarray(x, dtype=np.float32).copy()
    lib().or_97_line(_p(a), a.size, parity)
    return a


def v53_ref(hres: np.ndarray, lh: int):
    a = np.ascontiguousarray(hres, dtype=np.int32)
    n, w = a.shape
    out = np.zeros_like(a)
    lib().or_v53_ref(_p(out), _p(a), lh, n - lh, w)
    return out


PIXEL_BYTES = {0: 4, 1: 6, 2: 3, 3: 4, 4: 8, 5: 6, 6: 3, 8: 8}


def rct_pack(c0, c1, c2, bit_depth: int, fmt: int, check_out: int = -1):
    c = [np.ascontiguousarray(x, dtype=np.int32) for x in (c0, c1, c2)]
    h, w = c[0].shape
    out = np.zeros((h, w * PIXEL_BYTES[fmt]), np.uint8)
    rc = lib().or_rct_pack(_p(out), w, _p(c[0]), _p(c[1]), _p(c[2]), w, w, h, bit_depth, fmt,
                           check_out)
    assert rc == 0, rc
    return out


def ict_pack(c0, c1, c2, bit_depth: int, fmt: int, check_out: int = -1):
    c = [np.ascontiguousarray(x, dtype=np.float32) for x in (c0, c1, c2)]
    h, w = c[0].shape
    out = np.zeros((h, w * PIXEL_BYTES[fmt]), np.uint8)
    rc = lib().or_ict_pack(_p(out), w, _p(c[0]), _p(c[1]), _p(c[2]), w, w, h, bit_depth, fmt,
                           check_out)
    assert rc == 0, rc
    return out


def ict_float(c0, c1, c2, bit_depth: int):
    c = [np.ascontiguousarray(x, dtype=np.float32) for x in (c0, c1, c2)]
    r, g, b = (np.zeros_like(c[0]) for _ in range(3))
    lib().or_ict_float(_p(r), _p(g), _p(b), _p(c[0]), _p(c[1]), _p(c[2]), c[0].size, bit_depth)
    return r, g, b


def merge_pack(comps, bit_depth: int, fmt: int, check_out: int = -1, is_float: bool = False):
    dt = np.float32 if is_float else np.int32
    arrs = [None if x is None else np.ascontiguousarray(x, dtype=dt) for x in comps]
    first = next(a for a in arrs if a is not None)
    h, w = first.shape
    ptrs = (ctypes.c_void_p * 4)(*[None if a is None else a.ctypes.data for a in arrs] +
                                 [None] * (4 - len(arrs)))
    out = np.zeros((h, w * PIXEL_BYTES[fmt]), np.uint8)
    rc = lib().or_merge_pack(_p(out), w, ptrs, len(arrs), w, w, h, bit_depth, fmt, check_out,
                             int(is_float))
    assert rc == 0, rc
    return out


def dc_shift(plane: np.ndarray, bit_depth: int, is_float: bool):
    a = np.ascontiguousarray(plane).view(np.uint32)
    h, w = a.shape
    out = np.zeros((h, w), np.int32)
    lib().or_dc_shift(_p(out), w, _p(a), w, w, h, bit_depth, int(is_float))
    return out


def one_component(words: np.ndarray, mode: int, comp_id: int, bit_depth: int) -> np.ndarray:
    """Grey preview words of BMI_tile_One_Component: mode 0 from packed
    32-bit result words, 1 from a 5/3 plane, 2 the 9/7 fixed-point kernel."""
    a = np.ascontiguousarray(words).view(np.int32)
    h, w = a.shape
    out = np.zeros((h, w), np.int32)
    lib().or_one_component(_p(out), w, _p(a), w, w, h, mode, comp_id, bit_depth)
    return out


def ignore_components(planes, steps, check_out: int, out_format: int, ignore_mask: int):
    """mIgnoreComponents (codec_base.h:215-228): the components whose bit is
    set are not decoded (InitPicture skips them, cudainterface.cpp:545-560,
    648-653); the remaining ones fill the component slots in order
    (bmi_cpu_dwt_decoder.cpp:1180-1290), and a check-out component that is
    ignored means no check-out (:1222-1225).  Returns (planes, steps,
    check_out) of the remaining components; RGB formats keep their output
    channel check-out."""
    keep = [c for c in range(planes.shape[0]) if not (ignore_mask >> c) & 1]
    if not keep:
        raise ValueError("every component ignored")
    rgb = out_format in (0, 1, 2, 8)
    if not rgb and check_out >= 0:
        check_out = keep.index(check_out) if check_out in keep else -1
    st = None if steps is None else [steps[c] for c in keep]
    return planes[keep], st, check_out


def decode_tile(planes: np.ndarray, *, levels: int, filt: str, out_format: int, bit_depth: int = 8,
                check_out: int = -1, x0=0, y0=0, toplevel=-1, steps=None,
                coef_kind=COEF_FLOAT, r53_mode=R53_REFERENCE, ignore_mask: int = 0):
    """Reference CPU pipeline for one tile: IDWT per component, then
    RCT/ICT (RGB formats), merge (*_COMPONENTS) or the planes (NONE).
    Returns (recon planes list, packed output or None)."""
    if ignore_mask:
        planes, steps, check_out = ignore_components(planes, steps, check_out, out_format,
                                                     ignore_mask)
    ncomp = planes.shape[0]
    rgb = out_format in (0, 1, 2, 8)
    recon = []
    for c in range(ncomp):
        if filt == "53":
            recon.append(idwt53(planes[c], levels, x0=x0, y0=y0, toplevel=toplevel, mode=r53_mode))
        else:
            st = None if steps is None else steps[c]
            recon.append(idwt97(planes[c], levels, st, x0=x0, y0=y0, toplevel=toplevel,
                                coef_kind=coef_kind))
    if toplevel >= 0 and toplevel < levels:
        s = levels - toplevel
        a0, a1 = -(-x0 // (1 << s)), -(-(x0 + planes.shape[2]) // (1 << s))
        b0, b1 = -(-y0 // (1 << s)), -(-(y0 + planes.shape[1]) // (1 << s))
        recon = [r[:b1 - b0, :a1 - a0] for r in recon]
    if out_format == 7:
        return recon, None
    if rgb:
        fn = ict_pack if filt == "97" else rct_pack
        return recon, fn(recon[0], recon[1], recon[2], bit_depth, out_format, check_out)
    if check_out >= 0:
        comps = [recon[check_out]]
        co = 0
    else:
        comps = recon
        co = -1
    return recon, merge_pack(comps, bit_depth, out_format, co, is_float=(filt == "97"))


# ------------------------------------------------------------------- T1 --

def t1_decode_block(data: bytes, w: int, h: int, orient: int, numbps: int, missing: int,
                    num_pass: int) -> np.ndarray:
    """decode_code_block: T1 words (int32, sign bit 31) of one code-block."""
    buf = np.frombuffer(bytes(data) + b"\0", dtype=np.uint8)
    out = np.zeros((h, w), dtype=np.int32)
    rc = lib().or_t1_decode_block(_p(buf), len(data), w, h, orient, numbps, missing, num_pass,
                                  _p(out), w)
    assert rc == 0, rc
    return out


def t1_decode_block_ex(data: bytes, w: int, h: int, orient: int, numbps: int, missing: int,
                       num_pass: int, style: int, seg_lens=()) -> np.ndarray:
    """decode_code_block with code-block style switches: `data` holds the
    codeword segments of seg_lens (one segment of all bytes unless the
    style terminates passes)."""
    buf = np.frombuffer(bytes(data) + b"\0", dtype=np.uint8)
    segs = np.asarray(list(seg_lens) + [0], dtype=np.int32)
    out = np.zeros((h, w), dtype=np.int32)
    rc = lib().or_t1_decode_block_ex(_p(buf), len(data), w, h, orient, numbps, missing, num_pass,
                                     style, _p(segs), len(seg_lens), _p(out), w)
    if rc != 0:
        raise ValueError(f"or_t1_decode_block_ex rejected the block ({rc})")
    return out


def t1_num_segments(num_pass: int, style: int) -> int:
    return lib().or_t1_num_segments(num_pass, style)


def t1_words_to_int(words: np.ndarray, numbps: int) -> np.ndarray:
    a = np.ascontiguousarray(words, dtype=np.int32)
    h, w = a.shape
    out = np.empty_like(a)
    lib().or_t1_words_to_int(_p(a), w, h, w, numbps, _p(out), w)
    return out


def t1_tables():
    qe = np.zeros(47, np.uint16)
    nm, nl, sw = (np.zeros(47, np.uint8) for _ in range(3))
    lib().or_t1_tables(_p(qe), _p(nm), _p(nl), _p(sw))
    return qe, nm, nl, sw


def t1_zc(nb9, orient: int) -> int:
    a = np.asarray(nb9, dtype=np.uint8)
    return lib().or_t1_zc(_p(a), orient)


def t1_sc(nb9) -> int:
    a = np.asarray(nb9, dtype=np.uint8)
    return lib().or_t1_sc(_p(a))


# ------------------------------------------------- multithreaded CPU baseline --

BASELINE_PATH = ORACLE_DIR / "libcpu_baseline.so"
_BASE = None


def baseline_lib() -> ctypes.CDLL:
    """oracle/libcpu_baseline.so: the multithreaded CPU restatement of the
    reference CPU back end's 9/7 + ICT pipeline (timed as the CPU baseline)."""
    global _BASE
    if _BASE is None:
        if not BASELINE_PATH.exists():
            build()
        L = ctypes.CDLL(str(BASELINE_PATH))
        P = ctypes.c_void_p
        i = ctypes.c_int
        L.cpu_decode97_ict.argtypes = [P, i, i, i, i, i, P, P, P, i, i]
        L.cpu_decode97_ict.restype = i
        _BASE = L
    return _BASE


def cpu_decode97_ict(planes: np.ndarray, levels: int, threads: int, steps=None, out=None,
                     work=None):
    """Decode float coefficient planes (ncomp, h, w) at tile offset (0, 0):
    9/7 synthesis then (ncomp >= 3) ICT to 24-bit BGR, on `threads` host
    threads.  Returns (planes (ncomp, h, w) float32, bgr (h, 3w) uint8 or
    None).  `out` / `work` may pass preallocated result arrays."""
    a = np.ascontiguousarray(planes, dtype=np.float32)
    nc, h, w = a.shape
    if work is None:
        work = np.empty((nc, h, w), np.float32)
    if out is None and nc >= 3:
        out = np.empty((h, 3 * w), np.uint8)
    cptr = (ctypes.c_void_p * nc)(*[a[c].ctypes.data for c in range(nc)])
    wptr = (ctypes.c_void_p * nc)(*[work[c].ctypes.data for c in range(nc)])
    st = None
    if steps is not None:
        st = np.ascontiguousarray(steps, dtype=np.float32)
    rc = baseline_lib().cpu_decode97_ict(
        cptr, w, w, h, levels, nc, st.ctypes.data if st is not None else None, wptr,
        out.ctypes.data if out is not None else None, 3 * w, threads)
    assert rc == 0, rc
    return work, out
