"""Helpers for the GPU tests: run the HIP back end (plain and fused paths)
on numpy coefficient planes through the C-ABI and return host arrays."""
from __future__ import annotations

import numpy as np


def torch_mod():
    import torch
    return torch


def to_device(arr: np.ndarray):
    """numpy 32-bit array -> contiguous int32 CUDA tensor with the same bits."""
    torch = torch_mod()
    a = np.ascontiguousarray(arr)
    assert a.dtype.itemsize == 4
    return torch.from_numpy(a.view(np.int32).copy()).to("cuda")


def out_buffer(fmt_bytes: int, w: int, h: int):
    torch = torch_mod()
    return torch.zeros((h, w * fmt_bytes), dtype=torch.uint8, device="cuda")


def make_case(pkg, planes: np.ndarray, *, filt: str, levels: int, out_format: int,
              bit_depth: int = 8, check_out: int = -1, x0: int = 0, y0: int = 0,
              top_level: int = -1, coef_kind: int = 1, r53_mode: int = 0, steps=None,
              pitch_pad: int = 0):
    """Device buffers + BmiTileDesc for one tile.  Returns (desc, keep) where
    keep holds the tensors (coef, out, planes)."""
    torch = torch_mod()
    ncomp, h, w = planes.shape
    pitch = w + pitch_pad
    coefs = []
    for c in range(ncomp):
        buf = np.zeros((h, pitch), dtype=planes.dtype)
        buf[:, :w] = planes[c]
        coefs.append(to_device(buf))
    ow, oh = pkg.output_size(w, h, x0, y0, levels, top_level)
    out = None
    planes_out = []
    if out_format == pkg.FMT_NONE:
        for c in range(ncomp):
            planes_out.append(torch.zeros((oh, ow), dtype=torch.int32, device="cuda"))
    else:
        out = out_buffer(pkg.PIXEL_BYTES[out_format], ow, oh)
    desc = pkg.tile_desc(
        coef_ptrs=[t.data_ptr() for t in coefs], coef_pitch=pitch, width=w, height=h, x0=x0, y0=y0,
        levels=levels, method=pkg.W5X3 if filt == "53" else pkg.W9X7, ncomp=ncomp,
        coef_kind=coef_kind, bit_depth=bit_depth, out_format=out_format, check_out=check_out,
        r53_mode=r53_mode, steps=steps, out_ptr=(out.data_ptr() if out is not None else 0),
        out_pitch_px=ow, planes_ptrs=[p.data_ptr() for p in planes_out], planes_pitch=ow,
        top_level=top_level)
    keep = dict(coefs=coefs, out=out, planes=planes_out, ow=ow, oh=oh, pitch=pitch, w=w, h=h)
    return desc, keep


def fetch(keep, filt: str):
    """Host copies of the results: (planes list or None, packed or None)."""
    torch = torch_mod()
    torch.cuda.synchronize()
    planes = None
    if keep["planes"]:
        planes = [p.cpu().numpy().view(np.float32 if filt == "97" else np.int32) for p in keep["planes"]]
    packed = keep["out"].cpu().numpy() if keep["out"] is not None else None
    return planes, packed


def run_fused(pkg, planes, tuning=None, **kw):
    desc, keep = make_case(pkg, planes, **kw)
    plan = pkg.BatchPlan([desc], tuning)
    plan.execute(0)
    res = fetch(keep, kw["filt"])
    assert plan.error_flag() == 0
    plan.close()
    return res


def run_plain(pkg, planes, **kw):
    torch = torch_mod()
    desc, keep = make_case(pkg, planes, **kw)
    w, h = keep["w"], keep["h"]
    temp = torch.zeros(w * h + 64, dtype=torch.int32, device="cuda")
    pkg.plain_decode(desc, temp.data_ptr(), temp.numel() * 4, 0)
    torch.cuda.synchronize()
    planes_out, packed = fetch(keep, kw["filt"])
    if planes_out is None and kw["out_format"] == pkg.FMT_NONE:
        pass
    if kw["out_format"] == pkg.FMT_NONE:
        # the plain path works in place: results are in the coefficient planes
        ow, oh = keep["ow"], keep["oh"]
        dt = np.float32 if kw["filt"] == "97" else np.int32
        planes_out = [c.cpu().numpy().view(dt)[:oh, :ow] for c in keep["coefs"]]
    return planes_out, packed


def diff_report(a: np.ndarray, b: np.ndarray) -> str:
    a64, b64 = a.astype(np.float64), b.astype(np.float64)
    d = np.abs(a64 - b64)
    idx = np.unravel_index(np.argmax(d), d.shape) if d.size else ()
    return f"max|d|={d.max() if d.size else 0} at {idx}, n_diff={(d > 0).sum()} of {d.size}"
