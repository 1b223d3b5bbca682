"""The drop-in boundary driven the way the reference's CudaManager drives it.

The reference's GPU host thread (decoder/cudainterface/src/cudainterface.cpp)
keeps one picture-wide pitched coefficient buffer and one transposed scratch
buffer per component (MallocGpuBuf(GPU_ALLIGN_BUF), :545-570), uploads every
code-block stripe of every subband to its place in that buffer with
Gpu2DCopy (UploadStripeBuf, :1070-1115: x offset in bytes = (tile + subband
offset) * 4, y = tile + subband offset + stripe row), synchronises, runs
BMI_comp_idwt per tile-component (or BMI_resolution_idwt level by level) in
place at the tile's offset, then BMI_tile_ICT / BMI_tile_MCT /
BMI_tile_Merge into a picture-wide result buffer, and reads single
components back with BMI_Comp_DC_SHIFT into the scratch buffer and
Gpu2DCopyWithDiffPitch (GetOneComponent, :1894-2093), or the grey preview
BMI_tile_One_Component.  The subband tables, parities and steps are built
as InitPicture builds them (:598-720).

Everything here goes through those C entry points with ctypes mirrors of
TileInfo_c / CompInfo_c / SubbandInfo_c / CudaBuf, on a 301x203 picture of
2x2 tiles (161x113 tiles, so two tiles start at odd offsets), and is
compared with the oracle: float and integer planes, packed pixels, the DC
shift and the grey preview, at full and reduced resolution."""
import ctypes

import numpy as np
import pytest

from gpu_helpers import diff_report

pytestmark = pytest.mark.gpu

W, H, TW, TH, L = 301, 203, 161, 113, 3


@pytest.fixture(scope="module")
def L_(pkg):
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    lib = pkg.lib()
    info = pkg.GpuInfo()
    assert lib.GpuInit(ctypes.byref(info)) == 0
    assert info.active == 1 and info.multiProcessorCount > 0 and info.name
    yield lib
    assert lib.GpuRemove(ctypes.byref(info)) == 0


def _ok(lib, what):
    rc = lib.BMI_last_error()
    assert rc == 0, f"{what}: error {rc} ({lib.BMI_error_string(rc).decode()})"


def _malloc(pkg, lib, typ, width_bytes, height):
    b = pkg.CudaBuf()
    b.mType, b.mWidth, b.mHeight = typ, width_bytes, height
    lib.MallocGpuBuf(ctypes.byref(b))
    _ok(lib, "MallocGpuBuf")
    assert b.mBuf
    if typ == pkg.GPU_ALLIGN_BUF:
        assert b.mPitch >= width_bytes and b.mSize == b.mPitch * height
    return b


def _tiles():
    out = []
    for ty in range(0, H, TH):
        for tx in range(0, W, TW):
            out.append((tx, ty, min(TW, W - tx), min(TH, H - ty)))
    return out


class Picture:
    """Host side of one decode: per-component picture buffers, the tile /
    component / subband trees, and the coefficients of every tile."""

    def __init__(self, pkg, lib, workloads, filt, seed, result_fmt_bytes, result_aligned=False,
                 result_size=(W, H)):
        self.pkg, self.lib, self.filt = pkg, lib, filt
        self.orig = [_malloc(pkg, lib, pkg.GPU_ALLIGN_BUF, W * 4, H) for _ in range(3)]
        self.temp = [_malloc(pkg, lib, pkg.GPU_ALLIGN_BUF, H * 4, W) for _ in range(3)]
        rw, rh = result_size
        self.result = _malloc(pkg, lib, pkg.GPU_ALLIGN_BUF if result_aligned else pkg.GPU_BUF,
                              rw * result_fmt_bytes, rh)
        self.stream = lib.CreateANewStream()
        assert self.stream
        self.tiles, self.keep = [], []
        for ti, (x0, y0, tw, th) in enumerate(_tiles()):
            planes = workloads.tile_planes(tw, th, L, filt, 3, seed=seed, tile=ti, x0=x0, y0=y0,
                                           ll53=(-60, 60), ll97=(0.0, 40.0),
                                           signmag=(filt == "97"))
            steps = np.stack([np.linspace(0.5 + 0.1 * c + 0.05 * ti, 1.5, 3 * L + 1)
                              for c in range(3)]).astype(np.float32)
            if filt == "53":
                steps[:] = 1.0
            tile = pkg.TileInfo_c()
            tile.mOff.x, tile.mOff.y, tile.mSize.x, tile.mSize.y = x0, y0, tw, th
            # parities as InitPicture (cudainterface.cpp:607-615)
            xo, yo = x0, y0
            for d in range(L - 1, -1, -1):
                tile.mXParity[d], tile.mYParity[d] = xo & 1, yo & 1
                xo, yo = (xo + 1) // 2, (yo + 1) // 2
            comps = (pkg.CompInfo_c * 3)()
            sbs = []
            for c in range(3):
                sb = (pkg.SubbandInfo_c * (3 * L + 1))()
                for (bid, _dl, bx, by, bw, bh) in workloads.band_rects(tw, th, x0, y0, L):
                    sb[bid].mId = bid
                    sb[bid].mOff.x, sb[bid].mOff.y = bx, by
                    sb[bid].mSize.x, sb[bid].mSize.y = bw, bh
                    sb[bid].mfAbsStepNorm = float(steps[c, bid])
                comps[c].mSubbandInfo = sb
                comps[c].mNumOfSubband = 3 * L + 1
                comps[c].mCompBuf = ctypes.pointer(self.orig[c])
                sbs.append(sb)
            tile.mCompInfo = comps
            tile.mResultBuf = ctypes.pointer(self.result)
            self.tiles.append(dict(x0=x0, y0=y0, w=tw, h=th, planes=planes, steps=steps,
                                   tile=tile))
            self.keep += [comps, sbs]

    def upload_stripes(self, workloads, rows=32):
        """UploadStripeBuf: each subband of each tile-component in stripes of
        `rows` lines, staged in pinned host memory, copied asynchronously
        on the stream with Gpu2DCopy at byte offsets, then GpuSync."""
        pkg, lib = self.pkg, self.lib
        total = sum(t["planes"].nbytes for t in self.tiles)
        self.stage = _malloc(pkg, lib, pkg.HOST_BUF, total, 1)
        base, off = self.stage.mBuf, 0
        for t in self.tiles:
            for c in range(3):
                for (bid, _dl, bx, by, bw, bh) in workloads.band_rects(t["w"], t["h"], t["x0"],
                                                                       t["y0"], L):
                    for r0 in range(0, bh, rows):
                        n = min(rows, bh - r0)
                        chunk = np.ascontiguousarray(t["planes"][c, by + r0:by + r0 + n, bx:bx + bw])
                        if chunk.size == 0:
                            continue
                        ctypes.memmove(base + off, chunk.ctypes.data, chunk.nbytes)
                        stripe = pkg.CudaBuf()
                        stripe.mType, stripe.mBuf = pkg.HOST_BUF, base + off
                        stripe.mWidth, stripe.mHeight, stripe.mPitch = bw * 4, n, bw * 4
                        self.keep.append(stripe)
                        s = ctypes.c_void_p(self.stream)
                        lib.Gpu2DCopy(ctypes.byref(stripe), ctypes.byref(self.orig[c]), 0, 0,
                                      (t["x0"] + bx) * 4, t["y0"] + by + r0, bw * 4, n,
                                      ctypes.byref(s), pkg.COPY_H2D, 0, 0)
                        _ok(lib, "Gpu2DCopy")
                        off += chunk.nbytes
        s = ctypes.c_void_p(self.stream)
        lib.GpuSync(ctypes.byref(s))
        _ok(lib, "GpuSync")

    def download(self, buf, row_bytes, rows, pitch=None):
        """Gpu2DCopyWithDiffPitch of a device buffer into a packed host
        array (destination without pitch), synchronous."""
        pkg, lib = self.pkg, self.lib
        host = np.zeros(row_bytes * rows, np.uint8)
        dst = pkg.CudaBuf()
        dst.mType, dst.mBuf = pkg.HOST_BUF, host.ctypes.data
        p = pitch if pitch is not None else buf.mPitch
        lib.Gpu2DCopyWithDiffPitch(ctypes.byref(buf), ctypes.byref(dst), 0, 0, 0, 0, p, row_bytes,
                                   row_bytes, rows, None, pkg.COPY_D2H, 0, 1)
        _ok(lib, "Gpu2DCopyWithDiffPitch")
        return host.reshape(rows, row_bytes)

    def idwt(self, method, toplevel=-1, per_level=(1, 2)):
        """BMI_comp_idwt per tile-component; the tiles in per_level run
        BMI_resolution_idwt level by level instead (all levels only)."""
        lib = self.lib
        for ti, t in enumerate(self.tiles):
            for c in range(3):
                args = (ctypes.byref(t["tile"]), ctypes.byref(t["tile"].mCompInfo[c]),
                        ctypes.byref(self.orig[c]), ctypes.byref(self.temp[c]), self.stream, 2,
                        method)
                if ti in per_level and toplevel == -1:
                    for lev in range(L):
                        assert lib.BMI_resolution_idwt(*args, lev) == 0
                else:
                    assert lib.BMI_comp_idwt(*args, toplevel) == 0

    def sync(self):
        s = ctypes.c_void_p(self.stream)
        self.lib.GpuSync(ctypes.byref(s))
        _ok(self.lib, "GpuSync")

    def close(self):
        lib = self.lib
        self.sync()
        for b in self.orig + self.temp + [self.result] + ([self.stage] if hasattr(self, "stage") else []):
            lib.FreeGpuBuf(ctypes.byref(b))
        s = ctypes.c_void_p(self.stream)
        lib.DestroyStream(ctypes.byref(s))


def _expect(orc, pic, t, fmt, toplevel=-1, check_out=-1):
    return orc.decode_tile(t["planes"], levels=L, filt=pic.filt, out_format=fmt, x0=t["x0"],
                           y0=t["y0"], toplevel=toplevel, steps=t["steps"],
                           coef_kind=orc.COEF_SIGNMAG if pic.filt == "97" else 1,
                           check_out=check_out)


def test_abi_struct_mirrors_are_filled_as_the_reference_fills_them(pkg):
    """Sizes of the ctypes mirrors used below (also checked field by field
    against the C header in test_abi.py)."""
    assert ctypes.sizeof(pkg.SubbandInfo_c) == 24
    assert ctypes.sizeof(pkg.TileInfo_c) == 48


def test_97_tiles_in_picture_buffers(pkg, L_, workloads, orc):
    """9/7: sign-magnitude stripes uploaded into the picture buffers, two
    tiles by BMI_comp_idwt and two level by level by BMI_resolution_idwt, in
    place at their offsets (two at odd offsets); float planes bit for bit,
    then BMI_tile_ICT into the 24-bit picture and BMI_Comp_DC_SHIFT of every
    component, against the oracle."""
    lib = L_
    pic = Picture(pkg, lib, workloads, "97", seed=51, result_fmt_bytes=3)
    try:
        pic.upload_stripes(workloads)
        pic.idwt(pkg.W9X7)
        pic.sync()
        got_planes = [pic.download(pic.orig[c], W * 4, H).view(np.uint32) for c in range(3)]
        exp_pix = np.zeros((H, W * 3), np.uint8)
        exp_planes = []
        for t in pic.tiles:
            recon, px = _expect(orc, pic, t, 2)
            exp_planes.append(recon)
            x0, y0, tw, th = t["x0"], t["y0"], t["w"], t["h"]
            for c in range(3):
                g = got_planes[c][y0:y0 + th, x0:x0 + tw]
                assert np.array_equal(g, recon[c].view(np.uint32)), \
                    f"tile {(x0, y0)} comp {c}: {diff_report(g.view(np.float32), recon[c])}"
            exp_pix[y0:y0 + th, 3 * x0:3 * (x0 + tw)] = px
        for t in pic.tiles:
            lib.BMI_tile_ICT(ctypes.byref(t["tile"]), 8, 2, L, -1, pkg.FMT_24BITS_RGB, pic.stream)
            _ok(lib, "BMI_tile_ICT")
        pic.sync()
        got = pic.download(pic.result, W * 3, H)
        assert np.array_equal(got, exp_pix), diff_report(got, exp_pix)
        # GetOneComponentInt: DC shift of each tile into the scratch buffer
        for c in range(3):
            for t in pic.tiles:
                lib.BMI_Comp_DC_SHIFT(ctypes.byref(pic.orig[c]), ctypes.byref(pic.temp[c]),
                                      pkg.W9X7, t["x0"], t["y0"], t["x0"], t["y0"], t["w"],
                                      t["h"], 8, 4, pic.stream)
                _ok(lib, "BMI_Comp_DC_SHIFT")
            pic.sync()
            got = pic.download(pic.temp[c], W * 4, H, pitch=pic.orig[c].mPitch).view(np.int32)
            for t, recon in zip(pic.tiles, exp_planes):
                e = orc.dc_shift(recon[c], 8, True)
                g = got[t["y0"]:t["y0"] + t["h"], t["x0"]:t["x0"] + t["w"]]
                assert np.array_equal(g, e), f"DC shift comp {c}: {diff_report(g, e)}"
    finally:
        pic.close()


def test_53_tiles_mct_merge_preview_and_dc_shift(pkg, L_, workloads, orc):
    """5/3: integer stripes; BMI_comp_idwt / BMI_resolution_idwt in place;
    BMI_tile_MCT to 32-bit ARGB, the grey preview of a component taken from
    that result (hasMCT) and from the planes, BMI_tile_Merge to three 8-bit
    components and BMI_Comp_DC_SHIFT, all against the oracle."""
    lib = L_
    pic = Picture(pkg, lib, workloads, "53", seed=52, result_fmt_bytes=4, result_aligned=True)
    try:
        pic.upload_stripes(workloads, rows=16)
        pic.idwt(pkg.W5X3)
        pic.sync()
        assert pic.result.mPitch == pic.orig[0].mPitch  # the preview indexes it with the plane pitch
        exp_argb = np.zeros((H, W), np.uint32)
        exp_planes = []
        got_planes = [pic.download(pic.orig[c], W * 4, H).view(np.int32) for c in range(3)]
        for t in pic.tiles:
            recon, px = _expect(orc, pic, t, 0)
            exp_planes.append(recon)
            for c in range(3):
                g = got_planes[c][t["y0"]:t["y0"] + t["h"], t["x0"]:t["x0"] + t["w"]]
                assert np.array_equal(g, recon[c]), f"tile {t['x0'], t['y0']} comp {c}"
            exp_argb[t["y0"]:t["y0"] + t["h"], t["x0"]:t["x0"] + t["w"]] = \
                px.view(np.uint32).reshape(t["h"], t["w"])
            lib.BMI_tile_MCT(ctypes.byref(t["tile"]), 8, 2, L, -1, pkg.FMT_32BITS_ARGB, pic.stream)
            _ok(lib, "BMI_tile_MCT")
        pic.sync()
        got = pic.download(pic.result, W * 4, H).view(np.uint32)
        assert np.array_equal(got, exp_argb), diff_report(got, exp_argb)
        # grey preview of component 1 from the packed result, in place
        for t in pic.tiles:
            lib.BMI_tile_One_Component(ctypes.byref(t["tile"]), 1, pkg.W5X3, 1, 8, 2, L, -1,
                                       pic.stream)
            _ok(lib, "BMI_tile_One_Component(hasMCT)")
        pic.sync()
        got = pic.download(pic.result, W * 4, H).view(np.int32)
        e = orc.one_component(exp_argb.view(np.int32), 0, 1, 8)
        assert np.array_equal(got, e), diff_report(got, e)
        # grey preview of components 0 and 2 from the planes
        for comp in (0, 2):
            for t in pic.tiles:
                lib.BMI_tile_One_Component(ctypes.byref(t["tile"]), comp, pkg.W5X3, 0, 8, 2, L, -1,
                                           pic.stream)
                _ok(lib, "BMI_tile_One_Component")
            pic.sync()
            got = pic.download(pic.result, W * 4, H).view(np.int32)
            for t, recon in zip(pic.tiles, exp_planes):
                e = orc.one_component(recon[comp], 1, comp, 8)
                g = got[t["y0"]:t["y0"] + t["h"], t["x0"]:t["x0"] + t["w"]]
                assert np.array_equal(g, e), f"preview comp {comp}: {diff_report(g, e)}"
        # merge of the three planes without colour transform, 24-bit
        merged = _malloc(pkg, lib, pkg.GPU_BUF, W * 3, H)
        try:
            exp = np.zeros((H, W * 3), np.uint8)
            for t in pic.tiles:
                t["tile"].mResultBuf = ctypes.pointer(merged)
                lib.BMI_tile_Merge(ctypes.byref(t["tile"]), 3, L, -1, 8, 2, 0,
                                   pkg.FMT_24BITS_3COMPONENTS, pkg.W5X3, pic.stream)
                _ok(lib, "BMI_tile_Merge")
                _, px = _expect(orc, pic, t, pkg.FMT_24BITS_3COMPONENTS)
                exp[t["y0"]:t["y0"] + t["h"], 3 * t["x0"]:3 * (t["x0"] + t["w"])] = px
            pic.sync()
            got = pic.download(merged, W * 3, H)
            assert np.array_equal(got, exp), diff_report(got, exp)
        finally:
            lib.FreeGpuBuf(ctypes.byref(merged))
        for c in range(3):
            for t in pic.tiles:
                lib.BMI_Comp_DC_SHIFT(ctypes.byref(pic.orig[c]), ctypes.byref(pic.temp[c]),
                                      pkg.W5X3, t["x0"], t["y0"], t["x0"], t["y0"], t["w"],
                                      t["h"], 8, 4, pic.stream)
                _ok(lib, "BMI_Comp_DC_SHIFT")
            pic.sync()
            got = pic.download(pic.temp[c], W * 4, H, pitch=pic.orig[c].mPitch).view(np.int32)
            for t, recon in zip(pic.tiles, exp_planes):
                e = orc.dc_shift(recon[c], 8, False)
                g = got[t["y0"]:t["y0"] + t["h"], t["x0"]:t["x0"] + t["w"]]
                assert np.array_equal(g, e), f"DC shift comp {c}: {diff_report(g, e)}"
    finally:
        pic.close()


@pytest.mark.parametrize("filt", ["97", "53"])
def test_reduced_decode_level(pkg, L_, workloads, orc, filt):
    """mDwtDecodeTopLevel: BMI_comp_idwt stops after 2 of 3 levels; the
    colour transform (BMI_tile_ICT / BMI_tile_MCT) and BMI_tile_Merge at
    decodeLevel 2 write each tile at its halved offset ((x + 1) / 2) of the
    reduced picture; the DC shift reads at the tile offset and writes at the
    halved one, as GetOneComponent does at a reduced level."""
    lib = L_
    rw, rh = (W + 1) // 2, (H + 1) // 2
    pic = Picture(pkg, lib, workloads, filt, seed=53, result_fmt_bytes=3, result_size=(rw, rh))
    method = pkg.W9X7 if filt == "97" else pkg.W5X3
    try:
        pic.upload_stripes(workloads)
        pic.idwt(method, toplevel=2)
        exp = np.zeros((rh, rw * 3), np.uint8)
        exp_m = np.zeros((rh, rw * 3), np.uint8)
        recons = []
        for t in pic.tiles:
            recon, px = _expect(orc, pic, t, 2, toplevel=2)
            _, pm = _expect(orc, pic, t, pkg.FMT_24BITS_3COMPONENTS, toplevel=2)
            recons.append(recon)
            rx, ry = (t["x0"] + 1) // 2, (t["y0"] + 1) // 2
            h, w = px.shape[0], px.shape[1] // 3
            exp[ry:ry + h, 3 * rx:3 * (rx + w)] = px
            exp_m[ry:ry + h, 3 * rx:3 * (rx + w)] = pm
            fn = lib.BMI_tile_ICT if filt == "97" else lib.BMI_tile_MCT
            fn(ctypes.byref(t["tile"]), 8, 2, L, 2, pkg.FMT_24BITS_RGB, pic.stream)
            _ok(lib, "tile colour at decodeLevel 2")
        pic.sync()
        got = pic.download(pic.result, rw * 3, rh)
        assert np.array_equal(got, exp), diff_report(got, exp)
        for t in pic.tiles:
            lib.BMI_tile_Merge(ctypes.byref(t["tile"]), 3, L, 2, 8, 2, 0,
                               pkg.FMT_24BITS_3COMPONENTS, method, pic.stream)
            _ok(lib, "BMI_tile_Merge at decodeLevel 2")
        pic.sync()
        got = pic.download(pic.result, rw * 3, rh)
        assert np.array_equal(got, exp_m), diff_report(got, exp_m)
        c = 1
        for t, recon in zip(pic.tiles, recons):
            h, w = recon[c].shape
            lib.BMI_Comp_DC_SHIFT(ctypes.byref(pic.orig[c]), ctypes.byref(pic.temp[c]), method,
                                  t["x0"], t["y0"], (t["x0"] + 1) // 2, (t["y0"] + 1) // 2, w, h,
                                  8, 4, pic.stream)
            _ok(lib, "BMI_Comp_DC_SHIFT at decodeLevel 2")
        pic.sync()
        got = pic.download(pic.temp[c], rw * 4, rh, pitch=pic.orig[c].mPitch).view(np.int32)
        for t, recon in zip(pic.tiles, recons):
            rx, ry = (t["x0"] + 1) // 2, (t["y0"] + 1) // 2
            h, w = recon[c].shape
            e = orc.dc_shift(recon[c], 8, filt == "97")
            g = got[ry:ry + h, rx:rx + w]
            assert np.array_equal(g, e), diff_report(g, e)
    finally:
        pic.close()


def test_boundary_rejects_regions_outside_buffers(pkg, L_, workloads):
    """A tile, result or DC-shift region reaching past a buffer that
    MallocGpuBuf sized is an error, not an out-of-bounds kernel."""
    lib = L_
    pic = Picture(pkg, lib, workloads, "53", seed=54, result_fmt_bytes=3, result_size=(64, 64))
    try:
        t = pic.tiles[3]  # at (161, 113): its pixels lie outside a 64x64 result
        lib.BMI_tile_MCT(ctypes.byref(t["tile"]), 8, 2, L, -1, pkg.FMT_24BITS_RGB, pic.stream)
        assert lib.BMI_last_error() == pkg.BMI_E_INVALID
        # destination rows (addressed with the plane pitch) past the end of
        # the scratch buffer
        below = pic.temp[0].mSize // pic.orig[0].mPitch
        lib.BMI_Comp_DC_SHIFT(ctypes.byref(pic.orig[0]), ctypes.byref(pic.temp[0]), pkg.W5X3,
                              0, 0, 0, below - 8, 16, 16, 8, 4, pic.stream)
        assert lib.BMI_last_error() == pkg.BMI_E_INVALID
        lib.BMI_Comp_DC_SHIFT(ctypes.byref(pic.orig[0]), ctypes.byref(pic.temp[0]), pkg.W5X3,
                              0, 0, 0, 0, 16, 16, 8, 4, pic.stream)
        assert lib.BMI_last_error() == 0
        big = pkg.TileInfo_c.from_buffer_copy(t["tile"])
        big.mOff.y = H  # the whole tile below the picture buffer
        rc = lib.BMI_comp_idwt(ctypes.byref(big), ctypes.byref(t["tile"].mCompInfo[0]),
                               ctypes.byref(pic.orig[0]), ctypes.byref(pic.temp[0]), pic.stream,
                               2, pkg.W5X3, -1)
        assert rc == pkg.BMI_E_INVALID
    finally:
        pic.close()
