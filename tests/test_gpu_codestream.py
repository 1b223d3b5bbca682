"""GPU decode of JPEG 2000 codestreams end to end: the product's host tier-2
parser (C++), one GPU tier-1 launch over every code-block of every tile and
one batched IDWT / colour plan over all tiles.

  - every fixture (streams written by OpenJPEG, tests/golden/j2k) decodes
    bit-exactly to the oracle's whole-image decode in the packed output
    format (reference 5/3 mode, reference numbps);
  - lossless streams decode back to the source image exactly with the
    T.800 5/3 synthesis (component planes, inverse RCT and DC shift applied
    here);
  - lossy streams with BMI_CS_GUARD_T800 equal the oracle's T.800-placement
    planes bit for bit;
  - streams rewritten with SOP / EPH markers, tile-parts and truncated
    tiles decode as the oracle does; a decoder object re-runs to the same
    output;
  - streams of every code-block style switch (bypass, reset, termall,
    causal, segmentation symbols) decode exactly."""
import pathlib

import numpy as np
import pytest

import j2k_tools as jt
import t2_oracle as t2

J2K = pathlib.Path(__file__).resolve().parent / "golden" / "j2k"
FIXTURES = sorted(p for p in J2K.iterdir() if p.suffix in (".j2k", ".jp2"))
LOSSLESS = [p for p in FIXTURES if "97" not in p.name]
LOSSY = [p for p in FIXTURES if "97" in p.name]


@pytest.fixture(scope="module")
def images():
    with np.load(J2K / "images.npz", allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _gpu_packed(pkg, data, **kw):
    out = pkg.decode_codestream(data, **kw)
    return out.cpu().numpy()


def _gpu_planes(pkg, data, **kw):
    planes = pkg.decode_codestream(data, out_format=pkg.FMT_NONE, **kw)
    return [p.cpu().numpy() for p in planes]


@pytest.mark.gpu
@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: p.stem)
def test_gpu_codestream_matches_oracle(pkg, orc, path):
    import torch
    assert torch.cuda.is_available()
    data = path.read_bytes()
    print(f"case {path.stem}", flush=True)
    cs = pkg.Codestream(data)
    fmt, co = pkg.default_format(cs.info)
    got = _gpu_packed(pkg, data)
    exp = t2.decode_image(data, orc, out_format=fmt, check_out=co)
    assert got.shape == exp.shape
    assert np.array_equal(got, exp), (path.stem, np.count_nonzero(got != exp))


@pytest.mark.gpu
@pytest.mark.parametrize("path", LOSSLESS, ids=lambda p: p.stem)
def test_gpu_lossless_codestream_is_exact(pkg, images, path):
    data = path.read_bytes()
    print(f"case {path.stem}", flush=True)
    cs = pkg.Codestream(data)
    planes = _gpu_planes(pkg, data, r53_mode=pkg.R53_TEXTBOOK)
    p = [x.astype(np.int64) for x in planes]
    if cs.info.mct and cs.info.ncomp >= 3:
        y0, y1, y2 = p[:3]
        g = y0 - ((y1 + y2) >> 2)
        p[:3] = [y2 + g, g, y1 + g]
    got = np.stack([x + (1 << (cs.info.bit_depth - 1)) for x in p], axis=-1)
    src = images[path.stem]
    assert np.array_equal(got, src.reshape(got.shape)), path.stem


@pytest.mark.gpu
@pytest.mark.parametrize("path", LOSSY, ids=lambda p: p.stem)
def test_gpu_codestream_guard_t800(pkg, orc, path):
    data = path.read_bytes()
    print(f"case {path.stem}", flush=True)
    got = _gpu_planes(pkg, data, flags=pkg.CS_GUARD_T800)
    exp = t2.decode_image(data, orc, out_format=7, guard_t800=True)
    for c, (g, e) in enumerate(zip(got, exp)):
        assert np.array_equal(g.view(np.uint32), e.view(np.uint32)), (path.stem, c)


@pytest.mark.gpu
@pytest.mark.parametrize("kw", [dict(sop=True, eph=True, parts=3), dict(parts=2, comment_in_tile=True),
                                dict(keep_packets=5)], ids=["sop_eph_parts", "parts_com", "truncated"])
def test_gpu_codestream_rewritten(pkg, orc, kw):
    for name in ("rgb97_tiles_pcrl.j2k", "grey_layers_lrcp.j2k"):
        data = jt.rewrite((J2K / name).read_bytes(), **kw)
        print(f"case {name} {kw}", flush=True)
        cs = pkg.Codestream(data)
        fmt, co = pkg.default_format(cs.info)
        got = _gpu_packed(pkg, data)
        exp = t2.decode_image(data, orc, out_format=fmt, check_out=co)
        assert np.array_equal(got, exp), (name, kw)


@pytest.mark.gpu
def test_gpu_codestream_decoder_reruns(pkg):
    import torch
    data = (J2K / "rgb97_256.j2k").read_bytes()
    dec = pkg.CodestreamDecoder(data)
    dec.run(0)
    torch.cuda.synchronize()
    first = dec.out.cpu().numpy().copy()
    dec.out.zero_()
    for _ in range(3):
        dec.run(0)
    torch.cuda.synchronize()
    assert np.array_equal(dec.out.cpu().numpy(), first)
    assert dec.plan is not None and dec.plan.error_flag() == 0
    dec.close()


@pytest.mark.gpu
def test_gpu_own_writer_codestream_equals_t1_path(pkg, orc, workloads):
    """The benchmark's codestream (tier-1 workload wrapped by
    workloads.t1_codestream) decodes to the oracle's result on its blocks."""
    img = workloads.t1_image(512, 384, 5, 3, seed=5, filt="97")
    data = workloads.t1_codestream(img)
    got = _gpu_packed(pkg, data, out_format=pkg.FMT_24BITS_RGB)
    exp = t2.decode_image(data, orc, out_format=pkg.FMT_24BITS_RGB)
    assert np.array_equal(got, exp)
    planes = np.zeros((3, 384, 512), np.int32)
    for b in img["blocks"]:
        words = orc.t1_decode_block(img["data"][b["offset"]:b["offset"] + b["len"]], b["w"], b["h"],
                                    b["orient"], b["numbps"], b["missing"], b["num_pass"])
        planes[b["comp"], b["y"]:b["y"] + b["h"], b["x"]:b["x"] + b["w"]] = words
    _, packed = orc.decode_tile(planes, levels=5, filt="97", out_format=pkg.FMT_24BITS_RGB,
                                steps=img["steps"], coef_kind=orc.COEF_SIGNMAG)
    assert np.array_equal(got, packed)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["rgb97_tiles_pcrl.j2k", "rgb_tiles_offsets.j2k", "grey_layers_lrcp.j2k",
                                  "rgb97_256.j2k"])
@pytest.mark.parametrize("top", [0, 1, 3])
def test_gpu_codestream_reduced_resolution(pkg, orc, name, top):
    """Reduced-resolution and thumbnail decodes of whole files (only the
    code-blocks of the needed bands go through tier 1) equal the oracle's
    decode at that resolution."""
    import torch
    data = (J2K / name).read_bytes()
    print(f"case {name} top {top}", flush=True)
    dec = pkg.CodestreamDecoder(data, top_level=top)
    dec.run(0)
    torch.cuda.synchronize()
    got = dec.out.cpu().numpy()
    assert dec.nblocks < dec.cs.info.total_blocks
    exp = t2.decode_image(data, orc, out_format=dec.out_format, check_out=dec.check_out, toplevel=top)
    assert got.shape == exp.shape
    assert np.array_equal(got, exp), (name, top)
    dec.close()


@pytest.mark.gpu
@pytest.mark.parametrize("style", [0x01, 0x04, 0x05, 0x08, 0x22, 0x3F], ids=lambda s: f"style{s:#04x}")
def test_gpu_style_codestreams(pkg, orc, workloads, style):
    """Streams of every code-block style switch from the project's encoder
    (OpenJPEG decodes them to the source: tests/test_codestream.py): the
    GPU decode is exact against the source with the T.800 5/3 synthesis and
    equals the oracle's packed decode in the reference mode; three layers,
    so codeword segments span packets."""
    rng = np.random.default_rng(style)
    for (h, w, nc, L, cb) in [(70, 90, 3, 3, 32), (130, 67, 1, 4, 64)]:
        src = rng.integers(0, 256, (h, w, nc), dtype=np.uint8)
        if nc == 1:
            src = src[..., 0]
        img = workloads.lossless_image(src, L, cb=cb, style=style)
        data = workloads.t1_codestream(img, cb=cb, layers=3)
        print(f"case style {style:#x} {h}x{w}x{nc}", flush=True)
        planes = _gpu_planes(pkg, data, r53_mode=pkg.R53_TEXTBOOK)
        p = [x.astype(np.int64) for x in planes]
        if nc == 3:
            y0, y1, y2 = p
            g = y0 - ((y1 + y2) >> 2)
            p = [y2 + g, g, y1 + g]
        got = np.stack([x + 128 for x in p], axis=-1)
        assert np.array_equal(got, src.reshape(got.shape)), (style, h, w)
        cs = pkg.Codestream(data)
        fmt, co = pkg.default_format(cs.info)
        packed = _gpu_packed(pkg, data)
        exp = t2.decode_image(data, orc, out_format=fmt, check_out=co)
        assert np.array_equal(packed, exp), (style, h, w)
