"""GPU parity: the HIP back end (plain drop-in path and fused path) against
the CPU oracle on the same seeded inputs.  5/3 and every packed output are
compared bit for bit; 9/7 float planes are compared bit for bit as well
(the kernels use the oracle's exact operation order), which is stricter
than the 1-LSB bound on 8/12-bit samples."""
import numpy as np
import pytest

from gpu_helpers import diff_report, make_case, run_fused, run_plain

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu(pkg):
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    pkg.lib()


SHAPES = [
    # (w, h, levels, x0, y0)
    (1, 1, 1, 0, 0),
    (2, 2, 1, 0, 0),
    (3, 5, 2, 0, 0),
    (7, 9, 3, 0, 0),
    (5, 3, 5, 0, 0),
    (64, 64, 5, 0, 0),
    (65, 63, 4, 0, 0),
    (97, 131, 4, 0, 0),
    (130, 257, 6, 0, 0),
    (300, 17, 3, 0, 0),
    (129, 66, 3, 1, 0),
    (66, 129, 3, 0, 1),
    (77, 91, 4, 3, 5),
    (250, 250, 5, 0, 0),
]


def _check_planes(got, exp, filt, what):
    for c, (g, e) in enumerate(zip(got, exp)):
        if filt == "97":
            ok = np.array_equal(g.view(np.uint32), e.view(np.uint32))
        else:
            ok = np.array_equal(g, e)
        assert ok, f"{what} comp {c}: {diff_report(g, e)}"


@pytest.mark.parametrize("path", ["plain", "fused"])
@pytest.mark.parametrize("filt", ["53", "97"])
@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_planes_match_oracle(pkg, workloads, orc, path, filt, shape):
    w, h, L, x0, y0 = shape
    planes = workloads.tile_planes(w, h, L, filt, 2, seed=w * 1000 + h, x0=x0, y0=y0)
    modes = [orc.R53_REFERENCE, orc.R53_TEXTBOOK] if filt == "53" else [0]
    for mode in modes:
        exp, _ = orc.decode_tile(planes, levels=L, filt=filt, out_format=7, x0=x0, y0=y0,
                                 r53_mode=mode)
        run = run_plain if path == "plain" else run_fused
        got, _ = run(pkg, planes, filt=filt, levels=L, out_format=pkg.FMT_NONE, x0=x0, y0=y0,
                     r53_mode=mode)
        _check_planes(got, exp, filt, f"{path} {filt} {shape} mode {mode}")


FORMATS = [
    (2, -1, 8), (0, -1, 8), (1, -1, 12), (8, -1, 8), (2, 0, 8), (2, 1, 8), (2, 2, 8),
    (6, -1, 8), (3, -1, 8), (5, -1, 12), (4, -1, 8), (6, 0, 8), (3, 1, 8),
]


@pytest.mark.parametrize("path", ["plain", "fused"])
@pytest.mark.parametrize("filt", ["53", "97"])
@pytest.mark.parametrize("fmt", FORMATS, ids=lambda f: f"fmt{f[0]}_co{f[1]}_b{f[2]}")
def test_packed_output_matches_oracle(pkg, workloads, orc, path, filt, fmt):
    out_format, check_out, bd = fmt
    w, h, L = 123, 77, 4
    ncomp = 3 if out_format in (0, 1, 2, 8, 5, 6) else 4
    # LL near zero so the level-shifted output is not saturated
    planes = workloads.tile_planes(w, h, L, filt, ncomp, seed=11, ll53=(-60, 60),
                                   ll97=(0.0, 40.0), lap53=6.0 * (1 << (bd - 8)))
    _, exp = orc.decode_tile(planes, levels=L, filt=filt, out_format=out_format, bit_depth=bd,
                             check_out=check_out)
    run = run_plain if path == "plain" else run_fused
    if path == "plain" and out_format in (3, 4, 5, 6) and check_out > 0:
        pytest.skip("the drop-in merge has no check-out argument")
    _, got = run(pkg, planes, filt=filt, levels=L, out_format=out_format, bit_depth=bd,
                 check_out=check_out)
    assert np.array_equal(got, exp), diff_report(got, exp)


ODD_EDGES = [
    # (x0, w, out_x): the tile's first and last columns, and where its first
    # column lands in the picture
    (1, 301, 5), (3, 250, 1), (0, 301, 2), (1, 130, 3), (5, 461, 7), (1, 301, 4), (2, 99, 1),
]


@pytest.mark.parametrize("filt,ncomp,fmt,check_out,bd", [
    ("97", 3, 2, -1, 8), ("53", 3, 2, -1, 8), ("97", 1, 6, 0, 8), ("53", 1, 6, 0, 8),
    ("97", 3, 1, -1, 12), ("53", 3, 1, -1, 12)],
    ids=["97rgb", "53rgb", "97grey", "53grey", "97rgb48", "53rgb48"])
@pytest.mark.parametrize("edge", ODD_EDGES, ids=lambda e: "x%d_w%d_ox%d" % e)
def test_odd_edge_pixels_inside_a_larger_picture(pkg, workloads, orc, filt, ncomp, fmt, check_out,
                                                 bd, edge):
    """A tile at an odd column offset or of odd width, written into a larger
    24-bit or 48-bit picture at (out_x, out_y): lanes whose column pair
    straddles the tile's left or right edge store only the pixel inside the
    tile, on the paired store path (out_x of the same parity as the tile's
    first column; rows of both parities of the picture's odd pitch) and on
    the per-pixel one; every byte outside the tile stays untouched."""
    import torch
    x0, w, out_x = edge
    h, L, y0, out_y = 37, 4, 1, 3
    planes = workloads.tile_planes(w, h, L, filt, ncomp, seed=w + x0, x0=x0, y0=y0, ll53=(-60, 60),
                                   ll97=(0.0, 40.0), lap53=6.0 * (1 << (bd - 8)))
    _, exp = orc.decode_tile(planes, levels=L, filt=filt, out_format=fmt, bit_depth=bd, x0=x0, y0=y0,
                             check_out=check_out)
    pb = pkg.PIXEL_BYTES[fmt]
    for seg in (0, 6):  # the cost model's segments, and fixed 6-row ones (two-level included)
        desc, keep = make_case(pkg, planes, filt=filt, levels=L, out_format=fmt, bit_depth=bd,
                               check_out=check_out, x0=x0, y0=y0)
        ow, oh = keep["ow"], keep["oh"]
        pw, ph = ow + out_x + 5, oh + out_y + 2
        canvas = torch.zeros((ph, pw * pb), dtype=torch.uint8, device="cuda")
        desc.out, desc.out_pitch_px, desc.out_x, desc.out_y = canvas.data_ptr(), pw, out_x, out_y
        tun = pkg.BmiTuning(seg, 0, 0, 0, 0, 1 if seg else 0, 0)
        plan = pkg.BatchPlan([desc], tun)
        plan.execute(0)
        torch.cuda.synchronize()
        assert plan.error_flag() == 0
        plan.close()
        got = canvas.cpu().numpy()
        want = np.zeros_like(got)
        want[out_y:out_y + oh, pb * out_x:pb * (out_x + ow)] = exp.reshape(oh, pb * ow)
        assert np.array_equal(got, want), f"seg {seg}: {diff_report(got, want)}"


@pytest.mark.parametrize("filt", ["53", "97"])
@pytest.mark.parametrize("top", [1, 2, 3])
def test_reduced_resolution_top_level(pkg, workloads, orc, filt, top):
    w, h, L = 150, 101, 4
    planes = workloads.tile_planes(w, h, L, filt, 3, seed=5, ll53=(-60, 60), ll97=(0.0, 40.0))
    exp_planes, exp = orc.decode_tile(planes, levels=L, filt=filt, out_format=2, toplevel=top)
    for run in (run_plain, run_fused):
        _, got = run(pkg, planes, filt=filt, levels=L, out_format=2, top_level=top)
        assert np.array_equal(got, exp), f"{run.__name__}: {diff_report(got, exp)}"


@pytest.mark.parametrize("filt,coef_kind", [("53", 1), ("97", 1), ("97", 0)])
@pytest.mark.parametrize("fmt,ncomp,check_out", [(2, 3, -1), (0, 3, 1), (6, 1, 0), (3, 4, -1), (7, 3, -1)])
def test_thumbnail_top_level_0(pkg, workloads, orc, filt, coef_kind, fmt, ncomp, check_out):
    """top_level 0: the LL band alone (9/7 dequantised with its step),
    colour-converted and packed like a final level, at odd tile offsets."""
    w, h, L, x0, y0 = 301, 157, 4, 3, 5
    planes = workloads.tile_planes(w, h, L, filt, ncomp, seed=9, x0=x0, y0=y0, ll53=(-60, 60),
                                   ll97=(0.0, 40.0), signmag=(coef_kind == 0))
    steps = [np.linspace(0.5 + c * 0.1, 2.0, 3 * L + 1).astype(np.float32) for c in range(ncomp)]
    exp_planes, exp = orc.decode_tile(planes, levels=L, filt=filt, out_format=fmt, toplevel=0,
                                      x0=x0, y0=y0, steps=steps, coef_kind=coef_kind,
                                      check_out=check_out)
    got_planes, got = run_fused(pkg, planes, filt=filt, levels=L, out_format=fmt, top_level=0,
                                x0=x0, y0=y0, steps=steps, coef_kind=coef_kind, check_out=check_out)
    if fmt == 7:
        for g, e in zip(got_planes, exp_planes):
            assert np.array_equal(np.asarray(g).view(np.uint32), np.asarray(e).view(np.uint32))
    else:
        assert got.shape == exp.shape
        assert np.array_equal(got, exp), diff_report(got, exp)


def test_signmag_steps_match_oracle(pkg, workloads, orc):
    """T1 sign-magnitude words with per-band mfAbsStepNorm (the reference's
    BMI_comp_idwt 9/7 contract)."""
    w, h, L = 200, 150, 5
    planes = workloads.tile_planes(w, h, L, "97", 3, seed=21, signmag=True)
    steps = [np.linspace(0.25 + c * 0.1, 2.0, 3 * L + 1).astype(np.float32) for c in range(3)]
    exp, _ = orc.decode_tile(planes, levels=L, filt="97", out_format=7, steps=steps,
                             coef_kind=orc.COEF_SIGNMAG)
    for run in (run_plain, run_fused):
        got, _ = run(pkg, planes, filt="97", levels=L, out_format=pkg.FMT_NONE, steps=steps,
                     coef_kind=pkg.COEF_SIGNMAG)
        _check_planes(got, exp, "97", run.__name__)


def _items(pkg, planes, tuning, **kw):
    from gpu_helpers import make_case
    desc, _keep = make_case(pkg, planes, **kw)
    plan = pkg.BatchPlan([desc], tuning)
    n = plan.num_items
    plan.close()
    return n


@pytest.mark.parametrize("seg_rows", [2, 6, 16, 64, 256])
@pytest.mark.parametrize("persistent", [0, 1])
def test_fused_launch_shapes_agree(pkg, workloads, orc, seg_rows, persistent):
    """Fixed segment heights (fixed_seg = 1 turns the cost model off, so
    each height reaches the kernel: the item count shows it) give the
    oracle's pixels, for step and persistent launches."""
    w, h, L = 301, 199, 5
    planes = workloads.tile_planes(w, h, L, "97", 3, seed=8, ll97=(0.0, 40.0))
    _, exp = orc.decode_tile(planes, levels=L, filt="97", out_format=2)
    t = pkg.BmiTuning(seg_rows, 4, 4, persistent, 0, 1)
    _, got = run_fused(pkg, planes, tuning=t, filt="97", levels=L, out_format=2)
    assert np.array_equal(got, exp), diff_report(got, exp)
    # final level alone: ceil(199 / seg_rows) segments of each of 3 strips
    # (ceil(301 / 120)); the four coarser levels add theirs
    n = _items(pkg, planes, t, filt="97", levels=L, out_format=2)
    final = 3 * (-(-h // seg_rows))
    assert n >= final, (n, final)
    if seg_rows == 2:
        assert n > _items(pkg, planes, pkg.BmiTuning(64, 4, 4, persistent, 0, 1), filt="97",
                          levels=L, out_format=2)


@pytest.mark.parametrize("w", [300, 301])
def test_ict_saturation_matches_oracle(pkg, workloads, orc, w):
    """9/7 + ICT to 8-bit BGR where most pixels clamp at 0 or 255 (the
    packed-conversion store path; an odd width puts odd rows at odd byte
    offsets), bit for bit against the oracle."""
    h, L = 200, 3
    planes = workloads.tile_planes(w, h, L, "97", 3, seed=77, ll97=(-300.0, 500.0))
    _, exp = orc.decode_tile(planes, levels=L, filt="97", out_format=2)
    assert (exp == 0).mean() > 0.3 and (exp == 255).mean() > 0.1 and ((exp > 0) & (exp < 255)).mean() > 0.02
    _, got = run_fused(pkg, planes, filt="97", levels=L, out_format=2)
    assert np.array_equal(got, exp), diff_report(got, exp)


@pytest.mark.parametrize("seg_rows", [2, 4, 6, 10, 38])
@pytest.mark.parametrize("y0", [0, 1, 6])
def test_alternating_segment_directions_match_oracle(pkg, workloads, orc, seg_rows, y0):
    """9/7 level steps run odd segments bottom-up and even ones top-down;
    float planes for several segment heights and row origins (an odd origin
    starts the tile on a high-pass row, the first and last segments cross the
    symmetric extension) equal the oracle's bit for bit."""
    w, h, L = 157, 203, 4
    planes = workloads.tile_planes(w, h, L, "97", 2, seed=seg_rows * 10 + y0, y0=y0)
    exp, _ = orc.decode_tile(planes, levels=L, filt="97", out_format=7, y0=y0)
    for persistent in (0, 1):
        t = pkg.BmiTuning(seg_rows, 0, 0, persistent, 1, 1)
        got, _ = run_fused(pkg, planes, tuning=t, filt="97", levels=L, out_format=pkg.FMT_NONE,
                           y0=y0)
        _check_planes(got, exp, "97", f"seg {seg_rows} y0 {y0} persistent {persistent}")


@pytest.mark.parametrize("mode", [2, 0])
def test_ict_saturation_in_mixed_batch(pkg, workloads, orc, mode):
    """The saturating 9/7 + ICT case in one plan beside a 5/3 RCT tile and a
    grey 9/7 tile: mode 2 runs every step as one dispatching launch (its
    packed 24-bit 9/7 path clamps at 0 and 255), mode 0 lets the plan choose.
    Every tile equals the oracle bit for bit."""
    import torch
    from gpu_helpers import fetch, make_case
    specs = [
        dict(w=301, h=200, L=3, filt="97", nc=3, fmt=2, co=-1,
             kw=dict(ll97=(-300.0, 500.0))),
        dict(w=150, h=97, L=4, filt="53", nc=3, fmt=2, co=-1, kw=dict(ll53=(-200, 200))),
        dict(w=77, h=130, L=3, filt="97", nc=1, fmt=6, co=0, kw=dict(ll97=(-300.0, 500.0))),
    ]
    cases, exps = [], []
    for i, s in enumerate(specs):
        planes = workloads.tile_planes(s["w"], s["h"], s["L"], s["filt"], s["nc"], seed=90 + i,
                                       **s["kw"])
        exps.append(orc.decode_tile(planes, levels=s["L"], filt=s["filt"], out_format=s["fmt"],
                                    check_out=s["co"])[1])
        cases.append(make_case(pkg, planes, filt=s["filt"], levels=s["L"], out_format=s["fmt"],
                               check_out=s["co"]))
    assert (exps[0] == 0).mean() > 0.3 and (exps[0] == 255).mean() > 0.1
    plan = pkg.BatchPlan([d for d, _ in cases], pkg.BmiTuning(32, 4, 4, mode, 0, 0))
    plan.execute(0)
    torch.cuda.synchronize()
    assert plan.error_flag() == 0
    plan.close()
    for i, ((_, keep), exp) in enumerate(zip(cases, exps)):
        _, got = fetch(keep, specs[i]["filt"])
        assert np.array_equal(got, exp), f"tile {i}: {diff_report(got, exp)}"


FUSE_SHAPES = [
    # (w, h, levels, x0, y0, seg_rows): the last two 9/7 levels in one pass
    (4, 4, 2, 0, 0, 0), (5, 7, 2, 1, 1, 0), (9, 6, 3, 0, 1, 2), (33, 17, 3, 1, 0, 4),
    (131, 97, 4, 3, 5, 6), (257, 131, 5, 0, 0, 0), (301, 199, 5, 1, 2, 10), (130, 257, 6, 0, 0, 38),
    (600, 40, 3, 0, 0, 0), (40, 600, 3, 0, 0, 64),
    # items of two strips (115 column pairs): widths at and around one and
    # two items, the second strip partly or wholly past the tile edge
    (232, 33, 2, 0, 0, 0), (233, 64, 3, 1, 0, 0), (461, 50, 3, 3, 1, 0), (120, 90, 3, 2, 0, 0),
    (231, 20, 2, 1, 3, 4), (470, 24, 3, 0, 0, 0),
]


@pytest.mark.parametrize("shape", FUSE_SHAPES, ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("mode", ["reference", "textbook"])
@pytest.mark.parametrize("fmt,bd", [(2, 8), (1, 12)], ids=["bgr24", "bgr48_12bit"])
def test_two_level_53_final_step_matches_oracle(pkg, workloads, orc, shape, mode, fmt, bd):
    """The final two 5/3 levels in one pass: 24-bit and 48-bit (12-bit
    samples) RGB equal the oracle and the two-launch form bit for bit, in
    both synthesis flavours (the reference's bottom-edge rule applies to
    each of the two levels on its own), at odd offsets, small levels and
    fixed segment heights; the fused plan has one launch fewer."""
    from gpu_helpers import make_case
    w, h, L, x0, y0, seg = shape
    r53 = orc.R53_REFERENCE if mode == "reference" else orc.R53_TEXTBOOK
    planes = workloads.tile_planes(w, h, L, "53", 3, seed=3 * w + h, x0=x0, y0=y0, ll53=(-60, 60),
                                   lap53=6.0 * (1 << (bd - 8)))
    _, exp = orc.decode_tile(planes, levels=L, filt="53", out_format=fmt, bit_depth=bd, x0=x0, y0=y0,
                             r53_mode=r53)
    on = pkg.BmiTuning(seg, 0, 0, 0, 0, 1 if seg else 0, 0)
    off = pkg.BmiTuning(seg, 0, 0, 0, 0, 1 if seg else 0, -1)
    for name, t in (("fused", on), ("two launches", off)):
        _, got = run_fused(pkg, planes, tuning=t, filt="53", levels=L, out_format=fmt, bit_depth=bd,
                           x0=x0, y0=y0, r53_mode=r53)
        assert np.array_equal(got, exp), f"{name}: {diff_report(got, exp)}"
    d, _keep = make_case(pkg, planes, filt="53", levels=L, out_format=fmt, bit_depth=bd, x0=x0, y0=y0,
                         r53_mode=r53)
    p_on, p_off = pkg.BatchPlan([d], on), pkg.BatchPlan([d], off)
    assert p_on.num_launches == p_off.num_launches - 1
    p_on.close()
    p_off.close()


@pytest.mark.parametrize("seg", [2, 4])
@pytest.mark.parametrize("fuse", [0, -1], ids=["two_level", "one_level"])
def test_53_segments_around_the_bottom_edge_rule(pkg, workloads, orc, seg, fuse):
    """5/3 in the reference's flavour with segments of one or two row pairs:
    odd segments run bottom-up except where a pipeline's range holds the
    pair of the bottom-edge rule (the last segment, or the last two when
    the last holds one pair, and the inner level's own pair).  Every height
    of 2 to 41 rows at each row offset 0..3 equals the oracle; heights of
    2 or 3 rows at odd offsets have an empty coarsest level (no rows),
    which gets no work items."""
    L, w = 3, 21
    t = pkg.BmiTuning(seg, 0, 0, 0, 0, 1, fuse)
    for h in range(2, 42):
        for y0 in range(4):
            planes = workloads.tile_planes(w, h, L, "53", 3, seed=h * 4 + y0, y0=y0, ll53=(-60, 60))
            _, exp = orc.decode_tile(planes, levels=L, filt="53", out_format=2, y0=y0,
                                     r53_mode=orc.R53_REFERENCE)
            _, got = run_fused(pkg, planes, tuning=t, filt="53", levels=L, out_format=2, y0=y0,
                               r53_mode=orc.R53_REFERENCE)
            assert np.array_equal(got, exp), f"h {h} y0 {y0}: {diff_report(got, exp)}"


@pytest.mark.parametrize("filt", ["53", "97"])
def test_empty_coarse_levels(pkg, workloads, orc, filt):
    """Tiles of 1 to 4 samples a side at offsets 0..3 with 3 levels: their
    coarse levels can have no columns or no rows.  Such a level gets no
    work items (its span would send the kernels' reflections outside the
    bands); the fused and the plain path equal the oracle."""
    L = 3
    for w, h in ((1, 5), (2, 5), (3, 6), (5, 2), (6, 3), (2, 2), (3, 3), (4, 1), (1, 4)):
        for x0, y0 in ((0, 0), (1, 1), (3, 2), (2, 3), (1, 3)):
            planes = workloads.tile_planes(w, h, L, filt, 3, seed=w * 7 + h + x0, x0=x0, y0=y0,
                                           ll53=(-60, 60), ll97=(0.0, 40.0))
            _, exp = orc.decode_tile(planes, levels=L, filt=filt, out_format=2, x0=x0, y0=y0)
            for run in (run_fused, run_plain):
                _, got = run(pkg, planes, filt=filt, levels=L, out_format=2, x0=x0, y0=y0)
                assert np.array_equal(got, exp), f"{run.__name__} {w}x{h} at ({x0},{y0}): " + \
                    diff_report(got, exp)


@pytest.mark.parametrize("shape", FUSE_SHAPES, ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("kind", ["float", "signmag"])
def test_two_level_final_step_matches_oracle(pkg, workloads, orc, shape, kind):
    """The final two 9/7 levels synthesised in one pass (the lower one in an
    LDS ring of rows): planes and 24-bit BGR equal the oracle and the
    two-launch form bit for bit, at odd offsets, small levels (2..5 rows),
    fixed segment heights and sign-magnitude input with per-band steps; the
    fused plan has one launch fewer."""
    from gpu_helpers import make_case
    w, h, L, x0, y0, seg = shape
    sm = kind == "signmag"
    planes = workloads.tile_planes(w, h, L, "97", 3, seed=w + 7 * h, x0=x0, y0=y0,
                                   ll97=(0.0, 40.0), signmag=sm)
    steps = [np.linspace(0.5 + 0.1 * c, 1.7, 3 * L + 1).astype(np.float32) for c in range(3)]
    ck = pkg.COEF_SIGNMAG if sm else pkg.COEF_FLOAT
    exp_planes, exp = orc.decode_tile(planes, levels=L, filt="97", out_format=2, x0=x0, y0=y0,
                                      steps=steps, coef_kind=ck)
    on = pkg.BmiTuning(seg, 0, 0, 0, 0, 1 if seg else 0, 0)
    off = pkg.BmiTuning(seg, 0, 0, 0, 0, 1 if seg else 0, -1)
    for fmt, want in ((2, exp), (pkg.FMT_NONE, None)):
        got = {}
        for name, t in (("fused", on), ("two launches", off)):
            got[name] = run_fused(pkg, planes, tuning=t, filt="97", levels=L, out_format=fmt,
                                  x0=x0, y0=y0, steps=steps, coef_kind=ck)
        if want is not None:
            for name in got:
                assert np.array_equal(got[name][1], want), f"{name}: {diff_report(got[name][1], want)}"
        else:
            for name in got:
                _check_planes(got[name][0], exp_planes, "97", name)
    d, _keep = make_case(pkg, planes, filt="97", levels=L, out_format=2, x0=x0, y0=y0, steps=steps,
                         coef_kind=ck)
    p_on, p_off = pkg.BatchPlan([d], on), pkg.BatchPlan([d], off)
    assert p_on.num_launches == p_off.num_launches - 1
    p_on.close()
    p_off.close()
