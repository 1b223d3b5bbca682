"""Parity at the benchmark workloads' own sizes, against the oracle.

- 512x512x1, 5-level 5/3 (LL U{0..4095}, detail round(Laplace(6)), seed 0):
  pre-clamp int32 planes and the 8-bit grey merge, both 5/3 modes, plain and
  fused paths.
- 4096x2160x3, 5-level 9/7 float32 (LL N(128,40), detail Laplace(4)*2^-l,
  seed 0): float planes bit for bit and 24-bit BGR, the fused path with its
  default launch shape (the one the benchmark times) and the plain path.
- The same image through tier 1: every one of its 6,708 64x64 code-blocks
  decoded on the GPU into the coefficient planes, then the fused IDWT + ICT.
- A served mix holding an 8192x8192 12-bit request beside smaller ones of
  the serving workload's kinds (odd sides, 8/12-bit, 5/3 and 9/7).

The oracle's multi-component decodes run one host thread per component
(the C oracle releases the GIL)."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from gpu_helpers import diff_report, make_case, run_fused, run_plain

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu(pkg):
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    pkg.lib()


def _oracle_decode(orc, planes, *, levels, filt, out_format, bit_depth=8, steps=None,
                   coef_kind=1, check_out=-1, r53_mode=0):
    """oracle.decode_tile with the per-component IDWTs on threads."""
    n = planes.shape[0]

    def one(c):
        if filt == "53":
            return orc.idwt53(planes[c], levels, mode=r53_mode)
        return orc.idwt97(planes[c], levels, None if steps is None else steps[c],
                          coef_kind=coef_kind)

    with ThreadPoolExecutor(n) as ex:
        recon = list(ex.map(one, range(n)))
    if out_format == 7:
        return recon, None
    if out_format in (0, 1, 2, 8):
        fn = orc.ict_pack if filt == "97" else orc.rct_pack
        return recon, fn(recon[0], recon[1], recon[2], bit_depth, out_format, check_out)
    comps = [recon[check_out]] if check_out >= 0 else recon
    return recon, orc.merge_pack(comps, bit_depth, out_format, 0 if check_out >= 0 else -1,
                                 is_float=(filt == "97"))


def _same_planes(got, exp, filt, what):
    for c, (g, e) in enumerate(zip(got, exp)):
        g = np.asarray(g)
        if filt == "97":
            ok = np.array_equal(g.view(np.uint32), np.asarray(e).view(np.uint32))
        else:
            ok = np.array_equal(g, e)
        assert ok, f"{what} comp {c}: {diff_report(g, np.asarray(e))}"


# ------------------------------------------------------------ 512x512 5/3 --

@pytest.mark.parametrize("path", ["plain", "fused"])
@pytest.mark.parametrize("mode", [0, 1], ids=["r53_reference", "r53_textbook"])
def test_config_512x512_53_planes_and_grey(pkg, workloads, orc, path, mode):
    cfg = workloads.config1()
    planes, L = cfg["planes"], cfg["levels"]
    assert planes.shape == (1, 512, 512) and planes.dtype == np.int32
    print(f"512x512x1 5/3 L{L} {path} r53 mode {mode}", flush=True)
    run = run_plain if path == "plain" else run_fused
    exp_planes, _ = orc.decode_tile(planes, levels=L, filt="53", out_format=7, r53_mode=mode)
    got_planes, _ = run(pkg, planes, filt="53", levels=L, out_format=pkg.FMT_NONE, r53_mode=mode)
    _same_planes(got_planes, exp_planes, "53", f"512x512 {path} mode {mode}")
    # the pre-clamp planes leave the 8-bit range (LL up to 4095 at 8 bits)
    assert exp_planes[0].max() > 255
    # 8-bit grey: merge of component 0 into all three slots
    _, exp = orc.decode_tile(planes, levels=L, filt="53", out_format=6, check_out=0,
                             r53_mode=mode)
    _, got = run(pkg, planes, filt="53", levels=L, out_format=6, check_out=0, r53_mode=mode)
    assert got.shape == exp.shape == (512, 512 * 3)
    assert np.array_equal(got, exp), diff_report(got, exp)


# ------------------------------------------------------ 4096x2160x3 9/7 --

@pytest.fixture(scope="module")
def image97(workloads, orc):
    cfg = workloads.config2()
    planes = cfg["planes"]
    assert planes.shape == (3, 2160, 4096)
    recon, bgr = _oracle_decode(orc, planes, levels=5, filt="97", out_format=2)
    return dict(planes=planes, recon=recon, bgr=bgr)


@pytest.mark.parametrize("path", ["fused", "plain"])
def test_config_4096x2160x3_97_full_size(pkg, orc, image97, path):
    print(f"4096x2160x3 9/7 L5 {path}", flush=True)
    run = run_plain if path == "plain" else run_fused
    planes = image97["planes"]
    got_planes, _ = run(pkg, planes, filt="97", levels=5, out_format=pkg.FMT_NONE)
    _same_planes(got_planes, image97["recon"], "97", f"4096x2160 {path}")
    _, got = run(pkg, planes, filt="97", levels=5, out_format=2)
    exp = image97["bgr"]
    assert got.shape == exp.shape == (2160, 4096 * 3)
    assert np.array_equal(got, exp), diff_report(got, exp)


def test_config_4096x2160x3_97_all_launch_modes(pkg, image97):
    """The benchmark's timed forms on the full image: direct launches,
    graph replays, the dispatching kernel and the persistent queue."""
    import torch
    planes, exp = image97["planes"], image97["bgr"]
    for name, tuning, graph in (("direct", None, False), ("graph", None, True),
                                ("dispatching", pkg.BmiTuning(0, 0, 0, 2, 0, 0), False),
                                ("persistent", pkg.BmiTuning(32, 4, 4, 1, 0, 0), False)):
        print(f"4096x2160x3 9/7 launch mode {name}", flush=True)
        desc, keep = make_case(pkg, planes, filt="97", levels=5, out_format=2)
        plan = pkg.BatchPlan([desc], tuning)
        for _ in range(3 if graph else 1):
            keep["out"].zero_()
            plan.execute_graph(0) if graph else plan.execute(0)
            torch.cuda.synchronize()
            got = keep["out"].cpu().numpy()
            assert np.array_equal(got, exp), f"{name}: {diff_report(got, exp)}"
        assert plan.error_flag() == 0
        plan.close()


# ------------------------------------------------ tier 1, 6,708 blocks --

def test_config_t1_4096x2160x3_full_image(pkg, workloads, orc):
    """Every 64x64 code-block of the 4096x2160x3 9/7 image (quantised
    Laplace(4) detail, N(128,40) LL, the project's encoder, three passes per
    plane, no bypass) decoded on the GPU in the benchmark's launch order,
    bit for bit against the oracle's decode_code_block, then the fused IDWT
    + ICT against the oracle pipeline on those words."""
    import torch
    W, H, L = 4096, 2160, 5
    img = workloads.t1_image(W, H, L, 3, seed=0, filt="97", threads=16)
    assert len(img["blocks"]) == 6708
    print(f"T1 image {W}x{H}x3: {len(img['blocks'])} blocks, {len(img['data'])} bytes", flush=True)
    planes = [torch.zeros((H, W), dtype=torch.int32, device="cuda") for _ in range(3)]
    blocks = []
    for bd in img["blocks"]:
        b = pkg.BmiT1Block()
        b.data_offset, b.len = bd["offset"], bd["len"]
        b.width, b.height, b.orient = bd["w"], bd["h"], bd["orient"]
        b.numbps, b.missing, b.num_pass, b.kind = bd["numbps"], bd["missing"], bd["num_pass"], pkg.T1_WORDS
        b.out = planes[bd["comp"]].data_ptr() + 4 * (bd["y"] * W + bd["x"])
        b.out_pitch = W
        blocks.append(b)
    pkg.t1_check_blocks(blocks)
    arr = (pkg.BmiT1Block * len(blocks))(*blocks)
    pkg.t1_order_blocks(arr)  # longest code streams first, as the benchmark runs them
    dblocks = torch.from_numpy(np.frombuffer(bytes(arr), np.uint8).copy()).cuda()
    ddata = torch.from_numpy(np.frombuffer(img["data"], np.uint8).copy()).cuda()
    pkg.t1_decode_blocks(dblocks.data_ptr(), len(blocks), ddata.data_ptr(), 0)
    torch.cuda.synchronize()

    def dec(bd):
        return orc.t1_decode_block(img["data"][bd["offset"]:bd["offset"] + bd["len"]], bd["w"],
                                   bd["h"], bd["orient"], bd["numbps"], bd["missing"],
                                   bd["num_pass"])

    with ThreadPoolExecutor(16) as ex:
        wds = list(ex.map(dec, img["blocks"]))
    words = np.zeros((3, H, W), np.int32)
    for bd, wd in zip(img["blocks"], wds):
        words[bd["comp"], bd["y"]:bd["y"] + bd["h"], bd["x"]:bd["x"] + bd["w"]] = wd
    for c in range(3):
        got = planes[c].cpu().numpy()
        assert np.array_equal(got, words[c]), f"T1 comp {c}: {diff_report(got, words[c])}"
    print("T1 words equal the oracle's; IDWT + ICT", flush=True)
    out = torch.zeros((H, W * 3), dtype=torch.uint8, device="cuda")
    desc = pkg.tile_desc(coef_ptrs=[p.data_ptr() for p in planes], coef_pitch=W, width=W, height=H,
                         levels=L, method=pkg.W9X7, ncomp=3, coef_kind=pkg.COEF_SIGNMAG,
                         steps=img["steps"], out_format=2, out_ptr=out.data_ptr(), out_pitch_px=W)
    plan = pkg.BatchPlan([desc])
    plan.execute(0)
    torch.cuda.synchronize()
    plan.close()
    _, exp = _oracle_decode(orc, words.view(np.uint32), levels=L, filt="97", out_format=2,
                            steps=img["steps"], coef_kind=orc.COEF_SIGNMAG)
    got = out.cpu().numpy()
    assert np.array_equal(got, exp), diff_report(got, exp)


# ---------------------------------------------- served mix up to 8192 px --

SERVED = [  # (side, bits, filter)
    (8192, 12, "97"),
    (2999, 8, "53"),
    (1501, 12, "53"),
    (701, 8, "97"),
    (257, 12, "97"),
    (4097, 8, "53"),
]


def test_config_served_mix_with_8192_12bit(pkg, workloads, orc):
    """Requests of the serving workload's kinds, the largest 8192x8192x3
    12-bit 9/7 (48-bit BGR out), submitted together to the fused server:
    every result equals the oracle's; each one served alone equals the mix."""
    import torch
    reqs = []
    for i, (side, bits, filt) in enumerate(SERVED):
        s = 1 << (bits - 8)
        planes = workloads.tile_planes(side, side, 5, filt, 3, seed=41, img=i,
                                       ll53=(-60 * s, 60 * s), lap53=6.0 * s,
                                       ll97=(0.0, 40.0 * s), lap97=4.0 * s)
        fmt = 2 if bits == 8 else 1
        _, exp = _oracle_decode(orc, planes, levels=5, filt=filt, out_format=fmt, bit_depth=bits)
        reqs.append(dict(planes=planes, filt=filt, bits=bits, fmt=fmt, exp=exp, side=side))
        print(f"served request {i}: {side}x{side}x3 {bits}-bit {filt} expected", flush=True)
    assert reqs[0]["exp"].shape == (8192, 8192 * 6)
    # the 12-bit outputs use the full 16-bit container range
    assert reqs[0]["exp"].view(np.uint16).max() > 255

    def serve(alone):
        srv = pkg.Server(16 << 30, num_streams=4, max_batch=32)
        cases = [make_case(pkg, r["planes"], filt=r["filt"], levels=5, out_format=r["fmt"],
                           bit_depth=r["bits"]) for r in reqs]
        torch.cuda.synchronize()
        if alone:
            res = [srv.wait(srv.submit(d), 120000) for d, _ in cases]
        else:
            tickets = [srv.submit(d) for d, _ in cases]
            res = [srv.wait(t, 120000) for t in tickets]
        st = srv.stats()
        srv.close()
        outs = [keep["out"].cpu().numpy() for _, keep in cases]
        del cases
        torch.cuda.empty_cache()
        return outs, res, st

    mix, res, st = serve(False)
    assert st.completed == len(reqs) and st.failed == 0
    for i, r in enumerate(reqs):
        assert res[i].status == 0
        assert np.array_equal(mix[i], r["exp"]), f"req {i} in mix: {diff_report(mix[i], r['exp'])}"
    print(f"mix: {st.batches} batches for {len(reqs)} requests", flush=True)
    alone, res1, _ = serve(True)
    for i in range(len(reqs)):
        assert res1[i].status == 0
        assert np.array_equal(alone[i], mix[i]), f"req {i} alone vs mix"
