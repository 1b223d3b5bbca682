"""The ignore-component mask (the reference's mIgnoreComponents,
decoder/common/include/codec_base.h:215-228) on the GPU: ignored components
are not decoded (their coefficient pointers are null here, so a read of
them would fault) and the remaining ones fill the output slots in order
(bmi_cpu_dwt_decoder.cpp:1180-1290).  Fused plans, the plain drop-in
sequence, the request server and whole codestreams (whose ignored
components' code-blocks are not decoded either) against the oracle."""
import pathlib

import numpy as np
import pytest

from gpu_helpers import diff_report, fetch, make_case

import t2_oracle as t2

pytestmark = pytest.mark.gpu

J2K = pathlib.Path(__file__).resolve().parent / "golden" / "j2k"

CASES = [
    # (filt, ncomp, mask, out_format, check_out)
    ("97", 4, 0b1000, 2, -1),    # RGBA, alpha ignored: ICT of 0..2
    ("53", 4, 0b1000, 0, -1),    # 32-bit ARGB from RGB + ignored alpha
    ("53", 4, 0b0010, 6, -1),    # merge of components 0, 2, 3
    ("97", 4, 0b0101, 3, -1),    # merge of components 1, 3
    ("53", 3, 0b0100, 6, 1),     # merge, check-out of component 1 (kept)
    ("97", 3, 0b0010, 6, 1),     # check-out of an ignored component: none
    ("97", 4, 0b0110, 7, -1),    # planes of components 0 and 3
]


def _case(pkg, workloads, filt, nc, mask, fmt, co, seed):
    w, h, L = 157, 93, 4
    planes = workloads.tile_planes(w, h, L, filt, nc, seed=seed, ll53=(-60, 60), ll97=(0.0, 40.0))
    steps = [np.linspace(0.6 + 0.1 * c, 1.4, 3 * L + 1).astype(np.float32) for c in range(nc)]
    return planes, steps, dict(filt=filt, levels=L, out_format=fmt, check_out=co, steps=steps)


def _null_ignored(desc, mask, nc):
    for c in range(nc):
        if (mask >> c) & 1:
            desc.coef[c] = None
    desc.ignore_mask = mask
    return desc


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}_n{c[1]}_m{c[2]:04b}_f{c[3]}_co{c[4]}")
def test_ignore_mask_fused_and_plain(pkg, workloads, orc, case):
    import torch
    filt, nc, mask, fmt, co = case
    planes, steps, kw = _case(pkg, workloads, filt, nc, mask, fmt, co, seed=mask * 10 + nc)
    exp_planes, exp = orc.decode_tile(planes, levels=kw["levels"], filt=filt, out_format=fmt,
                                      check_out=co, steps=steps, ignore_mask=mask)
    kept = [c for c in range(nc) if not (mask >> c) & 1]
    # fused plan
    desc, keep = make_case(pkg, planes, **kw)
    _null_ignored(desc, mask, nc)
    plan = pkg.BatchPlan([desc])
    plan.execute(0)
    torch.cuda.synchronize()
    plan.close()
    got_planes, got = fetch(keep, filt)
    if fmt == pkg.FMT_NONE:
        for i, c in enumerate(kept):
            g, e = got_planes[c], np.asarray(exp_planes[i])
            assert np.array_equal(g.view(np.uint32), e.view(np.uint32)), f"fused comp {c}"
    else:
        assert np.array_equal(got, exp), f"fused: {diff_report(got, exp)}"
    # plain drop-in sequence (in place in the coefficient planes)
    if fmt == pkg.FMT_NONE:
        return
    desc, keep = make_case(pkg, planes, **kw)
    _null_ignored(desc, mask, nc)
    temp = torch.zeros(157 * 93 + 64, dtype=torch.int32, device="cuda")
    pkg.plain_decode(desc, temp.data_ptr(), temp.numel() * 4, 0)
    torch.cuda.synchronize()
    _, got = fetch(keep, filt)
    if fmt in (3, 4, 5, 6) and co > 0:
        return  # the drop-in merge has no check-out argument
    assert np.array_equal(got, exp), f"plain: {diff_report(got, exp)}"


def test_ignore_mask_rejections(pkg, workloads):
    planes, steps, kw = _case(pkg, workloads, "97", 3, 0, 2, -1, seed=5)
    desc, _keep = make_case(pkg, planes, **kw)
    desc.ignore_mask = 0b111  # nothing left
    with pytest.raises(pkg.BmiError):
        pkg.BatchPlan([desc])
    desc.ignore_mask = 0b001  # RGB from two components
    with pytest.raises(pkg.BmiError) as e:
        pkg.BatchPlan([desc])
    assert "(-5)" in str(e.value)


def test_ignore_mask_served(pkg, workloads, orc):
    import torch
    reqs = []
    srv = pkg.Server(1 << 30, num_streams=2, max_batch=8)
    for i, (filt, nc, mask, fmt, co) in enumerate(CASES[:5]):
        planes, steps, kw = _case(pkg, workloads, filt, nc, mask, fmt, co, seed=100 + i)
        exp = orc.decode_tile(planes, levels=kw["levels"], filt=filt, out_format=fmt, check_out=co,
                              steps=steps, ignore_mask=mask)[1]
        desc, keep = make_case(pkg, planes, **kw)
        _null_ignored(desc, mask, nc)
        reqs.append((desc, keep, exp))
    torch.cuda.synchronize()
    tickets = [srv.submit(d) for d, _, _ in reqs]
    res = [srv.wait(t, 60000) for t in tickets]
    srv.close()
    for (d, keep, exp), r in zip(reqs, res):
        assert r.status == 0
        got = keep["out"].cpu().numpy()
        assert np.array_equal(got, exp), diff_report(got, exp)


@pytest.mark.parametrize("name,mask,fmt,co", [("rgba_mct.j2k", 0b1000, 2, -1),
                                              ("rgb_64_mct.j2k", 0b0110, 6, 0),
                                              ("rgb97_tiles_pcrl.j2k", 0b0001, 6, -1)])
def test_ignore_mask_codestream(pkg, orc, name, mask, fmt, co):
    """Whole files: the ignored components' code-blocks are left out of the
    tier-1 launch; the picture equals the oracle's decode with the mask."""
    data = (J2K / name).read_bytes()
    full = pkg.CodestreamDecoder(data, out_format=fmt, check_out=co)
    dec = pkg.CodestreamDecoder(data, out_format=fmt, check_out=co, ignore_mask=mask)
    nc = dec.cs.info.ncomp
    assert dec.nblocks < full.nblocks  # fewer code-blocks decoded
    full.close()
    dec.run(0)
    import torch
    torch.cuda.synchronize()
    got = dec.out.cpu().numpy()
    dec.close()
    exp = t2.decode_image(data, orc, out_format=fmt, check_out=co, ignore_mask=mask)
    assert got.shape == exp.shape
    assert np.array_equal(got, exp), (name, diff_report(got, exp), nc)
