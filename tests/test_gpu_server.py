"""Request server on the GPU: continuous batching over an HBM budget.

Each result must equal the oracle's and be the same whether the request
was decoded alone or in a mix with others (bit for bit: the fused engine's
output does not depend on what else is in its batch), for device-resident
and host-copied requests, in fused and plain serving modes.  Malformed
and oversized requests are rejected without disturbing the others."""
import numpy as np
import pytest

from gpu_helpers import diff_report, make_case

pytestmark = pytest.mark.gpu


def _requests(workloads, n, seed):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(n):
        w, h = (int(v) for v in rng.integers(64, 400, 2))
        filt = "53" if rng.random() < 0.5 else "97"
        bits = 8 if rng.random() < 0.5 else 12
        levels = int(rng.integers(3, 6))
        planes = workloads.tile_planes(w, h, levels, filt, 3, seed=seed, img=i,
                                       ll53=(-60 << (bits - 8), 60 << (bits - 8)),
                                       lap53=6.0 * (1 << (bits - 8)),
                                       ll97=(0.0, 40.0 * (1 << (bits - 8))),
                                       lap97=4.0 * (1 << (bits - 8)))
        fmt = 2 if bits == 8 else 1  # 24-bit BGR / 48-bit BGR
        reqs.append(dict(w=w, h=h, filt=filt, bits=bits, levels=levels, planes=planes, fmt=fmt))
    return reqs


def _expect(orc, r):
    return orc.decode_tile(r["planes"], levels=r["levels"], filt=r["filt"], out_format=r["fmt"],
                           bit_depth=r["bits"])[1]


@pytest.fixture(scope="module")
def reqs(pkg, workloads, orc):
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    rs = _requests(workloads, 24, seed=3)
    for r in rs:
        r["exp"] = _expect(orc, r)
    return rs


def _serve(pkg, reqs, *, mode=0, max_batch=64, streams=4, budget=1 << 30, alone=False):
    """Submit every request (all at once, or one at a time when alone) on
    device buffers; return (outputs, results, stats)."""
    import torch
    srv = pkg.Server(budget, num_streams=streams, max_batch=max_batch, mode=mode)
    cases = [make_case(pkg, r["planes"], filt=r["filt"], levels=r["levels"], out_format=r["fmt"],
                       bit_depth=r["bits"]) for r in reqs]
    torch.cuda.synchronize()
    results = []
    if alone:
        for d, _ in cases:
            results.append(srv.wait(srv.submit(d), 60000))
    else:
        tickets = [srv.submit(d) for d, _ in cases]
        results = [srv.wait(t, 60000) for t in tickets]
    st = srv.stats()
    srv.close()
    outs = [keep["out"].cpu().numpy() for _, keep in cases]
    return outs, results, st


def test_mix_equals_alone_equals_oracle(pkg, reqs):
    mix, res, st = _serve(pkg, reqs)
    alone, _, st1 = _serve(pkg, reqs, alone=True)
    for i, r in enumerate(reqs):
        assert np.array_equal(mix[i], r["exp"]), f"req {i} in mix: {diff_report(mix[i], r['exp'])}"
        assert np.array_equal(alone[i], mix[i]), f"req {i} alone vs mix"
        assert res[i].status == 0
    assert st.completed == len(reqs) and st.failed == 0
    # requests submitted together were batched (continuous batching)
    assert max(x.batch_size for x in res) > 1
    assert st.batches < len(reqs)
    assert st1.batches == len(reqs)


def test_small_budget_admits_by_footprint(pkg, reqs):
    """A budget that holds only a few requests at a time: every request
    still completes correctly and the arena high-water mark stays within
    the budget."""
    budget = 24 << 20
    out, res, st = _serve(pkg, reqs, budget=budget)
    for i, r in enumerate(reqs):
        assert np.array_equal(out[i], r["exp"]), f"req {i}"
    assert st.peak_bytes <= budget
    assert st.batches > 1


def test_plain_mode_matches(pkg, reqs):
    out, res, st = _serve(pkg, reqs, mode=pkg.SERVE_PLAIN)
    for i, r in enumerate(reqs):
        assert np.array_equal(out[i], r["exp"]), f"req {i}: {diff_report(out[i], r['exp'])}"
    assert all(x.batch_size == 1 for x in res)


def test_host_io_requests(pkg, reqs):
    """Coefficients in pinned host memory and packed pixels written back to
    a host image at an offset."""
    import torch
    srv = pkg.Server(1 << 30, num_streams=2)
    jobs = []
    for r in reqs[:6]:
        pb = pkg.PIXEL_BYTES[r["fmt"]]
        coefs = [torch.from_numpy(r["planes"][c].view(np.int32).copy()).pin_memory()
                 for c in range(3)]
        canvas = torch.zeros((r["h"] + 5, (r["w"] + 7) * pb), dtype=torch.uint8).pin_memory()
        d = pkg.tile_desc(coef_ptrs=[t.data_ptr() for t in coefs], coef_pitch=r["w"],
                          width=r["w"], height=r["h"], levels=r["levels"],
                          method=pkg.W5X3 if r["filt"] == "53" else pkg.W9X7, ncomp=3,
                          bit_depth=r["bits"], out_format=r["fmt"], out_ptr=canvas.data_ptr(),
                          out_pitch_px=r["w"] + 7, out_x=3, out_y=2)
        jobs.append((srv.submit(d, host_io=True), coefs, canvas, r, pb))
    for t, _, canvas, r, pb in jobs:
        assert srv.wait(t, 60000).status == 0
        got = canvas.numpy()[2:2 + r["h"], 3 * pb:(3 + r["w"]) * pb]
        assert np.array_equal(got, r["exp"]), diff_report(got, r["exp"])
        assert not canvas.numpy()[:2].any() and not canvas.numpy()[:, :3 * pb].any()
    srv.close()


def test_rejections_do_not_disturb_others(pkg, reqs):
    srv = pkg.Server(16 << 20, num_streams=2)
    r = reqs[0]
    d, keep = make_case(pkg, r["planes"], filt=r["filt"], levels=r["levels"], out_format=r["fmt"],
                        bit_depth=r["bits"])
    good = srv.submit(d)
    bad_levels = make_case(pkg, r["planes"], filt=r["filt"], levels=r["levels"],
                           out_format=r["fmt"], bit_depth=r["bits"])[0]
    bad_levels.levels = 9
    with pytest.raises(pkg.BmiError):
        srv.submit(bad_levels)
    none_fmt = make_case(pkg, r["planes"], filt=r["filt"], levels=r["levels"],
                         out_format=pkg.FMT_NONE)[0]
    with pytest.raises(pkg.BmiError):
        srv.submit(none_fmt)
    # footprint above the whole budget (host copies of a 4096x4096 image)
    huge = pkg.tile_desc(coef_ptrs=[1, 1, 1], coef_pitch=4096, width=4096, height=4096, levels=5,
                         method=pkg.W9X7, ncomp=3, out_format=2, out_ptr=1, out_pitch_px=4096)
    with pytest.raises(pkg.BmiError):
        srv.submit(huge, host_io=True)
    res = srv.wait(good, 60000)
    assert res.status == 0
    assert np.array_equal(keep["out"].cpu().numpy(), r["exp"])
    st = srv.stats()
    assert st.submitted == 1 and st.completed == 1 and st.failed == 0
    srv.close()
