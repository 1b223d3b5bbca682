"""BMI_decode_tiles_noalloc: the batch decode with its scratch in the
library's fixed per-GPU device array, called through the C-ABI, against
the allocating entry point (BMI_batch_plan_create + execute) on the same
tiles, bit for bit: two calls in a row, the second with a smaller input
(so stale contents of the array from the first call would show), then an
oversized batch rejected before any device work."""
import ctypes

import numpy as np
import pytest

from gpu_helpers import diff_report, fetch, make_case

pytestmark = pytest.mark.gpu


def _cases(pkg, workloads, specs, seed):
    out = []
    for i, (w, h, L, filt, nc, fmt, co) in enumerate(specs):
        planes = workloads.tile_planes(w, h, L, filt, nc, seed=seed, tile=i, ll53=(-60, 60),
                                       ll97=(0.0, 40.0))
        out.append((planes, dict(filt=filt, levels=L, out_format=fmt, check_out=co)))
    return out


def _run(pkg, cases, noalloc, tuning=None):
    import torch
    made = [make_case(pkg, p, **kw) for p, kw in cases]
    descs = [d for d, _ in made]
    if noalloc:
        pkg.decode_tiles_noalloc(descs, tuning, 0)
    else:
        plan = pkg.BatchPlan(descs, tuning)
        plan.execute(0)
        torch.cuda.synchronize()
        plan.close()
    torch.cuda.synchronize()
    return [fetch(keep, kw["filt"]) for (_, keep), (_, kw) in zip(made, cases)]


def _same(a, b, what):
    (pa, xa), (pb, xb) = a, b
    if xa is not None:
        assert np.array_equal(xa, xb), f"{what}: {diff_report(xa, xb)}"
    if pa is not None:
        for c, (u, v) in enumerate(zip(pa, pb)):
            assert np.array_equal(u.view(np.uint32), v.view(np.uint32)), f"{what} comp {c}"


def test_noalloc_matches_allocating_entry_point(pkg, workloads, orc):
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    cap = pkg.lib().BMI_decode_tiles_noalloc_capacity()
    assert cap == 64 << 20
    first = _cases(pkg, workloads, [
        (1500, 900, 5, "97", 3, 2, -1),
        (701, 503, 4, "53", 1, 6, 0),
        (301, 199, 3, "53", 3, 2, -1),
        (257, 131, 6, "97", 3, 7, -1),
    ], seed=61)
    second = _cases(pkg, workloads, [(129, 67, 3, "97", 3, 2, -1)], seed=62)
    for name, cases in (("first call", first), ("second, smaller call", second)):
        print(f"no-alloc {name}: {len(cases)} tiles", flush=True)
        got = _run(pkg, cases, noalloc=True)
        exp = _run(pkg, cases, noalloc=False)
        for i, (g, e) in enumerate(zip(got, exp)):
            _same(g, e, f"{name} tile {i}")
    # the smaller call also equals the oracle
    planes, kw = second[0]
    _, want = orc.decode_tile(planes, levels=kw["levels"], filt="97", out_format=2)
    got = _run(pkg, second, noalloc=True)
    assert np.array_equal(got[0][1], want), diff_report(got[0][1], want)
    # the persistent work queue in the array (its counters are cleared per call)
    t = pkg.BmiTuning(32, 4, 4, 1, 0, 0)
    for _ in range(2):
        _same(_run(pkg, first[:2], noalloc=True, tuning=t)[0], _run(pkg, first[:2], noalloc=False)[0],
              "persistent")


def test_noalloc_rejects_batches_over_capacity(pkg, workloads):
    """A batch whose plan needs more than the array is refused before any
    device work.  The tile's buffers are real device memory, so even a
    wrong acceptance could not make the kernels touch memory outside them:
    an 8192x8192 9/7 tile in two-launch form (two 4096x4096 float scratch
    planes, 134 MB)."""
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    n = 8192
    coef = torch.zeros((n, n), dtype=torch.float32, device="cuda")
    out = torch.zeros((n, n * 3), dtype=torch.uint8, device="cuda")
    d = pkg.tile_desc(coef_ptrs=[coef.data_ptr()], coef_pitch=n, width=n, height=n, levels=5,
                      method=pkg.W9X7, ncomp=1, out_format=6, check_out=0, out_ptr=out.data_ptr(),
                      out_pitch_px=n)
    arr = (pkg.BmiTileDesc * 1)(d)
    two_launch = pkg.BmiTuning(0, 0, 0, 0, 0, 0, -1)
    rc = pkg.lib().BMI_decode_tiles_noalloc(arr, 1, ctypes.byref(two_launch), None)
    assert rc == pkg.BMI_E_NOMEM
    torch.cuda.synchronize()
