"""Seeded random tiles (tests/random_cases.py) decoded by the fused engine
and compared with the oracle bit for bit: planes (9/7 as float bits),
24-bit and 48-bit RGB, 32-bit ARGB, merged grey and four-component merges, at every offset parity, 1..5
levels, both 5/3 flavours, several segment heights, the two-level step on
and off.  Every case's kernel addresses are checked on the CPU first
(test_kernel_addresses.py::test_random_case_addresses)."""
import numpy as np
import pytest

from gpu_helpers import diff_report, run_fused
from random_cases import random_cases

pytestmark = pytest.mark.gpu


def _planes_equal(got, exp, filt):
    for g, e in zip(got, exp):
        if filt == "97":
            if not np.array_equal(g.view(np.uint32), e.view(np.uint32)):
                return False
        elif not np.array_equal(g, e):
            return False
    return True


def test_random_tiles_match_oracle(pkg, workloads, orc):
    failures = []
    for i, c in enumerate(random_cases()):
        w, h, levels, x0, y0, filt_i, ncomp, fmt, bd, r53, seg, fuse = c
        filt = str(filt_i)
        planes = workloads.tile_planes(w, h, levels, filt, ncomp, seed=1000 + i, x0=x0, y0=y0,
                                       ll53=(-60, 60), ll97=(0.0, 40.0), lap53=6.0 * (1 << (bd - 8)))
        check_out = 0 if (ncomp == 1 and fmt == 6) else -1
        r53m = orc.R53_REFERENCE if r53 == 0 else orc.R53_TEXTBOOK
        exp_planes, exp = orc.decode_tile(planes, levels=levels, filt=filt, out_format=fmt, bit_depth=bd,
                                          x0=x0, y0=y0, check_out=check_out, r53_mode=r53m)
        tun = pkg.BmiTuning(seg, 0, 0, 0, 0, 1 if seg else 0, fuse)
        got_planes, got = run_fused(pkg, planes, tuning=tun, filt=filt, levels=levels,
                                    out_format=pkg.FMT_NONE if fmt == 7 else fmt, bit_depth=bd,
                                    x0=x0, y0=y0, check_out=check_out, r53_mode=r53m)
        if fmt == 7:
            ok = _planes_equal(got_planes, exp_planes, filt)
            detail = "planes differ"
        else:
            ok = np.array_equal(got, exp)
            detail = "" if ok else diff_report(got, exp)
        if not ok:
            failures.append((c, detail))
    assert not failures, f"{len(failures)} cases differ; first: {failures[:2]}"
