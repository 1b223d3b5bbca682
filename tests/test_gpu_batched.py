"""Batched decode on the GPU: the 256-tile mixed workload (sides 64..1024,
odd sizes, 3..6 levels, 5/3 and 9/7, 1 or 3 components) in one plan,
compared tile by tile with the oracle; hipGraph replays, the persistent
queue, the autotuner and rebinding against the same expectations."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from gpu_helpers import diff_report, fetch, make_case

pytestmark = pytest.mark.gpu


def _fmt(t):
    # 3 components: inverse RCT/ICT to 24-bit BGR; 1 component: grey merge
    return (2, -1) if t["ncomp"] == 3 else (6, 0)


@pytest.fixture(scope="module")
def batch(pkg, workloads, orc):
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    tiles = workloads.config3_params(0)
    planes = [workloads.tile_planes(t["w"], t["h"], t["levels"], t["filt"], t["ncomp"], seed=0,
                                    tile=t["tile"]) for t in tiles]

    def expect(i):
        t = tiles[i]
        fmt, co = _fmt(t)
        return orc.decode_tile(planes[i], levels=t["levels"], filt=t["filt"], out_format=fmt,
                               check_out=co)[1]

    with ThreadPoolExecutor(16) as ex:
        exp = list(ex.map(expect, range(len(tiles))))
    cases = []
    for t, p in zip(tiles, planes):
        fmt, co = _fmt(t)
        cases.append(make_case(pkg, p, filt=t["filt"], levels=t["levels"], out_format=fmt,
                               check_out=co))
    print(f"batch: {len(tiles)} tiles, {sum(t['w'] * t['h'] for t in tiles)} pixels", flush=True)
    return dict(tiles=tiles, planes=planes, exp=exp, cases=cases)


def _clear(batch):
    for _, keep in batch["cases"]:
        keep["out"].zero_()


def _check_all(batch, what):
    bad = []
    for i, (_, keep) in enumerate(batch["cases"]):
        _, got = fetch(keep, batch["tiles"][i]["filt"])
        if not np.array_equal(got, batch["exp"][i]):
            bad.append(f"tile {i} {batch['tiles'][i]}: {diff_report(got, batch['exp'][i])}")
    assert not bad, f"{what}: {len(bad)} tiles differ; first: {bad[:3]}"


def test_batch_step_mode_matches_oracle(pkg, batch):
    _clear(batch)
    plan = pkg.BatchPlan([d for d, _ in batch["cases"]])
    plan.execute(0)
    _check_all(batch, "step mode")
    plan.close()


@pytest.mark.parametrize("mode", [2, 3])
def test_batch_launch_grouping_matches_oracle(pkg, batch, mode):
    """Every step as one dispatching launch (2) and one launch per step and
    kernel variant (3) decode as the default grouping does."""
    _clear(batch)
    plan = pkg.BatchPlan([d for d, _ in batch["cases"]], pkg.BmiTuning(32, 4, 4, mode, 0))
    assert plan.num_launches >= 1
    plan.execute(0)
    _check_all(batch, f"launch grouping {mode}")
    plan.close()


@pytest.mark.parametrize("from_coarse", [0, 1], ids=["steps_from_finest", "steps_from_coarsest"])
@pytest.mark.parametrize("fuse2", [-1, 0, 256], ids=["two_launches", "two_level_all", "two_level_256"])
def test_batch_step_alignment_and_fusion_match_oracle(pkg, batch, from_coarse, fuse2):
    """Level steps aligned at every tile's finest level (the default: all
    final levels, and so all two-level launches, in the last step) or at
    the coarsest, with the last two 9/7 levels of every tile, of tiles whose
    final level spans 256 samples both ways, or of none in one pass."""
    _clear(batch)
    t = pkg.BmiTuning(32, 0, 0, 0, 4, 0, fuse2, from_coarse)
    plan = pkg.BatchPlan([d for d, _ in batch["cases"]], t)
    plan.execute(0)
    _check_all(batch, f"from_coarse={from_coarse} fuse2={fuse2}")
    plan.close()


def test_batch_persistent_queue_matches_oracle(pkg, batch):
    _clear(batch)
    plan = pkg.BatchPlan([d for d, _ in batch["cases"]], pkg.BmiTuning(32, 4, 4, 1, 0))
    plan.execute(0)
    _check_all(batch, "persistent")
    assert plan.error_flag() == 0
    plan.close()


def test_batch_graph_replays(pkg, batch):
    """Each replay of the captured graph, with eager work and
    synchronisation between replays, gives the eager result."""
    import torch
    plan = pkg.BatchPlan([d for d, _ in batch["cases"]])
    scratch = torch.empty(1 << 20, device="cuda")
    for rep in range(3):
        _clear(batch)
        scratch.normal_()  # unrelated eager work between replays
        torch.cuda.synchronize()
        plan.execute_graph(0)
        _check_all(batch, f"graph replay {rep}")
    plan.close()


def test_autotune_result_decodes_identically(pkg, batch):
    descs = [d for d, _ in batch["cases"]]
    cands = [pkg.BmiTuning(16, 4, 0, 0, 4), pkg.BmiTuning(64, 2, 0, 0, 16),
             pkg.BmiTuning(32, 4, 0, 0, 8)]
    best, us = pkg.autotune(descs, cands, reps=2)
    assert us > 0 and best.seg_rows in (16, 64, 32)
    _clear(batch)
    plan = pkg.BatchPlan(descs, best)
    plan.execute(0)
    _check_all(batch, f"autotuned seg_rows={best.seg_rows} min_waves={best.min_waves_per_cu}")
    plan.close()


def test_rebind_serves_a_new_frame(pkg, workloads, orc):
    """One plan + graph, coefficients of tile 1 swapped for a new frame of
    the same geometry: the replay decodes the new frame."""
    import torch
    shapes = [(200, 120, 4, "97", 3), (333, 77, 3, "53", 3)]
    cases, planes = [], []
    for i, (w, h, L, f, c) in enumerate(shapes):
        p = workloads.tile_planes(w, h, L, f, c, seed=40 + i, ll53=(-60, 60), ll97=(0.0, 40.0))
        planes.append(p)
        cases.append(make_case(pkg, p, filt=f, levels=L, out_format=2))
    plan = pkg.BatchPlan([d for d, _ in cases])
    plan.execute_graph(0)
    torch.cuda.synchronize()
    w, h, L, f, c = shapes[1]
    newp = workloads.tile_planes(w, h, L, f, c, seed=99, ll53=(-60, 60))
    nd, nkeep = make_case(pkg, newp, filt=f, levels=L, out_format=2)
    plan.rebind(1, nd)
    plan.execute_graph(0)
    _, got = fetch(nkeep, f)
    _, exp = orc.decode_tile(newp, levels=L, filt=f, out_format=2)
    assert np.array_equal(got, exp), diff_report(got, exp)
    _, got0 = fetch(cases[0][1], "97")
    _, exp0 = orc.decode_tile(planes[0], levels=4, filt="97", out_format=2)
    assert np.array_equal(got0, exp0), diff_report(got0, exp0)
    plan.close()
