"""Benchmark of the fused IDWT + ICT back end on one MI355X.

Workload (one step = one image): decode a 4096x2160 3-component image whose
coefficients sit in HBM -- 5-level irreversible 9/7 inverse DWT on float32
subbands (LL ~ N(128,40), detail Laplace(4)*2^-level), inverse ICT, level
shift, clamp, packed 24-bit BGR -- with the fused engine (one launch per
level step, the last two levels in one launch).  `value` is the throughput
of a stream of such images: plans of up to 24 images (one plan holds all
the timed images when there are no more, as with 20 steps), 2 plans in
flight on their own streams when there are, timed with direct launches and
with hipGraph replays (the faster is reported); the same images decoded
one at a time on one stream (latency) are reported beside it.  The timed images rotate over buffer sets
whose total exceeds the 256 MB last-level cache, so their data comes from
HBM.  One timed output is compared with the oracle before anything is
printed.  Prints one JSON line.

Also measured, beside `value`: the plain reference-shaped GPU path (the
drop-in BMI_comp_idwt per component + BMI_tile_ICT, 2 launches per level),
the dominant kernel against HBM roofline, the 256-tile batch, the request
server, tier 1 and the file-to-pixels path, and the multithreaded CPU
restatement of the reference back end on the host (all cores and one).
"""
from __future__ import annotations

import argparse
import ctypes
import importlib
import importlib.util
import json
import os
import pathlib
import sys
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent
PKG_DIR = ROOT / "jpeg2000-decoder_amd"

METRIC = "decoded Mpixel/s + IDWT HBM GB/s (% peak); served p50/p99 latency vs plain"
HBM_PEAK_GBS = 8000.0


def load_package():
    name = "jpeg2000_decoder_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(
        name, PKG_DIR / "__init__.py", submodule_search_locations=[str(PKG_DIR)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _cpu_model() -> str:
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(planes: np.ndarray, w: int, h: int, expect: np.ndarray, budget_s: float = 10.0):
    """The multithreaded CPU restatement of the reference CPU back end
    (oracle/cpu_pipeline.c: rows / column bands / row bands over the
    threads, -O3 x86-64-v3, bit-identical to the scalar oracle) decoding the
    same image, on all host cores of the GPU's share (at most 16) and on
    one core, each repeated for about budget_s of wall time.  Its output is
    compared with the oracle's before it is reported."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle  # checker / CPU baseline only
    nproc = os.cpu_count() or 1
    threads = min(16, nproc)
    work = np.empty(planes.shape, np.float32)
    out = np.empty((h, 3 * w), np.uint8)
    res = {}
    for name, t in (("all", threads), ("one", 1)):
        reps, t0 = 0, time.perf_counter()
        while True:
            oracle.cpu_decode97_ict(planes, 5, t, out=out, work=work)
            reps += 1
            el = time.perf_counter() - t0
            if el > budget_s or reps >= 200:
                break
        res[name] = (reps, el, w * h * reps / el / 1e6)
    assert np.array_equal(out, expect), "CPU baseline output differs from the oracle"
    reps, el, mpix = res["all"]
    reps1, el1, mpix1 = res["one"]
    return {"value": round(mpix, 3), "unit": "Mpixel/s", "cores": threads, "kind": "port",
            "nproc": nproc, "cpu_model": _cpu_model(),
            "one_core": {"value": round(mpix1, 3), "unit": "Mpixel/s", "cores": 1,
                         "sample": f"{reps1} decodes in {el1:.1f} s"},
            "sample": f"{reps} full {w}x{h}x3 9/7 5-level IDWT + ICT decodes by the "
                      f"multithreaded CPU restatement of the reference back end on {threads} "
                      f"threads ({el:.1f} s); output equal to the scalar oracle's"}


def _device_planes(torch, dev, w, h, filt, ncomp, gen):
    """Synthetic coefficient planes generated on the device (timing of the
    decode does not depend on the values): 5/3 small integers, 9/7 floats
    of the configured distributions' scale."""
    if filt == "53":
        return [torch.randint(-64, 64, (h, w), dtype=torch.int32, device=dev, generator=gen)
                for _ in range(ncomp)]
    return [(torch.randn((h, w), device=dev, generator=gen) * 8.0).view(torch.int32)
            for _ in range(ncomp)]


def batched_leg(pkg, wl, torch, dev, stream, steps, warmup, inflight=2):
    """256 tiles (sides 64..1024, 3..6 levels, 5/3 / 9/7, 1 or 3
    components) in one plan: autotuned launch shape, hipGraph replay and
    direct launches timed.  Reported: batches decoded as a stream,
    `inflight` plans (each with its own buffer set) on their own streams,
    over max(60, 6 * steps) batches (a stream of a few batches mostly
    measures its start and its drain, when fewer than `inflight` plans
    run); one batch at a time (`steps` batches) beside it, and the plain
    path decoding the same tiles one per call."""
    tiles = wl.config3_params(0)
    gen = torch.Generator(device=dev)
    gen.manual_seed(0)
    sets = []
    pixels = comp_samples = out_bytes = 0
    for k in range(max(1, inflight)):
        descs, keep = [], []
        for t in tiles:
            coef = _device_planes(torch, dev, t["w"], t["h"], t["filt"], t["ncomp"], gen)
            fmt, co = (2, -1) if t["ncomp"] == 3 else (6, 0)
            out = torch.zeros((t["h"], t["w"] * 3), dtype=torch.uint8, device=dev)
            descs.append(pkg.tile_desc(coef_ptrs=[c.data_ptr() for c in coef], coef_pitch=t["w"],
                                       width=t["w"], height=t["h"], levels=t["levels"],
                                       method=pkg.W5X3 if t["filt"] == "53" else pkg.W9X7,
                                       ncomp=t["ncomp"], bit_depth=8, out_format=fmt, check_out=co,
                                       out_ptr=out.data_ptr(), out_pitch_px=t["w"]))
            keep.append((coef, out))
            if k == 0:
                pixels += t["w"] * t["h"]
                comp_samples += t["w"] * t["h"] * t["ncomp"]
                out_bytes += t["w"] * t["h"] * 3
        sets.append((descs, keep))
    descs, keep = sets[0]
    sh = stream.cuda_stream
    t0 = time.perf_counter()
    best, best_us = pkg.autotune(descs, None, reps=3, stream=sh)
    tune_s = time.perf_counter() - t0
    plans = [pkg.BatchPlan(d, best) for d, _ in sets]
    plan = plans[0]
    one_ms = {}
    for name, fn in (("hipGraph", lambda: plan.execute_graph(sh)),
                     ("direct launches", lambda: plan.execute(sh))):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        one_ms[name] = e0.elapsed_time(e1) / steps
    # a stream of batches: plan i % inflight on stream i % inflight
    streams = [torch.cuda.Stream(dev) for _ in plans]
    mode_ms = {}
    for name, graph in (("hipGraph", True), ("direct launches", False)):
        for rep in range(2):
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for st_ in streams:
                st_.wait_stream(stream)
            n = max(60, 6 * steps) if rep else max(warmup, len(plans))
            for i in range(n):
                p_, ss = plans[i % len(plans)], streams[i % len(plans)].cuda_stream
                p_.execute_graph(ss) if graph else p_.execute(ss)
            for st_ in streams:
                stream.wait_stream(st_)
            e1.record(stream)
            torch.cuda.synchronize()
        mode_ms[name] = e0.elapsed_time(e1) / n
    launch_mode = min(mode_ms, key=mode_ms.get)
    ms = mode_ms[launch_mode]
    ms_one = min(one_ms.values())
    compulsory = comp_samples * 4 + out_bytes
    res = {"tiles": len(tiles), "pixels": pixels, "ms_per_batch": round(ms, 4),
           "Mpixel_s": round(pixels / (ms * 1e-3) / 1e6, 1),
           "compulsory_bytes": compulsory,
           "GBs_compulsory": round(compulsory / (ms * 1e-3) / 1e9, 1),
           "frac_of_peak": round(compulsory / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
           "work_items": plan.num_items, "launches": plan.num_launches,
           "launch_mode": launch_mode, "batches_in_flight": len(plans),
           "batches_timed_in_stream": max(60, 6 * steps),
           "us_per_batch_by_launch_mode": {k: round(v * 1e3, 1) for k, v in mode_ms.items()},
           "ms_per_batch_one_at_a_time": round(ms_one, 4),
           "frac_of_peak_one_at_a_time": round(compulsory / (ms_one * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
           "us_one_at_a_time_by_launch_mode": {k: round(v * 1e3, 1) for k, v in one_ms.items()},
           "tuning": {"seg_rows": best.seg_rows, "waves_per_group": best.waves_per_group,
                      "min_waves_per_cu": best.min_waves_per_cu, "persistent": best.persistent,
                      "fuse2": best.fuse2, "from_coarse": best.from_coarse},
           "autotune_s": round(tune_s, 2)}
    for p_ in plans:
        p_.close()
    del sets[1:]
    # plain path: the drop-in sequence, one tile per call, on copies (in place)
    work = [[torch.empty_like(c) for c in coef] for coef, _ in keep]
    temps = torch.empty(1024 * 1024 + 64, dtype=torch.int32, device=dev)
    pdescs = []
    for d, wk in zip(descs, work):
        pd = pkg.BmiTileDesc.from_buffer_copy(d)
        for c, t in enumerate(wk):
            pd.coef[c] = t.data_ptr()
        pdescs.append(pd)
    times = []
    for it in range(2):
        for wk, (coef, _) in zip(work, keep):
            for a, b in zip(wk, coef):
                a.copy_(b)
        p0, p1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        p0.record(stream)
        for pd in pdescs:
            pkg.plain_decode(pd, temps.data_ptr(), temps.numel() * 4, sh)
        p1.record(stream)
        torch.cuda.synchronize()
        if it:
            times.append(p0.elapsed_time(p1))
    pms = float(np.mean(times))
    res["plain_ms_per_batch"] = round(pms, 3)
    res["plain_Mpixel_s"] = round(pixels / (pms * 1e-3) / 1e6, 1)
    res["speedup_vs_plain"] = round(pms / ms, 2)
    res["speedup_vs_plain_one_at_a_time"] = round(pms / ms_one, 2)
    return res


def serving_leg(pkg, wl, torch, dev, rates, nreq, budget_gib=64):
    """Poisson arrivals at each rate; sides log-uniform in 256..8192, 8 or
    12 bits, 5/3 or 9/7, 3 components, 5 levels; device-resident
    coefficient and output buffers; admitted against an HBM budget.  The
    same arrival trace goes through the fused server and the plain server
    (one request at a time, drop-in call sequence)."""
    import threading
    gen = torch.Generator(device=dev)
    gen.manual_seed(1)
    side = 8192
    pools = {f: _device_planes(torch, dev, side, side, f, 3, gen) for f in ("53", "97")}
    out = {}
    for mode_name, mode in (("fused", pkg.SERVE_FUSED), ("plain", pkg.SERVE_PLAIN)):
        rows = []
        for rate in rates:
            reqs = wl.config4_requests(7, nreq, rate)
            outs = [torch.empty((r["h"], r["w"] * (3 if r["bits"] == 8 else 6)), dtype=torch.uint8,
                                device=dev) for r in reqs]
            descs = []
            for r, o in zip(reqs, outs):
                coef = pools[r["filt"]]
                descs.append(pkg.tile_desc(coef_ptrs=[c.data_ptr() for c in coef], coef_pitch=side,
                                           width=r["w"], height=r["h"], levels=5,
                                           method=pkg.W5X3 if r["filt"] == "53" else pkg.W9X7,
                                           ncomp=3, bit_depth=r["bits"],
                                           out_format=2 if r["bits"] == 8 else 1,
                                           out_ptr=o.data_ptr(), out_pitch_px=r["w"]))
            torch.cuda.synchronize()
            srv = pkg.Server(budget_gib << 30, num_streams=4, max_batch=32, mode=mode)
            tickets = [None] * len(reqs)
            results = [None] * len(reqs)

            def waiter():
                for i in range(len(reqs)):
                    while tickets[i] is None:
                        time.sleep(1e-4)
                    results[i] = srv.wait(tickets[i], 120000)

            th = threading.Thread(target=waiter)
            th.start()
            t_start = time.perf_counter()
            for i, r in enumerate(reqs):
                delay = r["arrival"] - (time.perf_counter() - t_start)
                if delay > 0:
                    time.sleep(delay)
                tickets[i] = srv.submit(descs[i])
            th.join()
            st = srv.stats()
            srv.close()
            lat = np.array([(x.done_us - x.submit_us) / 1e3 for x in results])
            span = (max(x.done_us for x in results) - min(x.submit_us for x in results)) / 1e6
            mpix = sum(r["w"] * r["h"] for r in reqs) / span / 1e6
            rows.append({"rate_req_s": rate, "requests": len(reqs),
                         "p50_ms": round(float(np.percentile(lat, 50)), 3),
                         "p99_ms": round(float(np.percentile(lat, 99)), 3),
                         "Mpixel_s": round(mpix, 1), "batches": int(st.batches),
                         "mean_batch": round(len(reqs) / max(1, st.batches), 2),
                         "failed": int(st.failed), "peak_GiB": round(st.peak_bytes / 2**30, 2)})
            log(f"serving {mode_name} rate {rate}: p50 {rows[-1]['p50_ms']} ms, "
                f"p99 {rows[-1]['p99_ms']} ms, {rows[-1]['Mpixel_s']} Mpixel/s")
            del outs
        out[mode_name] = rows
    out["workload"] = (f"{nreq} requests per rate, Poisson arrivals, square sides log-uniform in "
                       "256..8192, 8/12-bit, 5/3 or 9/7, 3 components, 5 levels, device-resident "
                       f"buffers, {budget_gib} GiB budget, 4 streams, batches of at most 32")
    return out


def t1_leg(pkg, wl, torch, dev, stream, steps, warmup):
    """Full tile decode including tier 1: the 4096x2160x3 9/7 image,
    quantised and split into 64x64 code-blocks encoded by the project's
    EBCOT encoder; per step the GPU decodes every code-block (one per
    wavefront) into the coefficient planes, then the fused IDWT + ICT packs
    24-bit BGR.  Code streams and the block table are resident in HBM."""
    W, H, L = 4096, 2160, 5
    t0 = time.perf_counter()
    img = wl.t1_image(W, H, L, 3, seed=0, filt="97", threads=16)
    enc_s = time.perf_counter() - t0
    planes = [torch.zeros((H, W), dtype=torch.int32, device=dev) for _ in range(3)]
    blocks = []
    for bd in img["blocks"]:
        b = pkg.BmiT1Block()
        b.data_offset, b.len = bd["offset"], bd["len"]
        b.width, b.height, b.orient = bd["w"], bd["h"], bd["orient"]
        b.numbps, b.missing, b.num_pass, b.kind = bd["numbps"], bd["missing"], bd["num_pass"], 0
        b.out = planes[bd["comp"]].data_ptr() + 4 * (bd["y"] * W + bd["x"])
        b.out_pitch = W
        blocks.append(b)
    pkg.t1_check_blocks(blocks)
    arr = (pkg.BmiT1Block * len(blocks))(*blocks)
    pkg.t1_order_blocks(arr)   # longest code-block data first (BMI_t1_order_blocks)
    dblocks = torch.from_numpy(np.frombuffer(bytes(arr), dtype=np.uint8).copy()).to(dev)
    ddata = torch.from_numpy(np.frombuffer(img["data"], dtype=np.uint8).copy()).to(dev)
    out = torch.zeros((H, W * 3), dtype=torch.uint8, device=dev)
    desc = pkg.tile_desc(coef_ptrs=[p.data_ptr() for p in planes], coef_pitch=W, width=W, height=H,
                         levels=L, method=pkg.W9X7, ncomp=3, coef_kind=pkg.COEF_SIGNMAG,
                         steps=img["steps"], out_format=pkg.FMT_24BITS_RGB,
                         out_ptr=out.data_ptr(), out_pitch_px=W)
    plan = pkg.BatchPlan([desc])
    sh = stream.cuda_stream
    nb = len(blocks)

    def step():
        pkg.t1_decode_blocks(dblocks.data_ptr(), nb, ddata.data_ptr(), sh)
        plan.execute_graph(sh)

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    ev[0].record(stream)
    for _ in range(steps):
        step()
    ev[1].record(stream)
    ev[2].record(stream)
    for _ in range(steps):
        pkg.t1_decode_blocks(dblocks.data_ptr(), nb, ddata.data_ptr(), sh)
    ev[3].record(stream)
    torch.cuda.synchronize()
    ms = ev[0].elapsed_time(ev[1]) / steps
    t1_ms = ev[2].elapsed_time(ev[3]) / steps
    plan.close()
    ncoef = W * H * 3
    return {"workload": f"{W}x{H}x3 9/7 L={L}, {nb} code-blocks of 64x64 (own EBCOT encoder, "
                        "quantised Laplace(4)*2^-l / N(128,40) LL, 3 passes per plane, no bypass)",
            "code_bytes": len(img["data"]) - 16,
            "bits_per_coef": round((len(img["data"]) - 16) * 8 / ncoef, 3),
            "ms_per_image": round(ms, 3), "Mpixel_s": round(W * H / (ms * 1e-3) / 1e6, 1),
            "t1_ms": round(t1_ms, 3), "t1_Mcoef_s": round(ncoef / (t1_ms * 1e-3) / 1e6, 1),
            "t1_blocks_s": round(nb / (t1_ms * 1e-3), 0),
            "encode_s_host": round(enc_s, 2), "_img": img, "_out": out}


def codestream_leg(pkg, wl, torch, dev, stream, steps, warmup, img, ref_out):
    """From a JPEG 2000 file in host memory to pixels: the tier-1 leg's
    code-blocks wrapped in a codestream (one tile, one layer, LRCP, default
    precincts; workloads.t1_codestream).  Host tier 2 (BMI_codestream_open +
    BMI_codestream_tile_blocks into pinned memory), one H2D copy of the
    code-block segments, then GPU tier 1 and the fused IDWT + ICT; each part
    timed on its own and the output checked against the tier-1 leg's."""
    W, H = img["w"], img["h"]
    t0 = time.perf_counter()
    data = wl.t1_codestream(img)
    write_s = time.perf_counter() - t0
    dec = pkg.CodestreamDecoder(data, out_format=pkg.FMT_24BITS_RGB)
    sh = stream.cuda_stream
    _, _, w, h, nb, sb = dec.cs.tile_info(0)
    # the decoder's device copy holds the tile's share padded to 16 bytes
    pinned = torch.full((dec.d_segs.numel(),), 0xFF, dtype=torch.uint8, pin_memory=True)
    pn = pinned.numpy()
    ptrs = [dec.planes[0][c].data_ptr() for c in range(3)]
    pitch = dec.planes[0].shape[2]
    host = []
    for _ in range(7):
        a = time.perf_counter()
        cs = pkg.Codestream(data)
        b = time.perf_counter()
        cs.tile_blocks(0, ptrs, pitch, 0, segs=pn)
        c = time.perf_counter()
        cs.close()
        host.append((b - a, c - b))
    host.sort(key=lambda x: x[0] + x[1])
    parse_s, blocks_s = host[len(host) // 2]
    assert pn[:sb].tobytes() == dec.d_segs[:sb].cpu().numpy().tobytes()
    for _ in range(warmup):
        dec.run(sh)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    ev[0].record(stream)
    for _ in range(steps):
        dec.d_segs.copy_(pinned, non_blocking=True)
    ev[1].record(stream)
    ev[2].record(stream)
    for _ in range(steps):
        dec.run(sh)
    ev[3].record(stream)
    torch.cuda.synchronize()
    h2d_ms = ev[0].elapsed_time(ev[1]) / steps
    gpu_ms = ev[2].elapsed_time(ev[3]) / steps
    same = bool(torch.equal(dec.out, ref_out))
    dec.close()
    host_ms = (parse_s + blocks_s) * 1e3
    e2e = host_ms + h2d_ms + gpu_ms
    return {"workload": f"{W}x{H}x3 9/7 L=5 .j2k, {nb} code-blocks, {len(data)} bytes "
                        "(one tile, one layer, LRCP)",
            "host_parse_ms": round(parse_s * 1e3, 3), "host_tile_blocks_ms": round(blocks_s * 1e3, 3),
            "h2d_ms": round(h2d_ms, 3), "h2d_GBs": round(sb / (h2d_ms * 1e-3) / 1e9, 1),
            "gpu_ms": round(gpu_ms, 3), "gpu_Mpixel_s": round(W * H / (gpu_ms * 1e-3) / 1e6, 1),
            "serial_file_to_pixels_ms": round(e2e, 3),
            "serial_file_to_pixels_Mpixel_s": round(W * H / (e2e * 1e-3) / 1e6, 1),
            "output_equals_t1_leg": same, "write_s_host": round(write_s, 2)}


def t1_cpu_baseline(img, budget_s: float = 8.0):
    """The oracle's tier-1 decoder (scalar C restatement of
    decode_code_block) on the same code-blocks, one host thread per core,
    until about budget_s of wall time."""
    from concurrent.futures import ThreadPoolExecutor
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle  # checker / CPU baseline only
    oracle.lib()
    data = img["data"]
    blocks = img["blocks"]
    threads = min(16, os.cpu_count() or 1)

    def one(bd):
        oracle.t1_decode_block(data[bd["offset"]:bd["offset"] + bd["len"]], bd["w"], bd["h"],
                               bd["orient"], bd["numbps"], bd["missing"], bd["num_pass"])
        return bd["w"] * bd["h"]

    t0 = time.perf_counter()
    ncoef = nblk = 0
    with ThreadPoolExecutor(threads) as ex:
        for i in range(0, len(blocks), 256):
            chunk = blocks[i:i + 256]
            ncoef += sum(ex.map(one, chunk))
            nblk += len(chunk)
            if time.perf_counter() - t0 > budget_s:
                break
    el = time.perf_counter() - t0
    return {"value": round(ncoef / el / 1e6, 2), "unit": "Mcoef/s", "cores": threads,
            "kind": "port", "blocks_s": round(nblk / el, 1),
            "sample": f"{nblk} of the {len(blocks)} code-blocks decoded by the C oracle's "
                      f"decode_code_block on {threads} host threads ({el:.1f} s)"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--width", type=int, default=4096)
    ap.add_argument("--height", type=int, default=2160)
    ap.add_argument("--sets", type=int, default=8,
                    help="coefficient / output buffer sets the timed images rotate over "
                         "(8 x 133 MB: more than the 256 MB last-level cache)")
    ap.add_argument("--images-per-plan", type=int, default=0,
                    help="images per plan in the throughput (reported) form; 0: the timed "
                         "images, at most 24 per plan")
    ap.add_argument("--batch-inflight", type=int, default=3,
                    help="batched leg: plans in flight on their own streams")
    ap.add_argument("--inflight", type=int, default=2,
                    help="plans in flight at once, each on its own stream")
    ap.add_argument("--legs", default="single,plain,batched,serving,t1,cpu",
                    help="comma list of legs to run (profiling runs pick one)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain-path comparison")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--seg-rows", type=int, default=0)
    ap.add_argument("--groups-per-cu", type=int, default=0)
    ap.add_argument("--waves-per-group", type=int, default=0)
    ap.add_argument("--persistent", type=int, default=0)
    ap.add_argument("--min-waves", type=int, default=0)
    ap.add_argument("--no-fuse2", action="store_true",
                    help="keep the last two synthesis levels as separate launches")
    ap.add_argument("--no-batched", action="store_true", help="skip the 256-tile batched leg")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving leg")
    ap.add_argument("--no-t1", action="store_true", help="skip the tier-1 leg")
    ap.add_argument("--serve-rates", default="50,400,3200,12800,51200")
    ap.add_argument("--serve-requests", type=int, default=120)
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run a 256 MiB device copy, the reference for scaling HBM counters")
    args = ap.parse_args()
    if args.images_per_plan <= 0:
        # one plan holds the timed images (up to 24): a longer launch
        # amortises its last round of waves (two-level launch per image:
        # 37 us in plans of 8, 33 in plans of 16, 32 in plans of 24)
        args.images_per_plan = max(1, min(24, args.steps))
    legs = set(x for x in args.legs.split(",") if x)
    for flag, leg in (("no_cpu", "cpu"), ("no_plain", "plain"), ("no_batched", "batched"),
                      ("no_serving", "serving"), ("no_t1", "t1")):
        if getattr(args, flag):
            legs.discard(leg)

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the library decodes on a single MI355X "
              "(tiles and images shard across GPUs with no exchange between them). "
              "Run with --gpus 1.", file=sys.stderr)
        return 2

    import torch
    pkg = load_package()
    pkg.lib()
    wl = importlib.import_module("jpeg2000_decoder_amd.workloads")
    assert torch.cuda.is_available(), "bench needs a HIP device"
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    sh = stream.cuda_stream
    W, H, L = args.width, args.height, 5
    compulsory = W * H * 3 * 4 + W * H * 3
    tuning = pkg.BmiTuning(args.seg_rows, args.waves_per_group, args.groups_per_cu, args.persistent,
                           args.min_waves, 0, -1 if args.no_fuse2 else 0)
    rec_single = {}
    roofline = None
    planes = None
    expect = None
    mpix = ms_step = None
    plans = []
    if "single" in legs:
        # Buffer sets (coefficients, output; each plan has its own scratch):
        # set 0 is the workload itself (seed 0), checked against the oracle;
        # sets 1.. are same-sized images of device-generated floats (the
        # decode's time does not depend on the values).  The timed images
        # rotate over all sets, more memory than the last-level cache holds.
        nsets = max(args.images_per_plan * args.inflight, args.sets)
        log("generating config: %dx%dx3 9/7 L=%d, %d buffer sets" % (W, H, L, nsets))
        planes = wl.tile_planes(W, H, L, "97", 3, seed=0)
        gen = torch.Generator(device=dev)
        gen.manual_seed(1)
        sets = []
        for k in range(nsets):
            if k == 0:
                coef = [torch.from_numpy(planes[c].view(np.int32).copy()).to(dev) for c in range(3)]
            else:
                coef = [(torch.randn((H, W), device=dev, generator=gen) * 8.0).view(torch.int32)
                        for _ in range(3)]
            out = torch.zeros((H, W * 3), dtype=torch.uint8, device=dev)
            desc = pkg.tile_desc(coef_ptrs=[t.data_ptr() for t in coef], coef_pitch=W, width=W,
                                 height=H, levels=L, method=pkg.W9X7, ncomp=3,
                                 coef_kind=pkg.COEF_FLOAT, bit_depth=8,
                                 out_format=pkg.FMT_24BITS_RGB, out_ptr=out.data_ptr(),
                                 out_pitch_px=W)
            sets.append(dict(coef=coef, out=out, desc=desc))
        # latency form: one image per plan, one stream
        for st in sets:
            st["plan"] = pkg.BatchPlan([st["desc"]], tuning)
            plans.append(st["plan"])
        plan = sets[0]["plan"]
        log(f"one-image plan: {plan.num_items} work items in {plan.num_launches} launches, "
            f"scratch {plan.scratch_bytes / 2**20:.1f} MiB")

        def timed(step, n, warm, steps):
            for i in range(warm):
                step(i % n)
            torch.cuda.synchronize()
            ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            t0 = time.perf_counter()
            ev0.record(stream)
            for i in range(steps):
                step(i % n)
            ev1.record(stream)
            torch.cuda.synchronize()
            return ev0.elapsed_time(ev1), time.perf_counter() - t0

        lat = {}
        for name, graph in (("direct launches", False), ("hipGraph", True)):
            if graph and args.no_graph:
                continue
            for rot in (False, True):
                n = len(sets) if rot else 1
                fn = ((lambda i: sets[i]["plan"].execute_graph(sh)) if graph
                      else (lambda i: sets[i]["plan"].execute(sh)))
                # every set's plan runs (and captures its graph) before timing
                ms, _ = timed(fn, n, max(args.warmup, n), args.steps)
                key = f"{name}, {'rotating over %d sets' % n if rot else 'one set'}"
                lat[key] = ms / args.steps
                log(f"one image at a time ({key}): {lat[key] * 1e3:.1f} us/image")
        for st in sets:
            assert st["plan"].error_flag() == 0, "fused kernel reported a dependency timeout"
        lat_rot = min(v for k, v in lat.items() if "rotating" in k)
        lat_one = min(v for k, v in lat.items() if "one set" in k)

        # throughput form (reported): plans of `images_per_plan` images, that
        # many plans in flight on as many streams (each plan always on its
        # own stream, so a plan never overlaps itself); K images timed
        P, S = args.images_per_plan, args.inflight
        groups = [[sets[g * P + j] for j in range(P)] for g in range(nsets // P)]
        tplans = [pkg.BatchPlan([st["desc"] for st in g], tuning) for g in groups]
        plans.extend(tplans)
        streams = [torch.cuda.Stream(dev) for _ in range(S)]
        nexec = max(1, -(-args.steps // P))
        images = nexec * P
        thr = {}
        for name, graph in (("direct launches", False), ("hipGraph", True)):
            if graph and args.no_graph:
                continue
            for rep in range(2):  # the first pass warms up (graph capture included)
                torch.cuda.synchronize()
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                t0 = time.perf_counter()
                e0.record(stream)
                for st_ in streams:
                    st_.wait_stream(stream)
                for i in range(nexec if rep else max(1, args.warmup // P + 1)):
                    tp = tplans[i % len(tplans)]
                    ss = streams[(i % len(tplans)) % S].cuda_stream
                    tp.execute_graph(ss) if graph else tp.execute(ss)
                for st_ in streams:
                    stream.wait_stream(st_)
                e1.record(stream)
                torch.cuda.synchronize()
            thr[name] = (e0.elapsed_time(e1) / images, time.perf_counter() - t0)
            log(f"stream of images ({name}): {thr[name][0] * 1e3:.1f} us/image")
        for tp in tplans:
            assert tp.error_flag() == 0, "fused kernel reported a dependency timeout"
        launch_mode = min(thr, key=lambda k: thr[k][0])
        ms_step, t_wall = thr[launch_mode]
        mpix = W * H / (ms_step * 1e-3) / 1e6
        log(f"fused: {ms_step * 1e3:.1f} us/image as a stream ({P} images per plan, {S} in flight), "
            f"{lat_rot * 1e3:.1f} us one at a time; {mpix:.0f} Mpixel/s")

        # Dominant kernel: the last launch of a plan (the final level with
        # the level below it fused: 9/7 + ICT + 24-bit pack), timed alone
        # with events on its stream after the plan's coarser levels, for the
        # throughput plans (P images per launch, rotating over the plans)
        # and for one image of one set.
        def last_launch_us(plan_list):
            nl = plan_list[0].num_launches
            k_ms = []
            for it in range(args.warmup + args.steps):
                p_ = plan_list[it % len(plan_list)]
                p_.execute_launches(0, nl - 1, sh)
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                p_.execute_launches(nl - 1, 1, sh)
                e1.record(stream)
                if it >= args.warmup:
                    k_ms.append((e0, e1))
            torch.cuda.synchronize()
            return float(np.mean([a_.elapsed_time(b_) for a_, b_ in k_ms])) * 1e3

        k_us = last_launch_us(tplans)
        k_us_one = last_launch_us([sets[0]["plan"]])
        # algorithmic bytes of that kernel per image: it reads every
        # coefficient of the two finest levels' detail bands and the
        # level-2 output (their LL), W*H*3 words in all, and writes every
        # 24-bit pixel once
        k_img = W * H * 3 * 4 + W * H * 3
        k_bytes = P * k_img
        k_gbs = k_bytes / (k_us * 1e-6) / 1e9
        k_gbs_one = k_img / (k_us_one * 1e-6) / 1e9
        # measured HBM bytes of the same launch (rocprofv3 FETCH_SIZE and
        # WRITE_SIZE, calibrated), scaled from the profiled launch's image
        # count to this one's
        traffic = None
        prof = ROOT / "profiles" / "fused_idwt97_ict_4096x2160x3.json"
        if prof.exists() and (W, H) == (4096, 2160):
            pj = json.loads(prof.read_text())
            for kk in pj["kernels"]:
                if kk.get("final_level") and kk.get("images_per_launch"):
                    traffic = int((kk["read_bytes"] + kk["write_bytes"]) * P / kk["images_per_launch"])
        roofline = {"bound": "hbm", "achieved": round(k_gbs, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(k_gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
                    "traffic_source": "profiles/fused_idwt97_ict_4096x2160x3.json" if traffic else None,
                    "kernel": "fused2_step_kernel<137> (final two 9/7 levels, ICT, 24-bit pack; two-strip items)",
                    "kernel_us": round(k_us, 2), "images_per_launch": P,
                    "algorithmic_bytes_per_launch": k_bytes,
                    "kernel_us_one_image": round(k_us_one, 2),
                    "achieved_one_image": round(k_gbs_one, 1),
                    "timing": f"events around the launch on its stream, rotating over {len(tplans)} "
                              f"plans of {P} images; and one image of one set"}
        pipeline_gbs = compulsory / (ms_step * 1e-3) / 1e9
        log(f"final two-level kernel: {k_us:.1f} us for {P} images ({k_gbs:.0f} GB/s), "
            f"{k_us_one:.1f} us for one image; stream of images {pipeline_gbs:.0f} GB/s compulsory")

        # parity of a timed output: set 0's pixels (last written by a timed
        # throughput decode) against the oracle's decode of its coefficients
        sys.path.insert(0, str(ROOT / "oracle"))
        import oracle  # checker only
        expect = oracle.decode_tile(planes, levels=L, filt="97", out_format=2)[1]
        got = sets[0]["out"].cpu().numpy()
        parity = bool(np.array_equal(got, expect))
        log(f"timed output equals the oracle: {parity}")
        assert parity, "GPU output of the timed decode differs from the oracle"
        set_bytes = compulsory + plan.scratch_bytes
        rec_single = {"images_per_plan": P, "plans_in_flight": min(S, nexec), "images_timed": images,
                      "us_per_image_by_launch_mode": {k: round(v[0] * 1e3, 2) for k, v in thr.items()},
                      "one_image_at_a_time_us": {k: round(v * 1e3, 2) for k, v in lat.items()},
                      "latency_us_rotating": round(lat_rot * 1e3, 2),
                      "latency_us_one_set": round(lat_one * 1e3, 2),
                      "Mpixel_s_one_at_a_time": round(W * H / (lat_rot * 1e-3) / 1e6, 1),
                      "buffer_sets": len(sets), "bytes_per_set": set_bytes,
                      "work_items_one_image": plan.num_items, "launches": plan.num_launches,
                      "compulsory_bytes": compulsory,
                      "decode_GBs_compulsory": round(pipeline_gbs, 1),
                      "decode_frac_of_peak": round(pipeline_gbs / HBM_PEAK_GBS, 4),
                      "decode_frac_of_peak_one_at_a_time": round(
                          compulsory / (lat_rot * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                      "timed_output_equals_oracle": parity}
    plain = None
    if "plain" in legs and "single" in legs:
        log("plain path (BMI_comp_idwt x3 + BMI_tile_ICT)")
        coef = sets[0]["coef"]
        work = [torch.empty_like(t) for t in coef]
        temp = torch.empty(W * H + 64, dtype=torch.int32, device=dev)
        pout = torch.zeros((H, W * 3), dtype=torch.uint8, device=dev)
        pdesc = pkg.tile_desc(coef_ptrs=[t.data_ptr() for t in work], coef_pitch=W, width=W,
                              height=H, levels=L, method=pkg.W9X7, ncomp=3,
                              coef_kind=pkg.COEF_FLOAT, bit_depth=8, out_format=pkg.FMT_24BITS_RGB,
                              out_ptr=pout.data_ptr(), out_pitch_px=W)
        times = []
        for it in range(args.warmup + args.steps):
            for a, b in zip(work, coef):
                a.copy_(b)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            pkg.plain_decode(pdesc, temp.data_ptr(), temp.numel() * 4, sh)
            e1.record(stream)
            torch.cuda.synchronize()
            if it >= args.warmup:
                times.append(e0.elapsed_time(e1))
        pms = float(np.mean(times))
        plain_ok = bool(np.array_equal(pout.cpu().numpy(), expect))
        plain = {"ms_per_step": round(pms, 4), "Mpixel_s": round(W * H / (pms * 1e-3) / 1e6, 1),
                 "fused_speedup": round(pms / ms_step, 2),
                 "fused_speedup_one_at_a_time": round(pms / lat_rot, 2),
                 "launches_per_step": 3 * (2 * L) + 1, "output_equals_oracle": plain_ok}
        log(f"plain: {pms * 1e3:.1f} us/step, fused speedup {pms / ms_step:.2f}x, "
            f"equal to oracle {plain_ok}")
        del work, temp, pout

    batched = None
    if "batched" in legs:
        log("batched: 256 mixed tiles, autotuned, hipGraph")
        batched = batched_leg(pkg, wl, torch, dev, stream, max(5, args.steps // 5), 2,
                              inflight=args.batch_inflight)
        log(f"batched: {batched['ms_per_batch']} ms, {batched['Mpixel_s']} Mpixel/s, "
            f"{batched['frac_of_peak'] * 100:.1f}% of peak; plain {batched['plain_Mpixel_s']} "
            f"Mpixel/s ({batched['speedup_vs_plain']}x)")

    serving = None
    if "serving" in legs:
        rates = [float(r) for r in args.serve_rates.split(",") if r]
        serving = serving_leg(pkg, wl, torch, dev, rates, args.serve_requests)

    t1 = None
    codestream = None
    if "t1" in legs:
        log("tier 1: GPU code-block decode + fused IDWT/ICT")
        t1 = t1_leg(pkg, wl, torch, dev, stream, max(5, args.steps // 5), 2)
        log(f"t1: {t1['t1_ms']} ms ({t1['t1_Mcoef_s']} Mcoef/s), image {t1['ms_per_image']} ms")
        log("codestream: host tier 2 + H2D + GPU tier 1 + IDWT/ICT")
        codestream = codestream_leg(pkg, wl, torch, dev, stream, max(5, args.steps // 5), 2,
                                    t1["_img"], t1["_out"])
        log(f"codestream: host {codestream['host_parse_ms']}+{codestream['host_tile_blocks_ms']} ms, "
            f"h2d {codestream['h2d_ms']} ms, gpu {codestream['gpu_ms']} ms, "
            f"equal={codestream['output_equals_t1_leg']}")

    if args.calib_copy:
        # a plain 256 MiB device copy: the reference for scaling HBM counters
        a = torch.empty(64 << 20, dtype=torch.int32, device=dev)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        log("calibration copy: 3 x 268435456 bytes")
        del a, b

    cpu = None
    if "cpu" in legs and planes is not None:
        log("cpu baseline (multithreaded CPU restatement, all cores and one core)")
        cpu = cpu_baseline(planes, W, H, expect)
        log(f"cpu: {cpu['value']} Mpixel/s on {cpu['cores']} of {cpu['nproc']} cores "
            f"({cpu['cpu_model']}); one core {cpu['one_core']['value']} Mpixel/s")
        if t1 is not None:
            cpu["t1"] = t1_cpu_baseline(t1["_img"])
            log(f"cpu t1: {cpu['t1']['value']} Mcoef/s")
    if t1 is not None:
        t1.pop("_img", None)
        t1.pop("_out", None)

    rec = {
        "metric": METRIC,
        "value": round(mpix, 1) if mpix else None,
        "unit": "Mpixel/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4) if ms_step else None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": dict({"workload": f"{W}x{H}x3 9/7 5-level IDWT + inverse ICT -> 24-bit BGR per "
                                    "image, coefficients resident in HBM, fused engine (one launch "
                                    "per level step, the last two levels in one, "
                                    f"{launch_mode if mpix else '-'}); a stream of such images, "
                                    + (f"all {args.steps} timed images in one plan"
                                       if args.steps <= args.images_per_plan else
                                       f"{args.images_per_plan} per plan, {args.inflight} plans in "
                                       "flight on their own streams")
                                    + f", rotating over {max(args.sets, args.images_per_plan * args.inflight)} "
                                    "buffer sets; ms_per_step is per image"}, **rec_single),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "plain_gpu": plain,
        "batched": batched,
        "serving": serving,
        "t1": t1,
        "codestream": codestream,
    }
    for p in plans:
        p.close()
    print(json.dumps(rec), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
