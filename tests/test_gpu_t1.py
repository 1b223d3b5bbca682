"""GPU tier-1 decoding (one code-block per wavefront) against the oracle's
restatement of decode_code_block, bit for bit: random code-blocks of every
size class, orientation and kind (9/7 words, 5/3 integers), truncated pass
counts, the single-bit-plane case and empty blocks; then whole images whose
code-blocks are decoded on the GPU straight into the coefficient planes and
fed to the fused IDWT + colour transform."""
import ctypes

import numpy as np
import pytest

from gpu_helpers import diff_report, fetch, make_case

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(77)


@pytest.fixture(scope="module", autouse=True)
def _gpu(pkg):
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    pkg.lib()


def upload_blocks(pkg, blocks):
    """Host BmiT1Block list -> device tensor holding the table."""
    import torch
    arr = (pkg.BmiT1Block * len(blocks))(*blocks)
    raw = np.frombuffer(bytes(arr), dtype=np.uint8).copy()
    return torch.from_numpy(raw).to("cuda")


def upload_data(data: bytes):
    import torch
    return torch.from_numpy(np.frombuffer(data + b"\xff" * 16, dtype=np.uint8).copy()).to("cuda")


def _random_blocks(workloads):
    cases = []
    sizes = [(1, 1), (1, 64), (64, 1), (2, 3), (3, 5), (4, 4), (5, 7), (8, 8), (13, 29),
             (31, 17), (63, 64), (64, 63), (64, 64), (32, 64), (64, 32)]
    for i in range(180):
        h, w = sizes[i % len(sizes)]
        orient = i % 4
        kind = (i // 4) % 2  # 0 words (9/7), 1 integers (5/3)
        amp = [0.6, 3.0, 12.0, 80.0, 700.0][i % 5]
        q = np.round(RNG.laplace(0, amp, (h, w))).astype(np.int32)
        if i % 7 == 3:
            q[RNG.random((h, w)) < 0.8] = 0
        if i % 11 == 5:
            q[:] = 0
        K = int(np.abs(q).max()).bit_length()
        mb = max(K, 1) + int(RNG.integers(0, 3))
        limit = 0
        if K > 1 and i % 3 == 1:
            limit = int(RNG.integers(1, 3 * K - 2))
        data, npass, miss = workloads.t1_encode_block(q, mb, orient, limit)
        numbps = ((mb << 11) | int(RNG.integers(0, 2048))) if kind == 0 else mb
        cases.append(dict(q=q, h=h, w=w, orient=orient, kind=kind, numbps=numbps, missing=miss,
                          num_pass=npass, data=data))
    return cases


def test_random_code_blocks_match_oracle(pkg, workloads, orc):
    import torch
    cases = _random_blocks(workloads)
    pitch = 64 * 16
    rows = (len(cases) + 15) // 16
    out = torch.full((rows * 64, pitch), -7, dtype=torch.int32, device="cuda")
    blocks, chunks, off = [], [], 0
    for i, c in enumerate(cases):
        bx, by = 64 * (i % 16), 64 * (i // 16)
        b = pkg.BmiT1Block()
        b.data_offset, b.len = off, len(c["data"])
        b.width, b.height, b.orient = c["w"], c["h"], c["orient"]
        b.numbps, b.missing, b.num_pass, b.kind = c["numbps"], c["missing"], c["num_pass"], c["kind"]
        b.out = out.data_ptr() + 4 * (by * pitch + bx)
        b.out_pitch = pitch
        blocks.append(b)
        chunks.append(c["data"])
        off += len(c["data"])
    pkg.t1_check_blocks(blocks)
    dblocks = upload_blocks(pkg, blocks)
    ddata = upload_data(b"".join(chunks))
    torch.cuda.synchronize()
    print(f"decoding {len(blocks)} code-blocks, {off} bytes", flush=True)
    pkg.t1_decode_blocks(dblocks.data_ptr(), len(blocks), ddata.data_ptr(), 0)
    got_all = out.cpu().numpy()
    bad = []
    for i, c in enumerate(cases):
        bx, by = 64 * (i % 16), 64 * (i // 16)
        got = got_all[by:by + c["h"], bx:bx + c["w"]]
        exp = orc.t1_decode_block(c["data"], c["w"], c["h"], c["orient"], c["numbps"],
                                  c["missing"], c["num_pass"])
        if c["kind"] == 1:
            exp = orc.t1_words_to_int(exp, c["numbps"])
        if not np.array_equal(got, exp):
            bad.append((i, c["h"], c["w"], c["orient"], c["kind"], c["num_pass"]))
        # nothing written outside the block
        if c["w"] < 64:
            assert (got_all[by:by + c["h"], bx + c["w"]:bx + 64] == -7).all(), i
    assert not bad, f"{len(bad)} blocks differ, first {bad[:5]}"


def test_style_code_blocks_match_oracle(pkg, workloads, orc):
    """Every code-block style combination (T.800 Table A.19: bypass, reset,
    termall, causal, predictable termination, segmentation symbols) on
    blocks of several shapes, whole and truncated: GPU words == oracle
    words.  Segmented styles carry their segment-length tables inside the
    data buffer (seg_offset)."""
    import torch
    rng = np.random.default_rng(2024)
    sizes = [(64, 64), (13, 29), (4, 4), (1, 64), (64, 1), (33, 64), (7, 5)]
    cases = []
    for style in range(64):
        for j in range(3):
            h, w = sizes[(style + 3 * j) % len(sizes)]
            q = np.round(rng.laplace(0, [4.0, 40.0, 400.0][j], (h, w))).astype(np.int32)
            K = int(np.abs(q).max()).bit_length()
            mb = max(K, 1) + int(rng.integers(0, 2))
            limit = int(rng.integers(1, 3 * K - 2)) if (K > 4 and j == 2) else 0
            orient = int(rng.integers(0, 4))
            data, npass, miss, segs = workloads.t1_encode_block(q, mb, orient, limit, style=style,
                                                                segments=True)
            kind = (style + j) % 2
            numbps = ((mb << 11) | int(rng.integers(0, 2048))) if kind == 0 else mb
            cases.append(dict(h=h, w=w, orient=orient, kind=kind, numbps=numbps, missing=miss,
                              num_pass=npass, data=data, style=style, segs=segs))
    tables = [np.asarray(c["segs"], np.int32) for c in cases]
    tab_bytes = -(-sum(4 * len(t) for t in tables) // 16) * 16
    pitch = 64 * 16
    rows = (len(cases) + 15) // 16
    out = torch.full((rows * 64, pitch), -7, dtype=torch.int32, device="cuda")
    blocks, chunks, off, toff = [], [], tab_bytes, 0
    for i, c in enumerate(cases):
        bx, by = 64 * (i % 16), 64 * (i // 16)
        b = pkg.BmiT1Block()
        b.data_offset, b.len = off, len(c["data"])
        b.width, b.height, b.orient = c["w"], c["h"], c["orient"]
        b.numbps, b.missing, b.num_pass, b.kind = c["numbps"], c["missing"], c["num_pass"], c["kind"]
        b.out = out.data_ptr() + 4 * (by * pitch + bx)
        b.out_pitch = pitch
        b.style = c["style"]
        if c["style"] & 5:
            b.nseg, b.seg_offset = len(c["segs"]), toff
            toff += 4 * len(c["segs"])
        else:
            b.nseg = 1
        blocks.append(b)
        chunks.append(c["data"])
        off += len(c["data"])
    table = b"".join(tables[i].tobytes() for i, c in enumerate(cases) if c["style"] & 5)
    pkg.t1_check_blocks(blocks)
    dblocks = upload_blocks(pkg, blocks)
    ddata = upload_data(table + b"\0" * (tab_bytes - len(table)) + b"".join(chunks))
    torch.cuda.synchronize()
    print(f"decoding {len(blocks)} style code-blocks, {off} bytes", flush=True)
    pkg.t1_decode_blocks(dblocks.data_ptr(), len(blocks), ddata.data_ptr(), 0)
    got_all = out.cpu().numpy()
    bad = []
    for i, c in enumerate(cases):
        bx, by = 64 * (i % 16), 64 * (i // 16)
        got = got_all[by:by + c["h"], bx:bx + c["w"]]
        exp = orc.t1_decode_block_ex(c["data"], c["w"], c["h"], c["orient"], c["numbps"],
                                     c["missing"], c["num_pass"], c["style"], c["segs"])
        if c["kind"] == 1:
            exp = orc.t1_words_to_int(exp, c["numbps"])
        if not np.array_equal(got, exp):
            bad.append((i, c["style"], c["h"], c["w"], c["orient"], c["kind"], c["num_pass"]))
    assert not bad, f"{len(bad)} blocks differ, first {bad[:5]}"


def test_check_blocks_rejects_styles(pkg):
    b = pkg.BmiT1Block()
    b.width, b.height, b.out, b.out_pitch, b.kind, b.num_pass = 8, 8, 1, 8, 0, 12
    b.style = 0x40                       # not a Part 1 switch
    with pytest.raises(pkg.BmiError):
        pkg.t1_check_blocks([b])
    b.style, b.nseg = 0x04, 11           # termall: 12 passes open 12 segments
    with pytest.raises(pkg.BmiError):
        pkg.t1_check_blocks([b])
    b.nseg, b.seg_offset = 12, 2         # table not 4-byte aligned
    with pytest.raises(pkg.BmiError):
        pkg.t1_check_blocks([b])
    b.seg_offset = 8
    pkg.t1_check_blocks([b])


def test_check_blocks_rejects(pkg):
    b = pkg.BmiT1Block()
    b.width, b.height, b.out, b.out_pitch, b.kind = 65, 4, 1, 128, 0
    with pytest.raises(pkg.BmiError):
        pkg.t1_check_blocks([b])
    b.width, b.kind, b.numbps = 8, 1, 0
    with pytest.raises(pkg.BmiError):
        pkg.t1_check_blocks([b])


def _t1_into_planes(pkg, img):
    """Decode every code-block of a t1_image on the GPU into fresh device
    coefficient planes (one per component, pitch = width)."""
    import torch
    w, h, nc = img["w"], img["h"], img["ncomp"]
    planes = [torch.zeros((h, w), dtype=torch.int32, device="cuda") for _ in range(nc)]
    kind = pkg.T1_WORDS if img["filt"] == "97" else pkg.T1_INT
    blocks = []
    for bd in img["blocks"]:
        b = pkg.BmiT1Block()
        b.data_offset, b.len = bd["offset"], bd["len"]
        b.width, b.height, b.orient = bd["w"], bd["h"], bd["orient"]
        b.numbps, b.missing, b.num_pass, b.kind = bd["numbps"], bd["missing"], bd["num_pass"], kind
        b.out = planes[bd["comp"]].data_ptr() + 4 * (bd["y"] * w + bd["x"])
        b.out_pitch = w
        blocks.append(b)
    pkg.t1_check_blocks(blocks)
    dblocks = upload_blocks(pkg, blocks)
    ddata = torch.from_numpy(np.frombuffer(img["data"], dtype=np.uint8).copy()).to("cuda")
    pkg.t1_decode_blocks(dblocks.data_ptr(), len(blocks), ddata.data_ptr(), 0)
    torch.cuda.synchronize()
    return planes, (dblocks, ddata)


def _oracle_words(orc, img):
    w, h, nc = img["w"], img["h"], img["ncomp"]
    words = np.zeros((nc, h, w), np.int32)
    for bd in img["blocks"]:
        wd = orc.t1_decode_block(img["data"][bd["offset"]:bd["offset"] + bd["len"]], bd["w"],
                                 bd["h"], bd["orient"], bd["numbps"], bd["missing"],
                                 bd["num_pass"])
        if img["filt"] == "53":
            wd = orc.t1_words_to_int(wd, bd["numbps"])
        words[bd["comp"], bd["y"]:bd["y"] + bd["h"], bd["x"]:bd["x"] + bd["w"]] = wd
    return words


@pytest.mark.parametrize("filt", ["97", "53"])
def test_image_t1_then_idwt_ict(pkg, workloads, orc, filt):
    """Full tile decode on the GPU: T1 into the Mallat planes, then the fused
    IDWT + RCT/ICT to 24-bit BGR; against the oracle's T1 + pipeline."""
    import torch
    w, h, L = 301, 227, 4
    img = workloads.t1_image(w, h, L, 3, seed=5, filt=filt)
    planes, keep = _t1_into_planes(pkg, img)
    words = _oracle_words(orc, img)
    for c in range(3):
        got = planes[c].cpu().numpy()
        assert np.array_equal(got, words[c]), f"comp {c}: {diff_report(got, words[c])}"
    out = torch.zeros((h, w * 3), dtype=torch.uint8, device="cuda")
    desc = pkg.tile_desc(coef_ptrs=[p.data_ptr() for p in planes], coef_pitch=w, width=w,
                         height=h, levels=L, method=pkg.W9X7 if filt == "97" else pkg.W5X3,
                         ncomp=3, coef_kind=pkg.COEF_SIGNMAG, steps=img["steps"],
                         out_format=pkg.FMT_24BITS_RGB, out_ptr=out.data_ptr(), out_pitch_px=w)
    plan = pkg.BatchPlan([desc])
    plan.execute(0)
    torch.cuda.synchronize()
    plan.close()
    src = words.view(np.uint32) if filt == "97" else words
    _, exp = orc.decode_tile(src, levels=L, filt=filt, out_format=2, steps=img["steps"],
                             coef_kind=orc.COEF_SIGNMAG)
    got = out.cpu().numpy()
    assert np.array_equal(got, exp), diff_report(got, exp)


def test_arbitrary_bytes_decode_like_the_oracle(pkg, orc):
    """Code streams of random bytes (with runs of 0xFF and marker-like byte
    pairs) and random pass counts: the GPU decoder reads only inside the
    segment, always terminates, and produces the oracle's words."""
    import torch
    rng = np.random.default_rng(91)
    cases = []
    for i in range(96):
        h, w = int(rng.integers(1, 65)), int(rng.integers(1, 65))
        n = int(rng.integers(0, 600))
        data = rng.integers(0, 256, n).astype(np.uint8)
        if n and i % 3 == 0:
            data[rng.random(n) < 0.2] = 0xFF
        if n > 4 and i % 5 == 1:
            data[n // 2] = 0xFF
            data[n // 2 + 1] = 0x90 + (i % 0x60)
        kind = i % 2
        mb = int(rng.integers(1, 16))
        numbps = ((mb << 11) | int(rng.integers(0, 2048))) if kind == 0 else mb
        missing = int(rng.integers(0, max(1, mb)))
        # up to two passes more than the band's bit-planes hold
        num_pass = int(rng.integers(0, 3 * (mb - missing) + 1))
        cases.append(dict(h=h, w=w, data=data.tobytes(), kind=kind, numbps=numbps, missing=missing,
                          num_pass=num_pass, orient=i % 4))
    pitch = 64 * 16
    out = torch.full((64 * ((len(cases) + 15) // 16), pitch), 5, dtype=torch.int32, device="cuda")
    blocks, chunks, off = [], [], 0
    for i, c in enumerate(cases):
        b = pkg.BmiT1Block()
        b.data_offset, b.len = off, len(c["data"])
        b.width, b.height, b.orient = c["w"], c["h"], c["orient"]
        b.numbps, b.missing, b.num_pass, b.kind = c["numbps"], c["missing"], c["num_pass"], c["kind"]
        b.out = out.data_ptr() + 4 * (64 * (i // 16) * pitch + 64 * (i % 16))
        b.out_pitch = pitch
        blocks.append(b)
        chunks.append(c["data"])
        off += len(c["data"])
    pkg.t1_check_blocks(blocks)
    dblocks = upload_blocks(pkg, blocks)
    ddata = upload_data(b"".join(chunks))
    pkg.t1_decode_blocks(dblocks.data_ptr(), len(blocks), ddata.data_ptr(), 0)
    got_all = out.cpu().numpy()
    for i, c in enumerate(cases):
        y0, x0 = 64 * (i // 16), 64 * (i % 16)
        got = got_all[y0:y0 + c["h"], x0:x0 + c["w"]]
        exp = orc.t1_decode_block(c["data"], c["w"], c["h"], c["orient"], c["numbps"],
                                  c["missing"], c["num_pass"])
        if c["kind"] == 1:
            exp = orc.t1_words_to_int(exp, c["numbps"])
        assert np.array_equal(got, exp), (i, c["h"], c["w"], c["num_pass"], diff_report(got, exp))
