"""Committed golden fixtures (tests/golden/golden.npz, written by
tests/golden/make_golden.py): the oracle must reproduce them on the CPU,
and the GPU paths (plain, fused, tier 1) must reproduce them bit for bit
without consulting the oracle at run time."""
import pathlib

import numpy as np
import pytest

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden" / "golden.npz"


@pytest.fixture(scope="module")
def G():
    with np.load(GOLDEN, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _tiles(G):
    for k in G:
        if k.startswith("tile") and k.endswith("_in"):
            key = k[:-3]
            filt = key[4:6]
            w, h = (int(v) for v in key.split("_")[1].split("x"))
            L = int(key.split("_")[2][1:])
            x0, y0 = int(key.split("_")[3][1]), int(key.split("_")[3][2])
            yield key, filt, w, h, L, x0, y0


def test_oracle_reproduces_v53_quirk(orc, G):
    for n in range(2, 13):
        got = orc.v53_ref(G[f"v53_in_{n}"], (n + 1) // 2)
        assert np.array_equal(got, G[f"v53_out_{n}"]), n


def test_oracle_reproduces_tiles(orc, G):
    for key, filt, w, h, L, x0, y0 in _tiles(G):
        planes = G[key + "_in"]
        for mode in ([0, 1] if filt == "53" else [0]):
            rec, packed = orc.decode_tile(planes, levels=L, filt=filt, out_format=2, x0=x0,
                                          y0=y0, r53_mode=mode)
            exp = G[f"{key}_m{mode}_planes"]
            assert np.array_equal(np.stack(rec).view(np.uint32), exp.view(np.uint32)), key
            assert np.array_equal(packed, G[f"{key}_m{mode}_bgr"]), key


def test_oracle_reproduces_t1(orc, G):
    for i, (w, h, orient, numbps, miss, npass, kind) in enumerate(G["t1_params"]):
        words = orc.t1_decode_block(G[f"t1_{i}_data"].tobytes(), int(w), int(h), int(orient),
                                    int(numbps), int(miss), int(npass))
        if kind == 1:
            words = orc.t1_words_to_int(words, int(numbps))
        assert np.array_equal(words, G[f"t1_{i}_words"]), i


@pytest.mark.gpu
def test_gpu_paths_reproduce_tiles(pkg, G):
    import torch
    assert torch.cuda.is_available()
    from gpu_helpers import run_fused, run_plain
    for key, filt, w, h, L, x0, y0 in _tiles(G):
        planes = G[key + "_in"]
        for mode in ([0, 1] if filt == "53" else [0]):
            for run in (run_plain, run_fused):
                _, got = run(pkg, planes, filt=filt, levels=L, out_format=2, x0=x0, y0=y0,
                             r53_mode=mode)
                assert np.array_equal(got, G[f"{key}_m{mode}_bgr"]), (key, run.__name__)


@pytest.mark.gpu
def test_gpu_t1_reproduces_blocks(pkg, G):
    import torch
    assert torch.cuda.is_available()
    params = G["t1_params"]
    out = torch.zeros((64, 64 * len(params)), dtype=torch.int32, device="cuda")
    blocks, chunks, off = [], [], 0
    for i, (w, h, orient, numbps, miss, npass, kind) in enumerate(params):
        d = G[f"t1_{i}_data"].tobytes()
        b = pkg.BmiT1Block()
        b.data_offset, b.len = off, len(d)
        b.width, b.height, b.orient = int(w), int(h), int(orient)
        b.numbps, b.missing, b.num_pass, b.kind = int(numbps), int(miss), int(npass), int(kind)
        b.out = out.data_ptr() + 4 * 64 * i
        b.out_pitch = 64 * len(params)
        blocks.append(b)
        chunks.append(d)
        off += len(d)
    arr = (pkg.BmiT1Block * len(blocks))(*blocks)
    dblocks = torch.from_numpy(np.frombuffer(bytes(arr), np.uint8).copy()).cuda()
    ddata = torch.from_numpy(np.frombuffer(b"".join(chunks) + b"\xff" * 16, np.uint8).copy()).cuda()
    pkg.t1_decode_blocks(dblocks.data_ptr(), len(blocks), ddata.data_ptr(), 0)
    got = out.cpu().numpy()
    for i, (w, h, *_rest) in enumerate(params):
        assert np.array_equal(got[:h, 64 * i:64 * i + w], G[f"t1_{i}_words"]), i
