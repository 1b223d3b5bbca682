"""Write the JPEG 2000 codestream fixtures under tests/golden/j2k/.

The streams come from an independent encoder, OpenJPEG 2.5 (through
Pillow's JPEG 2000 plugin), so the parser, tier-1 and IDWT are checked on
codestreams this project did not write.  Every fixture stores the encoded
file and the source image (j2k/<name>.j2k or .jp2, images in
j2k/images.npz); lossless streams must decode back to the source exactly
(with the ITU-T T.800 5/3 synthesis); for lossy ones OpenJPEG's own decode
is stored instead (<name>_opj), to be met within a stated tolerance.

Run from the repository root:  python tests/golden/make_j2k.py
"""
from __future__ import annotations

import io
import pathlib

import numpy as np

OUT = pathlib.Path(__file__).resolve().parent / "j2k"


def smooth(rng, shape, scale=40.0):
    """Band-limited test image (sums of a few random cosines plus noise)."""
    h, w = shape[:2]
    nc = shape[2] if len(shape) > 2 else 1
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    out = np.zeros((h, w, nc))
    for c in range(nc):
        v = np.full((h, w), 128.0)
        for _ in range(4):
            fx, fy, ph = rng.uniform(0, 0.08), rng.uniform(0, 0.08), rng.uniform(0, 6.3)
            v += scale * np.cos(fx * x + fy * y + ph)
        v += rng.normal(0, 6, (h, w))
        out[..., c] = v
    out = np.clip(np.round(out), 0, 255).astype(np.uint8)
    return out if len(shape) > 2 else out[..., 0]


# name -> (shape, kind, encoder options); kind 'noise' | 'smooth' | 'u16'
CASES = {
    "grey_37x53": ((37, 53), "noise", dict()),
    "rgb_64_mct": ((64, 64, 3), "noise", dict(mct=1)),
    "rgb_tiles_offsets": ((100, 90, 3), "noise",
                          dict(mct=1, tile_size=(32, 32), tile_offset=(7, 11), offset=(9, 13))),
    "grey_layers_lrcp": ((90, 70), "smooth", dict(quality_layers=[40, 20, 0], quality_mode="rates")),
    "grey_layers_rlcp": ((70, 60), "smooth", dict(progression="RLCP", quality_layers=[50, 10, 0])),
    "rgb_rpcl_prec32": ((90, 70, 3), "noise", dict(progression="RPCL", precinct_size=(32, 32), mct=1)),
    "grey_pcrl_cb16x32": ((130, 99), "noise",
                          dict(progression="PCRL", precinct_size=(64, 64), codeblock_size=(16, 32))),
    "rgb_cprl_4lev": ((130, 99, 3), "smooth",
                      dict(progression="CPRL", precinct_size=(32, 64), num_resolutions=4, mct=1)),
    "rgba_mct": ((64, 48, 4), "noise", dict(mct=1)),
    "grey16": ((40, 50), "u16", dict()),
    "rgb_nomct_jp2": ((33, 40, 3), "noise", dict(mct=0, no_jp2=False)),
    "one_pixel": ((1, 1), "noise", dict(num_resolutions=1)),
    "tiny_3x2_2lev": ((3, 2), "noise", dict(num_resolutions=2)),
    # lossy 9/7
    "rgb97_layers_rpcl": ((96, 80, 3), "smooth",
                          dict(irreversible=True, mct=1, quality_layers=[20, 5], quality_mode="rates",
                               progression="RPCL", precinct_size=(64, 64))),
    "rgb97_tiles_pcrl": ((200, 150, 3), "smooth",
                         dict(irreversible=True, mct=1, tile_size=(128, 96), progression="PCRL",
                              precinct_size=(32, 32), quality_layers=[30, 10, 3])),
    "grey97": ((96, 80), "smooth", dict(irreversible=True, quality_layers=[8], quality_mode="rates")),
    "rgb97_256": ((192, 256, 3), "smooth", dict(irreversible=True, mct=1, quality_layers=[12])),
}


def main() -> None:
    from PIL import Image
    OUT.mkdir(exist_ok=True)
    rng = np.random.default_rng(2000)
    images = {}
    for name, (shape, kind, opts) in CASES.items():
        opts = dict(opts)
        jp2 = not opts.pop("no_jp2", True)
        if kind == "u16":
            img = rng.integers(0, 65536, shape, dtype=np.uint16)
            pil = Image.fromarray(img, mode="I;16")
        else:
            img = rng.integers(0, 256, shape, dtype=np.uint8) if kind == "noise" else smooth(rng, shape)
            pil = Image.fromarray(img)
        b = io.BytesIO()
        pil.save(b, format="JPEG2000", no_jp2=not jp2, **opts)
        data = b.getvalue()
        (OUT / (name + (".jp2" if jp2 else ".j2k"))).write_bytes(data)
        if opts.get("irreversible"):
            images[name + "_opj"] = np.asarray(Image.open(io.BytesIO(data)))
        else:
            images[name] = img
        print(f"{name}: {len(data)} bytes")
    np.savez_compressed(OUT / "images.npz", **images)


if __name__ == "__main__":
    main()
