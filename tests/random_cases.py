"""Seeded random tile cases shared by the CPU address check
(test_kernel_addresses.py) and the GPU parity sweep (test_gpu_random.py):
sizes from one sample to a few hundred, offsets 0..3, 1..5 levels, both
filters and 5/3 flavours, 1 or 3 components, planes / 24-bit / 48-bit /
merged outputs, segment heights and the two-level step on or off."""
import numpy as np

SIDES = [1, 2, 3, 5, 8, 13, 21, 34, 55, 89, 144, 233, 300]
# (ncomp, out_format, bit_depth): 7 = no packing (planes), 2 = 24-bit RGB,
# 1 = 48-bit RGB, 6 = 24-bit merge (one component as grey), 0 = 32-bit
# ARGB, 3 = 32-bit merge of four components
OUTPUTS = [(3, 2, 8), (3, 1, 12), (1, 7, 8), (3, 7, 8), (1, 6, 8), (3, 0, 8), (4, 3, 8)]


def random_cases(n=120, seed=11):
    rnd = np.random.default_rng(seed)
    cases = []
    for _ in range(n):
        w, h = int(rnd.choice(SIDES)), int(rnd.choice(SIDES))
        levels = int(rnd.integers(1, 6))
        x0, y0 = int(rnd.integers(0, 4)), int(rnd.integers(0, 4))
        filt = 53 if rnd.random() < 0.5 else 97
        ncomp, fmt, bd = OUTPUTS[int(rnd.integers(0, len(OUTPUTS)))]
        r53 = int(rnd.integers(0, 2))
        seg = int(rnd.choice([0, 2, 4, 6, 10]))
        fuse = int(rnd.choice([0, -1]))
        cases.append((w, h, levels, x0, y0, filt, ncomp, fmt, bd, r53, seg, fuse))
    return cases


def case_id(c):
    return "w%d_h%d_L%d_x%d_y%d_f%d_c%d_fmt%d_bd%d_r%d_s%d_fu%d" % c
