"""The C-ABI library loads and exports every function include/*.h declares
(no GPU calls), and the ctypes structs match the C layout."""
import ctypes
import pathlib
import re

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]


def declared_functions():
    names = []
    for h in sorted((ROOT / "include").glob("*.h")):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;",
                             text, flags=re.M):
            name = m.group(1)
            if name in ("if", "while", "sizeof", "return"):
                continue
            names.append(name)
    return sorted(set(names))


def test_header_declares_reference_entry_points():
    names = declared_functions()
    # decoder/cudainterface/include/gpulib.h:24-57
    for ref in ["GpuInit", "GpuRemove", "MallocGpuBuf", "FreeGpuBuf", "CreateANewStream",
                "DestroyStream", "GpuSync", "Gpu2DCopy", "Gpu2DCopyWithDiffPitch",
                "GpuLinearMemCopy", "BMI_comp_idwt", "BMI_resolution_idwt", "BMI_tile_MCT",
                "BMI_tile_ICT", "BMI_tile_Merge", "BMI_tile_One_Component", "BMI_Comp_DC_SHIFT"]:
        assert ref in names, ref


def test_library_exports_every_declared_symbol(pkg):
    lib = pkg.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_layouts(pkg):
    # offsets that C computes for the ABI structs (x86-64 / amdgcn host)
    assert ctypes.sizeof(pkg.GpuInfo) == 4 * 2 + 256 + 4 * 5
    d = pkg.BmiTileDesc
    assert d.coef.offset == 0
    assert d.coef_pitch.offset == 32
    assert d.steps.offset == 92
    assert d.out.offset == 400
    assert ctypes.sizeof(pkg.BmiTuning) == 32


def test_error_strings_without_gpu(pkg):
    lib = pkg.lib()
    assert lib.BMI_error_string(0) == b"ok"
    assert lib.BMI_error_string(-5) == b"unsupported mode or format"
    lib.BMI_set_r53_mode(1)
    assert lib.BMI_get_r53_mode() == 1
    lib.BMI_set_r53_mode(0)
    assert lib.BMI_get_r53_mode() == 0


def test_plan_rejects_bad_descriptors_without_gpu(pkg):
    """Validation happens on the host before any device call."""
    bad = pkg.tile_desc(coef_ptrs=[0], coef_pitch=4, width=8, height=8, levels=9,
                        method=pkg.W5X3, ncomp=1, out_format=pkg.FMT_NONE)
    with pytest.raises(pkg.BmiError):
        pkg.BatchPlan([bad])
    rgb_one_comp = pkg.tile_desc(coef_ptrs=[0], coef_pitch=8, width=8, height=8, levels=2,
                                 method=pkg.W5X3, ncomp=1, out_format=pkg.FMT_24BITS_RGB,
                                 out_ptr=1, out_pitch_px=8)
    with pytest.raises(pkg.BmiError):
        pkg.BatchPlan([rgb_one_comp])


def test_t1_order_blocks_longest_first(pkg):
    """BMI_t1_order_blocks: a stable permutation of the table, longest data
    first (host only)."""
    import ctypes
    lens = [5, 90, 0, 90, 33, 7, 33, 1]
    arr = (pkg.BmiT1Block * len(lens))()
    for i, n in enumerate(lens):
        arr[i].len = n
        arr[i].data_offset = i          # identifies the entry
    pkg.t1_order_blocks(arr)
    got = [(arr[i].len, arr[i].data_offset) for i in range(len(lens))]
    assert got == sorted(((n, i) for i, n in enumerate(lens)), key=lambda x: -x[0])
    assert pkg.lib().BMI_t1_order_blocks(None, 0) == 0
    assert pkg.lib().BMI_t1_order_blocks(None, 3) != 0


def _c_layout(tmp_path, structs):
    """sizeof / offsetof of the named fields as the C compiler lays them out
    from include/bmi_types.h and include/bmi_gpulib.h."""
    import subprocess
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "bmi_gpulib.h"',
             'int main(void) {']
    for sname, fields in structs.items():
        lines.append(f'printf("{sname} size %zu\\n", sizeof({sname}));')
        for f in fields:
            lines.append(f'printf("{sname} {f} %zu\\n", offsetof({sname}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", str(ROOT / "include"), str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    res = {}
    for ln in out.splitlines():
        s, f, v = ln.split()
        res[(s, f)] = int(v)
    return res


def test_ctypes_mirrors_match_c_layout(pkg, tmp_path):
    """Every field of the ctypes mirrors of the ABI structs (the reference's
    cudainterfacedatatype.h types and the additive descriptors) sits at the
    offset the C compiler gives it."""
    mirrors = {
        "CudaBuf": pkg.CudaBuf, "SubbandInfo_c": pkg.SubbandInfo_c,
        "CompInfo_c": pkg.CompInfo_c, "TileInfo_c": pkg.TileInfo_c, "GpuInfo": pkg.GpuInfo,
        "BmiTileDesc": pkg.BmiTileDesc, "BmiTuning": pkg.BmiTuning,
        "BmiServerConfig": pkg.BmiServerConfig,
    }
    anon = {"miAbsStep": "mfAbsStepNorm"}

    def fields(cls):
        out = []
        for name, _t in cls._fields_:
            if name == "_u":
                out += [n for n, _ in _t._fields_]
            else:
                out.append(name)
        return out

    structs = {s: [f for f in fields(c) if f not in anon] for s, c in mirrors.items()}
    lay = _c_layout(tmp_path, structs)
    for s, cls in mirrors.items():
        assert ctypes.sizeof(cls) == lay[(s, "size")], s
        for f in structs[s]:
            assert getattr(cls, f).offset == lay[(s, f)], (s, f)
