"""The Python binding's ctypes structures and numpy record dtypes have the
sizes and field offsets of the C header's structs (a small C program built
with the host compiler prints them), so a header change cannot drift silently
from the binding."""
import ctypes
import shutil
import subprocess
from pathlib import Path

import pytest

import clear_amd as ca

ROOT = Path(__file__).resolve().parents[1]

# (C struct, Python type, [(C field, Python field), ...])
CASES = [
    ("clr_wavefront_config", ca.WavefrontConfig,
     [(n, n) for n, _ in ca.WavefrontConfig._fields_]),
    ("clr_render_params", ca.RenderParams, [(n, n) for n, _ in ca.RenderParams._fields_]),
    ("clr_scene_dev", ca.SceneDev, [(n, n) for n, _ in ca.SceneDev._fields_]),
    ("clr_diff_params", ca.DiffParams, [(n, n) for n, _ in ca.DiffParams._fields_]),
    ("clr_scene_params", ca.SceneParams, [(n, n) for n, _ in ca.SceneParams._fields_]),
]
DTYPES = [("clr_face", ca.FACE_DTYPE), ("clr_bvh_node", ca.BVH_NODE_DTYPE)]


@pytest.fixture(scope="module")
def c_layout(tmp_path_factory):
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no host C compiler")
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "clear_amd.h"',
             '#include "clear_amd_scenes.h"', "int main(void) {"]
    for cname, _, fields in CASES:
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for cf, _ in fields:
            lines.append(f'printf("{cname}.{cf} %zu\\n", offsetof({cname}, {cf}));')
    for cname, _ in DTYPES:
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
    lines += ["return 0;", "}"]
    d = tmp_path_factory.mktemp("abi")
    src, exe = d / "layout.c", d / "layout"
    src.write_text("\n".join(lines) + "\n")
    subprocess.run([cc, "-std=c11", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    return {k: int(v) for k, v in (line.rsplit(" ", 1) for line in out.splitlines())}


@pytest.mark.parametrize("cname,ptype,fields", CASES, ids=[c[0] for c in CASES])
def test_ctypes_struct_matches_header(c_layout, cname, ptype, fields):
    assert ctypes.sizeof(ptype) == c_layout[f"{cname} size"]
    for cf, pf in fields:
        assert getattr(ptype, pf).offset == c_layout[f"{cname}.{cf}"], (cname, cf)


@pytest.mark.parametrize("cname,dtype", DTYPES, ids=[d[0] for d in DTYPES])
def test_record_dtype_matches_header(c_layout, cname, dtype):
    assert dtype.itemsize == c_layout[f"{cname} size"]
