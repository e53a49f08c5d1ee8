"""Host-side checks of the C-ABI library (no GPU needed).

The library must load, export every function include/*.h declares, report
status strings, and refuse to create a context where there is no device.
"""
import re
from pathlib import Path

import numpy as np
import pytest

import clear_amd as ca

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names.update(re.findall(r"\b(clr_\w+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    import ctypes

    L = ctypes.CDLL(str(ca.library_path()))
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, f"declared but not exported: {missing}"
    assert len(declared_functions()) >= 12


def test_binding_covers_every_declared_symbol():
    ca.lib()
    bound = set(ca.exported_symbols())
    assert set(declared_functions()) <= bound


def test_error_strings():
    L = ca.lib()
    assert L.clr_error_string(0) == b"ok"
    assert L.clr_error_string(1) == b"invalid argument"
    assert L.clr_error_string(99) == b"unknown status"


def test_null_arguments_rejected_without_device():
    import ctypes

    L = ca.lib()
    assert L.clr_create(None, 0) == 1
    assert L.clr_destroy(None) == 1
    assert L.clr_lbvh_build(None, None, None, 10, 10, None, None, None) == 1
    assert L.clr_trace_closest(None, None, None, 1, None, None, None, 0, None, None, None,
                               None) == 1


def test_generators_are_deterministic():
    a = ca.gen_cluster_soup(5000, seed=1)
    b = ca.gen_cluster_soup(5000, seed=1)
    c = ca.gen_cluster_soup(5000, seed=2)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    r1 = ca.gen_random_rays(1000, seed=2)
    assert np.array_equal(r1, ca.gen_random_rays(1000, seed=2))
    d = r1[3:6]
    np.testing.assert_allclose((d * d).sum(0), 1.0, atol=1e-5)
    assert np.isinf(r1[6]).all()
    assert (r1[:3] >= 0).all() and (r1[:3] < 1).all()


def test_cluster_soup_shape_statistics():
    s = ca.gen_cluster_soup(200000, seed=1)
    v = s.reshape(-1, 3, 3)
    e1 = np.linalg.norm(v[:, 1] - v[:, 0], axis=1)
    e2 = np.linalg.norm(v[:, 2] - v[:, 0], axis=1)
    assert 1e-3 * 0.999 <= e1.min() and e1.max() <= 1e-2 * 1.001
    assert 1e-3 * 0.999 <= e2.min() and e2.max() <= 1e-2 * 1.001
    c = v.mean(axis=1)
    # 64 clusters of sigma 0.02: the centroid cloud spans roughly the unit cube
    assert c.min() > -0.2 and c.max() < 1.2


def test_cluster_scene_layout_and_batch_soup():
    """The batch scene holds the cluster soup of the same seed as faces
    0..n-1 (then the light and floor quads), and soup_of restates each scene
    as a triangle soup whose LBVH equals the scene's own."""
    import oracle as orc

    hosts = [ca.HostScene(ca.SCENE_CLUSTERS, 32, 24, seed=100 + i, num_triangles=n)
             for i, n in enumerate([1000, 2500])]
    for i, h in enumerate(hosts):
        assert (h.width, h.height) == (32, 24)
        assert h.num_faces == [1000, 2500][i] + 4 and h.num_lights == 2
        soup = ca.gen_cluster_soup([1000, 2500][i], seed=100 + i).reshape(-1, 9)
        p = h.faces["p"][: len(soup)].astype(np.int64)
        np.testing.assert_array_equal(h.vertices[p][:, :, :3].reshape(-1, 9), soup)
        lp = h.faces["lightPtr"]
        assert (lp[-4:-2] != 0xFFFF).all() and (lp[:-4] == 0xFFFF).all() and (lp[-2:] == 0xFFFF).all()
    offs = np.array([0, hosts[0].num_faces, hosts[0].num_faces + hosts[1].num_faces], np.uint32)
    soup = ca.soup_of(hosts, offs)
    for s, h in enumerate(hosts):
        tri = soup[offs[s]:offs[s + 1]]
        v = np.zeros((len(tri) * 3, 4), np.float32)
        v[:, :3] = tri.reshape(-1, 3)
        f = np.zeros(len(tri), ca.FACE_DTYPE)
        f["p"] = np.arange(len(tri) * 3).reshape(-1, 3)
        a, b = orc.lbvh_build(h.vertices, h.faces, 10), orc.lbvh_build(v, f, 10)
        np.testing.assert_array_equal(a["leaves"], b["leaves"])
        np.testing.assert_array_equal(a["inodes"], b["inodes"])


def test_cpp_wrapper_header_compiles_standalone(tmp_path):
    """include/clear_amd.hpp (the C++ LBVHBuilder / render wrappers) needs
    only clear_amd.h and the C++ standard library."""
    import shutil
    import subprocess
    from pathlib import Path

    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    src = tmp_path / "use.cpp"
    src.write_text('#include "clear_amd.hpp"\n'
                   'void use(clr::Context& c, clr_float3* v, clr_face* f, clr_lbvh_inode* i, clr_lbvh_leaf* l) {\n'
                   '    clr::LBVHBuilder b(c, 16);\n    b.perform(nullptr, v, f, 16, 10, i, l);\n}\n')
    inc = Path(__file__).resolve().parent.parent / "include"
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", f"-I{inc}", str(src)], check=True)


def test_pack_by_kind_keeps_every_scene_array():
    """pack_by_kind uploads one buffer per kind: each scene's view holds
    exactly its own bytes, views of one kind lie in one buffer in scene order
    at 256-byte aligned offsets, and in-place writes through a view reach
    only that scene's bytes."""
    import torch

    hosts = [ca.HostScene(ca.SCENE_CLUSTERS, 32, 24, seed=100 + i, num_triangles=n)
             for i, n in enumerate([1000, 37, 2500])]
    t = ca.pack_by_kind([h.arrays for h in hosts], "cpu")
    for k in hosts[0].arrays:
        base = t[0][k].data_ptr()
        prev_end = base
        for s, h in enumerate(hosts):
            ref = h.arrays[k].reshape(-1).view(np.uint8)
            np.testing.assert_array_equal(t[s][k].numpy(), ref)
            off = t[s][k].data_ptr() - base
            assert off % 256 == 0 and t[s][k].data_ptr() >= prev_end
            prev_end = t[s][k].data_ptr() + ref.nbytes
    v = t[1]["vertices"].view(torch.float32)
    v += 1.0
    np.testing.assert_array_equal(t[0]["vertices"].numpy(), hosts[0].arrays["vertices"].reshape(-1).view(np.uint8))
    np.testing.assert_array_equal(t[2]["vertices"].numpy(), hosts[2].arrays["vertices"].reshape(-1).view(np.uint8))
