"""The C++ wrappers of include/clear_amd.hpp driven from compiled C++ on the GPU.

tests/cpp/wrapper_driver (built by __graft_entry__.build()) builds an LBVH
through clr::LBVHBuilder (the reference's LBVHBuilder calling shape,
LBVHBuilder.h:32-75) from vertices + Face records, renders the Cornell box
through clr::render with both renderers, and checks that a bad argument
raises clr::Error. Here its outputs are compared with the oracle: the nodes
byte for byte, the renders by the render bar of test_gpu_render.py (and the
two renderers bit for bit).
"""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import clear_amd as ca
import oracle as orc
from test_gpu_render import THREADS

pytestmark = pytest.mark.gpu

DRIVER = Path(__file__).resolve().parent / "cpp" / "wrapper_driver"


def test_wrapper_driver(tmp_path):
    assert DRIVER.exists(), "tests/cpp/wrapper_driver is built by __graft_entry__.build()"
    print("[cpp wrappers] running the driver", flush=True)
    proc = subprocess.run([str(DRIVER), str(tmp_path)], capture_output=True, text=True, timeout=240)
    print(proc.stdout, proc.stderr, flush=True)
    assert proc.returncode == 0, proc.stderr
    assert "error path ok" in proc.stdout and "wrapper driver ok" in proc.stdout

    n = 50000
    v, f = ca.soup_to_mesh(ca.gen_cluster_soup(n, seed=7))
    o = orc.lbvh_build(v, f, 10)
    inodes = np.fromfile(tmp_path / "lbvh_inodes.bin", np.uint8).reshape(-1, 48)
    leaves = np.fromfile(tmp_path / "lbvh_leaves.bin", np.uint8).reshape(-1, 48)
    np.testing.assert_array_equal(inodes, o["inodes"])
    np.testing.assert_array_equal(leaves, o["leaves"])

    W = H = 48
    host = ca.HostScene(ca.SCENE_CORNELL, W, H, detail=3)
    ob = orc.lbvh_build(host.vertices, host.faces, 10)
    st = orc.seed_states(W, H)
    px = np.zeros(W * H * 4, np.float32)
    orc.render(host.arrays, host.num_faces, ob["inodes"], ob["leaves"], W, H, 4, 5, st, px, threads=THREADS)
    mega = np.fromfile(tmp_path / "render_mega.bin", np.float32)
    wave = np.fromfile(tmp_path / "render_wave.bin", np.float32)
    np.testing.assert_array_equal(mega, wave)
    g, c = mega.reshape(-1, 4)[:, :3], px.reshape(-1, 4)[:, :3]
    ok = (np.abs(g - c) <= 1e-3 * np.abs(c) + 1e-6 * 4).all(axis=1)
    assert c.mean() > 1e-4 and ok.mean() >= 0.999
