"""A scene built from files (clr_scene_build_files) rendered on the GPU: OBJ /
MTL geometry and materials (matte, glass, an RGBA image texture whose alpha
cuts a texel out), an emissive second model and an OpenEXR environment. The
megakernel is held to the oracle's bar of test_gpu_render.py on the same scene
bytes; the wavefront renderer must give the megakernel's image bit for bit."""
import numpy as np
import pytest

import asset_files as af
import clear_amd as ca
from test_gpu_render import compare, gpu_render, oracle_render
from test_gpu_wavefront import same_as_megakernel

pytestmark = pytest.mark.gpu

LIGHT_OBJ = """# a quad beside and above the fixture, facing down
v 1.5 -0.5 1.5
v 2.5 -0.5 1.5
v 2.5 0.5 1.5
v 1.5 0.5 1.5
vn 0 0 -1
f 1//1 3//1 2//1
f 1//1 4//1 3//1
"""


def _scene(tmp_path, w, h):
    af.write_obj_fixture(tmp_path)
    (tmp_path / "light.obj").write_text(LIGHT_OBJ)
    y, x = np.mgrid[0:16, 0:32]
    env = {c: (0.3 + 0.2 * k + 0.5 * np.sin(0.4 * x + k) ** 2 * (y < 8)).astype(np.float16)
           for k, c in enumerate("RGB")}
    af.write_exr(tmp_path / "env.exr", env, af.ZIP)
    return ca.HostScene.from_files([tmp_path / "a.obj", tmp_path / "light.obj"],
                                   emissions=[None, (3.0, 2.5, 2.0)], env_path=tmp_path / "env.exr",
                                   env_multiplier=1.5, width=w, height=h, fov_y=50.0)


def test_scene_from_files(gpu_ctx, tmp_path):
    host = _scene(tmp_path, 96, 64)
    print(f"[assets] files scene faces={host.num_faces} lights={host.num_lights} 96x64 spp=8 depth=6",
          flush=True)
    assert host.num_lights == 3  # two emissive triangles and the environment
    compare("files scene", host, 8, 6, gpu_render(gpu_ctx, host, 8, 6), oracle_render(host, 8, 6))
    same_as_megakernel(gpu_ctx, host, 8, 6)
