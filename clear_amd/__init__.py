"""clear_amd — MI355X (gfx950) LBVH build, closest-hit traversal and path tracing.

Thin ctypes binding over libclear_amd.so (C-ABI in include/clear_amd.h).
Device memory and streams come from PyTorch (torch.cuda tensors); the compute
is the library's own HIP kernels. Importing this module does not need a GPU;
calling into it does, and it raises if the shared library was not built.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "libclear_amd.so"

FACE_DTYPE = np.dtype(
    [("p", "<u4", 3), ("vn", "<u4", 3), ("vt", "<u4", 3), ("uv", "<u4", 3),
     ("alphaTexPtr", "<u4"), ("matPtr", "<u2"), ("lightPtr", "<u2")]
)
assert FACE_DTYPE.itemsize == 56
BVH_NODE_DTYPE = np.dtype([("min", "<f4", 4), ("max", "<f4", 4), ("children", "<u4", 2), ("dum0", "<u4", 2)])
assert BVH_NODE_DTYPE.itemsize == 48
INODE_BYTES = 48
LEAF_BYTES = 48

CLR_OK = 0
_lib = None


class ClrError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"clear_amd error {status}: {message}")
        self.status = status


def library_path() -> Path:
    return LIB_PATH


def lib():
    """Loads libclear_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ClrError(-1, f"{LIB_PATH} is missing: run __graft_entry__.build() first")
    L = ctypes.CDLL(str(LIB_PATH))
    vp, u32, i32, f32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int, ctypes.c_float
    sig = {
        "clr_create": (i32, [ctypes.POINTER(vp), i32]),
        "clr_destroy": (i32, [vp]),
        "clr_error_string": (ctypes.c_char_p, [i32]),
        "clr_last_error": (ctypes.c_char_p, [vp]),
        "clr_scratch_bytes": (i32, [vp, ctypes.POINTER(ctypes.c_uint64)]),
        "clr_device_status": (i32, [vp, ctypes.POINTER(u32), i32]),
        "clr_set_profiling": (i32, [vp, i32]),
        "clr_lbvh_last_profile": (i32, [vp, ctypes.POINTER(f32), i32, ctypes.POINTER(i32)]),
        "clr_lbvh_build": (i32, [vp, vp, vp, u32, u32, vp, vp, vp]),
        "clr_lbvh_build_soup": (i32, [vp, vp, u32, u32, vp, vp, vp]),
        "clr_lbvh_reserve": (i32, [vp, u32, u32]),
        "clr_lbvh_build_soup_noalloc": (i32, [vp, vp, u32, u32, vp, vp, vp]),
        "clr_lbvh_noalloc_max_tris": (u32, []),
        "clr_lbvh_build_segmented": (i32, [vp, vp, vp, u32, u32, u32, vp, vp, vp]),
        "clr_lbvh_reserve_segmented": (i32, [vp, u32, u32, u32]),
        "clr_trace_closest": (i32, [vp, vp, vp, u32, vp, vp, vp, u32, vp, vp, vp, vp]),
        "clr_trace_closest_ordered": (i32, [vp, vp, vp, u32, vp, vp, vp, u32, vp, vp, vp, vp]),
        "clr_trace_closest_wide": (i32, [vp, vp, vp, u32, vp, vp, vp, u32, vp, vp, vp, vp]),
        "clr_trace_closest_bvh": (i32, [vp, vp, u32, vp, vp, vp, u32, vp, vp, vp, i32, vp]),
        "clr_gen_cluster_soup": (i32, [u32, i32, u32, f32, f32, f32, vp]),
        "clr_gen_random_rays": (i32, [u32, i32, vp]),
        "clr_bvh_build_midpoint": (i32, [vp, vp, u32, vp]),
        "clr_render": (i32, [vp, vp, vp, vp, vp, i32, vp]),
        "clr_seed_pixel_states": (i32, [vp, u32, u32, vp, vp]),
        "clr_render_adjoint": (i32, [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
        "clr_render_batch": (i32, [vp, vp, u32, vp, vp, vp, vp]),
    "clr_render_batch_mode": (i32, [vp, vp, u32, vp, vp, vp, i32, vp]),
        "clr_bvh_visualizer": (i32, [vp, vp, vp, vp, vp]),
        "clr_post_clear": (i32, [vp, u32, u32, vp, vp]),
        "clr_post_scaling": (i32, [vp, u32, u32, u32, vp, vp, vp]),
        "clr_post_gaussian_scattering": (i32, [vp, u32, u32, u32, vp, vp, vp]),
        "clr_post_bloom": (i32, [vp, u32, u32, u32, vp, vp, vp]),
        "clr_post_tone_mapping": (i32, [vp, u32, u32, u32, vp, vp, vp]),
        "clr_bmp_row_bytes": (u32, [u32]),
        "clr_save_bmp": (i32, [ctypes.c_char_p, vp, u32, u32]),
        "clr_scene_diff_params": (i32, [vp, ctypes.POINTER(u32), vp, vp, vp, ctypes.POINTER(u32),
                                        ctypes.POINTER(u32)]),
        "clr_wavefront_get_config": (i32, [vp, vp]),
        "clr_wavefront_set_config": (i32, [vp, vp]),
        "clr_wavefront_autotune": (i32, [vp, vp, vp, vp, vp]),
        "clr_wavefront_autotune_batch": (i32, [vp, vp, u32, vp, vp, vp]),
        "clr_scene_build": (i32, [i32, vp, ctypes.POINTER(vp)]),
        "clr_scene_free": (i32, [vp]),
        "clr_scene_info": (i32, [vp, ctypes.POINTER(u32), ctypes.POINTER(u32),
                                 ctypes.POINTER(u32), ctypes.POINTER(u32)]),
        "clr_scene_array": (i32, [vp, i32, ctypes.POINTER(vp), ctypes.POINTER(ctypes.c_uint64)]),
        "clr_image_load": (i32, [ctypes.c_char_p, i32, ctypes.POINTER(u32), ctypes.POINTER(u32),
                                 ctypes.POINTER(u32), vp, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]),
        "clr_scene_build_files": (i32, [ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(f32), u32,
                                        ctypes.c_char_p, f32, vp, ctypes.POINTER(vp)]),
        "clr_scene_last_error": (ctypes.c_char_p, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def exported_symbols() -> list[str]:
    """Names of the C-ABI functions this binding declares."""
    lib()
    return [n for n in dir(_lib) if n.startswith("clr_")]


def _ptr(t) -> ctypes.c_void_p:
    if t is None:
        return ctypes.c_void_p(0)
    if isinstance(t, int):
        return ctypes.c_void_p(t)
    if isinstance(t, np.ndarray):
        return ctypes.c_void_p(t.ctypes.data)
    return ctypes.c_void_p(t.data_ptr())


def _stream(stream=None) -> ctypes.c_void_p:
    import torch

    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


class Context:
    """Owns a clr_ctx (one device, one scratch pool)."""

    def __init__(self, device: int = 0):
        self._h = ctypes.c_void_p()
        rc = lib().clr_create(ctypes.byref(self._h), device)
        if rc != CLR_OK:
            raise ClrError(rc, lib().clr_error_string(rc).decode())
        self.device = device

    @property
    def handle(self):
        return self._h

    def check(self, rc: int) -> None:
        if rc != CLR_OK:
            detail = lib().clr_last_error(self._h).decode()
            raise ClrError(rc, f"{lib().clr_error_string(rc).decode()}: {detail}")

    def close(self) -> None:
        if self._h:
            lib().clr_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def scratch_bytes(self) -> int:
        v = ctypes.c_uint64()
        self.check(lib().clr_scratch_bytes(self._h, ctypes.byref(v)))
        return v.value

    def device_status(self, clear: bool = True) -> int:
        """Device-side failure flags since the last clear (0 = none); syncs."""
        v = ctypes.c_uint32()
        self.check(lib().clr_device_status(self._h, ctypes.byref(v), int(clear)))
        return v.value

    def set_profiling(self, on: bool) -> None:
        self.check(lib().clr_set_profiling(self._h, int(on)))

    def last_profile(self) -> list[float]:
        buf = (ctypes.c_float * 8)()
        n = ctypes.c_int()
        self.check(lib().clr_lbvh_last_profile(self._h, buf, 8, ctypes.byref(n)))
        return [buf[i] for i in range(n.value)]


def node_buffers(n: int, device="cuda"):
    """Allocates (internalNodes, leafNodes) as uint8 tensors of the 48-B records."""
    import torch

    inodes = torch.empty(max(n - 1, 1) * INODE_BYTES, dtype=torch.uint8, device=device)
    leaves = torch.empty(n * LEAF_BYTES, dtype=torch.uint8, device=device)
    return inodes, leaves


class LBVHBuilder:
    """Mirror of the reference's LBVHBuilder (LBVHBuilder.h:32-75).

    LBVHBuilder(ctx, numFaces) reserves scratch for numFaces triangles;
    perform(...) enqueues one build on the stream (no host sync).
    """

    def __init__(self, ctx: Context, num_faces: int, bits_per_dim: int = 10):
        self.ctx = ctx
        if num_faces > 0:
            ctx.check(lib().clr_lbvh_reserve(ctx.handle, num_faces, bits_per_dim))

    def perform(self, vertices, faces, num_faces: int, num_bits_per_dim: int,
                internal_nodes, leaf_nodes, stream=None) -> None:
        self.ctx.check(lib().clr_lbvh_build(
            self.ctx.handle, _ptr(vertices), _ptr(faces), num_faces, num_bits_per_dim,
            _ptr(internal_nodes), _ptr(leaf_nodes), _stream(stream)))

    def perform_soup(self, tris, num_tris: int, num_bits_per_dim: int,
                     internal_nodes, leaf_nodes, stream=None) -> None:
        self.ctx.check(lib().clr_lbvh_build_soup(
            self.ctx.handle, _ptr(tris), num_tris, num_bits_per_dim,
            _ptr(internal_nodes), _ptr(leaf_nodes), _stream(stream)))

    def perform_segmented(self, tris, seg_offsets, num_tris: int, num_bits_per_dim: int,
                          internal_nodes, leaf_nodes, stream=None) -> None:
        """clr_lbvh_build_segmented: seg_offsets is a device uint32/int32
        tensor of numSegments + 1 prefix offsets ending at num_tris."""
        self.ctx.check(lib().clr_lbvh_build_segmented(
            self.ctx.handle, _ptr(tris), _ptr(seg_offsets), seg_offsets.numel() - 1, num_tris,
            num_bits_per_dim, _ptr(internal_nodes), _ptr(leaf_nodes), _stream(stream)))


def trace_closest_bvh(ctx: Context, nodes, num_faces: int, vertices, faces, rays, num_rays: int, hit_face,
                      hit_t, stats=None, stream=None, ordered: bool = False) -> None:
    """clr_trace_closest_bvh: closest hit over a CPU midpoint BVH (reference
    walk, or the ordered traversal with ordered=True)."""
    ctx.check(lib().clr_trace_closest_bvh(ctx.handle, _ptr(nodes), num_faces, _ptr(vertices), _ptr(faces),
                                          _ptr(rays), num_rays, _ptr(hit_face), _ptr(hit_t), _ptr(stats),
                                          int(ordered), _stream(stream)))


def trace_closest(ctx: Context, internal_nodes, leaf_nodes, num_faces: int, vertices, faces,
                  rays, num_rays: int, hit_face, hit_t, stats=None, stream=None,
                  ordered: bool = False, wide: bool = False) -> None:
    """Closest hit per ray: the reference's walk (clr_trace_closest), with
    ordered=True the ordered traversal (clr_trace_closest_ordered), with
    wide=True its 4-wide form (clr_trace_closest_wide)."""
    fn = (lib().clr_trace_closest_wide if wide else
          lib().clr_trace_closest_ordered if ordered else lib().clr_trace_closest)
    ctx.check(fn(ctx.handle, _ptr(internal_nodes), _ptr(leaf_nodes), num_faces, _ptr(vertices),
                 _ptr(faces), _ptr(rays), num_rays, _ptr(hit_face), _ptr(hit_t), _ptr(stats),
                 _stream(stream)))


# ---- host scene generators ------------------------------------------------

def gen_cluster_soup(num_tris: int, seed: int = 1, num_clusters: int = 64, sigma: float = 0.02,
                     e_min: float = 1e-3, e_max: float = 1e-2) -> np.ndarray:
    """Gaussian-cluster soup, float32 [num_tris, 9] (host)."""
    out = np.empty((num_tris, 9), dtype=np.float32)
    rc = lib().clr_gen_cluster_soup(num_tris, seed, num_clusters, sigma, e_min, e_max,
                                    _ptr(out))
    if rc:
        raise ClrError(rc, "clr_gen_cluster_soup")
    return out


def gen_random_rays(num_rays: int, seed: int = 2) -> np.ndarray:
    """Random rays, float32 SoA [7, num_rays] (host)."""
    out = np.empty((7, num_rays), dtype=np.float32)
    rc = lib().clr_gen_random_rays(num_rays, seed, _ptr(out))
    if rc:
        raise ClrError(rc, "clr_gen_random_rays")
    return out


def build_midpoint_bvh(vertices: np.ndarray, faces: np.ndarray) -> np.ndarray:
    """clr_bvh_build_midpoint on host arrays: BVH_NODE_DTYPE [2n - 1]."""
    n = len(faces)
    out = np.zeros(max(2 * n - 1, 1), BVH_NODE_DTYPE)
    v = np.ascontiguousarray(vertices, dtype=np.float32)
    f = np.ascontiguousarray(faces)
    rc = lib().clr_bvh_build_midpoint(v.ctypes.data, f.ctypes.data, n, out.ctypes.data)
    if rc:
        raise ClrError(rc, "clr_bvh_build_midpoint")
    return out


def soup_to_mesh(soup: np.ndarray):
    """Triangle soup [N,9] -> (vertices float32 [3N,4], faces FACE_DTYPE [N]).

    Vertex 3i+k is corner k of triangle i; every other face attribute absent.
    """
    n = soup.shape[0]
    verts = np.zeros((3 * n, 4), dtype=np.float32)
    verts[:, :3] = soup.reshape(3 * n, 3)
    faces = np.zeros(n, dtype=FACE_DTYPE)
    idx = np.arange(3 * n, dtype=np.uint32).reshape(n, 3)
    faces["p"] = idx
    faces["vn"] = 0xFFFFFFFF
    faces["vt"] = 0xFFFFFFFF
    faces["uv"] = 0xFFFFFFFF
    faces["alphaTexPtr"] = 0xFFFFFFFF
    faces["matPtr"] = 0xFFFF
    faces["lightPtr"] = 0xFFFF
    return verts, faces


def decode_inodes(buf: np.ndarray, n: int):
    """48-B InternalNode records -> dict of numpy arrays (min, max, isChild, c)."""
    raw = np.frombuffer(np.ascontiguousarray(buf).tobytes(), dtype=np.uint8)[: (n - 1) * 48]
    rec = raw.reshape(-1, 48)
    f = rec[:, :32].copy().view(np.float32).reshape(-1, 8)
    return {
        "min": f[:, 0:3], "max": f[:, 4:7],
        "isChild": rec[:, 32:34].copy(),
        "c": rec[:, 36:44].copy().view(np.uint32).reshape(-1, 2),
    }


def decode_leaves(buf: np.ndarray, n: int):
    raw = np.frombuffer(np.ascontiguousarray(buf).tobytes(), dtype=np.uint8)[: n * 48]
    rec = raw.reshape(-1, 48)
    f = rec[:, :32].copy().view(np.float32).reshape(-1, 8)
    return {"min": f[:, 0:3], "max": f[:, 4:7],
            "objIdx": rec[:, 32:36].copy().view(np.uint32).reshape(-1)}


# ---- scenes and rendering -------------------------------------------------

SCENE_CORNELL, SCENE_SHOWCASE, SCENE_FURNACE, SCENE_TERRAIN, SCENE_ADJOINT, SCENE_CLUSTERS = 0, 1, 2, 3, 4, 5
SCENE_AREA_LIGHTS = 6
ARRAY_NAMES = ["vertices", "normals", "tangents", "uvs", "faces", "lights", "materials",
               "textures", "others", "bvh_nodes"]
RENDER_MEGAKERNEL, RENDER_WAVEFRONT = 0, 1


class SceneParams(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
                ("detail", ctypes.c_uint32), ("seed", ctypes.c_int),
                ("lensRadius", ctypes.c_float), ("focusDistance", ctypes.c_float),
                ("buildCpuBvh", ctypes.c_uint32), ("albedo", ctypes.c_float),
                ("fovY", ctypes.c_float), ("numTriangles", ctypes.c_uint32)]


class SceneDev(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in
                ["vertices", "normals", "tangents", "uvs", "faces", "lights", "materialsData",
                 "texturesData", "otherResources", "internalNodes", "leafNodes"]] + \
               [("numFaces", ctypes.c_uint32), ("numLights", ctypes.c_uint32), ("bvhNodes", ctypes.c_void_p)]


class RenderParams(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
                ("samplesPerPixel", ctypes.c_uint32), ("maxDepth", ctypes.c_uint32),
                ("flags", ctypes.c_uint32)]


class HostScene:
    """A procedural scene built by the library's host Scene API; `arrays` holds
    numpy copies of the reference's arrays and blobs (empty attribute arrays
    become 1-element dummies, as the reference's host does)."""

    def __init__(self, kind: int, width: int = 0, height: int = 0, detail: int = 0, seed: int = 0,
                 lens_radius: float = 0.0, focus_distance: float = 0.0, cpu_bvh: bool = False,
                 albedo: float = 0.5, fov_y: float = 0.0, num_triangles: int = 0):
        p = SceneParams(width, height, detail, seed, lens_radius, focus_distance, int(cpu_bvh),
                        albedo, fov_y, num_triangles)
        h = ctypes.c_void_p()
        rc = lib().clr_scene_build(kind, ctypes.byref(p), ctypes.byref(h))
        if rc:
            raise ClrError(rc, f"clr_scene_build({kind})")
        self._take(h)

    @classmethod
    def from_files(cls, obj_paths, emissions=None, env_path=None, env_multiplier: float = 1.0,
                   width: int = 256, height: int = 256, fov_y: float = 0.0, lens_radius: float = 0.0,
                   focus_distance: float = 0.0, cpu_bvh: bool = False) -> "HostScene":
        """clr_scene_build_files: OBJ models (with MTL materials and image
        textures) loaded by the reference's loadModel rules; model k emits
        emissions[k] (RGB) when nonzero; env_path: an OpenEXR environment."""
        paths = [str(x).encode() for x in obj_paths]
        arr = (ctypes.c_char_p * max(len(paths), 1))(*paths)
        em = np.zeros(3 * max(len(paths), 1), np.float32)
        if emissions is not None:
            for k, e in enumerate(emissions):
                if e is not None:
                    em[3 * k:3 * k + 3] = e
        p = SceneParams(width, height, 0, 0, lens_radius, focus_distance, int(cpu_bvh), 0.0, fov_y, 0)
        h = ctypes.c_void_p()
        rc = lib().clr_scene_build_files(arr, em.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), len(paths),
                                         str(env_path).encode() if env_path else None, env_multiplier,
                                         ctypes.byref(p), ctypes.byref(h))
        if rc:
            raise ClrError(rc, "clr_scene_build_files: " + lib().clr_scene_last_error().decode())
        obj = cls.__new__(cls)
        obj._take(h)
        return obj

    def _take(self, h) -> None:
        """Copies the host scene's arrays and frees it."""
        try:
            nf, nl, w, hh = (ctypes.c_uint32() for _ in range(4))
            lib().clr_scene_info(h, ctypes.byref(nf), ctypes.byref(nl), ctypes.byref(w), ctypes.byref(hh))
            self.num_faces, self.num_lights = nf.value, nl.value
            self.width, self.height = w.value, hh.value
            self.arrays = {}
            for i, name in enumerate(ARRAY_NAMES):
                ptr, nbytes = ctypes.c_void_p(), ctypes.c_uint64()
                lib().clr_scene_array(h, i, ctypes.byref(ptr), ctypes.byref(nbytes))
                n = nbytes.value
                buf = np.zeros(max(n, 16), dtype=np.uint8)
                if n:
                    ctypes.memmove(buf.ctypes.data, ptr.value, n)
                self.arrays[name] = buf
            nm = ctypes.c_uint32(0)
            lib().clr_scene_diff_params(h, ctypes.byref(nm), None, None, None, None, None)
            m = nm.value
            mo, ao, so = (np.zeros(max(m, 1), np.uint32) for _ in range(3))
            lp, eo = ctypes.c_uint32(), ctypes.c_uint32()
            lib().clr_scene_diff_params(h, ctypes.byref(nm), mo.ctypes.data, ao.ctypes.data,
                                        so.ctypes.data, ctypes.byref(lp), ctypes.byref(eo))
            self.diff = {"mat_offsets": mo[:m], "albedo_offsets": ao[:m], "sigma_offsets": so[:m],
                         "light_prop": lp.value, "emission_offset": eo.value}
        finally:
            lib().clr_scene_free(h)
        self.faces = self.arrays["faces"][: self.num_faces * 56].view(FACE_DTYPE)
        self.vertices = self.arrays["vertices"].view(np.float32).reshape(-1, 4)

    def to_device(self, device="cuda"):
        import torch

        return {k: torch.from_numpy(v).to(device) for k, v in self.arrays.items()}


COLOR_RGB8, COLOR_RGBA8, COLOR_RGBA16F, COLOR_GRAY8 = 0, 1, 2, 3


def load_image(path, gamma_correction: bool = False):
    """clr_image_load: (pixels, width, height, color type) of a PNG or OpenEXR
    file: uint8 [h, w, 4] for RGB8 (4th byte 0xFF) and RGBA8, float16
    [h, w, 4] for RGBA16F, uint8 [h, w] for gray8."""
    w, h, ct, nb = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint64()
    bpath = str(path).encode()
    rc = lib().clr_image_load(bpath, int(gamma_correction), ctypes.byref(w), ctypes.byref(h), ctypes.byref(ct),
                              None, 0, ctypes.byref(nb))
    if rc:
        raise ClrError(rc, "clr_image_load: " + lib().clr_scene_last_error().decode())
    buf = np.zeros(max(nb.value, 1), np.uint8)
    rc = lib().clr_image_load(bpath, int(gamma_correction), ctypes.byref(w), ctypes.byref(h), ctypes.byref(ct),
                              buf.ctypes.data, nb.value, ctypes.byref(nb))
    if rc:
        raise ClrError(rc, "clr_image_load: " + lib().clr_scene_last_error().decode())
    buf = buf[:nb.value]
    if ct.value == COLOR_RGBA16F:
        px = buf.view(np.float16).reshape(h.value, w.value, 4)
    elif ct.value == COLOR_GRAY8:
        px = buf.reshape(h.value, w.value)
    else:
        px = buf.reshape(h.value, w.value, 4)
    return px, w.value, h.value, ct.value


class DeviceScene:
    """Scene arrays resident on the GPU plus its acceleration structure: the
    LBVH built by clr_lbvh_build (accel="lbvh"), or the CPU midpoint BVH of
    Scene::build uploaded as clr_scene_dev.bvhNodes (accel="midpoint")."""

    def __init__(self, ctx: Context, host: HostScene, bits: int = 10, device="cuda", stream=None,
                 accel: str = "lbvh"):
        import torch

        self.ctx, self.host, self.accel = ctx, host, accel
        self.t = host.to_device(device)
        n = host.num_faces
        bvh = None
        self.inodes = self.leaves = self.bvh = None
        if accel == "lbvh":
            self.inodes, self.leaves = node_buffers(n, device)
            LBVHBuilder(ctx, n, bits).perform(self.t["vertices"], self.t["faces"], n, bits, self.inodes,
                                              self.leaves, stream)
        elif accel == "midpoint":
            nodes = host.arrays["bvh_nodes"]
            if nodes.nbytes < (2 * n - 1) * 48:  # the host scene was built without its CPU BVH
                nodes = build_midpoint_bvh(host.vertices, host.faces).view(np.uint8)
            self.bvh = torch.from_numpy(np.ascontiguousarray(nodes[:(2 * n - 1) * 48])).to(device)
            bvh = _ptr(self.bvh).value
        else:
            raise ValueError(f"unknown accel {accel!r}")
        self.dev = SceneDev(
            _ptr(self.t["vertices"]).value, _ptr(self.t["normals"]).value,
            _ptr(self.t["tangents"]).value, _ptr(self.t["uvs"]).value, _ptr(self.t["faces"]).value,
            _ptr(self.t["lights"]).value, _ptr(self.t["materials"]).value,
            _ptr(self.t["textures"]).value, _ptr(self.t["others"]).value,
            _ptr(self.inodes).value, _ptr(self.leaves).value, n, host.num_lights, bvh)
        self.width, self.height = host.width, host.height


def pack_by_kind(arrays: list, device="cuda") -> list:
    """Uploads many scenes' arrays (dicts of uint8 arrays by kind) as one
    device buffer per kind: every scene's vertices, then every scene's
    normals, and so on, each at a 256-byte aligned offset. Returns each
    scene's dict of device views. A batched render reads the arrays of many
    scenes at once, and packed by kind they span fewer pages than one
    allocation per scene and kind."""
    import torch

    out = [dict() for _ in arrays]
    for k in arrays[0]:
        offs, total = [], 0
        for a in arrays:
            offs.append(total)
            total += (a[k].nbytes + 255) // 256 * 256
        buf = np.zeros(max(total, 1), np.uint8)
        for a, o in zip(arrays, offs):
            buf[o:o + a[k].nbytes] = a[k].reshape(-1).view(np.uint8)
        dev = torch.from_numpy(buf).to(device)
        for s, (a, o) in enumerate(zip(arrays, offs)):
            out[s][k] = dev[o:o + a[k].nbytes] if a[k].nbytes else torch.from_numpy(a[k]).to(device)
    return out


def soup_of(hosts: list, offsets: np.ndarray) -> np.ndarray:
    """float32 [sum faces, 9]: every scene's triangles in face order (corner k
    of face f = vertex p[k]), scene s at rows offsets[s]..offsets[s+1]."""
    soup = np.empty((int(offsets[-1]), 9), np.float32)
    for s, x in enumerate(hosts):
        p = x.faces["p"].astype(np.int64)
        o = int(offsets[s])
        soup[o:o + x.num_faces] = x.vertices[p][:, :, :3].reshape(-1, 9)
    return soup


class SceneBatch:
    """Many scenes resident on the GPU whose LBVHs come from ONE segmented
    build (clr_lbvh_build_segmented) over their concatenated triangles; scene
    s's nodes sit at leaves[offsets[s]] and inodes[offsets[s] - s], exactly
    the nodes clr_lbvh_build gives for that scene alone. `devs` is the
    clr_scene_dev array clr_render_batch takes."""

    def __init__(self, ctx: Context, hosts: list, bits: int = 10, device="cuda", stream=None):
        import torch

        if not hosts:
            raise ValueError("empty scene batch")
        w, h = hosts[0].width, hosts[0].height
        if any(x.width != w or x.height != h for x in hosts):
            raise ValueError("all scenes of a batch render at one image size")
        self.ctx, self.hosts, self.width, self.height = ctx, hosts, w, h
        self.t = pack_by_kind([x.arrays for x in hosts], device)
        counts = np.array([x.num_faces for x in hosts], np.int64)
        self.offsets = np.zeros(len(hosts) + 1, np.uint32)
        self.offsets[1:] = np.cumsum(counts)
        total = int(self.offsets[-1])
        self.tris_host = soup_of(hosts, self.offsets)
        self.tris = torch.from_numpy(self.tris_host.reshape(-1)).to(device)
        self.offsets_dev = torch.from_numpy(self.offsets.view(np.int32)).to(device)
        self.inodes = torch.empty(max(total - len(hosts), 1) * INODE_BYTES, dtype=torch.uint8, device=device)
        self.leaves = torch.empty(total * LEAF_BYTES, dtype=torch.uint8, device=device)
        self.build(bits, stream)
        ib, lb = self.inodes.data_ptr(), self.leaves.data_ptr()
        self.devs = (SceneDev * len(hosts))()
        for s, (x, t) in enumerate(zip(hosts, self.t)):
            o = int(self.offsets[s])
            self.devs[s] = SceneDev(
                _ptr(t["vertices"]).value, _ptr(t["normals"]).value, _ptr(t["tangents"]).value,
                _ptr(t["uvs"]).value, _ptr(t["faces"]).value, _ptr(t["lights"]).value,
                _ptr(t["materials"]).value, _ptr(t["textures"]).value, _ptr(t["others"]).value,
                ib + (o - s) * INODE_BYTES if x.num_faces > 1 else None, lb + o * LEAF_BYTES,
                x.num_faces, x.num_lights)

    def build(self, bits: int = 10, stream=None) -> None:
        """(Re)builds every scene's LBVH in one segmented build."""
        b = LBVHBuilder(self.ctx, 0, bits)
        b.perform_segmented(self.tris, self.offsets_dev, int(self.offsets[-1]), bits, self.inodes, self.leaves,
                            stream)

    def __len__(self) -> int:
        return len(self.hosts)


def wavefront_autotune_batch(ctx: Context, batch: SceneBatch, spp: int, max_depth: int,
                             stream=None, flags: int = 0) -> WavefrontConfig:
    """clr_wavefront_autotune_batch: the launch-shape search timed on batched
    renders of `batch`; the result stays in the context."""
    rp = RenderParams(batch.width, batch.height, spp, max_depth, flags)
    c = WavefrontConfig()
    ctx.check(lib().clr_wavefront_autotune_batch(ctx.handle, batch.devs, len(batch), ctypes.byref(rp),
                                                 ctypes.byref(c), _stream(stream)))
    return c


def render_batch(ctx: Context, batch: SceneBatch, spp: int, max_depth: int, states, pixels,
                 stream=None, flags: int = 0, mode: int | None = None) -> None:
    """clr_render_batch: every scene of `batch` in one wavefront pass; states
    and pixels hold the scenes' images back to back. With `mode`
    (RENDER_WAVEFRONT or RENDER_MEGAKERNEL) clr_render_batch_mode, whose
    megakernel form takes 8x8 tiles of all scenes from one counter."""
    rp = RenderParams(batch.width, batch.height, spp, max_depth, flags)
    if mode is None:
        ctx.check(lib().clr_render_batch(ctx.handle, batch.devs, len(batch), ctypes.byref(rp), _ptr(states),
                                         _ptr(pixels), _stream(stream)))
    else:
        ctx.check(lib().clr_render_batch_mode(ctx.handle, batch.devs, len(batch), ctypes.byref(rp), _ptr(states),
                                              _ptr(pixels), int(mode), _stream(stream)))


def seed_pixel_states(ctx: Context, width: int, height: int, states, stream=None) -> None:
    ctx.check(lib().clr_seed_pixel_states(ctx.handle, width, height, _ptr(states), _stream(stream)))


RENDER_REFERENCE_TRAVERSAL = 1
RENDER_NO_RR = 2
RENDER_LIGHT_SAMPLING_ONLY = 4
RENDER_BSDF_SAMPLING_ONLY = 8


class DiffParams(ctypes.Structure):
    _fields_ = [("numMaterials", ctypes.c_uint32), ("matOffsets", ctypes.c_void_p),
                ("lightProp", ctypes.c_uint32)]


def render_adjoint(ctx: Context, scene: "DeviceScene", spp: int, max_depth: int, states, d_image,
                   mat_offsets: np.ndarray, light_prop: int, g_albedo, g_sigma, g_emission,
                   stream=None, flags: int = 0) -> None:
    """clr_render_adjoint: gradients of sum(d_image * image) with image = the
    radiance a clr_render call from `states` adds, divided by spp."""
    rp = RenderParams(scene.width, scene.height, spp, max_depth, flags)
    mo = np.ascontiguousarray(mat_offsets, dtype=np.uint32)
    dp = DiffParams(len(mo), mo.ctypes.data if len(mo) else None, light_prop)
    ctx.check(lib().clr_render_adjoint(ctx.handle, ctypes.byref(scene.dev), ctypes.byref(rp),
                                       ctypes.byref(dp), _ptr(states), _ptr(d_image), _ptr(g_albedo),
                                       _ptr(g_sigma), _ptr(g_emission), _stream(stream)))


def ctx_render_dev(ctx: Context, dev: SceneDev, width: int, height: int, spp: int, max_depth: int,
                   states, pixels, mode: int = RENDER_MEGAKERNEL, flags: int = 0, stream=None) -> None:
    """clr_render on a raw clr_scene_dev (e.g. one scene of a SceneBatch)."""
    rp = RenderParams(width, height, spp, max_depth, flags)
    ctx.check(lib().clr_render(ctx.handle, ctypes.byref(dev), ctypes.byref(rp), _ptr(states), _ptr(pixels),
                               mode, _stream(stream)))


def bvh_visualizer(ctx: Context, scene: DeviceScene, pixels, flags: int = 0, stream=None) -> None:
    """clr_bvh_visualizer: pixels += base colour + 0.025 * box hits per pixel."""
    rp = RenderParams(scene.width, scene.height, 1, 1, flags)
    ctx.check(lib().clr_bvh_visualizer(ctx.handle, ctypes.byref(scene.dev), ctypes.byref(rp), _ptr(pixels),
                                       _stream(stream)))


def render(ctx: Context, scene: DeviceScene, spp: int, max_depth: int, states, pixels,
           mode: int = RENDER_MEGAKERNEL, stream=None, flags: int = 0) -> None:
    rp = RenderParams(scene.width, scene.height, spp, max_depth, flags)
    ctx.check(lib().clr_render(ctx.handle, ctypes.byref(scene.dev), ctypes.byref(rp), _ptr(states),
                               _ptr(pixels), mode, _stream(stream)))


class WavefrontConfig(ctypes.Structure):
    """clr_wavefront_config: launch shapes of the wavefront renderer."""
    _fields_ = [(n, ctypes.c_uint32) for n in
                ["traceBlock", "traceWavesPerCU", "shadeBlock", "shadeWavesPerCU", "itersPerGraph",
                 "checkEvery", "maxSlots", "useGraph", "refillMin", "ldsDepth", "raysPerLane"]]

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_}


def wavefront_config(ctx: Context) -> WavefrontConfig:
    c = WavefrontConfig()
    ctx.check(lib().clr_wavefront_get_config(ctx.handle, ctypes.byref(c)))
    return c


def set_wavefront_config(ctx: Context, cfg: WavefrontConfig) -> None:
    ctx.check(lib().clr_wavefront_set_config(ctx.handle, ctypes.byref(cfg)))


def wavefront_autotune(ctx: Context, scene: "DeviceScene", spp: int, max_depth: int,
                       stream=None, flags: int = 0) -> WavefrontConfig:
    """Times launch-shape candidates on this scene (scratch image and states)
    and keeps the fastest in the context."""
    rp = RenderParams(scene.width, scene.height, spp, max_depth, flags)
    c = WavefrontConfig()
    ctx.check(lib().clr_wavefront_autotune(ctx.handle, ctypes.byref(scene.dev), ctypes.byref(rp),
                                           ctypes.byref(c), _stream(stream)))
    return c


# ---- post-processing (post_processing.cl, BMPExporter.cpp) ----------------

def post_clear(ctx: Context, width: int, height: int, buf, stream=None) -> None:
    ctx.check(lib().clr_post_clear(ctx.handle, width, height, _ptr(buf), _stream(stream)))


def post_scaling(ctx: Context, width: int, height: int, spp: int, src, out, stream=None) -> None:
    ctx.check(lib().clr_post_scaling(ctx.handle, width, height, spp, _ptr(src), _ptr(out), _stream(stream)))


def post_gaussian_scattering(ctx: Context, width: int, height: int, spp: int, src, out, stream=None) -> None:
    ctx.check(lib().clr_post_gaussian_scattering(ctx.handle, width, height, spp, _ptr(src), _ptr(out),
                                                 _stream(stream)))


def post_bloom(ctx: Context, width: int, height: int, spp: int, src, out, stream=None) -> None:
    ctx.check(lib().clr_post_bloom(ctx.handle, width, height, spp, _ptr(src), _ptr(out), _stream(stream)))


def post_tone_mapping(ctx: Context, width: int, height: int, row_bytes: int, src, dst, stream=None) -> None:
    ctx.check(lib().clr_post_tone_mapping(ctx.handle, width, height, row_bytes, _ptr(src), _ptr(dst),
                                          _stream(stream)))


def bmp_row_bytes(width: int) -> int:
    return int(lib().clr_bmp_row_bytes(width))


def save_bmp(path: str, pixels: np.ndarray, width: int, height: int) -> None:
    """clr_save_bmp of host tone-mapped rows (bmp_row_bytes(width) * height bytes)."""
    px = np.ascontiguousarray(pixels, dtype=np.uint8)
    if px.size < bmp_row_bytes(width) * height:
        raise ValueError("pixel buffer smaller than the BMP data")
    rc = lib().clr_save_bmp(str(path).encode(), px.ctypes.data, width, height)
    if rc:
        raise ClrError(rc, f"clr_save_bmp({path})")


def new_image(width: int, height: int, device="cuda"):
    """(states uint32 [H*W*4] seeded per pixel, pixels float32 [H*W*4] zeroed)."""
    import torch

    states = torch.empty(width * height * 4, dtype=torch.int32, device=device)
    pixels = torch.zeros(width * height * 4, dtype=torch.float32, device=device)
    return states, pixels
