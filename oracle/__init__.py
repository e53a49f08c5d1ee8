"""CPU oracle — TEST INFRASTRUCTURE ONLY.

ctypes bindings over oracle/liboracle.so, a serial C++ restatement of the
reference's LBVH build, traversal and renderer (each function cites the
reference file:line it follows). Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg may import this package, and only as the checker
or the timed CPU baseline; the product (clear_amd) never does.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "liboracle.so"
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} missing: run __graft_entry__.build()")
        L = ctypes.CDLL(str(LIB_PATH))
        vp, u32, i32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int
        for name, args in {
            "orc_tri_aabbs": [vp, vp, u32, vp],
            "orc_lbvh_build": [vp, vp, u32, u32, vp, vp, vp, vp, vp],
            "orc_trace_closest": [vp, vp, u32, vp, vp, vp, u32, vp, vp, vp],
            "orc_trace_bruteforce": [vp, vp, u32, vp, u32, vp, vp, vp],
            "pin_interleave": [vp, u32, u32, vp],
            "pin_stable_order": [vp, u32, vp],
            "pin_radix_tree": [vp, u32, u32, vp, vp, vp, vp, vp, vp],
            "pin_range_boxes": [vp, vp, vp, u32, vp],
        }.items():
            fn = getattr(L, name)
            fn.restype = i32
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return ctypes.c_void_p(0) if a is None else ctypes.c_void_p(a.ctypes.data)


def lbvh_build(verts: np.ndarray, faces: np.ndarray, bits: int):
    """Returns dict: inodes (bytes [n-1,48]), leaves ([n,48]), morton [n,3], sorted [n], scene[6]."""
    n = faces.shape[0]
    verts = np.ascontiguousarray(verts, dtype=np.float32)
    faces = np.ascontiguousarray(faces)
    inodes = np.zeros((max(n - 1, 1), 48), dtype=np.uint8)
    leaves = np.zeros((n, 48), dtype=np.uint8)
    morton = np.zeros((n, 3), dtype=np.uint32)
    order = np.zeros(n, dtype=np.uint32)
    scene = np.zeros(6, dtype=np.float32)
    rc = lib().orc_lbvh_build(_p(verts), _p(faces), n, bits, _p(inodes), _p(leaves), _p(morton),
                              _p(order), _p(scene))
    if rc:
        raise RuntimeError(f"orc_lbvh_build failed: {rc}")
    return {"inodes": inodes[: n - 1], "leaves": leaves, "morton": morton, "sorted": order,
            "scene": scene}


def lbvh_build_parallel(verts: np.ndarray, faces: np.ndarray, bits: int, threads: int = 0):
    """The same build on `threads` host threads (0: all): dict of inodes
    ([n-1, 48] bytes) and leaves ([n, 48]), equal to lbvh_build's."""
    fn = lib().orc_lbvh_build_parallel
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p,
                   ctypes.c_void_p, ctypes.c_int]
    n = faces.shape[0]
    verts = np.ascontiguousarray(verts, dtype=np.float32)
    faces = np.ascontiguousarray(faces)
    inodes = np.zeros((max(n - 1, 1), 48), dtype=np.uint8)
    leaves = np.zeros((n, 48), dtype=np.uint8)
    rc = fn(_p(verts), _p(faces), n, bits, _p(inodes), _p(leaves), threads)
    if rc:
        raise RuntimeError(f"orc_lbvh_build_parallel failed: {rc}")
    return {"inodes": inodes[: n - 1], "leaves": leaves}


def trace_closest(inodes, leaves, n, verts, faces, rays, stats=False):
    R = rays.shape[1]
    hit = np.zeros(R, dtype=np.uint32)
    t = np.zeros(R, dtype=np.float32)
    st = np.zeros((R, 2), dtype=np.uint32) if stats else None
    inodes = np.ascontiguousarray(inodes) if n > 1 else np.zeros((1, 48), np.uint8)
    rc = lib().orc_trace_closest(_p(inodes), _p(np.ascontiguousarray(leaves)), n,
                                 _p(np.ascontiguousarray(verts)), _p(np.ascontiguousarray(faces)),
                                 _p(np.ascontiguousarray(rays, dtype=np.float32)), R, _p(hit),
                                 _p(t), _p(st))
    if rc:
        raise RuntimeError(f"orc_trace_closest failed: {rc}")
    return (hit, t, st) if stats else (hit, t)


def trace_overflows(reset=True) -> int:
    """Rays whose reference walk (trace_closest) dropped a push because its
    64-entry stack was full, since the last reset."""
    fn = lib().orc_trace_overflows
    fn.restype = ctypes.c_uint64
    fn.argtypes = [ctypes.c_int]
    return int(fn(int(reset)))


def trace_ordered(inodes, leaves, n, verts, faces, rays, any_hit=False):
    """Ordered traversal restatement: (face, t, stats[R, 2])."""
    L = lib()
    fn = L.orc_trace_ordered
    fn.restype = ctypes.c_int
    vp = ctypes.c_void_p
    fn.argtypes = [vp, vp, ctypes.c_uint32, vp, vp, vp, ctypes.c_uint32, ctypes.c_int, vp, vp, vp]
    R = rays.shape[1]
    hit = np.zeros(R, dtype=np.uint32)
    t = np.zeros(R, dtype=np.float32)
    st = np.zeros((R, 2), dtype=np.uint32)
    inodes = np.ascontiguousarray(inodes) if n > 1 else np.zeros((1, 48), np.uint8)
    rc = fn(_p(inodes), _p(np.ascontiguousarray(leaves)), n, _p(np.ascontiguousarray(verts)),
            _p(np.ascontiguousarray(faces)), _p(np.ascontiguousarray(rays, dtype=np.float32)), R,
            int(any_hit), _p(hit), _p(t), _p(st))
    if rc:
        raise RuntimeError(f"orc_trace_ordered failed: {rc}")
    return hit, t, st


def trace_wide(inodes, leaves, n, verts, faces, rays, any_hit=False):
    """4-wide form of the ordered traversal: (face, t, stats[R, 2]) with
    visited records and tests."""
    L = lib()
    fn = L.orc_trace_wide
    fn.restype = ctypes.c_int
    vp = ctypes.c_void_p
    fn.argtypes = [vp, vp, ctypes.c_uint32, vp, vp, vp, ctypes.c_uint32, ctypes.c_int, vp, vp, vp]
    R = rays.shape[1]
    hit = np.zeros(R, dtype=np.uint32)
    t = np.zeros(R, dtype=np.float32)
    st = np.zeros((R, 2), dtype=np.uint32)
    inodes = np.ascontiguousarray(inodes) if n > 1 else np.zeros((1, 48), np.uint8)
    rc = fn(_p(inodes), _p(np.ascontiguousarray(leaves)), n, _p(np.ascontiguousarray(verts)),
            _p(np.ascontiguousarray(faces)), _p(np.ascontiguousarray(rays, dtype=np.float32)), R,
            int(any_hit), _p(hit), _p(t), _p(st))
    if rc:
        raise RuntimeError(f"orc_trace_wide failed: {rc}")
    return hit, t, st


def trace_bvh(nodes, verts, faces, rays):
    """Midpoint-BVH walk restatement (reference order): (face, t, stats[R, 2])."""
    L = lib()
    fn = L.orc_trace_bvh
    fn.restype = ctypes.c_int
    vp = ctypes.c_void_p
    fn.argtypes = [vp, ctypes.c_uint32, vp, vp, vp, ctypes.c_uint32, vp, vp, vp]
    R = rays.shape[1]
    hit = np.zeros(R, dtype=np.uint32)
    t = np.zeros(R, dtype=np.float32)
    st = np.zeros((R, 2), dtype=np.uint32)
    nodes = np.ascontiguousarray(nodes)
    rc = fn(_p(nodes), nodes.nbytes // 48, _p(np.ascontiguousarray(verts)), _p(np.ascontiguousarray(faces)),
            _p(np.ascontiguousarray(rays, dtype=np.float32)), R, _p(hit), _p(t), _p(st))
    if rc:
        raise RuntimeError(f"orc_trace_bvh failed: {rc}")
    return hit, t, st


def trace_bruteforce_minid(verts, faces, rays):
    """Brute force with the (t, smaller face) rule of the ordered traversal."""
    L = lib()
    fn = L.orc_trace_bruteforce_minid
    fn.restype = ctypes.c_int
    vp = ctypes.c_void_p
    fn.argtypes = [vp, vp, ctypes.c_uint32, vp, ctypes.c_uint32, vp, vp]
    R = rays.shape[1]
    n = faces.shape[0]
    hit = np.zeros(R, dtype=np.uint32)
    t = np.zeros(R, dtype=np.float32)
    fn(_p(np.ascontiguousarray(verts)), _p(np.ascontiguousarray(faces)), n,
       _p(np.ascontiguousarray(rays, dtype=np.float32)), R, _p(hit), _p(t))
    return hit, t


def trace_bruteforce(verts, faces, rays):
    R = rays.shape[1]
    n = faces.shape[0]
    hit = np.zeros(R, dtype=np.uint32)
    t = np.zeros(R, dtype=np.float32)
    ties = np.zeros(R, dtype=np.uint32)
    rc = lib().orc_trace_bruteforce(_p(np.ascontiguousarray(verts)),
                                    _p(np.ascontiguousarray(faces)), n,
                                    _p(np.ascontiguousarray(rays, dtype=np.float32)), R, _p(hit),
                                    _p(t), _p(ties))
    if rc:
        raise RuntimeError(f"orc_trace_bruteforce failed: {rc}")
    return hit, t, ties


def interleave(morton: np.ndarray, bits: int) -> np.ndarray:
    n = morton.shape[0]
    keys = np.zeros(n, dtype=np.uint64)
    lib().pin_interleave(_p(np.ascontiguousarray(morton, dtype=np.uint32)), n, bits, _p(keys))
    return keys


def stable_order(keys: np.ndarray) -> np.ndarray:
    n = keys.shape[0]
    out = np.zeros(n, dtype=np.uint32)
    lib().pin_stable_order(_p(np.ascontiguousarray(keys, dtype=np.uint64)), n, _p(out))
    return out


def radix_tree(sorted_keys: np.ndarray, key_bits: int):
    n = sorted_keys.shape[0]
    m = max(n - 1, 1)
    c0 = np.zeros(m, np.uint32); c1 = np.zeros(m, np.uint32)
    l0 = np.zeros(m, np.uint8); l1 = np.zeros(m, np.uint8)
    lo = np.zeros(m, np.uint32); hi = np.zeros(m, np.uint32)
    lib().pin_radix_tree(_p(np.ascontiguousarray(sorted_keys, dtype=np.uint64)), n, key_bits,
                         _p(c0), _p(c1), _p(l0), _p(l1), _p(lo), _p(hi))
    return {"c0": c0[: n - 1], "c1": c1[: n - 1], "leaf0": l0[: n - 1], "leaf1": l1[: n - 1],
            "lo": lo[: n - 1], "hi": hi[: n - 1]}


def range_boxes(leaf_boxes: np.ndarray, lo: np.ndarray, hi: np.ndarray) -> np.ndarray:
    n = leaf_boxes.shape[0]
    out = np.zeros((max(n - 1, 1), 6), dtype=np.float32)
    lib().pin_range_boxes(_p(np.ascontiguousarray(leaf_boxes, dtype=np.float32)),
                          _p(np.ascontiguousarray(lo, dtype=np.uint32)),
                          _p(np.ascontiguousarray(hi, dtype=np.uint32)), n, _p(out))
    return out[: n - 1]


def seed_states(width, height):
    st = np.zeros(width * height * 4, dtype=np.uint32)
    lib().orc_seed_states(ctypes.c_uint32(width), ctypes.c_uint32(height), _p(st))
    return st


def _scene_ptrs(arrays, num_faces, inodes, leaves, bvh):
    keep = [np.ascontiguousarray(arrays[k]) for k in
            ["vertices", "normals", "tangents", "uvs", "faces", "lights", "materials", "textures",
             "others"]]
    keep.append(np.ascontiguousarray(inodes) if inodes is not None and num_faces > 1
                else np.zeros((1, 48), np.uint8))
    keep.append(np.ascontiguousarray(leaves) if leaves is not None else np.zeros((1, 48), np.uint8))
    ptrs = [k.ctypes.data for k in keep]
    if bvh is not None:
        keep.append(np.ascontiguousarray(bvh))
        ptrs.append(keep[-1].ctypes.data)
    else:
        ptrs.append(None)
    return keep, (ctypes.c_void_p * 12)(*ptrs)


def render(arrays, num_faces, inodes, leaves, width, height, spp, max_depth, states, pixels,
           pix_begin=0, pix_end=None, threads=0, flags=0, bvh=None):
    """Runs the render restatement over pixels [pix_begin, pix_end) (row-major),
    accumulating into `pixels` (float32 [H*W*4]) and advancing `states`.
    With `bvh` (CPU midpoint BVH nodes) the reference-order walk
    (flags & RENDER_REFERENCE_TRAVERSAL) traverses it; the ordered traversal,
    whose hits do not depend on the structure, always uses the LBVH."""
    L = lib()
    if not hasattr(L.orc_render, "_typed"):
        L.orc_render.restype = ctypes.c_int
        L.orc_render.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                 ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p,
                                 ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, ctypes.c_uint32]
        L.orc_render._typed = True
    keep, ptrs = _scene_ptrs(arrays, num_faces, inodes, leaves, bvh)
    end = width * height if pix_end is None else pix_end
    rc = L.orc_render(ptrs, num_faces, width, height, spp, max_depth, _p(states), _p(pixels),
                      pix_begin, end, threads, flags)
    if rc:
        raise RuntimeError(f"orc_render failed: {rc}")


def _probe_fn(name, argtypes):
    fn = getattr(lib(), name)
    fn.restype = ctypes.c_int
    fn.argtypes = argtypes
    return fn


def work_counters(reset=True, count_any_hit=True):
    """Traversal work of oracle renders (the ordered walk's 4-wide form: a
    visit is one record of up to four grandchildren) since the last reset:
    dict of rays, visits, tests (closest-hit walks) and, for environment
    shadow rays, their count, closest-hit visits / tests and the visits /
    tests of an any-hit walk of the same rays. count_any_hit makes later
    renders count the any-hit walks."""
    fn = _probe_fn("orc_work_counters", [ctypes.c_void_p, ctypes.c_int, ctypes.c_int])
    out = np.zeros(8, np.uint64)
    fn(out.ctypes.data, int(reset), int(count_any_hit))
    keys = ["rays", "visits", "tests", "env_shadow_rays", "env_shadow_visits", "env_shadow_tests",
            "env_shadow_anyhit_visits", "env_shadow_anyhit_tests"]
    return {k: int(v) for k, v in zip(keys, out)}


def bxdf_probe(kind, p0, p1, R, wo, wi, u0, u1):
    fp = ctypes.POINTER(ctypes.c_float)
    fn = _probe_fn("orc_bxdf_probe", [ctypes.c_int, ctypes.c_float, ctypes.c_float, fp, fp, fp,
                                      ctypes.c_float, ctypes.c_float, fp])
    out = (ctypes.c_float * 11)()
    arr = lambda v: (ctypes.c_float * 3)(*[float(x) for x in v])
    fn(kind, p0, p1, arr(R), arr(wo), arr(wi), u0, u1, out)
    o = list(out)
    return {"f": o[0:3], "pdf": o[3], "wi": o[4:7], "spdf": o[7], "sf": o[8:11]}


def camera_ray(arrays, px, py, lu0, lu1):
    """Origin, direction and direction pdf of the scene camera's ray for film
    point (px, py) (raster units) and lens sample (lu0, lu1)."""
    fn = _probe_fn("orc_camera_ray", [ctypes.c_void_p, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                      ctypes.c_float, ctypes.c_void_p])
    keep, ptrs = _scene_ptrs(arrays, 1, None, None, None)
    out = np.zeros(7, np.float32)
    fn(ptrs, px, py, lu0, lu1, out.ctypes.data)
    return out[0:3].astype(np.float64), out[3:6].astype(np.float64), float(out[6])


def normal_map_probe(s_normal, u_dir, tangent, texel):
    """Shading frame (N, S, T) after normal mapping a surface frame by a texel."""
    fp = ctypes.POINTER(ctypes.c_float)
    fn = _probe_fn("orc_normal_map_probe", [fp, fp, fp, fp, fp])
    arr = lambda v: (ctypes.c_float * 3)(*[float(x) for x in v])
    out = (ctypes.c_float * 9)()
    fn(arr(s_normal), arr(u_dir), arr(tangent), arr(texel), out)
    o = np.array(list(out), np.float64)
    return o[0:3], o[3:6], o[6:9]


def texture_probe(record: np.ndarray, u, v):
    """evalColor of one texture record (bytes) at (u, v), as float32 [3]."""
    fn = _probe_fn("orc_texture_probe", [ctypes.c_void_p, ctypes.c_float, ctypes.c_float, ctypes.c_void_p])
    rec = np.ascontiguousarray(record, dtype=np.uint8)
    out = np.zeros(3, np.float32)
    fn(rec.ctypes.data, u, v, out.ctypes.data)
    return out


def fresnel_probe(ftype, eta, k, eta_ext, eta_int, cosi):
    fp = ctypes.POINTER(ctypes.c_float)
    fn = _probe_fn("orc_fresnel_probe", [ctypes.c_int, fp, fp, ctypes.c_float, ctypes.c_float,
                                         ctypes.c_float, fp])
    out = (ctypes.c_float * 3)()
    arr = lambda v: (ctypes.c_float * 3)(*[float(x) for x in v])
    fn(ftype, arr(eta), arr(k), eta_ext, eta_int, cosi, out)
    return list(out)


# Planted mistakes (orc_set_plant): each makes the restatement commit one
# error, so a test can show that its pin detects it.
PLANT_LIGHT_NO_INV_AREA = 1
PLANT_AREA_PDF_NO_INV_AREA = 2
PLANT_NEE_NO_COS_L = 3
PLANT_MIS_NO_COS_L = 4
PLANT_NEE_MIS_SWAPPED = 5
PLANT_IMPLICIT_NO_DIST2 = 6
PLANT_MIX_NO_AVERAGE = 7
PLANT_MIX_OWN_PDF_ONLY = 8
PLANT_NO_LIGHT_CHOICE_PROB = 9
PLANT_LE_NO_INV_PI = 10
PLANT_LE_TWO_SIDED = 11
PLANT_BECKMANN_NO_4 = 12
PLANT_MIX_PICK_FIRST = 13


class planted:
    """Context manager: the oracle commits mistake `k` inside the block."""

    def __init__(self, k: int):
        self.k = k

    def __enter__(self):
        _probe_fn("orc_set_plant", [ctypes.c_int])(self.k)
        return self

    def __exit__(self, *exc):
        _probe_fn("orc_set_plant", [ctypes.c_int])(0)
        return False


def camera_rays(arrays, film):
    """Camera rays for film points film[i] = (px, py, lu0, lu1): (origins
    [n, 3], directions [n, 3], direction pdfs [n]) in float64, and the camera's
    +z axis in world space."""
    fn = _probe_fn("orc_camera_rays", [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p])
    keep, ptrs = _scene_ptrs(arrays, 1, None, None, None)
    film = np.ascontiguousarray(film, dtype=np.float32).reshape(-1, 4)
    n = film.shape[0]
    out = np.zeros(7 * n + 3, np.float32)
    fn(ptrs, n, film.ctypes.data, out.ctypes.data)
    r = out[: 7 * n].reshape(n, 7).astype(np.float64)
    return r[:, 0:3], r[:, 3:6], r[:, 6], out[7 * n:].astype(np.float64)


def bsdf_probe(arrays, mat_offset, wo, wi, u):
    """BSDF mixture of the material at byte offset `mat_offset`, at a point
    with normal +z: dict of fs [n, 3], fs_pdf [n] at (wo, wi), and sample_fs
    at (wo, u = (uc, u0, u1)): direction [n, 3], pdf [n], value [n, 3], type [n]."""
    fn = _probe_fn("orc_bsdf_probe", [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p,
                                      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p])
    keep, ptrs = _scene_ptrs(arrays, 1, None, None, None)
    wo = np.ascontiguousarray(wo, dtype=np.float32).reshape(-1, 3)
    n = wo.shape[0]
    wi = np.ascontiguousarray(np.broadcast_to(np.asarray(wi, np.float32).reshape(-1, 3), (n, 3)))
    u = np.ascontiguousarray(np.broadcast_to(np.asarray(u, np.float32).reshape(-1, 3), (n, 3)))
    out = np.zeros((n, 12), np.float32)
    fn(ptrs, mat_offset, n, wo.ctypes.data, wi.ctypes.data, u.ctypes.data, out.ctypes.data)
    return {"f": out[:, 0:3], "pdf": out[:, 3], "wi": out[:, 4:7], "spdf": out[:, 7], "sf": out[:, 8:11],
            "type": out[:, 11].astype(np.uint32)}


def disk_probe(u1, u2):
    fn = _probe_fn("orc_disk_probe", [ctypes.c_float, ctypes.c_float, ctypes.POINTER(ctypes.c_float)])
    out = (ctypes.c_float * 2)()
    fn(u1, u2, out)
    return list(out)


def render_adjoint(arrays, num_faces, inodes, leaves, width, height, spp, max_depth, flags, mat_offsets,
                   light_prop, states, d_image, threads=0, bvh=None):
    """CPU adjoint: (g_albedo [M,3], g_sigma [M], g_emission [3])."""
    L = lib()
    fn = L.orc_render_adjoint
    fn.restype = ctypes.c_int
    vp, u32 = ctypes.c_void_p, ctypes.c_uint32
    fn.argtypes = [vp, u32, u32, u32, u32, u32, u32, u32, vp, u32, vp, vp, vp, vp, vp, ctypes.c_int]
    keep, ptrs = _scene_ptrs(arrays, num_faces, inodes, leaves, bvh)
    m = len(mat_offsets)
    mo = np.ascontiguousarray(mat_offsets, dtype=np.uint32)
    ga = np.zeros((max(m, 1), 3), np.float32)
    gs = np.zeros(max(m, 1), np.float32)
    ge = np.zeros(3, np.float32)
    st = np.ascontiguousarray(states, dtype=np.uint32)
    di = np.ascontiguousarray(d_image, dtype=np.float32)
    rc = fn(ptrs, num_faces, width, height, spp, max_depth, flags, m, _p(mo), light_prop, _p(st), _p(di),
            _p(ga), _p(gs), _p(ge), threads)
    if rc:
        raise RuntimeError(f"orc_render_adjoint failed: {rc}")
    return ga[:m], gs[:m], ge


# ---- post-processing restatement (post_oracle.cpp) --------------------------

def _post(name, *args):
    fn = getattr(lib(), name)
    fn.restype = ctypes.c_int
    return fn(*args)


def post_scaling(src, spp):
    """src float32 [H, W, 4] -> src * (1 / spp) (channel 3 untouched: 0)."""
    src = np.ascontiguousarray(src, dtype=np.float32)
    h, w = src.shape[:2]
    out = np.zeros_like(src)
    _post("orc_post_scaling", ctypes.c_uint32(w), ctypes.c_uint32(h), ctypes.c_uint32(spp), _p(src), _p(out))
    return out


def post_gaussian(src, spp, out=None):
    """out (zeros by default) += the reference's clamped 9x9 Gaussian scatter."""
    src = np.ascontiguousarray(src, dtype=np.float32)
    h, w = src.shape[:2]
    out = np.zeros_like(src) if out is None else out
    _post("orc_post_gaussian", ctypes.c_uint32(w), ctypes.c_uint32(h), ctypes.c_uint32(spp), _p(src), _p(out))
    return out


def post_bloom(src, spp, out=None):
    src = np.ascontiguousarray(src, dtype=np.float32)
    h, w = src.shape[:2]
    out = np.zeros_like(src) if out is None else out
    _post("orc_post_bloom", ctypes.c_uint32(w), ctypes.c_uint32(h), ctypes.c_uint32(spp), _p(src), _p(out))
    return out


def post_tone(src, row_bytes):
    """Tone-mapped bottom-up BGR rows, uint8 [H * row_bytes]."""
    src = np.ascontiguousarray(src, dtype=np.float32)
    h, w = src.shape[:2]
    dst = np.full(h * row_bytes, 0xAB, np.uint8)
    _post("orc_post_tone", ctypes.c_uint32(w), ctypes.c_uint32(h), ctypes.c_uint32(row_bytes), _p(src), _p(dst))
    return dst


def bvh_visualizer(arrays, num_faces, inodes, leaves, width, height, flags=0, bvh=None):
    """Visualizer restatement: float32 [H*W*4] (added onto zeros)."""
    L = lib()
    fn = L.orc_bvh_visualizer
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                   ctypes.c_void_p]
    keep, ptrs = _scene_ptrs(arrays, num_faces, inodes, leaves, bvh)
    px = np.zeros(width * height * 4, np.float32)
    rc = fn(ptrs, num_faces, width, height, flags, _p(px))
    if rc:
        raise RuntimeError(f"orc_bvh_visualizer failed: {rc}")
    return px


def centre_rays(arrays, width, height):
    """The visualizer's rays (lens centre through pixel centres), float32 [7, W*H]."""
    L = lib()
    fn = L.orc_centre_rays
    fn.restype = ctypes.c_int
    keep, ptrs = _scene_ptrs(arrays, 1, None, None, None)
    rays = np.zeros((7, width * height), np.float32)
    fn(ptrs, ctypes.c_uint32(width), ctypes.c_uint32(height), _p(rays))
    return rays
