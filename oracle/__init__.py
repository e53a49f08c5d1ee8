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
