"""Exact host validation of final placements (ctypes over
include/spectrapack_validate.h, in libspectrapack_compute.so; host code, no
GPU).  The role of engine/pack_geometry/src/validation.cpp:517-634."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libspectrapack_compute.so")
_lib = None


class Pose(C.Structure):
    _fields_ = [("translation_mm", C.c_double * 3), ("rotation_xyzw", C.c_double * 4)]


class Report(C.Structure):
    _fields_ = [("validity", C.c_int32), ("reserved", C.c_uint32), ("first", C.c_uint64), ("second", C.c_uint64),
                ("pair_tests", C.c_uint64), ("triangle_tests", C.c_uint64), ("exact_evaluations", C.c_uint64),
                ("code", C.c_char * 48)]


VALIDITY = {0: "valid", 1: "invalid", 2: "indeterminate"}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"host library not built: {LIB_PATH} (run cpp-max-packer_amd/build.py)")
        _lib = C.CDLL(LIB_PATH)
    return _lib


def _mesh(vertices, triangles):
    v = np.ascontiguousarray(np.asarray(vertices, dtype=np.float64).reshape(-1, 3))
    t = np.ascontiguousarray(np.asarray(triangles, dtype=np.uint32).reshape(-1, 3))
    return v, t


def _poses(poses):
    arr = (Pose * max(len(poses), 1))()
    for i, (t, q) in enumerate(poses):
        arr[i] = Pose((C.c_double * 3)(*[float(x) for x in t]), (C.c_double * 4)(*[float(x) for x in q]))
    return arr


def _report(r: Report):
    none = 0xFFFFFFFFFFFFFFFF
    return dict(validity=VALIDITY[r.validity], code=r.code.decode(), first=None if r.first == none else int(r.first),
                second=None if r.second == none else int(r.second), pair_tests=int(r.pair_tests),
                triangle_tests=int(r.triangle_tests), exact_evaluations=int(r.exact_evaluations))


def _container(box_dims, container):
    if box_dims is not None:
        return (C.c_double * 3)(*[float(x) for x in box_dims]), None, None, 0, 0, ()
    cv, ct = _mesh(*container)
    return (None, cv.ctypes.data_as(C.POINTER(C.c_double)), ct.ctypes.data_as(C.POINTER(C.c_uint32)), len(cv), len(ct),
            (cv, ct))


def validate_placements(part, poses, box_dims=None, container=None, pair_clearance=0.0, wall_clearance=0.0):
    """part: (vertices, triangles); poses: [(translation_mm, rotation_xyzw)];
    box_dims: the box [0, dims] container, or container: (vertices, triangles)."""
    pv, pt = _mesh(*part)
    box, cvp, ctp, ncv, nct, keep = _container(box_dims, container)
    ps = _poses(poses)
    r = Report()
    rc = lib().sp_validate_placements(pv.ctypes.data_as(C.POINTER(C.c_double)), C.c_uint32(len(pv)),
                                      pt.ctypes.data_as(C.POINTER(C.c_uint32)), C.c_uint32(len(pt)), ps,
                                      C.c_uint64(len(poses)), box, cvp, C.c_uint32(ncv), ctp, C.c_uint32(nct),
                                      C.c_double(pair_clearance), C.c_double(wall_clearance), C.byref(r))
    if rc:
        raise ValueError("invalid validation arguments")
    return _report(r)


class Validator:
    """validate_incremental's role for a placement loop: try_add keeps a pose only when it is valid."""

    def __init__(self, part, box_dims=None, container=None, pair_clearance=0.0, wall_clearance=0.0):
        pv, pt = _mesh(*part)
        box, cvp, ctp, ncv, nct, keep = _container(box_dims, container)
        self._h = C.c_void_p()
        rc = lib().sp_validator_create(pv.ctypes.data_as(C.POINTER(C.c_double)), C.c_uint32(len(pv)),
                                       pt.ctypes.data_as(C.POINTER(C.c_uint32)), C.c_uint32(len(pt)), box, cvp,
                                       C.c_uint32(ncv), ctp, C.c_uint32(nct), C.c_double(pair_clearance),
                                       C.c_double(wall_clearance), C.byref(self._h))
        if rc:
            raise ValueError("invalid validator arguments")

    def try_add(self, translation_mm, rotation_xyzw):
        p = _poses([(translation_mm, rotation_xyzw)])
        r = Report()
        if lib().sp_validator_try_add(self._h, p, C.byref(r)):
            raise ValueError("invalid pose")
        return _report(r)

    def close(self):
        if self._h:
            lib().sp_validator_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
