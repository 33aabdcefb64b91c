"""ctypes wrapper of the CPU oracle (liboracle.so) — test infrastructure.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import
this module, as the checker or the timed CPU baseline; the product path never
does.  Each wrapper names the oracle function it calls; the oracle's sources
cite the reference file:line they restate.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref_intcorr.so")

_lib = None
_ref = None


def build(reference: bool = False) -> None:
    """Compile liboracle.so (and, when asked, _ref/ from the reference tree)."""
    targets = ["all"] + (["ref"] if reference else [])
    subprocess.check_call(["make", "-s", "-C", HERE] + targets)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run oracle/Makefile)")
        _lib = C.CDLL(LIB_PATH)
    return _lib


def ref_lib():
    """The reference's own direct integer oracle, if oracle/_ref was built."""
    global _ref
    if _ref is None and os.path.exists(REF_PATH):
        _ref = C.CDLL(REF_PATH)
    return _ref


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def _u32(v):
    return np.ascontiguousarray(np.asarray(v, dtype=np.uint32))


def _i64(v):
    return np.ascontiguousarray(np.asarray(v, dtype=np.int64))


def direct_correlation(blocked, b_ext, kernel, k_ext, origin, use_reference=False):
    """Direct integer linear cross-correlation (tests/support/src/integer_correlation.cpp:76-130).

    Returns (range[6], extent[3], values) or raises ValueError/OverflowError."""
    L = ref_lib() if use_reference else lib()
    fn = L.ref_direct_correlation if use_reference else L.orc_direct_correlation
    b = np.ascontiguousarray(np.asarray(blocked, dtype=np.int32))
    k = np.ascontiguousarray(np.asarray(kernel, dtype=np.int32))
    be = np.asarray(b_ext, dtype=np.int32)
    ke = np.asarray(k_ext, dtype=np.int32)
    org = np.asarray(origin, dtype=np.int32)
    rng = np.zeros(6, np.int32)
    ext = np.zeros(3, np.int32)
    cap = 1
    if (be > 0).all() and (ke > 0).all():
        cap = int(np.prod(be.astype(np.int64) + ke - 1))
    vals = np.zeros(max(cap, 1), np.int32)
    rc = fn(_p(b, C.c_int), _p(be, C.c_int), _p(k, C.c_int), _p(ke, C.c_int), _p(org, C.c_int),
            _p(rng, C.c_int), _p(ext, C.c_int), _p(vals, C.c_int), C.c_uint64(vals.size))
    if rc == 1:
        raise ValueError("invalid_argument")
    if rc == 2:
        raise OverflowError("overflow_error")
    return rng, ext, vals[: int(np.prod(ext.astype(np.int64)))]


LIMITS_DEFAULT = (512 << 20, 0, 16777216, 200000000)
LIMITS_UNBOUNDED = ((1 << 62), 0, (1 << 62), (1 << 63))


class Failure(Exception):
    def __init__(self, code, stats):
        super().__init__(code)
        self.code = code
        self.stats = stats


def correlate(spec, env, kernel, limits=LIMITS_DEFAULT, fault=0, proximity=False):
    """Checked full-linear correlation (engine/pack_compute/src/correlation.cpp:238-592).

    spec = (env_shape, kernel_shape, env_first, kernel_first).  Returns a dict
    with translation_first, shape, values (float64, X-fastest), report, stats;
    raises Failure(code) on the reference's failure variants."""
    es, ks, ef, kf = (_u32(spec[0]), _u32(spec[1]), _i64(spec[2]), _i64(spec[3]))
    envs = np.ascontiguousarray(env, dtype=np.float64 if proximity else np.uint8).ravel()
    ker = np.ascontiguousarray(kernel, dtype=np.uint8).ravel()
    padded = (es.astype(np.int64) + ks - 1)
    cap = int(np.prod(padded)) if (es > 0).all() and (ks > 0).all() else 1
    if cap > (1 << 31):
        cap = 1
    vals = np.zeros(max(cap, 1), np.float64)
    lim = np.asarray(limits, dtype=np.uint64)
    first = np.zeros(3, np.int64)
    shape = np.zeros(3, np.uint32)
    rep = np.zeros(3, np.float64)
    st = np.zeros(3, np.uint64)
    code = C.create_string_buffer(64)
    fn = lib().orc_correlate_proximity if proximity else lib().orc_correlate_binary
    envp = _p(envs, C.c_double) if proximity else _p(envs, C.c_uint8)
    rc = fn(_p(es, C.c_uint32), _p(ks, C.c_uint32), _p(ef, C.c_int64), _p(kf, C.c_int64), envp,
            C.c_uint64(envs.size), _p(ker, C.c_uint8), C.c_uint64(ker.size), _p(lim, C.c_uint64), C.c_int(fault),
            _p(first, C.c_int64), _p(shape, C.c_uint32), _p(vals, C.c_double), C.c_uint64(vals.size),
            _p(rep, C.c_double), _p(st, C.c_uint64), code, 64)
    stats = dict(padded_cells=int(st[0]), working_bytes_peak=int(st[1]), direct_terms=int(st[2]))
    if rc == 1:
        raise Failure(code.value.decode(), stats)
    if rc == 2:
        raise MemoryError("oracle output capacity")
    n = int(np.prod(shape.astype(np.int64)))
    return dict(translation_first=first, shape=shape, values=vals[:n].copy(),
                report=dict(max_integer_residual=rep[0], mass_residual=rep[1], max_probe_error=rep[2]), stats=stats)


def fft_1d(x, forward=True):
    a = np.ascontiguousarray(np.asarray(x, dtype=np.complex128)).copy()
    lib().orc_fft_1d(_p(a.view(np.float64), C.c_double), C.c_uint32(a.size), C.c_int(1 if forward else 0))
    return a


def cube_catalog():
    out = np.zeros((24, 4), np.float64)
    lib().orc_cube_catalog(_p(out, C.c_double))
    return out


def cube45_catalog():
    out = np.zeros((96, 4), np.float64)
    lib().orc_cube45_catalog(_p(out, C.c_double))
    return out


def custom_catalog(raw):
    raw = np.ascontiguousarray(np.asarray(raw, dtype=np.float64).reshape(-1, 4))
    out = np.zeros_like(raw)
    n = C.c_uint32(0)
    if lib().orc_custom_catalog(_p(raw, C.c_double), C.c_uint32(raw.shape[0]), _p(out, C.c_double), C.byref(n)):
        raise ValueError("ORIENTATION_INVALID")
    return out[: n.value].copy()


def _mesh(vertices, triangles):
    v = np.ascontiguousarray(np.asarray(vertices, dtype=np.float64).reshape(-1, 3))
    t = np.ascontiguousarray(np.asarray(triangles, dtype=np.uint32).reshape(-1, 3))
    return v, t


def voxelize_object(vertices, triangles, q, pitch, origin=(0.0, 0.0, 0.0), raw=False):
    """Oriented object kernel (conservative_fields.cpp:919-1003): (first, shape, cells[z,y,x])."""
    v, t = _mesh(vertices, triangles)
    q = np.asarray(q, dtype=np.float64)
    org = np.asarray(origin, dtype=np.float64)
    first = np.zeros(3, np.int64)
    shape = np.zeros(3, np.uint32)
    cap = 1
    for _ in range(2):
        cells = np.zeros(cap, np.uint8)
        tags = np.zeros(cap, np.uint8) if raw else None
        rc = lib().orc_voxelize_object(_p(v, C.c_double), C.c_uint32(len(v)), _p(t, C.c_uint32), C.c_uint32(len(t)),
                                       _p(org, C.c_double), C.c_double(pitch), _p(q, C.c_double), _p(first, C.c_int64),
                                       _p(shape, C.c_uint32), _p(cells, C.c_uint8), C.c_uint64(cap),
                                       _p(tags, C.c_uint8) if raw else None)
        if rc == 1:
            raise ValueError("voxelize_object failed")
        if rc == 0:
            n = int(np.prod(shape.astype(np.int64)))
            c = cells[:n].reshape(int(shape[2]), int(shape[1]), int(shape[0]))
            if raw:
                return first.copy(), shape.copy(), c, tags[:n].reshape(c.shape)
            return first.copy(), shape.copy(), c
        cap = int(np.prod(shape.astype(np.int64)))
    raise RuntimeError("unreachable")


def object_bounds(vertices, triangles, q):
    v, t = _mesh(vertices, triangles)
    q = np.asarray(q, dtype=np.float64)
    lo = np.zeros(3)
    hi = np.zeros(3)
    fl = np.zeros(3, np.int32)
    if lib().orc_object_bounds(_p(v, C.c_double), C.c_uint32(len(v)), _p(t, C.c_uint32), C.c_uint32(len(t)),
                               _p(q, C.c_double), _p(lo, C.c_double), _p(hi, C.c_double), _p(fl, C.c_int)):
        raise ValueError("place failed")
    return lo, hi, dict(is_cuboid=bool(fl[0]), is_cardinal=bool(fl[1]), represented_world_exact=bool(fl[2]))


def box_container(first, shape, pitch, dims, clearance=0.0, origin=(0.0, 0.0, 0.0)):
    f, s = _i64(first), _u32(shape)
    cells = np.zeros(int(np.prod(s.astype(np.int64))), np.uint8)
    d = np.asarray(dims, dtype=np.float64)
    org = np.asarray(origin, dtype=np.float64)
    if lib().orc_box_container(_p(f, C.c_int64), _p(s, C.c_uint32), _p(org, C.c_double), C.c_double(pitch),
                               _p(d, C.c_double), C.c_double(clearance), _p(cells, C.c_uint8)):
        raise ValueError("box container failed")
    return cells.reshape(int(s[2]), int(s[1]), int(s[0]))


def mesh_container(vertices, triangles, first, shape, pitch, clearance=0.0, origin=(0.0, 0.0, 0.0)):
    v, t = _mesh(vertices, triangles)
    f, s = _i64(first), _u32(shape)
    cells = np.zeros(int(np.prod(s.astype(np.int64))), np.uint8)
    org = np.asarray(origin, dtype=np.float64)
    if lib().orc_mesh_container(_p(v, C.c_double), C.c_uint32(len(v)), _p(t, C.c_uint32), C.c_uint32(len(t)),
                                _p(f, C.c_int64), _p(s, C.c_uint32), _p(org, C.c_double), C.c_double(pitch),
                                C.c_double(clearance), _p(cells, C.c_uint8)):
        raise ValueError("mesh container failed")
    return cells.reshape(int(s[2]), int(s[1]), int(s[0]))


def _kernel_arrays(kernels):
    n = len(kernels)
    cells = [np.ascontiguousarray(k["cells"], dtype=np.uint8).ravel() for k in kernels]
    ptrs = (C.POINTER(C.c_uint8) * n)(*[_p(c, C.c_uint8) for c in cells])
    ks = _u32([k["shape"] for k in kernels]).ravel()
    kf = _i64([k["first"] for k in kernels]).ravel()
    return cells, ptrs, ks, kf


def place_step(env, env_first, kernels, direct=False, fields=False, excluded=()):
    """One plain step (gap_fill.cpp:363-509 rule).  env is [z,y,x] uint8.

    kernels: list of dicts with first (x,y,z), shape (x,y,z), cells [z,y,x].
    Returns (orientation, translation(x,y,z), legal_fields or None)."""
    env = np.ascontiguousarray(env, dtype=np.uint8)
    es = _u32(env.shape[::-1])
    ef = _i64(env_first)
    cells, ptrs, ks, kf = _kernel_arrays(kernels)
    legal_sizes = []
    for k in kernels:
        sh = np.asarray(k["shape"], dtype=np.int64)
        leg = es.astype(np.int64) - sh + 1
        legal_sizes.append(int(np.prod(leg)) if (leg > 0).all() else 0)
    out = np.zeros(max(sum(legal_sizes), 1), np.int32) if fields else None
    o = C.c_int(-1)
    t = np.zeros(3, np.int64)
    code = C.create_string_buffer(64)
    ex = np.ascontiguousarray(np.asarray([[o_, *t_] for o_, t_ in excluded] or [[0, 0, 0, 0]], dtype=np.int64))
    rc = lib().orc_place_step(_p(env, C.c_uint8), _p(ef, C.c_int64), _p(es, C.c_uint32), C.c_uint32(len(kernels)),
                              ptrs, _p(ks, C.c_uint32), _p(kf, C.c_int64), C.c_int(1 if direct else 0), C.byref(o),
                              _p(t, C.c_int64), _p(out, C.c_int32) if fields else None, code, 64, _p(ex, C.c_int64),
                              C.c_uint32(len(excluded)))
    if rc:
        raise Failure(code.value.decode(), {})
    flds = None
    if fields:
        flds, off = [], 0
        for k, n in zip(kernels, legal_sizes):
            if n == 0:
                flds.append(None)
                continue
            sh = np.asarray(k["shape"], dtype=np.int64)
            leg = es.astype(np.int64) - sh + 1
            flds.append(out[off:off + n].reshape(int(leg[2]), int(leg[1]), int(leg[0])))
            off += n
    return o.value, t.copy(), flds


def correlate_orientations(env, env_first, kernels, begin, end, threads=1):
    """CPU-baseline sample: FFT restatement correlations of orientations [begin, end)."""
    env = np.ascontiguousarray(env, dtype=np.uint8)
    es = _u32(env.shape[::-1])
    ef = _i64(env_first)
    cells, ptrs, ks, kf = _kernel_arrays(kernels)
    chk = C.c_int64(0)
    done = lib().orc_correlate_orientations(_p(env, C.c_uint8), _p(ef, C.c_int64), _p(es, C.c_uint32),
                                            C.c_uint32(len(kernels)), ptrs, _p(ks, C.c_uint32), _p(kf, C.c_int64),
                                            C.c_uint32(begin), C.c_uint32(end), C.c_uint32(threads), C.byref(chk))
    return done, chk.value


def clearance_stencil(pitch, clearance):
    """Pair-clearance stencil offsets (x, y, z) as an int64 [n, 3] array
    (conservative_fields.cpp:189-204, 398-434; validation_kernel.cpp:3492-3522)."""
    cap = 1 << 20
    out = np.zeros((cap, 3), np.int64)
    n = lib().orc_clearance_stencil(C.c_double(pitch), C.c_double(clearance), _p(out, C.c_int64), C.c_uint32(cap))
    if n < 0:
        raise Failure("FIELD_INPUT", {})
    return out[:n].copy()


def stamp(env, env_first, kernel, t, stencil=None):
    """Sets the kernel's cells at translation t; with a clearance stencil,
    every cell within the stencil of a kernel cell (a placed copy grown by
    the pair clearance, conservative_fields.cpp:1015-1122)."""
    if stencil is not None and len(stencil):
        env = np.ascontiguousarray(env, dtype=np.uint8).copy()
        ez, ey, ex = env.shape
        z, y, x = np.nonzero(np.asarray(kernel["cells"]) != 0)
        base = np.stack([x, y, z], axis=1) + (np.asarray(t, np.int64) + np.asarray(kernel["first"], np.int64) -
                                              np.asarray(env_first, np.int64))
        for d in np.asarray(stencil, np.int64):
            p = base + d
            ok = (p >= 0).all(axis=1) & (p[:, 0] < ex) & (p[:, 1] < ey) & (p[:, 2] < ez)
            p = p[ok]
            env[p[:, 2], p[:, 1], p[:, 0]] = 1
        return env
    env = np.ascontiguousarray(env, dtype=np.uint8).copy()
    es = _u32(env.shape[::-1])
    ef = _i64(env_first)
    kc = np.ascontiguousarray(kernel["cells"], dtype=np.uint8).ravel()
    ks = _u32(kernel["shape"])
    kf = _i64(kernel["first"])
    tt = _i64(t)
    lib().orc_stamp(_p(env, C.c_uint8), _p(ef, C.c_int64), _p(es, C.c_uint32), _p(kc, C.c_uint8),
                    _p(ks, C.c_uint32), _p(kf, C.c_int64), _p(tt, C.c_int64))
    return env


def proximity(occupied):
    occ = np.ascontiguousarray(occupied, dtype=np.uint8)
    sh = _u32(occ.shape[::-1])
    vals = np.zeros(occ.size, np.float64)
    sq = np.zeros(occ.size, np.uint64)
    lib().orc_proximity(_p(occ, C.c_uint8), _p(sh, C.c_uint32), _p(vals, C.c_double), _p(sq, C.c_uint64))
    return vals.reshape(occ.shape), sq.reshape(occ.shape)


def orient3d(a, b, c, d):
    arr = [np.asarray(x, dtype=np.float64) for x in (a, b, c, d)]
    return lib().orc_orient3d(*[_p(x, C.c_double) for x in arr])


def segment_triangle_crossing(p, q, tri):
    p = np.asarray(p, dtype=np.float64)
    q = np.asarray(q, dtype=np.float64)
    tri = np.ascontiguousarray(np.asarray(tri, dtype=np.float64).reshape(9))
    return lib().orc_segment_triangle_crossing(_p(p, C.c_double), _p(q, C.c_double), _p(tri, C.c_double))
