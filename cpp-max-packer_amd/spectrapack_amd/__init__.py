"""Python binding of the MI355X spectral placement step (ctypes over the C ABI
declared in include/spectrapack_hip.h).

This is the host-side mirror used by tests and the benchmark: names, argument
meaning and failure codes follow engine/pack_compute's correlation interface
(engine/pack_compute/include/spectrapack/compute/correlation.hpp:17-117) and the
plain gap-fill step (engine/pack_solver/src/gap_fill.cpp:363-509).  Every call
runs the HIP library; there is no CPU fallback — loading fails loudly when the
library is missing.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libspectrapack_hip.so")

_lib = None

SP_PRECISION_AUTO, SP_PRECISION_FP32, SP_PRECISION_FP64 = 0, 32, 64


class Spec(C.Structure):
    _fields_ = [("environment_shape", C.c_uint32 * 3), ("kernel_shape", C.c_uint32 * 3),
                ("environment_first", C.c_int64 * 3), ("kernel_first", C.c_int64 * 3)]


class Limits(C.Structure):
    _fields_ = [("max_working_bytes", C.c_uint64), ("reserved_bytes", C.c_uint64),
                ("max_padded_cells", C.c_uint64), ("max_direct_terms", C.c_uint64), ("metadata_bytes", C.c_uint64)]


class Stats(C.Structure):
    _fields_ = [("padded_cells", C.c_uint64), ("working_bytes_peak", C.c_uint64), ("direct_terms", C.c_uint64)]


class Report(C.Structure):
    _fields_ = [("max_integer_residual", C.c_double), ("mass_residual", C.c_double),
                ("max_probe_error", C.c_double), ("error_bound", C.c_double), ("precision", C.c_int32),
                ("fft_shape", C.c_uint32 * 3)]


class Estimate(C.Structure):
    _fields_ = [("padded_shape", C.c_uint32 * 3), ("padded_cells", C.c_uint64), ("working_bytes", C.c_uint64),
                ("device_bytes", C.c_uint64)]


class SessionDesc(C.Structure):
    _fields_ = [("environment_shape", C.c_uint32 * 3), ("environment_first", C.c_int64 * 3),
                ("orientations", C.c_uint32), ("precision", C.c_int32), ("capture_graph", C.c_uint32)]


class StepResult(C.Structure):
    _fields_ = [("orientation", C.c_int32), ("run_flags", C.c_uint32), ("translation", C.c_int64 * 3),
                ("key", C.c_int64 * 3), ("max_residual", C.c_double), ("numeric_flags", C.c_uint32),
                ("precision", C.c_int32), ("error_bound", C.c_double)]


class RankingDesc(C.Structure):
    _fields_ = [("lattice_origin_z_mm", C.c_double), ("pitch_mm", C.c_double), ("container_min_z_mm", C.c_double),
                ("container_max_z_mm", C.c_double)]


class RankedCandidate(C.Structure):
    _fields_ = [("translation", C.c_int64 * 3), ("score", C.c_double), ("binary_overlap", C.c_double),
                ("orientation", C.c_uint32), ("direct_zero_overlap", C.c_uint32)]


RANKED_DTYPE = np.dtype([("translation", np.int64, 3), ("score", np.float64), ("binary_overlap", np.float64),
                         ("orientation", np.uint32), ("direct_zero_overlap", np.uint32)])
assert RANKED_DTYPE.itemsize == C.sizeof(RankedCandidate)


class Placement(C.Structure):
    _fields_ = [("translation", C.c_int64 * 3), ("orientation", C.c_uint32), ("reserved", C.c_uint32),
                ("score", C.c_double)]


class PlacedCopy(C.Structure):
    _fields_ = [("copy_id", C.c_uint64), ("translation", C.c_int64 * 3), ("orientation", C.c_uint32),
                ("removed", C.c_uint32)]


VALIDATE_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_uint32, C.POINTER(C.c_int64))

PAGE_SIZE = 32
PAGE_ORDINARY, PAGE_LOW_OVERLAP = 0, 1


class Failure(Exception):
    """A CorrelationFailure / RepresentationFailure variant (code string + stats)."""

    def __init__(self, status, code, message="", stats=None):
        super().__init__(f"{code}: {message}" if message else code)
        self.status = status
        self.code = code
        self.message = message
        self.stats = stats or {}


def library_path() -> str:
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP library not built: {LIB_PATH} (run cpp-max-packer_amd/build.py)")
        L = C.CDLL(LIB_PATH)
        L.sp_hip_status_code.restype = C.c_char_p
        L.sp_hip_last_error.restype = C.c_char_p
        L.sp_hip_session_stream.restype = C.c_void_p
        L.sp_hip_session_device_bytes.restype = C.c_uint64
        L.sp_hip_fp32_error_bound.restype = C.c_double
        L.sp_hip_fp32_error_bound.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
        _lib = L
    return _lib


def status_code(status: int) -> str:
    return lib().sp_hip_status_code(int(status)).decode()


def fp32_error_bound(fft_cells: int, env_mass: int, kernel_mass: int) -> float:
    return lib().sp_hip_fp32_error_bound(int(fft_cells), int(env_mass), int(kernel_mass))


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


class Context:
    """The sp_hip_ctx handle (the role FftExecution plays: correlation.hpp:78-98)."""

    def __init__(self, device: int = 0, bytes_cap: int = 0):
        self.device = int(device)
        self._h = C.c_void_p()
        rc = lib().sp_hip_ctx_create(C.c_int(device), C.c_uint64(bytes_cap), C.byref(self._h))
        if rc:
            raise Failure(rc, status_code(rc), "context creation failed")

    @property
    def handle(self):
        return self._h

    def last_error(self) -> str:
        return lib().sp_hip_last_error(self._h).decode()

    def close(self):
        if self._h:
            lib().sp_hip_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _fail(self, rc, stats=None):
        raise Failure(rc, status_code(rc), self.last_error(), stats)

    def estimate(self, spec, fft_threads=1, metadata_bytes=0):
        return estimate(spec, fft_threads, metadata_bytes)

    def correlate(self, spec, environment, kernel, limits=None, precision=SP_PRECISION_AUTO, proximity=False,
                  noalloc=False):
        """correlate_binary_cpu / correlate_proximity_cpu on the device.

        limits: (max_working_bytes, reserved_bytes, max_padded_cells,
        max_direct_terms[, metadata_bytes]).  noalloc: the allocation-free
        entry point (binary only).  Returns dict(translation_first, shape,
        values, numeric, stats); raises Failure with the reference's code
        string otherwise."""
        s = _spec(spec)
        env = np.ascontiguousarray(environment, dtype=np.float64 if proximity else np.uint8).ravel()
        ker = np.ascontiguousarray(kernel, dtype=np.uint8).ravel()
        lim_t = tuple(limits) if limits is not None else (512 << 20, 0, 16777216, 200000000)
        lim = Limits(*(lim_t + (0,) * (5 - len(lim_t))))
        es = np.array(spec[0], dtype=np.int64)
        ks = np.array(spec[1], dtype=np.int64)
        n = int(np.prod(es + ks - 1)) if (es > 0).all() and (ks > 0).all() else 0
        out = np.zeros(max(n, 1), np.float64)
        first = (C.c_int64 * 3)()
        shape = (C.c_uint32 * 3)()
        rep = Report()
        st = Stats()
        if proximity:
            rc = lib().sp_hip_correlate_proximity(self._h, C.byref(s), _p(env, C.c_double), C.c_uint64(env.size),
                                                  _p(ker, C.c_uint8), C.c_uint64(ker.size), C.byref(lim),
                                                  _p(out, C.c_double), C.c_uint64(out.size), first, shape, C.byref(rep),
                                                  C.byref(st))
        else:
            fn = lib().sp_hip_correlate_binary_noalloc if noalloc else lib().sp_hip_correlate_binary
            rc = fn(self._h, C.byref(s), _p(env, C.c_uint8), C.c_uint64(env.size), _p(ker, C.c_uint8),
                    C.c_uint64(ker.size), C.byref(lim), C.c_int(precision), _p(out, C.c_double), C.c_uint64(out.size),
                    first, shape, C.byref(rep), C.byref(st))
        stats = dict(padded_cells=st.padded_cells, working_bytes_peak=st.working_bytes_peak,
                     direct_terms=st.direct_terms)
        if rc:
            self._fail(rc, stats)
        sh = tuple(shape)
        return dict(translation_first=np.array(tuple(first)), shape=np.array(sh),
                    values=out[: int(np.prod(sh))].copy(),
                    numeric=dict(max_integer_residual=rep.max_integer_residual, mass_residual=rep.mass_residual,
                                 max_probe_error=rep.max_probe_error, error_bound=rep.error_bound,
                                 precision=rep.precision, fft_shape=tuple(rep.fft_shape)),
                    stats=stats)


    # ---- rasteriser (conservative_fields.hpp:100-164) ----
    def voxelize_objects(self, vertices, triangles, quaternions, pitch, origin=(0.0, 0.0, 0.0)):
        """voxelize_object for every quaternion: list of dict(first, shape, cells[z,y,x])."""
        v = np.ascontiguousarray(vertices, dtype=np.float64).reshape(-1, 3)
        t = np.ascontiguousarray(triangles, dtype=np.uint32).reshape(-1, 3)
        q = np.ascontiguousarray(quaternions, dtype=np.float64).reshape(-1, 4)
        n = q.shape[0]
        org = np.asarray(origin, dtype=np.float64)
        firsts = np.zeros((n, 3), np.int64)
        shapes = np.zeros((n, 3), np.uint32)
        offs = np.zeros(n, np.uint64)
        req = C.c_uint64(0)
        args = (self._h, _p(v, C.c_double), C.c_uint32(len(v)), _p(t, C.c_uint32), C.c_uint32(len(t)),
                _p(org, C.c_double), C.c_double(pitch), _p(q, C.c_double), C.c_uint32(n), _p(firsts, C.c_int64),
                _p(shapes, C.c_uint32), _p(offs, C.c_uint64))
        rc = lib().sp_hip_voxelize_objects(*args, None, C.c_uint64(0), C.byref(req))
        if rc and rc != 37:
            self._fail(rc)
        cells = np.zeros(max(int(req.value), 1), np.uint8)
        rc = lib().sp_hip_voxelize_objects(*args, _p(cells, C.c_uint8), C.c_uint64(cells.size), C.byref(req))
        if rc:
            self._fail(rc)
        out = []
        for k in range(n):
            sh = tuple(int(x) for x in shapes[k])
            m = sh[0] * sh[1] * sh[2]
            o = int(offs[k])
            out.append(dict(first=tuple(int(x) for x in firsts[k]), shape=sh,
                            cells=cells[o:o + m].reshape(sh[2], sh[1], sh[0]).copy()))
        return out

    def box_container(self, first, shape, pitch, dims, clearance=0.0, origin=(0.0, 0.0, 0.0)):
        f = (C.c_int64 * 3)(*[int(x) for x in first])
        s = (C.c_uint32 * 3)(*[int(x) for x in shape])
        org = np.asarray(origin, dtype=np.float64)
        d = np.asarray(dims, dtype=np.float64)
        cells = np.zeros(int(np.prod([int(x) for x in shape])), np.uint8)
        rc = lib().sp_hip_voxelize_box_container(self._h, f, s, _p(org, C.c_double), C.c_double(pitch),
                                                 _p(d, C.c_double), C.c_double(clearance), _p(cells, C.c_uint8))
        if rc:
            self._fail(rc)
        return cells.reshape(int(shape[2]), int(shape[1]), int(shape[0]))

    def mesh_container(self, vertices, triangles, first, shape, pitch, clearance=0.0, origin=(0.0, 0.0, 0.0)):
        v = np.ascontiguousarray(vertices, dtype=np.float64).reshape(-1, 3)
        t = np.ascontiguousarray(triangles, dtype=np.uint32).reshape(-1, 3)
        f = (C.c_int64 * 3)(*[int(x) for x in first])
        s = (C.c_uint32 * 3)(*[int(x) for x in shape])
        org = np.asarray(origin, dtype=np.float64)
        cells = np.zeros(int(np.prod([int(x) for x in shape])), np.uint8)
        rc = lib().sp_hip_voxelize_mesh_container(self._h, _p(v, C.c_double), C.c_uint32(len(v)), _p(t, C.c_uint32),
                                                  C.c_uint32(len(t)), f, s, _p(org, C.c_double), C.c_double(pitch),
                                                  C.c_double(clearance), _p(cells, C.c_uint8))
        if rc:
            self._fail(rc)
        return cells.reshape(int(shape[2]), int(shape[1]), int(shape[0]))


def estimate(spec, fft_threads=1, metadata_bytes=0):
    """estimate_correlation_cpu (correlation.cpp:596-641) through the C ABI;
    needs no GPU.  working_bytes counts `metadata_bytes` for the caller's
    result objects; device_bytes is what one FP64 correlation takes here."""
    s = _spec(spec)
    e = Estimate()
    rc = lib().sp_hip_estimate(C.byref(s), C.c_uint32(fft_threads), C.c_uint64(metadata_bytes), C.byref(e))
    if rc:
        raise Failure(rc, status_code(rc))
    return dict(padded_shape=tuple(e.padded_shape), padded_cells=e.padded_cells, working_bytes=e.working_bytes,
                device_bytes=e.device_bytes)


def _spec(spec) -> Spec:
    es, ks, ef, kf = spec
    return Spec((C.c_uint32 * 3)(*[int(v) for v in es]), (C.c_uint32 * 3)(*[int(v) for v in ks]),
                (C.c_int64 * 3)(*[int(v) for v in ef]), (C.c_int64 * 3)(*[int(v) for v in kf]))


class Session:
    """Batched placement step with the occupancy resident in HBM."""

    def __init__(self, ctx: Context, env_shape, env_first, orientations, precision=SP_PRECISION_AUTO,
                 capture_graph=True):
        self.ctx = ctx
        self.env_shape = tuple(int(v) for v in env_shape)  # (x, y, z)
        self.env_first = tuple(int(v) for v in env_first)
        self.orientations = int(orientations)
        d = SessionDesc((C.c_uint32 * 3)(*self.env_shape), (C.c_int64 * 3)(*self.env_first),
                        C.c_uint32(self.orientations), C.c_int32(precision), C.c_uint32(1 if capture_graph else 0))
        self._h = C.c_void_p()
        rc = lib().sp_hip_session_create(ctx.handle, C.byref(d), C.byref(self._h))
        if rc:
            ctx._fail(rc)
        self.kernel_shapes = [None] * self.orientations

    def close(self):
        if self._h:
            lib().sp_hip_session_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc:
            self.ctx._fail(rc)

    def set_environment(self, cells):
        """cells: uint8 array [z, y, x] (X-fastest)."""
        a = np.ascontiguousarray(cells, dtype=np.uint8).ravel()
        self._check(lib().sp_hip_session_set_environment(self._h, _p(a, C.c_uint8), C.c_uint64(a.size)))

    def environment(self):
        ex, ey, ez = self.env_shape
        a = np.zeros(ex * ey * ez, np.uint8)
        self._check(lib().sp_hip_session_get_environment(self._h, _p(a, C.c_uint8), C.c_uint64(a.size)))
        return a.reshape(ez, ey, ex)

    def set_kernel(self, o, cells, shape, first):
        """cells [z, y, x] uint8; shape and first as (x, y, z)."""
        c = np.asarray(cells)
        if c.ndim == 3 and tuple(c.shape) != tuple(int(v) for v in shape)[::-1]:
            raise ValueError(f"kernel cells have shape {c.shape}, expected [z, y, x] = {tuple(shape)[::-1]}")
        a = np.ascontiguousarray(c, dtype=np.uint8).ravel()
        sh = (C.c_uint32 * 3)(*[int(v) for v in shape])
        fi = (C.c_int64 * 3)(*[int(v) for v in first])
        self._check(lib().sp_hip_session_set_kernel(self._h, C.c_uint32(o), _p(a, C.c_uint8), C.c_uint64(a.size),
                                                    sh, fi))
        self.kernel_shapes[o] = tuple(int(v) for v in shape)

    def set_autotune(self, on=True):
        self._check(lib().sp_hip_session_set_autotune(self._h, C.c_int(1 if on else 0)))

    def set_tile(self, pass_, radix, width):
        self._check(lib().sp_hip_session_set_tile(self._h, C.c_int(pass_), C.c_int(radix), C.c_int(width)))

    def tile(self, pass_):
        r, w, ms = C.c_int32(), C.c_int32(), C.c_float()
        self._check(lib().sp_hip_session_get_tile(self._h, C.c_int(pass_), C.byref(r), C.byref(w), C.byref(ms)))
        return r.value, w.value, ms.value

    def tile_candidates(self, pass_):
        pairs = (C.c_int32 * 64)()
        n = C.c_uint32()
        self._check(lib().sp_hip_session_tile_candidates(self._h, C.c_int(pass_), pairs, C.c_uint32(32), C.byref(n)))
        return [(pairs[2 * i], pairs[2 * i + 1]) for i in range(n.value)]

    def set_pair_clearance(self, clearance_mm, pitch_mm):
        self._check(lib().sp_hip_session_set_pair_clearance(self._h, C.c_double(clearance_mm), C.c_double(pitch_mm)))

    def set_incremental(self, on=True):
        self._check(lib().sp_hip_session_set_incremental(self._h, C.c_int(1 if on else 0)))

    def set_resident_kernels(self, on=True):
        self._check(lib().sp_hip_session_set_resident_kernels(self._h, C.c_int(1 if on else 0)))

    def set_fast_path(self, on=True):
        self._check(lib().sp_hip_session_set_fast_path(self._h, C.c_int(1 if on else 0)))

    def set_fused_passes(self, on=True):
        """FP32 256^3 steps without per-value output run the inverse y and
        inverse z passes as one pass per legal x column (default on)."""
        self._check(lib().sp_hip_session_set_fused_passes(self._h, C.c_int(1 if on else 0)))

    def fused_passes(self):
        """Whether the next step runs the fused inverse y + z pass."""
        a = C.c_int32()
        self._check(lib().sp_hip_session_get_fused_passes(self._h, C.byref(a)))
        return bool(a.value)

    def stamp(self, orientation, translation):
        t = (C.c_int64 * 3)(*[int(v) for v in translation])
        self._check(lib().sp_hip_session_stamp(self._h, C.c_uint32(orientation), t))

    def exclude(self, orientation, translation):
        t = (C.c_int64 * 3)(*[int(v) for v in translation])
        self._check(lib().sp_hip_session_exclude(self._h, C.c_uint32(orientation), t))

    def copies(self):
        """Every copy stamped since the environment was set (BlockedField
        footprints): [dict(copy_id, orientation, translation, removed)]."""
        n = C.c_uint64(0)
        rc = lib().sp_hip_session_copies(self._h, None, C.c_uint64(0), C.byref(n))
        if rc and rc != 37:
            self._check(rc)
        out = (PlacedCopy * max(n.value, 1))()
        self._check(lib().sp_hip_session_copies(self._h, out, C.c_uint64(n.value), C.byref(n)))
        return [dict(copy_id=int(out[i].copy_id), orientation=int(out[i].orientation),
                     translation=tuple(out[i].translation), removed=bool(out[i].removed)) for i in range(n.value)]

    def remove_copy(self, copy_id):
        """BlockedField::remove of one placed copy (conservative_fields.cpp:1848-1887)."""
        self._check(lib().sp_hip_session_remove_copy(self._h, C.c_uint64(copy_id)))

    def placed_counts(self):
        ex, ey, ez = self.env_shape
        a = np.zeros(ex * ey * ez, np.uint32)
        self._check(lib().sp_hip_session_placed_counts(self._h, _p(a, C.c_uint32), C.c_uint64(a.size)))
        return a.reshape(ez, ey, ex)

    def keep_raw(self, keep=True):
        self._check(lib().sp_hip_session_keep_raw(self._h, C.c_int(1 if keep else 0)))

    def set_active(self, orientations=(), region_mask=None):
        """Restrict the next steps to `orientations` (empty: all) and, with
        region_mask (uint8 [rz, ry, rx]), to kernel corners in nonzero regions.
        No arguments: the plain search again."""
        ol = np.ascontiguousarray(np.asarray(orientations, dtype=np.uint32).ravel())
        if region_mask is None:
            m, regions = None, None
        else:
            rm = np.ascontiguousarray(np.asarray(region_mask, dtype=np.uint8))
            if rm.ndim != 3:
                raise ValueError("region_mask must be [rz, ry, rx]")
            m = _p(rm, C.c_uint8)
            regions = (C.c_uint32 * 3)(rm.shape[2], rm.shape[1], rm.shape[0])
        self._check(lib().sp_hip_session_set_active(self._h, _p(ol, C.c_uint32) if ol.size else None,
                                                    C.c_uint32(ol.size), m, regions))

    def orientation_minima(self):
        """Per orientation: None, or the lowest feasible (uz, uy, ux) corner of the last step."""
        a = np.zeros(self.orientations, np.uint64)
        self._check(lib().sp_hip_session_orientation_minima(self._h, _p(a, C.c_uint64), C.c_uint32(a.size)))
        out = []
        for v in a.tolist():
            out.append(None if v == 0xFFFFFFFFFFFFFFFF else ((v >> 48) & 0xFFFF, (v >> 32) & 0xFFFF, (v >> 16) & 0xFFFF))
        return out

    def raw_values(self, o):
        ex, ey, ez = self.env_shape
        kx, ky, kz = self.kernel_shapes[o]
        n = max((ex - kx + 1) * (ey - ky + 1) * (ez - kz + 1), 1)
        a = np.zeros(n, np.float64)
        self._check(lib().sp_hip_session_raw_values(self._h, C.c_uint32(o), _p(a, C.c_double), C.c_uint64(a.size)))
        return a

    def keep_fields(self, keep=True):
        self._check(lib().sp_hip_session_keep_fields(self._h, C.c_int(1 if keep else 0)))

    @staticmethod
    def _result(r: StepResult):
        return dict(orientation=r.orientation, translation=tuple(r.translation), key=tuple(r.key),
                    max_residual=r.max_residual, numeric_flags=r.numeric_flags, run_flags=r.run_flags,
                    precision=r.precision,
                    error_bound=r.error_bound)

    def step(self, stamp=True):
        r = StepResult()
        self._check(lib().sp_hip_session_step(self._h, C.c_int(1 if stamp else 0), C.byref(r)))
        return self._result(r)

    def run(self, steps, stamp=True):
        self._check(lib().sp_hip_session_run(self._h, C.c_uint32(steps), C.c_int(1 if stamp else 0)))

    def last_result(self):
        r = StepResult()
        self._check(lib().sp_hip_session_last_result(self._h, C.byref(r)))
        return self._result(r)

    def overlap_field(self, o):
        ex, ey, ez = self.env_shape
        kx, ky, kz = self.kernel_shapes[o]
        n = max((ex - kx + 1) * (ey - ky + 1) * (ez - kz + 1), 1)
        a = np.zeros(n, np.int32)
        legal = (C.c_uint32 * 3)()
        self._check(lib().sp_hip_session_overlap_field(self._h, C.c_uint32(o), _p(a, C.c_int32), C.c_uint64(a.size),
                                                       legal))
        lx, ly, lz = tuple(legal)
        if lx * ly * lz == 0:
            return None
        return a[: lx * ly * lz].reshape(lz, ly, lx)

    # ---- ranked pages (spectral_pipeline.cpp:1343,1563,1740-1857) ----------
    def set_ranking(self, lattice_origin_z_mm, pitch_mm, container_min_z_mm, container_max_z_mm, object_max_z_mm):
        """Score constants of the ranked pages and each orientation's oriented-bounds max z (mm)."""
        d = RankingDesc(float(lattice_origin_z_mm), float(pitch_mm), float(container_min_z_mm),
                        float(container_max_z_mm))
        mz = np.ascontiguousarray(np.asarray(object_max_z_mm, dtype=np.float64).ravel())
        self._check(lib().sp_hip_session_set_ranking(self._h, C.byref(d), _p(mz, C.c_double), C.c_uint32(mz.size)))

    def rank(self):
        self._check(lib().sp_hip_session_rank(self._h))

    def ranked_pages(self, orientations=None, mode=PAGE_ORDINARY, cursors=None):
        """One page per orientation: {o: [candidate dict, ...]} in rank order.
        cursors: {o: candidate dict} -- each page starts after its cursor."""
        ol = list(range(self.orientations)) if orientations is None else [int(o) for o in orientations]
        n = len(ol)
        if n == 0:
            return {}
        oa = np.ascontiguousarray(np.asarray(ol, dtype=np.uint32))
        cur = (RankedCandidate * n)()
        has = np.zeros(n, np.uint8)
        for i, o in enumerate(ol):
            c = (cursors or {}).get(o)
            if c is not None:
                cur[i] = RankedCandidate((C.c_int64 * 3)(*[int(v) for v in c["translation"]]), float(c["score"]),
                                         float(c["binary_overlap"]), C.c_uint32(o), C.c_uint32(0))
                has[i] = 1
        pages = np.zeros(n * PAGE_SIZE, RANKED_DTYPE)
        counts = np.zeros(n, np.uint32)
        self._check(lib().sp_hip_session_ranked_pages(self._h, C.c_int(mode), _p(oa, C.c_uint32), C.c_uint32(n), cur,
                                                      _p(has, C.c_uint8), pages.ctypes.data_as(C.c_void_p),
                                                      _p(counts, C.c_uint32)))
        out = {}
        tr = pages["translation"].tolist()
        sc = pages["score"].tolist()
        ov = pages["binary_overlap"].tolist()
        dz = pages["direct_zero_overlap"].tolist()
        for i, o in enumerate(ol):
            b = i * PAGE_SIZE
            out[o] = [dict(translation=tuple(tr[b + j]), score=sc[b + j], binary_overlap=ov[b + j], orientation=o,
                           direct_zero_overlap=bool(dz[b + j])) for j in range(int(counts[i]))]
        return out

    def ranked_fill(self, max_copies=1 << 20, validate=None):
        """The ordinary greedy loop over ranked pages (native); returns
        [(orientation, translation, score)].  validate(o, t) -> bool is the
        caller's exact check (every head accepted when None)."""
        cap = int(min(max_copies, 1 << 20))
        out = (Placement * max(cap, 1))()
        n = C.c_uint32()
        cb = None
        if validate is not None:
            cb = VALIDATE_FN(lambda user, o, t: 1 if validate(int(o), (t[0], t[1], t[2])) else 0)
        self._check(lib().sp_hip_session_ranked_fill(self._h, C.c_uint32(cap), cb, None, out, C.c_uint32(cap),
                                                     C.byref(n)))
        return [(int(out[i].orientation), tuple(out[i].translation), float(out[i].score)) for i in range(n.value)]

    def proximity_field(self):
        """(values, squared) of the last rank, each [z, y, x]."""
        ex, ey, ez = self.env_shape
        v = np.zeros(ex * ey * ez, np.float64)
        q = np.zeros(ex * ey * ez, np.uint32)
        self._check(lib().sp_hip_session_proximity_field(self._h, _p(v, C.c_double), _p(q, C.c_uint32),
                                                         C.c_uint64(v.size)))
        return v.reshape(ez, ey, ex), q.reshape(ez, ey, ex)

    def proximity_values(self, o):
        """Proximity correlation of the last rank at orientation o's legal corners, [lz, ly, lx]."""
        ex, ey, ez = self.env_shape
        kx, ky, kz = self.kernel_shapes[o]
        lx, ly, lz = ex - kx + 1, ey - ky + 1, ez - kz + 1
        if min(lx, ly, lz) <= 0:
            return None
        a = np.zeros(lx * ly * lz, np.float64)
        self._check(lib().sp_hip_session_proximity_values(self._h, C.c_uint32(o), _p(a, C.c_double),
                                                          C.c_uint64(a.size)))
        return a.reshape(lz, ly, lx)

    def profile_step(self, stamp=True):
        """One step without the graph; per-kernel milliseconds from HIP events."""
        cap = 32
        ms = np.zeros(cap, np.float32)
        names = C.create_string_buffer(32 * cap)
        n = C.c_uint32(0)
        self._check(lib().sp_hip_session_profile_step(self._h, C.c_int(1 if stamp else 0), _p(ms, C.c_float), names,
                                                      C.c_uint32(cap), C.byref(n)))
        raw = names.raw
        return [(raw[32 * i:32 * i + 32].split(b"\0")[0].decode(), float(ms[i])) for i in range(n.value)]

    def product_grid(self):
        """(blocks, row tiles) of the last column-major product pass."""
        b, t = C.c_uint32(), C.c_uint64()
        self._check(lib().sp_hip_session_product_grid(self._h, C.byref(b), C.byref(t)))
        return int(b.value), int(t.value)

    def time_pass(self, name, launches=20):
        """ms per launch of one pass of the step, launched back to back in a graph."""
        ms = C.c_float()
        self._check(lib().sp_hip_session_time_pass(self._h, name.encode(), C.c_uint32(launches), C.byref(ms)))
        return float(ms.value)

    def stream(self):
        return lib().sp_hip_session_stream(self._h)

    def device_bytes(self):
        return lib().sp_hip_session_device_bytes(self._h)


class Proposal:
    """Learned orientation/region proposal (BF16 3D CNN on MFMA)."""

    def __init__(self, ctx: Context, max_batch: int = 32):
        self.ctx = ctx
        self.max_batch = int(max_batch)
        self._h = C.c_void_p()
        rc = lib().sp_hip_proposal_create(ctx.handle, C.c_uint32(self.max_batch), C.byref(self._h))
        if rc:
            ctx._fail(rc)
        self.parameter_count = int(lib().sp_hip_proposal_parameter_count())
        self.samples = 0

    def close(self):
        if self._h:
            lib().sp_hip_proposal_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc:
            self.ctx._fail(rc)

    def set_parameters(self, flat):
        a = np.ascontiguousarray(np.asarray(flat, dtype=np.float32).ravel())
        self._check(lib().sp_hip_proposal_set_parameters(self._h, _p(a, C.c_float), C.c_uint64(a.size)))

    def set_prior(self, prior, weight=1.0):
        """Per-orientation prior added to the logits (times weight) when a
        learned step ranks orientations; an empty prior clears it."""
        a = np.ascontiguousarray(np.asarray(prior, dtype=np.float32).ravel())
        self._check(lib().sp_hip_proposal_set_prior(self._h, _p(a, C.c_float) if a.size else None,
                                                   C.c_uint32(a.size), C.c_float(weight)))

    def parameters(self):
        a = np.zeros(self.parameter_count, np.float32)
        self._check(lib().sp_hip_proposal_get_parameters(self._h, _p(a, C.c_float), C.c_uint64(a.size)))
        return a

    def reserve(self, samples):
        self._check(lib().sp_hip_proposal_reserve(self._h, C.c_uint32(samples)))
        self.samples = int(samples)

    def load_samples(self, samples, first=0):
        a = np.ascontiguousarray(np.asarray(samples, dtype=np.uint8))
        n = a.shape[0]
        self._check(lib().sp_hip_proposal_load_samples(self._h, _p(a, C.c_uint8), C.c_uint32(first), C.c_uint32(n)))

    def record(self, session: "Session", slot):
        self._check(lib().sp_hip_proposal_record(self._h, session._h, C.c_uint32(slot)))

    def sample(self, slot):
        a = np.zeros(64 ** 3, np.uint8)
        self._check(lib().sp_hip_proposal_read_sample(self._h, C.c_uint32(slot), _p(a, C.c_uint8)))
        return a.reshape(64, 64, 64)

    def forward(self, slots, descriptors):
        sl = np.ascontiguousarray(np.asarray(slots, dtype=np.uint32).ravel())
        d = np.ascontiguousarray(np.asarray(descriptors, dtype=np.float32).reshape(sl.size, 32))
        logits = np.zeros((sl.size, 96), np.float32)
        heat = np.zeros((sl.size, 4096), np.float32)
        self._check(lib().sp_hip_proposal_forward(self._h, _p(sl, C.c_uint32), C.c_uint32(sl.size), _p(d, C.c_float),
                                                  _p(logits, C.c_float), _p(heat, C.c_float)))
        return logits, heat

    def train(self, slots, descriptors, target, mask, region, heat_weight=1.0, lr=1e-3, update=True,
              gradients=False):
        sl = np.ascontiguousarray(np.asarray(slots, dtype=np.uint32).ravel())
        n = sl.size
        d = np.ascontiguousarray(np.asarray(descriptors, dtype=np.float32).reshape(n, 32))
        t = np.ascontiguousarray(np.asarray(target, dtype=np.float32).reshape(n, 96))
        m = np.ascontiguousarray(np.asarray(mask, dtype=np.uint8).reshape(n, 96))
        r = np.ascontiguousarray(np.asarray(region, dtype=np.int32).reshape(n))
        losses = np.zeros(2, np.float32)
        g = np.zeros(self.parameter_count, np.float32) if gradients else None
        self._check(lib().sp_hip_proposal_train(self._h, _p(sl, C.c_uint32), C.c_uint32(n), _p(d, C.c_float),
                                                _p(t, C.c_float), _p(m, C.c_uint8), _p(r, C.c_int32),
                                                C.c_float(heat_weight), C.c_float(lr), C.c_int(1 if update else 0),
                                                _p(losses, C.c_float), _p(g, C.c_float) if gradients else None))
        return (float(losses[0]), float(losses[1])), g

    def lazy_step(self, session: "Session", descriptor, top_orientations=16, stamp=True):
        """One lazy learned step (sp_hip_proposal_lazy_step): exact minima kept
        per orientation, the top-ranked blocked orientations searched."""
        d = np.ascontiguousarray(np.asarray(descriptor, dtype=np.float32).ravel())
        if d.size != 32:
            raise ValueError("descriptor must hold 32 floats")
        r = StepResult()
        full = C.c_int(0)
        searched = C.c_uint32(0)
        self._check(lib().sp_hip_proposal_lazy_step(self._h, session._h, _p(d, C.c_float),
                                                    C.c_uint32(top_orientations), C.c_int(1 if stamp else 0),
                                                    C.byref(r), C.byref(full), C.byref(searched)))
        out = Session._result(r)
        out["used_full"] = bool(full.value)
        out["searched"] = int(searched.value)
        return out

    def lazy_reset(self):
        """Forget the kept minima (the next lazy step runs the full search)."""
        self._check(lib().sp_hip_proposal_lazy_reset(self._h))

    def step(self, session: "Session", descriptor, top_orientations=8, top_regions=512, stamp=True, fallback=True):
        d = np.ascontiguousarray(np.asarray(descriptor, dtype=np.float32).ravel())
        if d.size != 32:
            raise ValueError("descriptor must hold 32 floats")
        r = StepResult()
        used = C.c_int(0)
        self._check(lib().sp_hip_proposal_step(self._h, session._h, _p(d, C.c_float), C.c_uint32(top_orientations),
                                               C.c_uint32(top_regions), C.c_int(1 if stamp else 0),
                                               C.c_int(1 if fallback else 0), C.byref(r), C.byref(used)))
        out = Session._result(r)
        out["fallback"] = bool(used.value)
        return out

    def mfma_probe(self, a_bits, b_bits):
        a = np.ascontiguousarray(np.asarray(a_bits, dtype=np.uint16).reshape(32, 16))
        b = np.ascontiguousarray(np.asarray(b_bits, dtype=np.uint16).reshape(16, 32))
        c = np.zeros((32, 32), np.float32)
        self._check(lib().sp_hip_proposal_mfma_probe(self._h, _p(a, C.c_uint16), _p(b, C.c_uint16), _p(c, C.c_float)))
        return c
