"""ctypes bindings of ``libdisptools_hip.so`` (declared in ``include/disptools_hip.h``).

The library must have been built in-tree (``python -m disptools_amd.build``);
importing this module without it raises ``ImportError`` -- there is no
fallback implementation.
"""

import ctypes
import os
import re

import numpy as np

from . import _hip_runtime

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libdisptools_hip.so")
HEADER = os.path.join(os.path.dirname(PKG), "include", "disptools_hip.h")

if not os.path.exists(LIB_PATH):
    raise ImportError("%s is not built; run `python -m disptools_amd.build`" % LIB_PATH)

# one HIP runtime per process, shared with torch when it is installed
_hip_runtime.preload()

lib = ctypes.CDLL(LIB_PATH)

GRADIENT = 0
GREEDY = 1
MATCHING = 2
ALGORITHMS = {"gradient": GRADIENT, "greedy": GREEDY}
# the batch engine also takes matching solves (epsilon, delta, zeta, it_max; no mask)
ENGINE_ALGORITHMS = {"gradient": GRADIENT, "greedy": GREEDY, "matching": MATCHING}

REASONS = {
    0: "running",
    1: "error exploded",
    2: "eta below iota",
    3: "relative improvement below theta",
    4: "voxel error exploded",
    5: "reached desired tolerance",
    6: "maximum number of iterations",
}

# Defaults of disptools.displacements.displacement (displacements.py:246-260).
DEFAULT_PARAMS = dict(
    epsilon=9.99e-4, tolerance=0.2, eta=0.1, eta_max=0.4, alpha=1.2, beta=0.5,
    gamma=0.1, delta=1e-3, zeta=10.0, theta=1e-6, iota=1e-9, strict=False, it_max=50000,
)


class ErrorState(ctypes.Structure):
    _fields_ = [
        ("error", ctypes.c_bool),
        ("line", ctypes.c_int),
        ("file", ctypes.c_char * 512),
        ("function", ctypes.c_char * 512),
        ("message", ctypes.c_char * 1024),
        ("trace", ctypes.c_char * 4096),
    ]


class Params(ctypes.Structure):
    _fields_ = [(k, ctypes.c_float) for k in (
        "epsilon", "tolerance", "eta", "eta_max", "alpha", "beta", "gamma", "delta",
        "zeta", "theta", "iota")] + [("strict", ctypes.c_int), ("it_max", ctypes.c_longlong)]


class SolveInfo(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_longlong),
        ("evaluations", ctypes.c_longlong),
        ("rejections", ctypes.c_longlong),
        ("launches", ctypes.c_longlong),
        ("reason", ctypes.c_int),
        ("final_error", ctypes.c_double),
        ("final_max_error", ctypes.c_double),
        ("seconds", ctypes.c_double),
        ("trace_len", ctypes.c_longlong),
        ("field_io", ctypes.c_longlong),
        ("kernel_seconds", ctypes.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_sz = ctypes.c_size_t
_f = ctypes.c_float
_vp = ctypes.c_void_p

lib.disptools_get_error.restype = ctypes.POINTER(ErrorState)
lib.disptools_has_error.restype = ctypes.c_bool
lib.disptools_clear_error.restype = None
lib.get_float_type_size.restype = ctypes.c_int
lib.set_device.argtypes = [ctypes.c_int]
lib.set_device.restype = ctypes.c_bool
_ref_solver = [_sz, _sz, _sz, _f, _f, _f, _vp, _vp] + [_f] * 11 + [ctypes.c_bool, _sz, _vp]
for _name in ("generate_displacement_gradient_cuda", "generate_displacement_greedy_cuda"):
    getattr(lib, _name).argtypes = _ref_solver
    getattr(lib, _name).restype = None
lib.jacobian_dynamic.argtypes = [_sz, _sz, _sz, _f, _f, _f, _vp, _vp]
lib.jacobian_dynamic.restype = None
lib.regularise.argtypes = [_sz, _sz, _sz, _vp, _f]
lib.regularise.restype = None
lib.dt_jacobian_device.argtypes = [_sz, _sz, _sz, _f, _f, _f, _vp, _vp, _vp]
lib.dt_jacobian_device.restype = ctypes.c_int
lib.dt_solve_device.argtypes = [ctypes.c_int, _sz, _sz, _sz, _f, _f, _f, _vp, _vp,
                                ctypes.POINTER(Params), _vp, ctypes.POINTER(SolveInfo), _vp,
                                ctypes.c_longlong, _vp]
lib.dt_solve_device.restype = ctypes.c_int
lib.dt_solve_device_noalloc.argtypes = [ctypes.c_int, _sz, _sz, _sz, _f, _f, _f, _vp, _vp,
                                        ctypes.POINTER(Params), _vp, ctypes.POINTER(SolveInfo), _vp]
lib.dt_solve_device_noalloc.restype = ctypes.c_int
lib.dt_solve_host.argtypes = [ctypes.c_int, _sz, _sz, _sz, _f, _f, _f, _vp, _vp,
                              ctypes.POINTER(Params), _vp, ctypes.POINTER(SolveInfo), _vp,
                              ctypes.c_longlong]
lib.dt_solve_host.restype = ctypes.c_int
lib.volume_matching_3d.argtypes = _ref_solver
lib.volume_matching_3d.restype = None


class MatchingInfo(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_longlong),
        ("reason", ctypes.c_int),
        ("cost_passes", ctypes.c_longlong),
        ("gradient_passes", ctypes.c_longlong),
        ("launches", ctypes.c_longlong),
        ("final_cost", ctypes.c_double),
        ("final_max_error", ctypes.c_double),
        ("trace_len", ctypes.c_longlong),
        ("seconds", ctypes.c_double),
        ("kernel_seconds", ctypes.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


lib.dt_matching_device.argtypes = [_sz, _sz, _sz, _f, _f, _f, _vp, _f, _f, _f, ctypes.c_longlong, _vp,
                                   ctypes.POINTER(MatchingInfo), _vp, ctypes.c_longlong, _vp]
lib.dt_matching_device.restype = ctypes.c_int
lib.dt_matching_host.argtypes = [_sz, _sz, _sz, _f, _f, _f, _vp, _f, _f, _f, ctypes.c_longlong, _vp,
                                 ctypes.POINTER(MatchingInfo), _vp, ctypes.c_longlong]
lib.dt_matching_host.restype = ctypes.c_int


class EngineConfig(ctypes.Structure):
    _fields_ = [
        ("hbm_budget", ctypes.c_size_t),
        ("max_slots", ctypes.c_int),
        ("chunk", ctypes.c_int),
        ("use_graphs", ctypes.c_int),
        ("variant", ctypes.c_int),
        ("tz", ctypes.c_int),
    ]


class EngineResult(ctypes.Structure):
    _fields_ = [("ticket", ctypes.c_longlong), ("status", ctypes.c_int), ("info", SolveInfo)]


class EngineStats(ctypes.Structure):
    _fields_ = [
        ("chunks", ctypes.c_longlong),
        ("passes", ctypes.c_longlong),
        ("graph_builds", ctypes.c_longlong),
        ("admitted", ctypes.c_longlong),
        ("finished", ctypes.c_longlong),
        ("active", ctypes.c_int),
        ("retiring", ctypes.c_int),
        ("pool_bytes", ctypes.c_size_t),
        ("pool_used", ctypes.c_size_t),
        ("pool_peak", ctypes.c_size_t),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


ENGINE_FULL = -2
TUNE_MAX = 32


class TuneResult(ctypes.Structure):
    _fields_ = [("best_variant", ctypes.c_int), ("count", ctypes.c_int),
                ("variant", ctypes.c_int * TUNE_MAX), ("us_per_pass", ctypes.c_double * TUNE_MAX)]


lib.dt_scale_volume_device.argtypes = [_sz, _sz, _sz, _sz, _f, _f, _f, _vp, _sz, _sz, _sz, _f, ctypes.c_int, _vp,
                                       _vp]
lib.dt_scale_volume_device.restype = ctypes.c_int
lib.dt_displacement_levels.argtypes = [ctypes.c_int, _sz, _sz, _sz, _f, _f, _f, _vp, _vp, _vp, _sz, _sz, _sz,
                                       ctypes.c_int, _vp, _vp, _vp, _vp]
lib.dt_displacement_levels.restype = ctypes.c_int

lib.dt_tune_gradient.argtypes = [_sz, _sz, _sz, _f, _f, _f, ctypes.c_int, ctypes.POINTER(TuneResult)]
lib.dt_tune_gradient.restype = ctypes.c_int

lib.dt_engine_create.argtypes = [ctypes.POINTER(EngineConfig), _vp]
lib.dt_engine_create.restype = _vp
lib.dt_engine_destroy.argtypes = [_vp]
lib.dt_engine_destroy.restype = None
lib.dt_engine_footprint.argtypes = [_vp, ctypes.c_int, _sz, _sz, _sz, ctypes.c_int]
lib.dt_engine_footprint.restype = _sz
lib.dt_engine_submit.argtypes = [_vp, ctypes.c_int, _sz, _sz, _sz, _f, _f, _f, _vp, _vp,
                                 ctypes.POINTER(Params), _vp, ctypes.c_int]
lib.dt_engine_submit.restype = ctypes.c_longlong
lib.dt_engine_progress.argtypes = [_vp, ctypes.c_int]
lib.dt_engine_progress.restype = ctypes.c_int
lib.dt_engine_poll.argtypes = [_vp, ctypes.POINTER(EngineResult), ctypes.c_int]
lib.dt_engine_poll.restype = ctypes.c_int
lib.dt_engine_drain.argtypes = [_vp]
lib.dt_engine_drain.restype = ctypes.c_int
lib.dt_engine_get_stats.argtypes = [_vp, ctypes.POINTER(EngineStats)]
lib.dt_engine_get_stats.restype = ctypes.c_int


class HipError(RuntimeError):
    pass


def last_error():
    e = lib.disptools_get_error().contents
    return "%s (%s:%d): %s" % (e.function.decode(), e.file.decode(), e.line, e.message.decode())


def check():
    if lib.disptools_has_error():
        raise HipError(last_error())


def declared_symbols():
    """Function names declared in include/disptools_hip.h."""
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"^\s*#.*$", "", text, flags=re.M)
    names = re.findall(r"\b([A-Za-z_]\w*)\s*\(", text)
    keywords = {"if", "while", "for", "return", "sizeof", "defined"}
    return sorted({n for n in names if n not in keywords})


def make_params(**kw):
    p = dict(DEFAULT_PARAMS)
    unknown = set(kw) - set(p)
    if unknown:
        raise TypeError("unknown solver parameters: %s" % sorted(unknown))
    p.update(kw)
    out = Params()
    for k in ("epsilon", "tolerance", "eta", "eta_max", "alpha", "beta", "gamma", "delta",
              "zeta", "theta", "iota"):
        setattr(out, k, float(p[k]))
    out.strict = 1 if p["strict"] else 0
    out.it_max = int(p["it_max"])
    return out


def set_device(i):
    return bool(lib.set_device(int(i)))


def _host_ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def jacobian_host(field, spacing=(1.0, 1.0, 1.0)):
    """J = det(I + grad u) of a host field [3, nz, ny, nx] (float32)."""
    f = np.ascontiguousarray(field, np.float32)
    nz, ny, nx = f.shape[1:]
    J = np.empty((nz, ny, nx), np.float32)
    lib.jacobian_dynamic(nx, ny, nz, *map(float, spacing), _host_ptr(f), _host_ptr(J))
    check()
    return J


def solve_host(algorithm, target, mask, spacing=(1.0, 1.0, 1.0), field=None, trace_cap=0, **params):
    """Solve with host numpy buffers; returns (field, info dict, trace array)."""
    t = np.ascontiguousarray(target, np.float32)
    m = np.ascontiguousarray(mask, np.bool_)
    nz, ny, nx = t.shape
    f = (np.zeros((3, nz, ny, nx), np.float32) if field is None
         else np.array(field, np.float32, order="C", copy=True))
    info = SolveInfo()
    trace = np.zeros((max(trace_cap, 1), 4), np.float64)
    p = make_params(**params)
    rc = lib.dt_solve_host(ALGORITHMS[algorithm], nx, ny, nz, *map(float, spacing),
                           _host_ptr(t), _host_ptr(m), ctypes.byref(p), _host_ptr(f),
                           ctypes.byref(info), _host_ptr(trace) if trace_cap else None, trace_cap)
    if rc != 0:
        raise HipError(last_error())
    d = info.as_dict()
    return f, d, trace[:min(info.trace_len, trace_cap)].copy()


def solve_device(algorithm, nx, ny, nz, spacing, d_target, d_mask, d_field, stream=0,
                 trace_cap=0, **params):
    """Solve on device pointers (ints); returns (info dict, trace array)."""
    info = SolveInfo()
    trace = np.zeros((max(trace_cap, 1), 4), np.float64)
    p = make_params(**params)
    rc = lib.dt_solve_device(ALGORITHMS[algorithm], nx, ny, nz, *map(float, spacing),
                             d_target, d_mask, ctypes.byref(p), d_field, ctypes.byref(info),
                             _host_ptr(trace) if trace_cap else None, trace_cap, stream or None)
    if rc != 0:
        raise HipError(last_error())
    return info.as_dict(), trace[:min(info.trace_len, trace_cap)].copy()


def solve_device_noalloc(algorithm, nx, ny, nz, spacing, d_target, d_mask, d_field, stream=0, **params):
    """dt_solve_device_noalloc on device pointers (ints); returns the info dict."""
    info = SolveInfo()
    p = make_params(**params)
    rc = lib.dt_solve_device_noalloc(ALGORITHMS[algorithm], nx, ny, nz, *map(float, spacing), d_target, d_mask,
                                     ctypes.byref(p), d_field, ctypes.byref(info), stream or None)
    if rc != 0:
        raise HipError(last_error())
    return info.as_dict()


def matching_host(target, spacing=(1.0, 1.0, 1.0), field=None, it_max=50000, trace_cap=0, epsilon=9.99e-4,
                  delta=1e-3, zeta=10.0):
    """The matching algorithm with host numpy buffers; returns (field, info,
    trace rows (it, cost, max voxel error, step))."""
    t = np.ascontiguousarray(target, np.float32)
    nz, ny, nx = t.shape
    f = (np.zeros((3, nz, ny, nx), np.float32) if field is None
         else np.array(field, np.float32, order="C", copy=True))
    info = MatchingInfo()
    trace = np.zeros((max(trace_cap, 1), 4), np.float64)
    rc = lib.dt_matching_host(nx, ny, nz, *map(float, spacing), _host_ptr(t), float(epsilon), float(delta),
                              float(zeta), int(it_max), _host_ptr(f), ctypes.byref(info),
                              _host_ptr(trace) if trace_cap else None, trace_cap)
    if rc != 0:
        raise HipError(last_error())
    return f, info.as_dict(), trace[:min(info.trace_len, trace_cap)].copy()


def matching_device(nx, ny, nz, spacing, d_target, d_field, stream=0, it_max=50000, trace_cap=0,
                    epsilon=9.99e-4, delta=1e-3, zeta=10.0, **unused):
    """The matching algorithm on device pointers (ints); returns (info, trace).
    Solver parameters other than epsilon, delta, zeta and it_max are accepted
    and unused, as in the reference."""
    info = MatchingInfo()
    trace = np.zeros((max(trace_cap, 1), 4), np.float64)
    rc = lib.dt_matching_device(nx, ny, nz, *map(float, spacing), d_target, float(epsilon), float(delta),
                                float(zeta), int(it_max), d_field, ctypes.byref(info),
                                _host_ptr(trace) if trace_cap else None, trace_cap, stream or None)
    if rc != 0:
        raise HipError(last_error())
    return info.as_dict(), trace[:min(info.trace_len, trace_cap)].copy()


def displacement_levels(algorithm, J, mask, guess, spacing, levels, it_max, **params):
    """Multi-resolution solve on host arrays (dt_displacement_levels): J and
    mask [nz, ny, nx] (mask as float), guess [gz, gy, gx, 3] or None,
    it_max a list of `levels` limits from the coarsest level.  Returns the
    field [nz, ny, nx, 3] (float32) and the finest level's record."""
    alg = {"gradient": GRADIENT, "greedy": GREEDY, "matching": MATCHING}[algorithm]
    J = np.ascontiguousarray(J, dtype=np.float32)
    mask = np.ascontiguousarray(mask, dtype=np.float32)
    if mask.shape != J.shape:
        raise ValueError("the shapes of the target and the mask must agree")
    nz, ny, nx = J.shape
    gz = gy = gx = 0
    if guess is not None:
        guess = np.ascontiguousarray(guess, dtype=np.float32)
        if guess.ndim != 4 or guess.shape[3] != 3:
            raise ValueError("the initial guess must be [z, y, x, 3]")
        gz, gy, gx = guess.shape[:3]
    if len(it_max) != levels:
        raise ValueError("len(it_max) should equal the value of `levels`")
    limits = (ctypes.c_longlong * levels)(*[int(v) for v in it_max])
    p = make_params(**params)
    info = SolveInfo()
    field = np.empty((nz, ny, nx, 3), np.float32)
    rc = lib.dt_displacement_levels(alg, nx, ny, nz, *map(float, spacing), _host_ptr(J), _host_ptr(mask),
                                    _host_ptr(guess) if guess is not None else None, gx, gy, gz, int(levels),
                                    ctypes.cast(limits, ctypes.c_void_p), ctypes.byref(p), _host_ptr(field),
                                    ctypes.byref(info))
    if rc != 0:
        raise HipError(last_error())
    return field, info.as_dict()


def scale_volume_device(nx, ny, nz, ncomp, spacing, d_in, size_out, sigma, nearest, d_out, stream=0):
    """Gaussian smoothing (sigma in physical units) + resampling of a device
    volume [ncomp][nz][ny][nx] onto a grid of size_out = (mx, my, mz)."""
    mx, my, mz = size_out
    rc = lib.dt_scale_volume_device(nx, ny, nz, ncomp, *map(float, spacing), d_in, mx, my, mz, float(sigma),
                                    1 if nearest else 0, d_out, stream or None)
    if rc != 0:
        raise HipError(last_error())


def tune_gradient(nx, ny, nz, spacing=(1.0, 1.0, 1.0), passes=0):
    """Times the fused gradient-pass variants for this shape; returns
    (best variant, {variant: microseconds per pass})."""
    r = TuneResult()
    if lib.dt_tune_gradient(nx, ny, nz, *map(float, spacing), int(passes), ctypes.byref(r)) != 0:
        raise HipError(last_error())
    return int(r.best_variant), {int(r.variant[i]): float(r.us_per_pass[i]) for i in range(r.count)}


def jacobian_device(nx, ny, nz, spacing, d_field, d_J, stream=0):
    rc = lib.dt_jacobian_device(nx, ny, nz, *map(float, spacing), d_field, d_J, stream or None)
    if rc != 0:
        raise HipError(last_error())


class Engine:
    """Continuous-batching solver (``dt_engine_*``).

    ``submit`` admits a solve if a slot and pool memory are free (returns a
    ticket) or returns None (full: call ``progress`` and retry); ``progress``
    advances the passes; ``poll`` returns finished ``(ticket, status, info)``.
    Buffers are device pointers (host_io=False) or host pointers (True).
    """

    def __init__(self, hbm_budget=0, max_slots=0, chunk=0, use_graphs=False, variant=-1, tz=0,
                 stream=0):
        cfg = EngineConfig(int(hbm_budget), int(max_slots), int(chunk), 1 if use_graphs else 0,
                           int(variant), int(tz))
        h = lib.dt_engine_create(ctypes.byref(cfg), stream or None)
        if not h:
            raise HipError(last_error())
        self._h = h
        self._buf = (EngineResult * 64)()

    def close(self):
        if self._h:
            lib.dt_engine_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def footprint(self, algorithm, nx, ny, nz, host_io=False):
        return int(lib.dt_engine_footprint(self._h, ENGINE_ALGORITHMS[algorithm], nx, ny, nz, 1 if host_io else 0))

    def submit(self, algorithm, nx, ny, nz, spacing, target, mask, field, host_io=False, **params):
        p = make_params(**params)
        t = lib.dt_engine_submit(self._h, ENGINE_ALGORITHMS[algorithm], nx, ny, nz, *map(float, spacing),
                                 target, mask, ctypes.byref(p), field, 1 if host_io else 0)
        if t == ENGINE_FULL:
            return None
        if t < 0:
            raise HipError(last_error())
        return int(t)

    def progress(self, wait=True):
        r = lib.dt_engine_progress(self._h, 1 if wait else 0)
        if r < 0:
            raise HipError(last_error())
        return r

    def poll(self):
        out = []
        while True:
            k = lib.dt_engine_poll(self._h, self._buf, len(self._buf))
            for i in range(k):
                r = self._buf[i]
                out.append((int(r.ticket), int(r.status), r.info.as_dict()))
            if k < len(self._buf):
                return out

    def drain(self):
        if lib.dt_engine_drain(self._h) != 0:
            raise HipError(last_error())

    def stats(self):
        s = EngineStats()
        lib.dt_engine_get_stats(self._h, ctypes.byref(s))
        return s.as_dict()
