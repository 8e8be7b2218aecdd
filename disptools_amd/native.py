"""ctypes bindings of ``libdisptools_hip.so`` (declared in ``include/disptools_hip.h``).

The library must have been built in-tree (``python -m disptools_amd.build``);
importing this module without it raises ``ImportError`` -- there is no
fallback implementation.
"""

import ctypes
import os
import re

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libdisptools_hip.so")
HEADER = os.path.join(os.path.dirname(PKG), "include", "disptools_hip.h")

if not os.path.exists(LIB_PATH):
    raise ImportError("%s is not built; run `python -m disptools_amd.build`" % LIB_PATH)

lib = ctypes.CDLL(LIB_PATH)

GRADIENT = 0
GREEDY = 1
ALGORITHMS = {"gradient": GRADIENT, "greedy": GREEDY}

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
lib.dt_solve_host.argtypes = [ctypes.c_int, _sz, _sz, _sz, _f, _f, _f, _vp, _vp,
                              ctypes.POINTER(Params), _vp, ctypes.POINTER(SolveInfo), _vp,
                              ctypes.c_longlong]
lib.dt_solve_host.restype = ctypes.c_int


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


def jacobian_device(nx, ny, nz, spacing, d_field, d_J, stream=0):
    rc = lib.dt_jacobian_device(nx, ny, nz, *map(float, spacing), d_field, d_J, stream or None)
    if rc != 0:
        raise HipError(last_error())
