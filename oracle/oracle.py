"""ctypes bindings of the CPU checkers.

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the CPU-baseline leg of ``bench.py``; the product package never imports it.

Two libraries are bound:

* ``Oracle`` -- this repository's C99 restatement (``oracle/dt_oracle.c``),
  built by ``oracle/Makefile`` into ``oracle/build/liboracle_f{32,64}.so``;
* ``Reference`` -- the reference C library compiled from its own sources
  (``src/{disptools,jacobian,displacement_field_gradient,displacement_field_greedy,
  VolumeMatching3D}.c``)
  into ``oracle/_ref/libdtref_f{32,64}.so``.  It prints one log line per
  iteration to the C ``stdout`` (``displacement_field_gradient.c:272-277``,
  ``displacement_field_greedy.c:184-189``); ``Reference.solve`` captures and
  parses them into the same trace format as the oracle.
"""

import ctypes
import os
import re
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

# Termination reasons (same numbering as the HIP library).
RUNNING = 0
ERROR_EXPLODED = 1
ETA_SMALL = 2
NO_IMPROVEMENT = 3
VOXEL_EXPLODED = 4
TOLERANCE = 5
MAX_ITERATIONS = 6

REASON_NAMES = {
    RUNNING: "running",
    ERROR_EXPLODED: "error exploded",
    ETA_SMALL: "eta below iota",
    NO_IMPROVEMENT: "relative improvement below theta",
    VOXEL_EXPLODED: "voxel error exploded",
    TOLERANCE: "reached desired tolerance",
    MAX_ITERATIONS: "maximum number of iterations",
}

# Defaults of disptools.displacements.displacement (displacements.py:246-260).
DEFAULT_PARAMS = dict(
    epsilon=9.99e-4, tolerance=0.2, eta=0.1, eta_max=0.4, alpha=1.2, beta=0.5,
    gamma=0.1, delta=1e-3, zeta=10.0, theta=1e-6, iota=1e-9, strict=False,
)
PARAM_ORDER = ("epsilon", "tolerance", "eta", "eta_max", "alpha", "beta", "gamma",
               "delta", "zeta", "theta", "iota")


class _Result(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_long),
        ("reason", ctypes.c_int),
        ("evaluations", ctypes.c_long),
        ("rejections", ctypes.c_long),
        ("final_error", ctypes.c_double),
        ("final_max_error", ctypes.c_double),
        ("trace", ctypes.POINTER(ctypes.c_double)),
        ("trace_cap", ctypes.c_long),
        ("trace_len", ctypes.c_long),
    ]


class _MatchResult(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_long),
        ("reason", ctypes.c_int),
        ("cost_evaluations", ctypes.c_long),
        ("gradient_evaluations", ctypes.c_long),
        ("final_cost", ctypes.c_double),
        ("final_max_error", ctypes.c_double),
        ("trace", ctypes.POINTER(ctypes.c_double)),
        ("trace_cap", ctypes.c_long),
        ("trace_len", ctypes.c_long),
    ]


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def oracle_lib_path(bits=32):
    # DT_ORACLE_BUILD: another build of the same sources (the sanitizer test's)
    return os.path.join(os.environ.get("DT_ORACLE_BUILD", os.path.join(HERE, "build")), "liboracle_f%d.so" % bits)


def reference_lib_path(bits=32):
    return os.path.join(HERE, "_ref", "libdtref_f%d.so" % bits)


def reference_available(bits=32):
    return os.path.exists(reference_lib_path(bits))


class _Base:
    def __init__(self, path, bits):
        if not os.path.exists(path):
            raise FileNotFoundError("%s is not built (run `make -C oracle`)" % path)
        self.bits = bits
        self.dtype = np.float32 if bits == 32 else np.float64
        self.ctype = ctypes.c_float if bits == 32 else ctypes.c_double
        self.lib = ctypes.CDLL(path, mode=ctypes.RTLD_LOCAL)

    def _field(self, field):
        f = np.ascontiguousarray(field, dtype=self.dtype)
        if f.ndim != 4 or f.shape[0] != 3:
            raise ValueError("field must be [3, nz, ny, nx]")
        return f

    def _solver_args(self, params):
        p = dict(DEFAULT_PARAMS)
        p.update(params)
        return [self.ctype(p[k]) for k in PARAM_ORDER], bool(p["strict"])


class Oracle(_Base):
    """The C restatement; FP32 (``bits=32``) or FP64 (``bits=64``)."""

    def __init__(self, bits=32):
        super().__init__(oracle_lib_path(bits), bits)
        L, R = self.lib, self.ctype
        sz = ctypes.c_size_t
        vp = ctypes.c_void_p
        L.oracle_jacobian.argtypes = [sz, sz, sz, R, R, R, vp, vp]
        L.oracle_jacobian.restype = None
        L.oracle_compute_error.argtypes = [sz, sz, sz, vp, vp, vp, R, vp, ctypes.POINTER(R)]
        L.oracle_compute_error.restype = R
        L.oracle_gradient.argtypes = [sz, sz, sz, R, R, R, vp, vp, vp, R, R, vp]
        L.oracle_gradient.restype = R
        L.oracle_stencil.argtypes = [ctypes.c_int, vp, sz, sz, sz, sz, sz, sz,
                                     ctypes.c_int, ctypes.c_int, R]
        L.oracle_stencil.restype = ctypes.c_double
        solver = [sz, sz, sz, R, R, R, vp, vp] + [R] * 11 + [ctypes.c_bool, sz, vp,
                                                             ctypes.POINTER(_Result)]
        for name in ("oracle_solve_gradient", "oracle_solve_greedy"):
            getattr(L, name).argtypes = solver
            getattr(L, name).restype = ctypes.c_int
        L.oracle_set_threads.argtypes = [ctypes.c_int]
        L.oracle_set_threads.restype = None
        L.oracle_match_cost.argtypes = [sz, sz, sz, vp, vp, vp, vp, R, R, ctypes.POINTER(ctypes.c_double)]
        L.oracle_match_cost.restype = ctypes.c_double
        L.oracle_match_grad.argtypes = [sz, sz, sz, vp, vp, vp, vp, R, R, vp, vp, vp]
        L.oracle_match_grad.restype = None
        L.oracle_solve_matching.argtypes = [sz, sz, sz, R, R, R, vp, R, R, R, sz, vp,
                                            ctypes.POINTER(_MatchResult)]
        L.oracle_solve_matching.restype = ctypes.c_int

    def set_threads(self, n):
        self.lib.oracle_set_threads(int(n))

    def jacobian(self, field, spacing=(1.0, 1.0, 1.0)):
        f = self._field(field)
        nz, ny, nx = f.shape[1:]
        J = np.zeros((nz, ny, nx), self.dtype)
        self.lib.oracle_jacobian(nx, ny, nz, *spacing, _ptr(f), _ptr(J))
        return J

    def compute_error(self, target, jf, mask, tolerance):
        t = np.ascontiguousarray(target, self.dtype)
        j = np.ascontiguousarray(jf, self.dtype)
        m = np.ascontiguousarray(mask, np.bool_)
        nz, ny, nx = t.shape
        e = np.zeros_like(t)
        mx = self.ctype(0)
        E = self.lib.oracle_compute_error(nx, ny, nz, _ptr(t), _ptr(j), _ptr(m),
                                          tolerance, _ptr(e), ctypes.byref(mx))
        return float(E), float(mx.value), e

    def gradient(self, field, jf, e, delta, zeta, spacing=(1.0, 1.0, 1.0)):
        f = self._field(field)
        nz, ny, nx = f.shape[1:]
        j = np.ascontiguousarray(jf, self.dtype)
        ee = np.ascontiguousarray(e, self.dtype)
        g = np.zeros_like(f)
        G = self.lib.oracle_gradient(nx, ny, nz, *spacing, _ptr(f), _ptr(j), _ptr(ee),
                                     delta, zeta, _ptr(g))
        return g, float(G)

    def stencil(self, kind, field, point, comp, axis, inv):
        f = self._field(field)
        nz, ny, nx = f.shape[1:]
        x, y, z = point
        return self.lib.oracle_stencil(kind, _ptr(f), nx, ny, nz, x, y, z, comp, axis, inv)

    def solve(self, algorithm, target, mask, spacing=(1.0, 1.0, 1.0), field=None,
              it_max=50000, trace_cap=200000, **params):
        t = np.ascontiguousarray(target, self.dtype)
        m = np.ascontiguousarray(mask, np.bool_)
        nz, ny, nx = t.shape
        f = (np.zeros((3, nz, ny, nx), self.dtype) if field is None
             else self._field(field).copy())
        trace = np.zeros((max(trace_cap, 1), 4), np.float64)
        res = _Result()
        res.trace = trace.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        res.trace_cap = trace_cap
        args, strict = self._solver_args(params)
        fn = {"gradient": self.lib.oracle_solve_gradient,
              "greedy": self.lib.oracle_solve_greedy}[algorithm]
        rc = fn(nx, ny, nz, *spacing, _ptr(t), _ptr(m), *args, strict, it_max, _ptr(f),
                ctypes.byref(res))
        if rc != 0:
            raise MemoryError("oracle solver could not allocate scratch memory")
        info = dict(iterations=res.iterations, reason=res.reason,
                    evaluations=res.evaluations, rejections=res.rejections,
                    final_error=res.final_error, final_max_error=res.final_max_error,
                    trace=trace[:min(res.trace_len, trace_cap)].copy())
        return f, info


    # -- matching (src/VolumeMatching3D.c) -------------------------------------

    def match_cost(self, F, G, E, V, delta=1e-3, zeta=10.0):
        """(cost, max voxel error) of node positions F, G, E [nz, ny, nx] and
        cell targets V [nz, ny, nx] (the target at each cell's lowest node)."""
        a = [np.ascontiguousarray(x, self.dtype) for x in (F, G, E, V)]
        nz, ny, nx = a[0].shape
        mx = ctypes.c_double(0)
        c = self.lib.oracle_match_cost(nx, ny, nz, *map(_ptr, a), delta, zeta, ctypes.byref(mx))
        return float(c), float(mx.value)

    def match_grad(self, F, G, E, V, delta=1e-3, zeta=10.0):
        """Gradient (dF, dG, dE) of match_cost."""
        a = [np.ascontiguousarray(x, self.dtype) for x in (F, G, E, V)]
        nz, ny, nx = a[0].shape
        d = [np.zeros((nz, ny, nx), self.dtype) for _ in range(3)]
        self.lib.oracle_match_grad(nx, ny, nz, *map(_ptr, a), delta, zeta, *map(_ptr, d))
        return d

    def matching(self, target, spacing=(1.0, 1.0, 1.0), field=None, it_max=100, trace_cap=100000,
                 epsilon=9.99e-4, delta=1e-3, zeta=10.0):
        """The matching solve; returns (field, info) with the trace rows
        (it, cost, max voxel error, step) of the accepted iterations."""
        t = np.ascontiguousarray(target, self.dtype)
        nz, ny, nx = t.shape
        f = (np.zeros((3, nz, ny, nx), self.dtype) if field is None
             else self._field(field).copy())
        trace = np.zeros((max(trace_cap, 1), 4), np.float64)
        res = _MatchResult()
        res.trace = trace.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        res.trace_cap = trace_cap
        rc = self.lib.oracle_solve_matching(nx, ny, nz, *spacing, _ptr(t), epsilon, delta, zeta, it_max,
                                            _ptr(f), ctypes.byref(res))
        if rc != 0:
            raise MemoryError("oracle matching could not allocate scratch memory")
        info = dict(iterations=res.iterations, reason=res.reason, cost_evaluations=res.cost_evaluations,
                    gradient_evaluations=res.gradient_evaluations, final_cost=res.final_cost,
                    final_max_error=res.final_max_error,
                    trace=trace[:min(res.trace_len, trace_cap)].copy())
        return f, info


_LINE = re.compile(r"Iteration\s+(-?\d+):\s+total error\s+(\S+)\s+max voxel error\s+(\S+)\s+eta\s+(\S+)")


class Reference(_Base):
    """The reference C library built from its sources into ``oracle/_ref``."""

    def __init__(self, bits=32):
        super().__init__(reference_lib_path(bits), bits)
        L, R = self.lib, self.ctype
        sz = ctypes.c_size_t
        vp = ctypes.c_void_p
        L.jacobian_dynamic.argtypes = [sz, sz, sz, R, R, R, vp, vp]
        L.jacobian_dynamic.restype = None
        solver = [sz, sz, sz, R, R, R, vp, vp] + [R] * 11 + [ctypes.c_bool, sz, vp]
        for name in ("generate_displacement_gradient", "generate_displacement_greedy"):
            getattr(L, name).argtypes = solver
            getattr(L, name).restype = None
        self.has_matching = hasattr(L, "volume_matching_3d")
        if self.has_matching:
            L.volume_matching_3d.argtypes = solver
            L.volume_matching_3d.restype = None
        self._libc = ctypes.CDLL(None)

    def matching(self, target, mask=None, spacing=(1.0, 1.0, 1.0), field=None, it_max=100, **params):
        """volume_matching_3d (src/VolumeMatching3D.c:689-796); the log lines
        of the accepted iterations are parsed into info['trace']."""
        t = np.ascontiguousarray(target, self.dtype)
        nz, ny, nx = t.shape
        m = (np.ones(t.shape, np.bool_) if mask is None else np.ascontiguousarray(mask, np.bool_))
        f = (np.zeros((3, nz, ny, nx), self.dtype) if field is None
             else self._field(field).copy())
        args, strict = self._solver_args(params)
        call = lambda: self.lib.volume_matching_3d(nx, ny, nz, *spacing, _ptr(t), _ptr(m), *args, strict,
                                                   it_max, _ptr(f))
        # The reference takes its gradient buffers from malloc without clearing
        # them before the first gradient (VolumeMatching3D.c:422-424; they are
        # zeroed only after an accepted step, :532-552), so when malloc recycles
        # freed heap memory the first step picks up stale values.  While it
        # runs, glibc's perturb option (M_PERTURB, mallopt) fills every block
        # malloc hands out with the byte 255 ^ 0xff = 0: the reference then
        # computes what it computes on fresh memory, a zero-initialised gradient.
        M_PERTURB = -6
        self._libc.mallopt(M_PERTURB, 255)
        try:
            text = self._with_stdout(call, True)
        finally:
            self._libc.mallopt(M_PERTURB, 0)
        rows = [[float(mm.group(i)) for i in range(1, 5)] for mm in map(_LINE.search, text.splitlines()) if mm]
        return f, dict(trace=np.array(rows, np.float64).reshape(-1, 4), text=text)

    def jacobian(self, field, spacing=(1.0, 1.0, 1.0)):
        f = self._field(field)
        nz, ny, nx = f.shape[1:]
        J = np.zeros((nz, ny, nx), self.dtype)
        self.lib.jacobian_dynamic(nx, ny, nz, *spacing, _ptr(f), _ptr(J))
        return J

    def solve(self, algorithm, target, mask, spacing=(1.0, 1.0, 1.0), field=None,
              it_max=50000, capture=True, **params):
        """Run the reference solver.  With ``capture`` the verbose log lines are
        parsed into ``info['trace']`` (rows it, E, max, eta as printed with
        ``%6e``), ``info['lines']`` and the termination reason; without it the
        log goes to /dev/null (for timing)."""
        t = np.ascontiguousarray(target, self.dtype)
        m = np.ascontiguousarray(mask, np.bool_)
        nz, ny, nx = t.shape
        f = (np.zeros((3, nz, ny, nx), self.dtype) if field is None
             else self._field(field).copy())
        args, strict = self._solver_args(params)
        fn = {"gradient": self.lib.generate_displacement_gradient,
              "greedy": self.lib.generate_displacement_greedy}[algorithm]
        call = lambda: fn(nx, ny, nz, *spacing, _ptr(t), _ptr(m), *args, strict, it_max, _ptr(f))
        text = self._with_stdout(call, capture)
        info = {}
        if capture:
            info = self._parse(text, algorithm, params)
        return f, info

    def _with_stdout(self, call, capture):
        self._libc.fflush(None)
        saved = os.dup(1)
        if capture:
            tmp = tempfile.TemporaryFile(mode="w+b")
            target_fd = tmp.fileno()
        else:
            tmp = None
            target_fd = os.open(os.devnull, os.O_WRONLY)
        try:
            os.dup2(target_fd, 1)
            call()
            self._libc.fflush(None)
        finally:
            os.dup2(saved, 1)
            os.close(saved)
            if tmp is None:
                os.close(target_fd)
        if tmp is None:
            return ""
        tmp.seek(0)
        text = tmp.read().decode()
        tmp.close()
        return text

    @staticmethod
    def _parse(text, algorithm, params):
        p = dict(DEFAULT_PARAMS)
        p.update(params)
        rows, lines = [], []
        reason = MAX_ITERATIONS
        for line in text.splitlines():
            mm = _LINE.search(line)
            if mm:
                rows.append([float(mm.group(i)) for i in range(1, 5)])
                lines.append(line.strip())
            elif "error exploded" in line and "voxel" not in line:
                reason = ERROR_EXPLODED
            elif "voxel error exploded" in line:
                reason = VOXEL_EXPLODED
            elif "reached desired tolerance" in line:
                reason = TOLERANCE
            elif "Error not decreasing, terminating" in line:
                reason = NO_IMPROVEMENT
                if algorithm == "gradient":
                    if rows and rows[-1][3] < p["iota"]:
                        reason = ETA_SMALL
                elif rows and rows[-1][3] * p["beta"] < p["iota"]:
                    # greedy prints the same message after a rejected step whose
                    # reduced eta fell below iota (displacement_field_greedy.c:198-206)
                    reason = ETA_SMALL
        trace = np.array(rows, np.float64).reshape(-1, 4)
        iterations = int(trace[-1, 0]) if len(trace) else 0
        if reason == MAX_ITERATIONS and len(trace):
            iterations += 1
        return dict(trace=trace, lines=lines, reason=reason, iterations=iterations, text=text)


def format_trace_line(it, E, m, eta):
    """The reference's log line (displacement_field_gradient.c:272-277)."""
    return "Iteration %5d:  total error %6e  max voxel error %6e  eta %6e" % (it, E, m, eta)
