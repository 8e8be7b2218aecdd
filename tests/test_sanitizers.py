"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only).

* ``oracle/``: the C restatement built with ``-fsanitize=address,undefined``
  into a temporary directory and driven through its own Python bindings
  (jacobian, error, gradient, the three solvers, on minimal 3x3x3 volumes and
  small odd-sized ones, both precisions, two OpenMP threads);
* ``disptools_amd/csrc/_disptools.c``: the extension built the same way
  (against the product library) and driven through every argument check it
  makes itself -- wrong dtype, rank, shape, layout, writeability, spacing
  arity, algorithm, it_max -- each of which ends before any device call.

Each part runs in a child process that preloads the sanitizer runtimes; a
report aborts the child (``halt_on_error``), so the test asserts that the
child ran to its last line.  Leak detection is off: the Python interpreter
itself keeps allocations until exit.
"""

import os
import shutil
import subprocess
import sys
import sysconfig

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer", "-g", "-O1"]


def _runtime(name):
    path = subprocess.run(["gcc", "-print-file-name=%s" % name], capture_output=True, text=True).stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


@pytest.fixture(scope="module")
def san_env():
    if shutil.which("gcc") is None:
        pytest.skip("gcc not found")
    asan, ubsan = _runtime("libasan.so"), _runtime("libubsan.so")
    if not asan or not ubsan:
        pytest.skip("sanitizer runtimes not found")
    env = dict(os.environ)
    env["LD_PRELOAD"] = asan + ":" + ubsan
    env["ASAN_OPTIONS"] = "detect_leaks=0:halt_on_error=1:abort_on_error=1"
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    env["PYTHONPATH"] = ROOT + os.pathsep + env.get("PYTHONPATH", "")
    return env


def _child(code, env, cwd):
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=600, env=env, cwd=cwd)
    out = r.stdout + r.stderr
    assert r.returncode == 0 and "ERROR: AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]
    assert r.stdout.strip().endswith("done"), out[-4000:]


ORACLE_DRIVER = r"""
import numpy as np
from oracle.oracle import Oracle
rng = np.random.default_rng(0)
for bits in (32, 64):
    o = Oracle(bits)
    o.set_threads(2)
    for shape in [(3, 3, 3), (4, 5, 3), (9, 7, 11)]:
        f = rng.normal(0, 0.1, (3,) + shape)
        sp = (1.0, 1.3, 0.8)
        J = o.jacobian(f, sp)
        T = np.full(shape, 1.05)
        M = rng.random(shape) < 0.6
        E, mx, e = o.compute_error(T, J, M, 0.2)
        o.gradient(f, J, e, 1e-3, 10.0, sp)
    z, y, x = np.meshgrid(np.arange(10), np.arange(9), np.arange(8), indexing="ij")
    ball = (x - 4) ** 2 + (y - 4) ** 2 + (z - 5) ** 2 <= 9
    T = np.where(ball, 1.1, 1.0)
    for alg in ("gradient", "greedy"):
        o.solve(alg, T, ball, (1.0, 1.0, 1.5), it_max=40, trace_cap=8)
        o.solve(alg, np.full((3, 3, 3), 1.05), np.ones((3, 3, 3), bool), it_max=5, trace_cap=2)
    o.matching(T, (1.0, 1.0, 1.5), it_max=20, trace_cap=4)
print("done")
"""


def test_oracle_under_sanitizers(san_env, tmp_path):
    src = [os.path.join(ROOT, "oracle", f) for f in ("dt_oracle.c", "dt_oracle_matching.c")]
    for bits in (32, 64):
        out = str(tmp_path / ("liboracle_f%d.so" % bits))
        subprocess.run(["gcc", "-std=c99", "-fopenmp", "-fPIC", "-shared", "-Wall"] + SAN +
                       ["-DORACLE_FLOAT_SIZE=%d" % bits] + src + ["-o", out, "-lm"], check=True)
    env = dict(san_env)
    env["DT_ORACLE_BUILD"] = str(tmp_path)
    _child(ORACLE_DRIVER, env, ROOT)


EXT_DRIVER = r"""
import numpy as np
import _disptools as ext
assert ext.get_float_type_size() in (32, 64)
f32 = np.float32
ok3 = np.zeros((3, 5, 5, 5), f32)
J = np.ones((5, 5, 5), f32)
M = np.ones((5, 5, 5), bool)

def refused(fn, *args):
    try:
        fn(*args)
    except (ValueError, TypeError, RuntimeError):
        return
    raise AssertionError("accepted: %r" % (args[:1],))

# jacobian(spacing, field, out)
refused(ext.jacobian, (1.0, 1.0), ok3, J.copy())
refused(ext.jacobian, (1.0, 1.0, 1.0), ok3.astype(np.float64), J.copy())
refused(ext.jacobian, (1.0, 1.0, 1.0), ok3[0], J.copy())
refused(ext.jacobian, (1.0, 1.0, 1.0), np.zeros((3, 5, 5, 4), f32), J.copy())
refused(ext.jacobian, (1.0, 1.0, 1.0), np.zeros((5, 5, 5, 3), f32)[..., 0:3].transpose(3, 0, 1, 2), J.copy())
ro = J.copy(); ro.setflags(write=False)
refused(ext.jacobian, (1.0, 1.0, 1.0), ok3, ro)
refused(ext.jacobian, "xyz", ok3, J.copy())
# regularise(J, eps)
refused(ext.regularise, J.astype(np.float64), 1e-5)
refused(ext.regularise, J[0], 1e-5)
refused(ext.regularise, ro, 1e-5)
# displacement(spacing, J, mask, 12 floats, strict, it_max, field, algorithm, gpu_id)
P = (9.99e-4, 0.2, 0.1, 0.4, 1.2, 0.5, 0.1, 1e-3, 10.0, 1e-6, 1e-9)
def disp(sp=(1.0, 1.0, 1.0), j=J, m=M, it=100, field=None, alg="gradient"):
    return ext.displacement(sp, j, m, *P, False, it, ok3.copy() if field is None else field, alg, 0)
refused(disp, (1.0, 1.0, 1.0, 1.0))
refused(disp, (1.0, 1.0, 1.0), J.astype(np.float64))
refused(disp, (1.0, 1.0, 1.0), J, np.ones((5, 5, 4), bool))
refused(disp, (1.0, 1.0, 1.0), J, M.astype(np.float32))
refused(disp, (1.0, 1.0, 1.0), J, M, -1)
refused(disp, (1.0, 1.0, 1.0), J, M, 100, np.zeros((3, 5, 5, 6), f32))
refused(disp, (1.0, 1.0, 1.0), J, M, 100, None, "nope")
print("done")
"""


def test_extension_argument_checks_under_sanitizers(san_env, tmp_path):
    import numpy
    lib = os.path.join(ROOT, "disptools_amd", "libdisptools_hip.so")
    if not os.path.exists(lib):
        pytest.fail("libdisptools_hip.so is not built (python -m disptools_amd.build)")
    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    out = str(tmp_path / ("_disptools" + suffix))
    subprocess.run(["gcc", "-std=c99", "-fPIC", "-shared", "-Wall"] + SAN +
                   ["-I", sysconfig.get_paths()["include"], "-I", numpy.get_include(),
                    "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "disptools_amd", "csrc", "_disptools.c"), "-o", out,
                    "-L", os.path.dirname(lib), "-ldisptools_hip", "-Wl,-rpath," + os.path.dirname(lib)],
                   check=True)
    env = dict(san_env)
    env["PYTHONPATH"] = str(tmp_path) + os.pathsep + env["PYTHONPATH"]
    _child(EXT_DRIVER, env, str(tmp_path))
