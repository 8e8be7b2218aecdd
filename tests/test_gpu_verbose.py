"""Log lines of the reference entry points: the release build of the reference
prints one line per iteration and a termination message
(src/displacement_field_gradient.c:272-418, displacement_field_greedy.c:184-236,
VolumeMatching3D.c:538-544); the HIP entry points print the same lines from
the solve's trace, parsed here with the oracle's parser of the reference's
output."""

import ctypes
import os
import re
import tempfile

import numpy as np
import pytest

from conftest import sphere
from oracle.oracle import Reference

pytestmark = pytest.mark.gpu

LINE = re.compile(r"^Iteration\s+(\d+):  total error (\S+)  max voxel error (\S+)  eta (\S+)$")


def _capture(call):
    libc = ctypes.CDLL(None)
    libc.fflush(None)
    saved = os.dup(1)
    with tempfile.TemporaryFile(mode="w+b") as tmp:
        os.dup2(tmp.fileno(), 1)
        try:
            call()
            libc.fflush(None)
        finally:
            os.dup2(saved, 1)
            os.close(saved)
        tmp.seek(0)
        return tmp.read().decode()


def _ref_args(T, M, f, it_max, sp=(1.0, 1.0, 1.0)):
    fl = ctypes.c_float
    vp = ctypes.c_void_p
    return [T.shape[2], T.shape[1], T.shape[0], *map(fl, sp), T.ctypes.data_as(vp), M.ctypes.data_as(vp),
            fl(9.99e-4), fl(0.2), fl(0.1), fl(0.4), fl(1.2), fl(0.5), fl(0.1), fl(1e-3), fl(10.0), fl(1e-6),
            fl(1e-9), False, it_max, f.ctypes.data_as(vp)]


@pytest.mark.parametrize("alg", ["gradient", "greedy"])
def test_solver_log_lines(gpu_lib, alg):
    T, M = sphere(24, 7, inside=0.9)
    it_max = 300
    fn = getattr(gpu_lib.lib, "generate_displacement_%s_cuda" % alg)
    f = np.zeros((3,) + T.shape, np.float32)
    text = _capture(lambda: fn(*_ref_args(T, M, f, it_max)))
    fh, info, trace = gpu_lib.solve_host(alg, T, M, it_max=it_max, trace_cap=100000)
    assert np.array_equal(f, fh)
    lines = [l for l in text.splitlines() if l.startswith("Iteration")]
    assert all(LINE.match(l) for l in lines), lines[:3]
    assert len(lines) == len(trace)
    got = np.array([[float(v) for v in LINE.match(l).groups()] for l in lines])
    np.testing.assert_allclose(got, trace, rtol=1e-5, atol=0)
    if alg == "gradient":
        assert got[0, 0] == 0
    parsed = Reference._parse(text, alg, {})
    assert parsed["reason"] == info["reason"], (parsed["reason"], info["reason"], text[-300:])
    if alg == "greedy" and info["rejections"] > 0:
        assert "reducing step size" in text


def test_matching_log_lines(gpu_lib):
    T, M = sphere(16, 5)
    f = np.zeros((3,) + T.shape, np.float32)
    text = _capture(lambda: gpu_lib.lib.volume_matching_3d(*_ref_args(T, M, f, 30)))
    fh, info, trace = gpu_lib.matching_host(T, it_max=30, trace_cap=100)
    lines = [l for l in text.splitlines() if l.startswith("Iteration")]
    assert len(lines) == info["iterations"] == len(trace)
    got = np.array([[float(v) for v in LINE.match(l).groups()] for l in lines])
    np.testing.assert_allclose(got, trace, rtol=1e-5)


def test_quiet_when_disabled(gpu_lib, monkeypatch):
    monkeypatch.setenv("DISPTOOLS_VERBOSE", "0")
    T, M = sphere(16, 5)
    f = np.zeros((3,) + T.shape, np.float32)
    text = _capture(lambda: gpu_lib.lib.generate_displacement_gradient_cuda(*_ref_args(T, M, f, 50)))
    assert text == ""
