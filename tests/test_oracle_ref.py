"""The oracle against the reference C compiled from its own sources (oracle/_ref),
bit for bit at one OpenMP thread, and against the committed golden fixtures."""

import os

import numpy as np
import pytest

from conftest import sphere
from oracle.oracle import format_trace_line

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "golden.npz")


@pytest.mark.parametrize("shape,spacing", [
    ((7, 6, 5), (2.0, 3.0, 4.0)),
    ((3, 3, 3), (1.0, 1.0, 1.0)),
    ((4, 5, 3), (1.0, 1.0, 1.5)),
    ((11, 10, 9), (1.0, 1.0, 1.5)),
    ((16, 16, 16), (0.8, 0.8, 1.0)),
])
@pytest.mark.parametrize("bits", [32, 64])
def test_jacobian_bit_exact(request, bits, shape, spacing):
    ref = request.getfixturevalue("reference%d" % bits)
    orc = request.getfixturevalue("oracle%d" % bits)
    rng = np.random.default_rng(hash(shape) % 2 ** 32)
    f = rng.normal(0, 0.4, (3,) + shape).astype(orc.dtype)
    a = orc.jacobian(f, spacing)
    b = ref.jacobian(f, spacing)
    assert np.array_equal(a.view(np.uint8), b.view(np.uint8))


def _cases():
    T16, M16 = sphere(16, 5)
    T24, M24 = sphere(24, 7, inside=0.9)
    ones = np.ones((8, 8, 8), np.float32)
    nan_t = T16.copy()
    nan_t[8, 8, 8] = np.nan
    guess = np.random.default_rng(9).normal(0, 0.02, (3, 16, 16, 16)).astype(np.float32)
    return {
        "sphere16": (T16, M16, (1.0, 1.0, 1.0), None, 300, {}),
        "sphere24_aniso": (T24, M24, (1.0, 1.0, 1.5), None, 200, {}),
        "strict16": (T16, M16, (1.0, 1.0, 1.0), None, 100, {"strict": True}),
        "ones": (ones, ones > 0, (1.0, 1.0, 1.0), None, 50, {}),
        "nan": (nan_t, M16, (1.0, 1.0, 1.0), None, 50, {}),
        "guess16": (T16, M16, (1.0, 1.0, 1.0), guess, 60, {}),
        "zeta100": (T16, M16, (1.0, 1.0, 1.0), None, 80, {"zeta": 100.0, "tolerance": 0.0}),
    }


@pytest.mark.parametrize("alg", ["gradient", "greedy"])
@pytest.mark.parametrize("case", list(_cases().keys()))
def test_solver_bit_exact(reference32, oracle32, alg, case):
    T, M, sp, guess, it_max, params = _cases()[case]
    fo, io = oracle32.solve(alg, T, M, sp, field=guess, it_max=it_max, **params)
    fr, ir = reference32.solve(alg, T, M, sp, field=guess, it_max=it_max, **params)
    assert np.array_equal(fo.view(np.uint8), fr.view(np.uint8))
    lines = [format_trace_line(*row) for row in io["trace"]]
    assert lines == ir["lines"]
    assert io["reason"] == ir["reason"]


def test_solver_fp64_matches_reference(reference64, oracle64):
    T, M = sphere(16, 5)
    for alg in ("gradient", "greedy"):
        fo, io = oracle64.solve(alg, T, M, it_max=200)
        fr, ir = reference64.solve(alg, T, M, it_max=200)
        assert np.array_equal(fo, fr)
        assert io["reason"] == ir["reason"]


@pytest.mark.skipif(not os.path.exists(GOLDEN), reason="golden fixtures not generated")
def test_oracle_reproduces_golden_fixtures(oracle32):
    g = np.load(GOLDEN)
    names = sorted({k.split("/")[0] for k in g.files})
    assert names
    for name in names:
        kind = str(g[name + "/kind"])
        if kind == "jacobian":
            J = oracle32.jacobian(g[name + "/field"], tuple(g[name + "/spacing"]))
            assert np.array_equal(J, g[name + "/J"]), name
        else:
            params = dict(zip([str(k) for k in g[name + "/param_names"]],
                              [float(v) for v in g[name + "/param_values"]]))
            params["strict"] = bool(params.get("strict", 0.0))
            guess = g[name + "/guess"] if (name + "/guess") in g.files else None
            f, info = oracle32.solve(kind, g[name + "/target"], g[name + "/mask"],
                                     tuple(g[name + "/spacing"]), field=guess,
                                     it_max=int(g[name + "/it_max"]), **params)
            assert int(info["reason"]) == int(g[name + "/reason"]), name
            lines = [format_trace_line(*row) for row in info["trace"]]
            assert lines == [str(s) for s in g[name + "/lines"]], name
            if (name + "/field") in g.files:
                assert np.array_equal(f, g[name + "/field"]), name
            import hashlib
            assert hashlib.sha256(f.tobytes()).hexdigest() == str(g[name + "/field_sha256"]), name
