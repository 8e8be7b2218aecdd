"""Following a solve from the host.

The pass kernels store a completion word (passes run, solve done) and, once
done, the control block in host-coherent memory; the host enqueues passes
against it.  When no word arrives (DISPTOOLS_STATUS_WORD=0: the kernels
publish none) the host falls back to reading the control blocks from device
memory once the stream has drained.  Both ways give bit-identical fields and
records, for the single-volume solvers, the matching solver and the batch
engine (where the fallback also stands in for the published control block of
a finished instance)."""

import numpy as np
import pytest

from conftest import sphere
from test_gpu_engine import SAME_KEYS, _cases, _run_all

pytestmark = pytest.mark.gpu

MATCH_KEYS = ("iterations", "reason", "cost_passes", "gradient_passes", "final_cost", "final_max_error")


@pytest.mark.parametrize("alg", ["gradient", "greedy"])
def test_solve_without_status_words(gpu_lib, monkeypatch, alg):
    T, M = sphere(20, 6, inside=0.9)
    print("words", alg, flush=True)
    f1, i1, _ = gpu_lib.solve_host(alg, T, M, it_max=120)
    monkeypatch.setenv("DISPTOOLS_STATUS_WORD", "0")
    print("fallback", alg, flush=True)
    f0, i0, _ = gpu_lib.solve_host(alg, T, M, it_max=120)
    for k in SAME_KEYS:
        assert i0[k] == i1[k], (k, i0[k], i1[k])
    np.testing.assert_array_equal(f0, f1)


def test_matching_without_status_words(gpu_lib, monkeypatch):
    T, _ = sphere(16, 5)
    print("words matching", flush=True)
    f1, i1, _ = gpu_lib.matching_host(T, it_max=40)
    monkeypatch.setenv("DISPTOOLS_STATUS_WORD", "0")
    print("fallback matching", flush=True)
    f0, i0, _ = gpu_lib.matching_host(T, it_max=40)
    for k in MATCH_KEYS:
        assert i0[k] == i1[k], (k, i0[k], i1[k])
    np.testing.assert_array_equal(f0, f1)


def test_engine_without_status_words(gpu_lib, monkeypatch):
    cases = _cases()[:3]
    print("words engine", flush=True)
    with gpu_lib.Engine() as eng:
        f1, r1 = _run_all(eng, cases, stagger=2)
    monkeypatch.setenv("DISPTOOLS_STATUS_WORD", "0")
    print("fallback engine", flush=True)
    with gpu_lib.Engine() as eng:
        f0, r0 = _run_all(eng, cases, stagger=2)
    for i in range(len(cases)):
        assert r0[i][0] == r1[i][0] == 0
        for k in SAME_KEYS:
            assert r0[i][1][k] == r1[i][1][k], (i, k, r0[i][1][k], r1[i][1][k])
        np.testing.assert_array_equal(f0[i], f1[i])
