"""Matching solves (src/VolumeMatching3D.c:689-796) in the batch engine: each
result equals dt_matching_host's bit for bit (field, iterations, stop reason,
final cost), whether the instance runs alone, beside other matching solves of
other shapes, beside gradient and greedy solves, or from captured graphs; the
gradient and greedy solves beside them are unchanged by their company."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(10, 12, 14), (16, 16, 16), (7, 19, 12), (20, 18, 22)]   # (nz, ny, nx)
SPACING = (1.0, 1.1, 0.9)
MPARAMS = dict(it_max=40, epsilon=9.99e-4, delta=1e-3, zeta=10.0)


def _targets(seed=5):
    rng = np.random.default_rng(seed)
    return [(1.0 + 0.08 * rng.standard_normal(s)).astype(np.float32) for s in SHAPES]


def _reference(gpu_lib, T):
    f, info, _ = gpu_lib.matching_host(T, SPACING, **MPARAMS)
    return f, info


def _submit_matching(eng, T, F):
    nz, ny, nx = T.shape
    t = eng.submit("matching", nx, ny, nz, SPACING, T.ctypes.data, None, F.ctypes.data, host_io=True, **MPARAMS)
    assert t is not None
    return t


def _submit_solve(eng, alg, T, M, F):
    nz, ny, nx = T.shape
    t = eng.submit(alg, nx, ny, nz, SPACING, T.ctypes.data, M.ctypes.data, F.ctypes.data, host_io=True, it_max=60)
    assert t is not None
    return t


def _drain(eng):
    eng.drain()
    return {t: (st, info) for t, st, info in eng.poll()}


def _check(res, F, ref):
    st, info = res
    fr, ir = ref
    assert st == 0
    assert info["iterations"] == ir["iterations"]
    assert info["reason"] == ir["reason"]
    assert info["final_error"] == ir["final_cost"]
    assert np.array_equal(F, fr)


def test_engine_matching_alone_and_together(gpu_lib):
    Ts = _targets()
    refs = [_reference(gpu_lib, T) for T in Ts]
    for k, T in enumerate(Ts):
        print("matching alone", T.shape, flush=True)
        with gpu_lib.Engine() as eng:
            assert eng.footprint("matching", T.shape[2], T.shape[1], T.shape[0]) > 0
            F = np.zeros((3,) + T.shape, np.float32)
            t = _submit_matching(eng, T, F)
            _check(_drain(eng)[t], F, refs[k])
    print("matching together", flush=True)
    with gpu_lib.Engine() as eng:
        Fs = [np.zeros((3,) + T.shape, np.float32) for T in Ts]
        ts = [_submit_matching(eng, T, F) for T, F in zip(Ts, Fs)]
        res = _drain(eng)
        for t, F, ref in zip(ts, Fs, refs):
            _check(res[t], F, ref)


@pytest.mark.parametrize("use_graphs", [False, True])
def test_engine_matching_beside_solves(gpu_lib, use_graphs):
    Ts = _targets(seed=9)
    refs = [_reference(gpu_lib, T) for T in Ts]
    rng = np.random.default_rng(3)
    solves = []
    for k, shape in enumerate([(18, 14, 16), (12, 20, 10)]):
        T = (1.0 + 0.06 * rng.standard_normal(shape)).astype(np.float32)
        M = (rng.random(shape) < 0.8).astype(np.uint8)
        solves.append((("gradient", "greedy")[k], T, M))
    # the gradient and greedy solves alone
    alone = []
    with gpu_lib.Engine(use_graphs=use_graphs) as eng:
        for alg, T, M in solves:
            F = np.zeros((3,) + T.shape, np.float32)
            t = _submit_solve(eng, alg, T, M, F)
            st, info = _drain(eng)[t]
            assert st == 0
            alone.append((F, info))
    print("matching beside solves, graphs", use_graphs, flush=True)
    with gpu_lib.Engine(use_graphs=use_graphs) as eng:
        Fm = [np.zeros((3,) + T.shape, np.float32) for T in Ts]
        Fs = [np.zeros((3,) + T.shape, np.float32) for _, T, _ in solves]
        tm, ts = [], []
        for k in range(len(Ts)):
            tm.append(_submit_matching(eng, Ts[k], Fm[k]))
            if k < len(solves):
                alg, T, M = solves[k]
                ts.append(_submit_solve(eng, alg, T, M, Fs[k]))
            eng.progress(wait=False)
        res = _drain(eng)
        for t, F, ref in zip(tm, Fm, refs):
            _check(res[t], F, ref)
        for t, F, (Fa, ia) in zip(ts, Fs, alone):
            st, info = res[t]
            assert st == 0
            assert info["iterations"] == ia["iterations"] and info["reason"] == ia["reason"]
            assert np.array_equal(F, Fa)


def test_engine_matching_rejects_bad_arguments(gpu_lib):
    with gpu_lib.Engine() as eng:
        T = np.ones((4, 4, 4), np.float32)
        F = np.zeros((3, 4, 4, 4), np.float32)
        with pytest.raises(gpu_lib.HipError):
            eng.submit("matching", 4, 4, 4, SPACING, T.ctypes.data, None, F.ctypes.data, host_io=True,
                       it_max=5, epsilon=0.0)
        with pytest.raises(gpu_lib.HipError):
            eng.submit("matching", 4, 4, 1, SPACING, T.ctypes.data, None, F.ctypes.data, host_io=True, it_max=5)


@pytest.mark.parametrize("shape,it_max", [((2, 2, 2), 5), ((2, 9, 3), 25), ((33, 2, 5), 1), ((5, 7, 65), 0),
                                          ((3, 3, 40), 30)])
def test_engine_matching_edge_shapes_device_buffers(gpu_lib, shape, it_max):
    """Two-voxel axes, it_max 0 and 1, long thin volumes; device buffers
    (the result is written in place) and an initial field."""
    import torch
    print("matching edge", shape, it_max, flush=True)
    rng = np.random.default_rng(sum(shape) + it_max)
    T = (1.0 + 0.1 * rng.standard_normal(shape)).astype(np.float32)
    F0 = (0.05 * rng.standard_normal((3,) + shape)).astype(np.float32)
    params = dict(MPARAMS, it_max=it_max)
    fr, ir, _ = gpu_lib.matching_host(T, SPACING, field=F0, **params)
    dT = torch.from_numpy(T).to("cuda:0")
    dF = torch.from_numpy(F0).to("cuda:0")
    nz, ny, nx = shape
    with gpu_lib.Engine() as eng:
        t = eng.submit("matching", nx, ny, nz, SPACING, dT.data_ptr(), None, dF.data_ptr(), **params)
        assert t is not None
        st, info = _drain(eng)[t]
    torch.cuda.synchronize()
    assert st == 0
    assert info["iterations"] == ir["iterations"] and info["reason"] == ir["reason"]
    assert info["final_error"] == ir["final_cost"]
    assert np.array_equal(dF.cpu().numpy(), fr)
