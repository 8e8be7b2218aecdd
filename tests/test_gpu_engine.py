"""Batch engine (continuous batching) on the GPU.

* Each result is the same alone and in the mix: an instance solved by itself
  and the same instance solved among others of other sizes and algorithms,
  admitted at different times, give bit-identical fields and records.
* With the lone-solve tiling (tz=-1) and the same column tile the engine's
  result is bit-identical to dt_solve_device's.
* Against the FP64 oracle: the same termination reason and a final masked
  max |J - T| within 1e-4 (the solver parity bar).
* Admission: a slot limit and a pool budget smaller than the batch queue the
  rest (submit returns None) until results retire.
* Host buffers (pinned) give the same results as device buffers; hipGraph
  replay gives the same results as direct launches.
* A solve whose line search cannot terminate (beta = 1) is stopped by the
  host's pass limit and reported failed; the others finish normally.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_ERR_TOL = 1e-4


def _cases():
    from disptools_amd.workloads import volume_case
    from conftest import sphere
    out = []
    T, M = sphere(16, 5)
    out.append(("gradient", T, M, (1.0, 1.0, 1.0), 300))
    T, M = volume_case((21, 13, 34), 11)
    out.append(("gradient", T, M, (1.0, 1.2, 0.9), 200))
    T, M = sphere(24, 7, inside=0.9)
    out.append(("greedy", T, M, (1.0, 1.0, 1.5), 200))
    T, M = volume_case((40, 33, 27), 12)
    out.append(("gradient", T, M, (1.0, 1.0, 1.0), 300))
    T, M = volume_case((18, 26, 22), 13)
    out.append(("greedy", T, M, (1.0, 1.0, 1.0), 100))
    return out


class _Dev:
    """Device copies of one instance's inputs and a zero initial field."""

    def __init__(self, T, M):
        import torch
        d = torch.device("cuda:0")
        self.T = torch.from_numpy(T).to(d)
        self.M = torch.from_numpy(M.astype(np.uint8)).to(d)
        self.F = torch.zeros((3,) + T.shape, dtype=torch.float32, device=d)

    def field(self):
        return self.F.cpu().numpy()


def _submit(eng, case, dev, host_io=False, host=None):
    alg, T, M, sp, it_max = case
    nz, ny, nx = T.shape
    if host_io:
        t, m, f = host
        return eng.submit(alg, nx, ny, nz, sp, t.data_ptr(), m.data_ptr(), f.data_ptr(), host_io=True,
                          it_max=it_max)
    return eng.submit(alg, nx, ny, nz, sp, dev.T.data_ptr(), dev.M.data_ptr(), dev.F.data_ptr(),
                      it_max=it_max)


def _run_all(eng, cases, stagger=0):
    """Submits the cases (the first `stagger` now, the rest one per progress
    step while the first ones run) and drains; returns fields and records."""
    import torch
    devs = [_Dev(c[1], c[2]) for c in cases]
    torch.cuda.synchronize()
    tickets = {}
    pending = list(range(len(cases)))
    first = len(cases) if stagger <= 0 else stagger
    for i in pending[:first]:
        t = _submit(eng, cases[i], devs[i])
        assert t is not None
        tickets[t] = i
    pending = pending[first:]
    results = {}
    while pending or len(results) < len(cases):
        if pending:
            t = _submit(eng, cases[pending[0]], devs[pending[0]])
            if t is not None:
                tickets[t] = pending.pop(0)
        eng.progress(wait=True)
        for t, status, info in eng.poll():
            results[tickets[t]] = (status, info)
    torch.cuda.synchronize()
    return [devs[i].field() for i in range(len(cases))], [results[i] for i in range(len(cases))]


SAME_KEYS = ("iterations", "evaluations", "rejections", "reason", "final_error", "final_max_error", "field_io")


def test_engine_alone_equals_mixed(gpu_lib):
    cases = _cases()
    alone_f, alone_r = [], []
    for c in cases:
        print("alone", c[0], c[1].shape, flush=True)
        with gpu_lib.Engine() as eng:
            f, r = _run_all(eng, [c])
        alone_f.append(f[0])
        alone_r.append(r[0])
    print("mixed (staggered admission)", flush=True)
    with gpu_lib.Engine() as eng:
        mixed_f, mixed_r = _run_all(eng, cases, stagger=2)
        st = eng.stats()
    assert st["admitted"] == st["finished"] == len(cases)
    for i, c in enumerate(cases):
        assert mixed_r[i][0] == alone_r[i][0] == 0
        for k in SAME_KEYS:
            assert mixed_r[i][1][k] == alone_r[i][1][k], (i, k, mixed_r[i][1][k], alone_r[i][1][k])
        np.testing.assert_array_equal(mixed_f[i], alone_f[i])


@pytest.mark.parametrize("tz", [-1, 0, -2, 8])
def test_engine_equals_plain_solve_for_any_tiling(gpu_lib, tz):
    """The engine's result equals dt_solve_device's, bit for bit, whatever the
    tiling: lone-solve z-chunks (-1), the batch default (0: re-planned as the
    set of instances changes), chunks of at most 16 planes (-2) or 8 planes, and
    the lane's column tile against the one a single solve picks per shape
    (the pass sums are exact)."""
    cases = _cases()
    with gpu_lib.Engine(tz=tz) as eng:
        ef, er = _run_all(eng, cases)
    for i, (alg, T, M, sp, it_max) in enumerate(cases):
        print("plain", alg, T.shape, "engine tz", tz, flush=True)
        f, info, _ = gpu_lib.solve_host(alg, T, M, sp, it_max=it_max)
        for k in SAME_KEYS:
            assert er[i][1][k] == info[k], (i, k, er[i][1][k], info[k])
        np.testing.assert_array_equal(ef[i], f)


def _random_batch(seed, count):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        shape = tuple(int(v) for v in rng.integers(3, 91, size=3))
        alg = ["gradient", "greedy"][int(rng.integers(0, 2))]
        sp = tuple(float(v) for v in rng.uniform(0.6, 1.6, size=3))
        T = (1.0 + rng.uniform(0.02, 0.12) * rng.standard_normal(shape)).astype(np.float32)
        M = rng.random(shape) < rng.uniform(0.3, 1.0)
        out.append((alg, T, M, sp, int(rng.integers(10, 120))))
    return out


def test_engine_random_batch_equals_lone_solves(gpu_lib):
    """Sixteen seeded volumes of 3..90 voxels per axis, gradient and greedy,
    random spacing, masks and iteration limits, admitted five at once and the
    rest as slots free: every result equals its lone dt_solve_host, bit for
    bit."""
    cases = _random_batch(31, 16)
    with gpu_lib.Engine() as eng:
        ef, er = _run_all(eng, cases, stagger=5)
    for i, (alg, T, M, sp, it_max) in enumerate(cases):
        print("random batch", i, alg, T.shape, flush=True)
        f, info, _ = gpu_lib.solve_host(alg, T, M, sp, it_max=it_max)
        assert er[i][0] == 0
        for k in SAME_KEYS:
            assert er[i][1][k] == info[k], (i, k, er[i][1][k], info[k])
        np.testing.assert_array_equal(ef[i], f)


def test_engine_parity_with_fp64_oracle(gpu_lib, oracle32, oracle64):
    cases = _cases()
    with gpu_lib.Engine() as eng:
        ef, er = _run_all(eng, cases)
    for i, (alg, T, M, sp, it_max) in enumerate(cases):
        fo, io = oracle64.solve(alg, T.astype(np.float64), M, sp, it_max=it_max)
        assert er[i][1]["reason"] == io["reason"], (i, er[i][1]["reason"], io["reason"])
        _, mo, _ = oracle32.compute_error(T, oracle32.jacobian(fo.astype(np.float32), sp), M, 0.2)
        _, mg, _ = oracle32.compute_error(T, oracle32.jacobian(ef[i], sp), M, 0.2)
        print(alg, T.shape, "final masked max error: engine %.7g  fp64 oracle %.7g" % (mg, mo))
        assert abs(mg - mo) <= MAX_ERR_TOL


def test_engine_admission_limits(gpu_lib):
    import torch
    cases = _cases()
    with gpu_lib.Engine() as eng:
        ref_f, ref_r = _run_all(eng, cases)
    # two slots
    with gpu_lib.Engine(max_slots=2) as eng:
        devs = [_Dev(c[1], c[2]) for c in cases]
        torch.cuda.synchronize()
        assert _submit(eng, cases[0], devs[0]) is not None
        assert _submit(eng, cases[1], devs[1]) is not None
        assert _submit(eng, cases[2], devs[2]) is None
    # a pool that holds the largest instance but not all of them
    probe = gpu_lib.Engine(hbm_budget=1 << 26)
    sizes = [probe.footprint(c[0], c[1].shape[2], c[1].shape[1], c[1].shape[0]) for c in cases]
    probe.close()
    budget = max(sizes) + min(sizes)
    assert budget < sum(sizes)
    with gpu_lib.Engine(hbm_budget=budget) as eng:
        f, r = _run_all(eng, cases, stagger=1)
        st = eng.stats()
    assert st["pool_peak"] <= budget
    for i in range(len(cases)):
        np.testing.assert_array_equal(f[i], ref_f[i])
        assert r[i][1]["reason"] == ref_r[i][1]["reason"]


def test_engine_host_io_and_graphs(gpu_lib):
    import torch
    cases = _cases()
    with gpu_lib.Engine() as eng:
        ref_f, ref_r = _run_all(eng, cases)
    # pinned host buffers
    host = []
    for alg, T, M, sp, it_max in cases:
        t = torch.from_numpy(T).pin_memory()
        m = torch.from_numpy(M.astype(np.uint8)).pin_memory()
        f = torch.zeros((3,) + T.shape, dtype=torch.float32).pin_memory()
        host.append((t, m, f))
    with gpu_lib.Engine() as eng:
        tickets = {}
        for i, c in enumerate(cases):
            t = _submit(eng, c, None, host_io=True, host=host[i])
            assert t is not None
            tickets[t] = i
        eng.drain()
        res = {tickets[t]: (s, info) for t, s, info in eng.poll()}
    for i in range(len(cases)):
        np.testing.assert_array_equal(host[i][2].numpy(), ref_f[i])
        assert res[i][1]["iterations"] == ref_r[i][1]["iterations"]
    # graph replay, twice in a row in one engine with eager work between
    with gpu_lib.Engine(use_graphs=True) as eng:
        for rep in range(2):
            print("graphs, round", rep, flush=True)
            gf, gr = _run_all(eng, cases, stagger=3)
            torch.zeros(1 << 20, device="cuda:0").add_(1.0)
            torch.cuda.synchronize()
            for i in range(len(cases)):
                np.testing.assert_array_equal(gf[i], ref_f[i])
                for k in SAME_KEYS:
                    assert gr[i][1][k] == ref_r[i][1][k]
        assert eng.stats()["graph_builds"] >= 1


def test_engine_stops_a_solve_that_cannot_terminate(gpu_lib):
    import torch
    from conftest import sphere
    T, M = sphere(12, 4)
    good = _cases()[0]
    with gpu_lib.Engine() as eng:
        bad = _Dev(T, M)
        ok = _Dev(good[1], good[2])
        torch.cuda.synchronize()
        nz, ny, nx = T.shape
        # beta = 1: a rejected trial is retried with the same step forever
        # (alpha = 0.5 makes the search step shrink, so a rejection comes)
        tb = eng.submit("gradient", nx, ny, nz, (1.0, 1.0, 1.0), bad.T.data_ptr(), bad.M.data_ptr(),
                        bad.F.data_ptr(), it_max=3, beta=1.0, eta=50.0, eta_max=100.0)
        tg = _submit(eng, good, ok)
        eng.drain()
        res = {t: (s, info) for t, s, info in eng.poll()}
    assert res[tg][0] == 0
    assert res[tb][0] == -1
    assert np.isfinite(bad.field()).all()
