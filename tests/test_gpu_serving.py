"""SolveServer (wire format over the batch engine) on the GPU: every response
matches the FP64 oracle to the solver parity bar and equals the engine's own
result for the same input."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_server_handles_encoded_requests(gpu_lib, oracle32, oracle64):
    from conftest import sphere
    from disptools_amd.serving import SolveRequest, SolveServer, decode_response, encode_request
    from disptools_amd.workloads import volume_case
    reqs = []
    T, M = sphere(16, 5)
    reqs.append(SolveRequest(T, M, (1.0, 1.0, 1.0), "gradient", {"it_max": 300}, request_id="a"))
    T, M = sphere(24, 7, inside=0.9)
    reqs.append(SolveRequest(T, M, (1.0, 1.0, 1.5), "greedy", {"it_max": 200}, request_id="b"))
    T, M = volume_case((21, 13, 34), 11)
    reqs.append(SolveRequest(T, M, (1.0, 1.2, 0.9), "gradient", {"it_max": 200}, request_id="c"))
    with SolveServer(max_slots=2) as srv:
        out = {r.request_id: r for r in map(decode_response, srv.handle([encode_request(r) for r in reqs]))}
    assert set(out) == {"a", "b", "c"}
    for r in reqs:
        res = out[r.request_id]
        assert res.status == 0
        fo, io = oracle64.solve(r.algorithm, r.target.astype(np.float64), r.mask, r.spacing, **r.params)
        assert res.reason == io["reason"], (r.request_id, res.reason, io["reason"])
        _, mo, _ = oracle32.compute_error(r.target, oracle32.jacobian(fo.astype(np.float32), r.spacing), r.mask, 0.2)
        _, mg, _ = oracle32.compute_error(r.target, oracle32.jacobian(res.field, r.spacing), r.mask, 0.2)
        print(r.request_id, r.algorithm, "max error: server %.7g  fp64 oracle %.7g" % (mg, mo))
        assert abs(mg - mo) <= 1e-4


def test_server_solves_matching_requests(gpu_lib):
    """Matching requests beside engine requests: the response equals the
    matching entry point's result for the same input, bit for bit."""
    from conftest import sphere
    from disptools_amd.serving import SolveRequest, SolveServer, decode_response, encode_request
    T1, M1 = sphere(16, 5)
    T2, M2 = sphere(20, 6, inside=0.95)
    reqs = [SolveRequest(T1, M1, (1.0, 1.0, 1.0), "gradient", {"it_max": 200}, request_id="g"),
            SolveRequest(T2, M2, (1.0, 1.0, 1.2), "matching", {"it_max": 60, "epsilon": 1e-3}, request_id="m")]
    with SolveServer(max_slots=2) as srv:
        out = {r.request_id: r for r in map(decode_response, srv.handle([encode_request(r) for r in reqs]))}
    assert set(out) == {"g", "m"} and out["g"].status == 0 and out["m"].status == 0
    fm, im, _ = gpu_lib.matching_host(T2, (1.0, 1.0, 1.2), it_max=60, epsilon=1e-3)
    res = out["m"]
    np.testing.assert_array_equal(res.field, fm)
    assert (res.reason, res.iterations) == (im["reason"], im["iterations"])
    assert res.final_max_error == im["final_max_error"]
    assert res.evaluations == im["cost_passes"] + im["gradient_passes"]


def test_server_answers_a_refused_request_and_keeps_serving(gpu_lib):
    """A request the engine refuses (here: larger than the engine's pool)
    comes back as a failed response; the valid request beside it is solved."""
    from conftest import sphere
    from disptools_amd.serving import SolveRequest, SolveServer, check_request
    T1, M1 = sphere(16, 5)
    T2, M2 = sphere(40, 12)
    small = SolveRequest(T1, M1, (1.0, 1.0, 1.0), "gradient", {"it_max": 100}, request_id="small")
    big = SolveRequest(T2, M2, (1.0, 1.0, 1.0), "gradient", {"it_max": 100}, request_id="big")
    with pytest.raises(ValueError):
        check_request((2, 16, 16), "gradient", {})
    probe = gpu_lib.Engine(hbm_budget=1 << 26)
    fp_small = probe.footprint("gradient", 16, 16, 16, host_io=True)
    fp_big = probe.footprint("gradient", 40, 40, 40, host_io=True)
    probe.close()
    assert fp_small < fp_big
    with SolveServer(hbm_budget=fp_small + (fp_big - fp_small) // 2) as srv:
        out = {r.request_id: r for r in srv.run([big, small])}
    assert set(out) == {"big", "small"}
    assert out["big"].status == -1 and "hbm_budget" in out["big"].reason_text
    assert out["small"].status == 0 and out["small"].reason != 0
    f, info, _ = gpu_lib.solve_host("gradient", T1, M1, it_max=100)
    assert out["small"].reason == info["reason"]
