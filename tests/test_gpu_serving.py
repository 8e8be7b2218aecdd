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
