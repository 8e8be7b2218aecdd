"""Serving wire format: request / response encode-decode round trips (CPU)."""

import numpy as np
import pytest

from disptools_amd.serving import (SolveRequest, SolveResponse, decode_request, decode_response,
                                   encode_request, encode_response)


def _request(with_field):
    rng = np.random.default_rng(5)
    T = (1.0 + 0.05 * rng.normal(size=(4, 5, 6))).astype(np.float32)
    M = rng.random((4, 5, 6)) < 0.5
    F = rng.normal(size=(3, 4, 5, 6)).astype(np.float32) if with_field else None
    return SolveRequest(target=T, mask=M, spacing=(1.0, 1.2, 0.8), algorithm="greedy",
                        params={"epsilon": 1e-3, "it_max": 77, "strict": True, "zeta": 10.0},
                        field=F, request_id="r-7")


@pytest.mark.parametrize("with_field", [False, True])
def test_request_round_trip(with_field):
    r = _request(with_field)
    d = decode_request(encode_request(r))
    np.testing.assert_array_equal(d.target, r.target)
    np.testing.assert_array_equal(d.mask, r.mask)
    assert d.mask.dtype == np.bool_ and d.target.dtype == np.float32
    assert d.spacing == r.spacing and d.algorithm == "greedy" and d.request_id == "r-7"
    assert d.params == {"epsilon": 1e-3, "it_max": 77, "strict": True, "zeta": 10.0}
    if with_field:
        np.testing.assert_array_equal(d.field, r.field)
    else:
        assert d.field is None


def test_response_round_trip():
    f = np.arange(3 * 2 * 3 * 4, dtype=np.float32).reshape(3, 2, 3, 4)
    r = SolveResponse(request_id="x", status=0, field=f, reason=5, reason_text="reached desired tolerance",
                      iterations=12, evaluations=15, rejections=2, final_error=0.5, final_max_error=1e-3,
                      seconds=0.01)
    d = decode_response(encode_response(r))
    np.testing.assert_array_equal(d.field, f)
    assert (d.reason, d.iterations, d.evaluations, d.rejections) == (5, 12, 15, 2)
    assert d.final_max_error == 1e-3 and d.reason_text == r.reason_text


def test_malformed_messages():
    good = encode_request(_request(True))
    with pytest.raises(ValueError):
        decode_request(good[:-4])
    with pytest.raises(ValueError):
        decode_request(b"DTRS1\n{}\n")
    with pytest.raises(ValueError):
        encode_request(SolveRequest(target=np.ones((3, 3, 3), np.float32), mask=np.ones((3, 3, 2), bool)))
    with pytest.raises(ValueError):
        encode_request(SolveRequest(target=np.ones((3, 3, 3), np.float32), mask=np.ones((3, 3, 3), bool),
                                    params={"learning_rate": 1.0}))


def test_matching_request_and_unknown_algorithm():
    r = _request(False)
    r.algorithm = "matching"
    r.params = {"epsilon": 1e-3, "delta": 1e-3, "zeta": 10.0, "it_max": 50}
    d = decode_request(encode_request(r))
    assert d.algorithm == "matching" and d.params == r.params
    np.testing.assert_array_equal(d.target, r.target)
    r.algorithm = "simulated-annealing"
    with pytest.raises(ValueError):
        encode_request(r)
    bad = encode_request(_request(False)).replace(b'"algorithm": "greedy"', b'"algorithm": "nope"')
    with pytest.raises(ValueError):
        decode_request(bad)


def test_requests_the_solvers_refuse_are_rejected_before_queueing():
    """Axes below the algorithm's minimum (3 for the Jacobian stencils, 2 for
    matching cells) and a non-positive matching epsilon are refused when the
    request is decoded, so a bad request never reaches the engine queue."""
    from disptools_amd.serving import SolveRequest, check_request, decode_request, encode_request
    t = np.ones((2, 8, 8), np.float32)
    buf = encode_request(SolveRequest(target=t, mask=t > 0, algorithm="gradient"))
    with pytest.raises(ValueError, match="at least 3"):
        decode_request(buf)
    decode_request(encode_request(SolveRequest(target=t, mask=t > 0, algorithm="matching")))
    with pytest.raises(ValueError, match="at least 2"):
        check_request((1, 8, 8), "matching", {})
    with pytest.raises(ValueError, match="epsilon"):
        check_request((4, 4, 4), "matching", {"epsilon": 0.0})
    check_request((3, 3, 3), "greedy", {"epsilon": 0.0})


def test_matching_volume_size_is_checked_without_wrapping():
    """dt_matching_host rejects a volume whose voxel count would wrap a size_t
    product, before any device work."""
    import ctypes
    from disptools_amd import native
    info = native.MatchingInfo()
    big = 1 << 22
    rc = native.lib.dt_matching_host(big, big, big, 1.0, 1.0, 1.0, None, 1e-3, 1e-3, 10.0, 10, None,
                                     ctypes.byref(info), None, 0)
    assert rc == -1 and "too large" in native.last_error()
