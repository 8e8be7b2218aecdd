"""Serving formats and a request server on the batch engine.

A solve request carries what ``disptools.displacements.displacement`` takes
for one level (target Jacobian, mask, spacing, algorithm, solver parameters,
optional initial field); the response carries the field and the solve record.
Both travel as bytes:

    b"DTRQ1\\n" | b"DTRS1\\n"          magic line
    <JSON header>\\n                   shapes, spacing, parameters, byte counts
    raw arrays, little-endian         request: target f4, mask u1, [field f4]
                                      response: field f4

``SolveServer`` admits decoded gradient, greedy and matching requests into a
continuous-batching engine as slots and device memory free up and hands back
responses as solves finish (a matching request carries a mask which, as in
the reference, is unused).
"""

import json
from dataclasses import dataclass, field as dc_field
from typing import Dict, List, Optional, Tuple

import numpy as np

REQUEST_MAGIC = b"DTRQ1\n"
RESPONSE_MAGIC = b"DTRS1\n"

ALGORITHMS = ("gradient", "greedy", "matching")

# Parameters a request may set (the reference's solver arguments,
# displacements.py:246-260, plus it_max; matching reads epsilon, delta, zeta
# and it_max).
PARAM_NAMES = ("epsilon", "tolerance", "eta", "eta_max", "alpha", "beta", "gamma", "delta", "zeta",
               "theta", "iota", "strict", "it_max")


@dataclass
class SolveRequest:
    target: np.ndarray                    # [nz, ny, nx] float32
    mask: np.ndarray                      # [nz, ny, nx] bool
    spacing: Tuple[float, float, float] = (1.0, 1.0, 1.0)
    algorithm: str = "gradient"
    params: Dict[str, float] = dc_field(default_factory=dict)
    field: Optional[np.ndarray] = None    # [3, nz, ny, nx] float32 initial guess
    request_id: str = ""


@dataclass
class SolveResponse:
    request_id: str
    status: int                           # 0 solved, -1 failed
    field: np.ndarray                     # [3, nz, ny, nx] float32
    reason: int = 0
    reason_text: str = ""
    iterations: int = 0
    evaluations: int = 0
    rejections: int = 0
    final_error: float = 0.0
    final_max_error: float = 0.0
    seconds: float = 0.0


# Smallest axis each algorithm accepts: the Jacobian stencils read 3 samples
# per axis (src/headers/jacobian_macros.h:8-99); the matching cells span 2.
MIN_AXIS = {"gradient": 3, "greedy": 3, "matching": 2}


def check_request(shape, algorithm, params):
    """Rejects a request the solvers would refuse (ValueError), before it is
    queued: unknown algorithm or parameters, axes below the algorithm's
    minimum, a non-positive matching epsilon."""
    if algorithm not in ALGORITHMS:
        raise ValueError("unknown algorithm %r" % (algorithm,))
    unknown = set(params) - set(PARAM_NAMES)
    if unknown:
        raise ValueError("unknown solver parameters: %s" % sorted(unknown))
    if len(shape) != 3 or min(shape) < MIN_AXIS[algorithm]:
        raise ValueError("%s needs every axis of at least %d voxels, got shape %r"
                         % (algorithm, MIN_AXIS[algorithm], tuple(shape)))
    if algorithm == "matching" and "epsilon" in params and not float(params["epsilon"]) > 0.0:
        raise ValueError("matching needs epsilon > 0")


def _split(buf, magic):
    mv = memoryview(buf)
    if bytes(mv[:len(magic)]) != magic:
        raise ValueError("not a %r message" % magic.strip().decode())
    end = bytes(mv[len(magic):len(magic) + (1 << 20)]).find(b"\n")
    if end < 0:
        raise ValueError("message header not terminated")
    head = json.loads(bytes(mv[len(magic):len(magic) + end]).decode("utf-8"))
    return head, mv[len(magic) + end + 1:]


def _take(mv, off, nbytes, dtype, shape):
    if off + nbytes > len(mv):
        raise ValueError("message truncated")
    a = np.frombuffer(mv[off:off + nbytes], dtype=dtype).reshape(shape)
    return np.array(a), off + nbytes


def encode_request(req: SolveRequest) -> bytes:
    t = np.ascontiguousarray(req.target, "<f4")
    if t.ndim != 3:
        raise ValueError("target must be [nz, ny, nx]")
    m = np.ascontiguousarray(req.mask, np.uint8)
    if m.shape != t.shape:
        raise ValueError("mask shape differs from target shape")
    unknown = set(req.params) - set(PARAM_NAMES)
    if unknown:
        raise ValueError("unknown solver parameters: %s" % sorted(unknown))
    if req.algorithm not in ALGORITHMS:
        raise ValueError("algorithm must be one of %s" % (ALGORITHMS,))
    f = None
    if req.field is not None:
        f = np.ascontiguousarray(req.field, "<f4")
        if f.shape != (3,) + t.shape:
            raise ValueError("field must be [3, nz, ny, nx]")
    head = {"id": str(req.request_id), "shape": list(t.shape), "spacing": [float(v) for v in req.spacing],
            "algorithm": req.algorithm, "params": {k: (bool(v) if k == "strict" else
                                                         int(v) if k == "it_max" else float(v))
                                                     for k, v in req.params.items()},
            "has_field": f is not None}
    out = [REQUEST_MAGIC, json.dumps(head).encode("utf-8"), b"\n", t.tobytes(), m.tobytes()]
    if f is not None:
        out.append(f.tobytes())
    return b"".join(out)


def decode_request(buf) -> SolveRequest:
    head, mv = _split(buf, REQUEST_MAGIC)
    shape = tuple(int(v) for v in head["shape"])
    if len(shape) != 3 or min(shape) < 1:
        raise ValueError("bad shape %r" % (shape,))
    n = int(np.prod(shape))
    t, off = _take(mv, 0, 4 * n, "<f4", shape)
    m, off = _take(mv, off, n, np.uint8, shape)
    f = None
    if head.get("has_field"):
        f, off = _take(mv, off, 12 * n, "<f4", (3,) + shape)
    params = dict(head.get("params", {}))
    check_request(shape, head.get("algorithm"), params)
    return SolveRequest(target=t.astype(np.float32), mask=m.astype(bool), spacing=tuple(head["spacing"]),
                        algorithm=head["algorithm"], params=params,
                        field=None if f is None else f.astype(np.float32), request_id=head["id"])


def encode_response(res: SolveResponse) -> bytes:
    f = np.ascontiguousarray(res.field, "<f4")
    head = {"id": res.request_id, "status": int(res.status), "shape": list(f.shape[1:]),
            "reason": int(res.reason), "reason_text": res.reason_text, "iterations": int(res.iterations),
            "evaluations": int(res.evaluations), "rejections": int(res.rejections),
            "final_error": float(res.final_error), "final_max_error": float(res.final_max_error),
            "seconds": float(res.seconds)}
    return b"".join([RESPONSE_MAGIC, json.dumps(head).encode("utf-8"), b"\n", f.tobytes()])


def decode_response(buf) -> SolveResponse:
    head, mv = _split(buf, RESPONSE_MAGIC)
    shape = tuple(int(v) for v in head["shape"])
    n = int(np.prod(shape))
    f, _ = _take(mv, 0, 12 * n, "<f4", (3,) + shape)
    return SolveResponse(request_id=head["id"], status=head["status"], field=f.astype(np.float32),
                         reason=head["reason"], reason_text=head["reason_text"], iterations=head["iterations"],
                         evaluations=head["evaluations"], rejections=head["rejections"],
                         final_error=head["final_error"], final_max_error=head["final_max_error"],
                         seconds=head["seconds"])


class SolveServer:
    """Requests in, responses out, over a continuous-batching engine.

    ``submit`` queues a request; ``step`` admits queued requests into free
    slots (the inputs are staged from pinned host buffers), advances the
    passes and returns the responses of the solves that finished.  Gradient,
    greedy and matching solves run in lanes of their own, each lane's passes
    covering all of its admitted volumes; a matching solve reads epsilon,
    delta, zeta and it_max and ignores the mask, as in the reference.
    """

    def __init__(self, hbm_budget=0, max_slots=0, **engine_kw):
        import torch  # pinned host staging
        from . import native
        self._torch = torch
        self._native = native
        self.engine = native.Engine(hbm_budget=hbm_budget, max_slots=max_slots, **engine_kw)
        self._queue: List[Tuple[SolveRequest, tuple]] = []
        self._live: Dict[int, Tuple[SolveRequest, tuple]] = {}
        self._failed: List[SolveResponse] = []

    def close(self):
        self.engine.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def pending(self):
        return len(self._queue) + len(self._live) + len(self._failed)

    def submit(self, req: SolveRequest):
        check_request(np.shape(req.target), req.algorithm, req.params)
        torch = self._torch
        t = torch.from_numpy(np.ascontiguousarray(req.target, np.float32)).pin_memory()
        m = torch.from_numpy(np.ascontiguousarray(req.mask, np.uint8)).pin_memory()
        nz, ny, nx = t.shape
        if req.field is None:
            f = torch.zeros((3, nz, ny, nx), dtype=torch.float32).pin_memory()
        else:
            f = torch.from_numpy(np.array(req.field, np.float32, copy=True)).pin_memory()
        self._queue.append((req, (t, m, f)))

    def _admit(self):
        while self._queue:
            req, (t, m, f) = self._queue[0]
            nz, ny, nx = t.shape
            try:
                ticket = self.engine.submit(req.algorithm, nx, ny, nz, req.spacing, t.data_ptr(),
                                            m.data_ptr(), f.data_ptr(), host_io=True, **req.params)
            except self._native.HipError as e:
                # the engine refused this request (e.g. larger than the pool):
                # answer it as failed and go on with the others
                self._queue.pop(0)
                self._failed.append(SolveResponse(request_id=req.request_id, status=-1, field=f.numpy().copy(),
                                                  reason=0, reason_text=str(e)))
                continue
            if ticket is None:
                return
            self._live[ticket] = self._queue.pop(0)

    def step(self, wait=True) -> List[SolveResponse]:
        self._admit()
        out, self._failed = self._failed, []
        if self._live:
            self.engine.progress(wait=wait)
        for ticket, status, info in self.engine.poll():
            req, (t, m, f) = self._live.pop(ticket)
            out.append(SolveResponse(request_id=req.request_id, status=status, field=f.numpy().copy(),
                                     reason=info["reason"], reason_text=self._native.REASONS[info["reason"]],
                                     iterations=info["iterations"], evaluations=info["evaluations"],
                                     rejections=info["rejections"], final_error=info["final_error"],
                                     final_max_error=info["final_max_error"], seconds=info["seconds"]))
        self._admit()
        out.extend(self._failed)
        self._failed = []
        return out

    def run(self, requests) -> List[SolveResponse]:
        """Solves all `requests` (closed batch); responses in completion order."""
        for r in requests:
            self.submit(r)
        out = []
        while self.pending():
            out.extend(self.step(wait=True))
        return out

    def handle(self, messages) -> List[bytes]:
        """Wire-level form of ``run``: encoded requests in, encoded responses out."""
        return [encode_response(r) for r in self.run([decode_request(b) for b in messages])]
