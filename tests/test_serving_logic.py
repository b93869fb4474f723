"""Batch formation of the request loop (host logic, no GPU): arrival order, one kind and stiffness
bucket per batch, column capacity, and that requests left out stay pending in order."""
import math

import numpy as np

from gradientdomaintextureprocessing_amd.serving import Request, ResidentMesh, SolverServer


class _Mesh:
    max_cols = 6
    heat_time = 1e-3
    buckets = [1e-4, 1e-3, 1e-2]
    bucket = ResidentMesh.bucket


def _filter(cols, a):
    return Request("m", "filter", np.zeros((cols, 4)), a)


def _geo(n):
    return Request("m", "geodesic", list(range(n)), 1e-3)


def _lic(cols):
    return Request("m", "lic", np.zeros((cols, 4)), 0.0)


def test_bucket_is_nearest_in_log_scale():
    m = _Mesh()
    assert m.bucket(3e-4) == 1e-4 and m.bucket(3.3e-4) == 1e-3 and m.bucket(0.5) == 1e-2 and m.bucket(1e-6) == 1e-4
    assert all(abs(math.log10(m.bucket(a)) - math.log10(a)) <= 0.5 + 1e-12 for a in np.logspace(-4, -2, 50))


def test_batches_follow_arrival_order_kind_bucket_and_capacity():
    m = _Mesh()
    p = [_filter(3, 1e-3), _filter(1, 2e-3), _geo(2), _filter(3, 9e-3), _filter(3, 8e-4), _filter(1, 1.1e-3)]
    batch, rest = SolverServer._form_batch(m, p)
    # kind and bucket of the first request (filter, 1e-3); p[2] is a geodesic, p[3] is in bucket
    # 1e-2, p[4] (3 columns) no longer fits after 3 + 1, p[5] (1 column) does
    assert batch == [p[0], p[1], p[5]]
    cols = sum(r.columns for r in batch)
    assert cols <= m.max_cols
    assert all(r.kind == "filter" and m.bucket(r.weight) == 1e-3 for r in batch)
    assert [r for r in p if r not in batch] == rest  # the rest keeps arrival order
    # the next batch starts with the oldest pending request
    batch2, rest2 = SolverServer._form_batch(m, rest)
    assert batch2[0] is rest[0] and batch2[0].kind == "geodesic"


def test_every_request_is_eventually_batched():
    m = _Mesh()
    rng = np.random.default_rng(0)
    pending = [(_geo(int(rng.integers(1, 7))) if rng.random() < 0.4 else _filter(int(rng.integers(1, 4)), float(10 ** rng.uniform(-4, -2))))
               for _ in range(60)]
    seen = []
    while pending:
        b, pending = SolverServer._form_batch(m, pending)
        assert b and sum(r.columns for r in b) <= m.max_cols
        seen += b
    assert len(seen) == 60 and len(set(map(id, seen))) == 60


def test_lic_requests_batch_with_their_own_kind():
    m = _Mesh()
    p = [_lic(2), _filter(1, 1e-3), _lic(3), _geo(1), _lic(2)]
    batch, rest = SolverServer._form_batch(m, p)
    assert batch == [p[0], p[2]] and rest == [p[1], p[3], p[4]]  # 2 + 3 columns; the third LIC (2) no longer fits in 6
    assert [r.columns for r in batch] == [2, 3]
