"""Decoding service: a request's result is the same whether it is decoded
alone or in a mix with other requests (bit-identical stitched ln g, ln Z,
decision and round counts), the HBM arenas are returned when requests
finish, and admission waits for memory instead of failing."""

from __future__ import annotations

import numpy as np
import pytest

from qec_pf_wang_landau_amd import serve

pytestmark = pytest.mark.gpu

SIZES = {4: (4, 4), 6: (10, 8)}


def requests():
    out = serve.poisson_requests(rate=40.0, duration=0.2, seed=3, sizes=(4, 6), weights=(0.5, 0.5))
    assert len(out) >= 4
    return out[:6]


def by_rid(results):
    return {r.rid: r for r in results}


def same(a: serve.Result, b: serve.Result):
    assert a.status == b.status == 1
    assert np.array_equal(a.log_z, b.log_z)
    assert a.decision == b.decision
    assert np.array_equal(a.rounds, b.rounds)
    for x, y in zip(a.lng, b.lng):
        assert np.array_equal(x, y)
    for x, y in zip(a.E, b.E):
        assert np.array_equal(x, y)


def test_result_alone_equals_result_in_mix(cuda):
    mix_srv = serve.Server(hbm_bytes=64 << 20, slice_rounds=50, sizes=SIZES)
    mix = by_rid(mix_srv.serve(requests()))
    assert len(mix) == len(requests())
    assert mix_srv.slices > 1
    for req in requests():
        req.arrival = 0.0
        alone_srv = serve.Server(hbm_bytes=64 << 20, slice_rounds=77, sizes=SIZES)
        alone = by_rid(alone_srv.serve([req]))
        same(alone[req.rid], mix[req.rid])
    # every arena is empty again
    for a in mix_srv.arena.values():
        assert a.used == 0


def test_admission_waits_for_memory(cuda):
    """A budget that holds one 6x6 request at a time still decodes all of
    them, one after another, with the same results."""
    reqs = [serve.Request(10, 6, 0.05, seed=901), serve.Request(11, 6, 0.08, seed=902)]
    small = serve.Server(hbm_bytes=1, slice_rounds=100, sizes={6: SIZES[6]})
    got = by_rid(small.serve([serve.Request(r.rid, r.L, r.p, r.seed, 0.0) for r in reqs]))
    roomy = serve.Server(hbm_bytes=64 << 20, slice_rounds=100, sizes={6: SIZES[6]})
    ref = by_rid(roomy.serve([serve.Request(r.rid, r.L, r.p, r.seed, 0.0) for r in reqs]))
    for r in reqs:
        same(got[r.rid], ref[r.rid])
    # one at a time: the second was admitted after the first finished
    a, b = sorted(got.values(), key=lambda x: x.admitted)
    assert b.admitted >= a.done


def test_decisions_prefer_the_error_class_at_low_p(cuda):
    srv = serve.Server(hbm_bytes=64 << 20, slice_rounds=200, sizes={6: SIZES[6]})
    reqs = [serve.Request(i, 6, 0.03, seed=500 + i) for i in range(4)]
    res = srv.serve(reqs)
    assert all(r.status == 1 for r in res)
    assert np.mean([r.decision == 0 for r in res]) >= 0.75
    s = serve.latency_summary(res)
    assert s["decoded"] == 4 and s["p99_s"] >= s["p50_s"] > 0


class _FailingSetup(serve.Server):
    """A server whose set-up of one request raises (forced from the host)."""

    def __init__(self, bad_rid, **kw):
        super().__init__(**kw)
        self.bad_rid = bad_rid

    def _setup(self, req):
        if req.rid == self.bad_rid:
            raise RuntimeError("set-up refused for this test")
        return super()._setup(req)


def test_failed_setup_is_reported_and_the_service_continues(cuda):
    """A request whose set-up fails comes back with status -1 and the error,
    its memory is released, and the other requests decode exactly as on a
    server without the failure."""
    reqs = [serve.Request(20, 6, 0.05, seed=801), serve.Request(21, 6, 0.05, seed=802),
            serve.Request(22, 4, 0.05, seed=803)]
    srv = _FailingSetup(21, hbm_bytes=64 << 20, slice_rounds=100, sizes=SIZES)
    got = by_rid(srv.serve([serve.Request(r.rid, r.L, r.p, r.seed, 0.0) for r in reqs]))
    assert got[21].status == -1 and "set-up refused" in got[21].error and got[21].decision == -1
    ref = by_rid(serve.Server(hbm_bytes=64 << 20, slice_rounds=100, sizes=SIZES).serve(
        [serve.Request(r.rid, r.L, r.p, r.seed, 0.0) for r in reqs if r.rid != 21]))
    for rid in (20, 22):
        same(got[rid], ref[rid])
    for a in srv.arena.values():
        assert a.used == 0
    assert serve.latency_summary(list(got.values()))["failed"] == 1
