"""Continuous-batching server: every request's result is the same alone (plain kernel) and in the
mix (batched persistent kernel), including under HBM-budget back-pressure; latencies are reported."""

import math

import pytest
import torch

from mach_amd import serving

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def stream_reqs(device):
    return serving.make_stream(device, n_requests=24, qps=2000.0, seed=4, template_scale=0.25,
                               frames=(1, 2, 4, 8, 16, 32, 64))


@pytest.fixture(scope="module")
def burst_reqs(device):
    # every request arrives within a few microseconds: the server must batch (and fuse)
    return serving.make_stream(device, n_requests=24, qps=1e7, seed=4, template_scale=0.25,
                               frames=(1, 2, 4, 8, 16, 32, 64))


def _reset(reqs):
    for r in reqs:
        r.done_s = r.admit_s = float("nan")
        r.out = None


@pytest.mark.parametrize("fuse,direct,hold", [(True, None, None), (False, None, None), (True, 1 << 14, None),
                                               (True, 1 << 14, 1)])
def test_mix_equals_alone(device, burst_reqs, fuse, direct, hold):
    """Each result is bitwise the same alone and in the mix: with and without fusion, with large
    requests launched as their own plain kernels inside a batch (``direct`` voxel-frames threshold),
    and with requests held back for fusion partners (``hold`` points threshold; out-of-order
    completion and ring release)."""
    reqs = burst_reqs
    _reset(reqs)
    plain = serving.serve_plain(reqs, device, keep_outputs=True)
    alone = {r.rid: r.out for r in reqs}
    _reset(reqs)
    mixed = serving.serve_batched(reqs, device, budget_bytes=8 << 30, keep_outputs=True, fuse=fuse,
                                  plain_min_voxel_frames=direct, hold_min_points=hold)
    for r in reqs:
        assert torch.equal(r.out, alone[r.rid]), r.rid
    print(plain.as_dict(), mixed.as_dict(), flush=True)
    assert mixed.n_requests == plain.n_requests == len(reqs)
    assert mixed.n_batches < len(reqs)  # requests were actually batched
    assert (mixed.n_fused > 0) == fuse
    assert (mixed.n_direct > 0) == (direct is not None)
    assert mixed.n_direct < len(reqs)  # the small ones still went through the persistent kernel
    for rep in (plain, mixed):
        assert math.isfinite(rep.p50_ms) and rep.p99_ms >= rep.p50_ms > 0


def test_poisson_stream_equals_alone(device, stream_reqs):
    # arrivals spread over time: small batches, two in flight
    _reset(stream_reqs)
    serving.serve_plain(stream_reqs, device, keep_outputs=True)
    alone = {r.rid: r.out for r in stream_reqs}
    _reset(stream_reqs)
    rep = serving.serve_batched(stream_reqs, device, budget_bytes=8 << 30, keep_outputs=True)
    assert rep.n_requests == len(stream_reqs)
    for r in stream_reqs:
        assert torch.equal(r.out, alone[r.rid]), r.rid


def test_budget_backpressure_keeps_results(device, stream_reqs):
    _reset(stream_reqs)
    serving.serve_plain(stream_reqs, device, keep_outputs=True)
    alone = {r.rid: r.out for r in stream_reqs}
    biggest = max(r.data_bytes + r.out_bytes for r in stream_reqs)
    _reset(stream_reqs)
    # room for little more than the largest request: admission must wait for completions
    rep = serving.serve_batched(stream_reqs, device, budget_bytes=int(biggest * 2.2), keep_outputs=True)
    assert rep.n_requests == len(stream_reqs)
    for r in stream_reqs:
        assert torch.equal(r.out, alone[r.rid]), r.rid


def test_a_failing_request_does_not_stop_the_others(device, burst_reqs):
    """Injected failure: one request of the burst carries an invalid parameter (tukey_alpha = 2, which
    the library refuses before any launch). It completes as failed, with the library's message, in both
    paths; every other request's result is bitwise its result alone."""
    import dataclasses

    reqs = list(burst_reqs)
    bad = reqs[5]
    bad_geo = dataclasses.replace(bad.geometry, params=dict(bad.geometry.params, tukey_alpha=2.0))
    reqs[5] = dataclasses.replace(bad, geometry=bad_geo)
    _reset(reqs)
    plain = serving.serve_plain(reqs, device, keep_outputs=True)
    alone = {r.rid: r.out for r in reqs}
    assert plain.n_failed == 1 and "tukey_alpha" in reqs[5].error and reqs[5].out is None
    for r in reqs:
        r.error = None
    _reset(reqs)
    mixed = serving.serve_batched(reqs, device, budget_bytes=8 << 30, keep_outputs=True)
    assert mixed.n_failed == 1 and mixed.n_requests == len(reqs)
    assert "tukey_alpha" in reqs[5].error and reqs[5].out is None
    for i, r in enumerate(reqs):
        if i != 5:
            assert r.error is None and torch.equal(r.out, alone[r.rid]), r.rid


def test_a_request_over_the_budget_fails_alone(device, burst_reqs):
    """A request whose channel data or output is larger than its ring (three quarters of the HBM budget
    for data, a quarter for outputs) can never be admitted: it completes as failed, with a message, and
    every other request's result is bitwise its result alone."""
    reqs = list(burst_reqs)
    _reset(reqs)
    serving.serve_plain(reqs, device, keep_outputs=True)
    alone = {r.rid: r.out for r in reqs}
    big = max(reqs, key=lambda r: r.data_bytes)
    budget = int(big.data_bytes * 1.3)  # the data ring holds 0.975x the largest request's data
    data_cap = (budget * 3 // 4) // 256 * 256
    out_cap = (budget - budget * 3 // 4) // 256 * 256
    over = {r.rid for r in reqs if r.data_bytes > data_cap or r.out_bytes > out_cap}
    assert big.rid in over and len(over) < len(reqs)
    for r in reqs:
        r.error = None
    _reset(reqs)
    rep = serving.serve_batched(reqs, device, budget_bytes=budget, keep_outputs=True)
    assert rep.n_requests == len(reqs) and rep.n_failed == len(over)
    for r in reqs:
        if r.rid in over:
            assert r.out is None and "HBM budget" in r.error, r.rid
        else:
            assert r.error is None and torch.equal(r.out, alone[r.rid]), r.rid


def test_voxel_orders_outlive_the_order_cache(device, monkeypatch):
    """Batches run on the server's own streams and read each request's voxel order there; the
    library's order cache is bounded. With the cache cut to two entries (so orders are dropped from it
    while batches are in flight), every result must still be bitwise its result alone: the geometry
    holds its orders and the server keeps every launched request alive until its batch completes."""
    from mach_amd import _hip_impl

    reqs = serving.make_stream(device, n_requests=40, qps=1e6, seed=21, template_scale=0.5,
                               frames=(1, 2, 4, 8, 16, 32, 64))
    monkeypatch.setattr(_hip_impl, "_ORDER_CACHE_SIZE", 2)
    _hip_impl._ORDER_CACHE.clear()
    serving.serve_plain(reqs, device, keep_outputs=True)
    alone = {r.rid: r.out for r in reqs}
    _reset(reqs)
    rep = serving.serve_batched(reqs, device, budget_bytes=16 << 30, keep_outputs=True)
    assert rep.n_requests == len(reqs) and rep.n_failed == 0
    assert sum(1 for g in {id(r.geometry): r.geometry for r in reqs}.values() for o in g.orders.values()
               if o is not None) > 2  # more orders in use than the cache holds
    for r in reqs:
        assert torch.equal(r.out, alone[r.rid]), r.rid
