"""Host logic of the serving loop that needs no GPU: the device ring allocator, request fusion and
the batch selection that holds large requests for fusion partners."""

import random

import pytest

from mach_amd.serving import DeviceRing


def test_ring_never_overlaps_and_frees_in_order():
    ring = DeviceRing(256 * 100, "cpu")
    rng = random.Random(0)
    live = []
    for _ in range(20000):
        if live and rng.random() < 0.5:
            live.pop(0)
            ring.free_oldest()
        else:
            slot = ring.alloc(rng.randint(1, 30) * 256 - rng.randint(0, 255))
            if slot is None:
                assert live  # an empty ring always fits a request no larger than its capacity
                continue
            for a, n in live:
                assert slot[0] + slot[1] <= a or a + n <= slot[0]
            assert slot[0] % 256 == 0 and slot[0] + slot[1] <= ring.capacity
            live.append(slot)
        assert ring.used == sum(n for _, n in live)


def test_ring_undo_and_limits():
    ring = DeviceRing(1024, "cpu")
    a = ring.alloc(512)
    b = ring.alloc(256)
    assert a == (0, 512) and b == (512, 256)
    ring.undo_last()
    assert ring.alloc(256) == (512, 256)
    assert ring.alloc(512) is None  # only 256 bytes left at the end, 0 before the tail
    ring.free_oldest()
    assert ring.alloc(512) == (0, 512)  # wraps into the freed head
    with pytest.raises(ValueError):
        ring.alloc(4096)
    view = ring.view((0, 512), __import__("torch").float32, (4, 32))
    assert view.shape == (4, 32)


def test_zero_byte_reservations_hold_nothing():
    ring = DeviceRing(1024, "cpu")
    z = ring.alloc(0)
    assert z == (0, 0) and ring.used == 0 and not ring.live
    assert ring.alloc(1024) == (0, 1024)  # the empty slot does not make the ring look full
    ring.undo_last()
    assert ring.used == 0
    z = ring.alloc(0)
    ring.undo_last()  # undoing a zero-byte reservation releases nothing else
    a = ring.alloc(512)
    assert a == (0, 512)
    ring.free(z)  # zero-byte slots free as no-ops, in any order
    ring.free(a)
    assert ring.used == 0 and ring.alloc(1024) == (0, 1024)


def test_fusion_groups_by_geometry_and_frame_cap():
    import torch

    from mach_amd.serving import ServeRequest, _fusion_groups

    ga, gb = object(), object()

    def req(rid, geo, frames, dtype=torch.complex64):
        return ServeRequest(rid, 0.0, geo, torch.empty((1, 4, 8, frames), dtype=dtype), "x", 0)

    rs = [req(0, ga, 100), req(1, gb, 8), req(2, ga, 100), req(3, ga, 100), req(4, gb, 8),
          req(5, ga, 4, torch.float32)]
    groups = [[r.rid for r in g] for g in _fusion_groups(rs, True, 256)]
    assert groups == [[0, 2], [1, 4], [3], [5]]  # frame cap 256, dtype separates
    assert [[r.rid for r in g] for g in _fusion_groups(rs, False)] == [[i] for i in range(6)]


def test_ring_out_of_order_frees_never_overlap():
    """Releases in any order; space is reclaimed from the oldest end only."""
    ring = DeviceRing(256 * 64, "cpu")
    rng = random.Random(1)
    live = []  # slots not yet released
    for _ in range(20000):
        if live and rng.random() < 0.5:
            slot = live.pop(rng.randrange(len(live)))
            ring.free(slot)
        else:
            slot = ring.alloc(rng.randint(1, 12) * 256)
            if slot is None:
                assert live or ring.used > 0
                continue
            for a, n in live:
                assert slot[0] + slot[1] <= a or a + n <= slot[0]
            assert slot[0] + slot[1] <= ring.capacity
            live.append(slot)
        assert ring.used >= sum(n for _, n in live)
    for slot in list(live):
        ring.free(slot)
    assert ring.used == 0 and ring.alloc(ring.capacity) == (0, ring.capacity)
    with pytest.raises(KeyError):
        ring.free((0, ring.capacity + 256))


def test_select_holds_large_requests_for_partners():
    import torch

    from mach_amd.serving import ServeRequest, _select

    ga, gb, gc = object(), object(), object()

    def item(rid, geo, points, admit=0.0, frames=8):
        r = ServeRequest(rid, 0.0, geo, torch.empty((1, 4, 8, frames)), "x", points)
        r.admit_s = admit
        return (r, None)

    adm = [item(0, ga, 10), item(1, gb, 100), item(2, gc, 100), item(3, gb, 100), item(4, ga, 10)]
    go, held = _select(adm, gpu_busy=False, now=0.0, hold_min_points=50, max_hold_s=1.0, max_fused_frames=256)
    assert [it[0].rid for it in go] == [0, 1, 2, 3, 4] and held == []  # an idle GPU takes everything
    go, held = _select(adm, gpu_busy=True, now=0.1, hold_min_points=50, max_hold_s=1.0, max_fused_frames=256)
    assert [it[0].rid for it in go] == [0, 1, 3, 4]  # small ones and the largest large group (gb)
    assert [it[0].rid for it in held] == [2]
    go, held = _select(adm, gpu_busy=True, now=2.0, hold_min_points=50, max_hold_s=1.0, max_fused_frames=256)
    assert held == []  # waited long enough


def test_failure_isolation_resubmits_a_refused_batch_request_by_request():
    """A batch submission is all or nothing: when it is refused, each request is submitted alone and
    only the one that fails again is reported; direct launches fail one by one. (Injected failure on
    the host: the fake submit refuses any batch holding request 'bad'.)"""
    from mach_amd.serving import _submit_isolated

    calls = []

    def submit(rs):
        calls.append(list(rs))
        if "bad" in rs:
            raise RuntimeError("request 0: Error: tukey_alpha must be in range [0, 1], but got 2.000000")

    ran = []

    def run_direct(r):
        if r == "bad-direct":
            raise RuntimeError("HIP error in beamform kernel launch")
        ran.append(r)

    batched = [("a", ["A"]), ("bad", ["B1", "B2"]), ("c", ["C"])]
    direct = [("d", ["D"]), ("bad-direct", ["E"])]
    failed = _submit_isolated(submit, run_direct, batched, direct)
    assert calls[0] == ["a", "bad", "c"]  # the whole batch first
    assert calls[1:] == [["a"], ["bad"], ["c"]]  # then one by one
    assert [m for m, _ in failed] == [["B1", "B2"], ["E"]]
    assert "tukey_alpha" in failed[0][1] and "HIP error" in failed[1][1]
    assert ran == ["d"]
    # a batch that is accepted is submitted once
    calls.clear()
    assert _submit_isolated(submit, run_direct, [("a", ["A"]), ("c", ["C"])], []) == []
    assert calls == [["a", "c"]]


def test_report_counts_only_completed_requests():
    import torch

    from mach_amd.serving import ServeRequest, _report

    reqs = []
    for i, err in enumerate([None, "bad", None]):
        r = ServeRequest(i, 0.1 * i, None, torch.zeros(1), "plane_wave", 100 * (i + 1))
        r.done_s, r.error = 0.1 * i + 0.05 * (i + 1), err
        reqs.append(r)
    rep = _report("batched", reqs, 0.0)
    assert rep.n_failed == 1 and rep.n_requests == 3
    assert rep.points == 100 + 300
    assert abs(rep.p50_ms - 100.0) < 1e-9  # latencies 50 and 150 ms (the failed one is left out)


def test_fused_groups_reserve_their_buffers_in_the_rings():
    """A fusion group's concatenated channel data and output take ring space of their own (so the HBM
    budget covers them); without room, the members run separately and nothing stays reserved."""
    import types

    import torch

    from mach_amd.serving import DeviceRing, ServeRequest, _reserve_fusion

    geo = types.SimpleNamespace(scan_coords_m=torch.zeros((100, 3)))
    group = [ServeRequest(i, 0.0, geo, torch.zeros((1, 4, 8, f), dtype=torch.complex64), "x", 0)
             for i, f in enumerate((3, 5))]
    data_bytes = sum(r.data_bytes for r in group)  # 1*4*8*8 frames * 8 B
    data_ring, out_ring = DeviceRing(1 << 16, "cpu"), DeviceRing(1 << 16, "cpu")
    extra = []
    runs = _reserve_fusion(group, data_ring, out_ring, extra)
    assert len(runs) == 1 and runs[0][0] == group and runs[0][1] is not None
    assert [(ring is data_ring, slot[1]) for ring, slot in extra] == [(True, (data_bytes + 255) // 256 * 256),
                                                                      (False, (100 * 8 * 8 + 255) // 256 * 256)]
    assert data_ring.used >= data_bytes and out_ring.used >= 100 * 8 * 8
    # no room for the fused output: both members run alone, nothing is left reserved
    tiny_out = DeviceRing(256, "cpu")
    extra2 = []
    data2 = DeviceRing(1 << 16, "cpu")
    runs = _reserve_fusion(group, data2, tiny_out, extra2)
    assert [m for m, slots in runs] == [[group[0]], [group[1]]] and all(s is None for _, s in runs)
    assert extra2 == [] and data2.used == 0 and tiny_out.used == 0
    assert _reserve_fusion(group[:1], data2, tiny_out, extra2) == [(group[:1], None)]
