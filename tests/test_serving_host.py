"""Host logic of the serving loop that needs no GPU: the FIFO device ring allocator."""

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
