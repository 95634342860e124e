"""Autotuner host logic (no GPU): candidate layouts, encoding, shape keys and the JSON table."""

import json

from mach_amd import autotune


def test_candidates_divide_frames_and_encode():
    for frames in (1, 2, 3, 8, 12, 33, 64, 256):
        cands = autotune.candidates(frames)
        assert cands, frames
        for lay in cands:
            lf, k = autotune.decode(lay)
            ch = autotune.chunks_per_lane(lay)
            assert (lf in autotune.LANES or lf == 0) and k in autotune.FRAMES_PER_LANE
            assert frames % k == 0 and (lf != 0 or frames % (64 * k) == 0)
            assert ch == 1 or ((lf, k) in autotune.MULTI_GROUP and lf * k * ch <= frames)
            assert autotune.encode(lf, k, ch) == lay
    assert 0x404 in autotune.candidates(256)  # 64 lanes x 4 frames
    assert 0x2104 in autotune.candidates(256)  # 16 lanes x 4 frames x 4 groups = 256 frames
    assert 0x1104 in autotune.candidates(128) and 0x2104 not in autotune.candidates(128)
    assert 0x2 in autotune.candidates(256) and 0x2 not in autotune.candidates(8)  # band kernel


def test_lateral_layout_codes():
    """Lateral groups: 0x8000 | log2(voxels per lane) << 12 | 64 << 4 | frames per lane, offered only
    for single-transmit requests whose frames keep at least half of the 64 lanes busy."""
    assert autotune.encode_lateral(2, 4) == 0x9404 and autotune.encode_lateral(4, 2) == 0xA402
    for lay in (0x9404, 0xA402):
        assert autotune.decode(lay)[0] == 64 and autotune.chunks_per_lane(lay) == 1
    assert autotune.voxels_per_lane(0x9404) == 2 and autotune.voxels_per_lane(0xA404) == 4
    assert autotune.voxels_per_lane(0x2104) == 1
    lat256 = [lay for lay in autotune.candidates(256, lateral=True) if lay & autotune.LATERAL_FLAG]
    assert sorted(lat256) == [0x9402, 0x9404, 0xA402, 0xA404]
    assert not any(lay & autotune.LATERAL_FLAG for lay in autotune.candidates(256))
    lat64 = [lay for lay in autotune.candidates(64, lateral=True) if lay & autotune.LATERAL_FLAG]
    assert sorted(lat64) == [0x9402, 0xA402]  # 64 x 4 frames would leave 48 of 64 lanes idle
    assert not any(lay & autotune.LATERAL_FLAG for lay in autotune.candidates(32, lateral=True))


def test_shape_key_buckets_by_octave():
    a = autotune.shape_key(1, 2, 1, 128, 256, 65536)
    b = autotune.shape_key(1, 2, 1, 100, 256, 40000)
    c = autotune.shape_key(1, 2, 1, 128, 128, 65536)
    assert a == b and a != c


def test_table_roundtrip_and_alignment_guard(tmp_path):
    path = str(tmp_path / "t.json")
    t = autotune.TuningTable(path)
    key = autotune.shape_key(1, 1, 1, 128, 8, 4096)
    t.record(key, autotune.encode(2, 4), {0x24: 0.5, 0x11: 0.7}, {"workload": "x"})
    t.save()
    with open(path) as fh:
        data = json.load(fh)
    assert data["entries"][key]["layout"] == 0x24
    t2 = autotune.TuningTable(path)
    assert t2.lookup(key) == 0x24 and t2.lookup("missing") == 0
    old = autotune._TABLE
    try:
        autotune._TABLE = t2
        assert autotune.layout_for(1, 1, 1, 128, 8, 4096, data_ptr=256, out_ptr=256) == 0x24
        assert autotune.layout_for(1, 1, 1, 128, 8, 4096, data_ptr=8, out_ptr=256) == 0  # misaligned for 4 frames
    finally:
        autotune._TABLE = old


def test_tune_alternates_passes_keeps_fastest_and_skips_rejected(monkeypatch):
    cands = autotune.candidates(8)
    rejected = cands[1]
    calls = []
    # each layout's scripted times per pass; the first candidate is slow on its first (cold) pass only
    script = {lay: [1.0 + 0.01 * i] * 3 for i, lay in enumerate(cands)}
    script[cands[0]] = [9.0, 0.5, 0.6]

    def fake_time(run, lay, reps):
        calls.append(lay)
        if lay == rejected:
            raise RuntimeError("layout rejected")
        return script[lay][sum(1 for c in calls if c == lay) - 1]

    monkeypatch.setattr(autotune, "time_layout", fake_time)
    best, times = autotune.tune(lambda lay: None, 8, reps=1, band=False, rounds=3)
    assert best == cands[0] and times[cands[0]] == 0.5
    assert rejected not in times and calls.count(rejected) == 1
    n = len(cands)
    assert calls[:n] == cands  # pass 1 forwards
    assert calls[n:2 * n - 1] == [c for c in cands[::-1] if c != rejected]  # pass 2 backwards


def test_shipped_table_entries_are_candidates_of_their_shape():
    """Every layout in the shipped gfx950 table is one the tuner could have chosen for its key: a
    candidate for the key's frame count, with band layouts only for single-transmit signal requests
    and lateral layouts only for single-transmit requests."""
    import re

    with open(autotune.DEFAULT_TABLE) as fh:
        data = json.load(fh)
    assert data["arch"] == "gfx950" and data["entries"]
    for key, entry in data["entries"].items():
        m = re.fullmatch(r"d(\d+):i(\d):t(\d+):e(\d+):f(\d+):v(\d+)(?::o(\d))?", key)
        assert m, key
        n_tx, n_frames, output = int(m.group(3)), int(m.group(5)), int(m.group(7) or 0)
        single = n_tx == 1
        cands = autotune.candidates(n_frames, band=single and output == 0, lateral=single)
        lay = int(entry["layout"])
        assert lay & 0xFFFF in cands, (key, hex(lay))
        assert autotune.xcd_code(lay) in autotune.XCD_CODES  # the block order the tuner fixed, if any
        assert entry["layout_hex"] == hex(int(entry["layout"]))


def test_nearest_class_fallback():
    """A shape class without an entry takes the tuned layout of the nearest class: same storage,
    interpolation, transmit/element octaves, frame count and output; nearest voxel octave within three."""
    t = autotune.TuningTable(path="/nonexistent/table.json")
    key = autotune.shape_key
    t.record(key(2, 1, 5, 1024, 8, 2097152), 0x1014)
    t.record(key(2, 1, 5, 1024, 8, 65536), 0x24)
    t.record(key(2, 1, 5, 1024, 16, 2097152), 0x2014)
    t.record(key(2, 1, 5, 1024, 8, 262144, output=1), 0x44)
    assert t.lookup(key(2, 1, 5, 1024, 8, 2097152)) == 0x1014          # exact
    assert t.lookup(key(2, 1, 5, 1024, 8, 884736)) == 0x1014           # a voxel octave away
    assert t.lookup(key(2, 1, 5, 1024, 8, 524288)) == 0x1014           # 2 octaves from 2^21, 3 from 2^16
    assert t.lookup(key(2, 1, 5, 1024, 8, 262144)) == 0x24             # 2 octaves from 2^16, 3 from 2^21
    t.record(key(2, 1, 5, 1024, 8, 1048576), 0x44)
    assert t.lookup(key(2, 1, 5, 1024, 8, 262144)) == 0x44             # 2 octaves from 2^16 and 2^20: the larger
    assert t.lookup(key(2, 1, 5, 1024, 24, 2097152)) == 0              # other frame counts stay apart
    assert t.lookup(key(1, 1, 5, 1024, 8, 2097152)) == 0               # other storage
    assert t.lookup(key(2, 1, 5, 1024, 8, 2097152, output=1)) == 0x44  # output kinds stay apart
    assert t.lookup(key(2, 1, 5, 1024, 8, 1024)) == 0                  # too many octaves away
    assert t.lookup(key(2, 1, 75, 1024, 8, 2097152)) == 0              # other transmit octave


def test_xcd_bits_of_layout_codes():
    """Bits 16-19 of a layout code fix the launch's block order over the XCDs and change nothing else."""
    lay = autotune.with_xcd(0x1014, 2)
    assert lay == 0x21014 and autotune.xcd_code(lay) == 2 and autotune.with_xcd(lay, 0) == 0x1014
    assert autotune.decode(lay) == autotune.decode(0x1014) and autotune.chunks_per_lane(lay) == 2
    from mach_amd._hip_impl import wave_tile

    assert wave_tile(lay) == wave_tile(0x1014) == 64
    assert wave_tile(autotune.with_xcd(0, 3)) is None  # XCD bits alone: still the automatic layout
