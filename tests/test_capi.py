"""C-ABI library checks that need no GPU: it loads, exports every function include/*.h declares, and
rejects invalid arguments with the reference's messages before touching the device."""

import ctypes
import glob
import os
import re
import subprocess

import pytest

from mach_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for hdr in glob.glob(os.path.join(ROOT, "include", "*.h")):
        with open(hdr) as fh:
            text = fh.read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b(mach_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_the_entry_points():
    names = declared_functions()
    for required in ("mach_beamform", "mach_beamform_c64", "mach_beamform_f32", "mach_last_error",
                     "mach_abi_version", "mach_hip_runtime_version", "mach_device_arch", "mach_to_db",
                     "mach_batch_create", "mach_batch_submit", "mach_batch_destroy"):
        assert required in names


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True, check=True)
    exported = {line.split()[-1] for line in out.stdout.splitlines() if line.strip()}
    for name in declared_functions():
        assert name in exported, name
        assert getattr(lib, name) is not None
        assert name in _lib.SIGNATURES, f"{name} has no ctypes signature in mach_amd/_lib.py"


def test_abi_and_version():
    lib = _lib.load()
    assert lib.mach_abi_version() == 2
    v = lib.mach_hip_runtime_version().decode()
    major, minor, _ = v.split(".")
    assert int(major) >= 6 and int(minor) >= 0


def _args(**over):
    a = _lib.BeamformArgs()
    a.n_tx, a.n_rx, a.n_samples, a.n_frames, a.n_vox = 1, 4, 100, 2, 8
    a.data_type, a.interp_type = _lib.DATA_C64, 1
    a.f_number, a.sampling_freq_hz, a.sound_speed_m_s, a.tukey_alpha = 1.0, 1e6, 1500.0, 0.5
    a.accumulate = 1
    for k, v in over.items():
        setattr(a, k, v)
    return a


@pytest.mark.parametrize("alpha", [-0.1, 1.5, float("nan")])
def test_tukey_alpha_range_is_rejected(alpha):
    lib = _lib.load()
    st = lib.mach_beamform(ctypes.byref(_args(tukey_alpha=alpha)), None)
    assert st == 1
    assert _lib.last_error().startswith("Error: tukey_alpha must be in range [0, 1], but got")


def test_bad_codes_and_null_pointers_are_rejected():
    lib = _lib.load()
    assert lib.mach_beamform(None, None) == 1
    assert lib.mach_beamform(ctypes.byref(_args(data_type=7)), None) == 5
    assert "unsupported data_type" in _lib.last_error()
    assert lib.mach_beamform(ctypes.byref(_args(interp_type=3)), None) == 1
    assert "unsupported interp_type" in _lib.last_error()
    assert lib.mach_beamform(ctypes.byref(_args()), None) == 1  # out is NULL
    assert "out is NULL" in _lib.last_error()
    assert lib.mach_beamform(ctypes.byref(_args(n_samples=1 << 40)), None) == 1
    with pytest.raises(RuntimeError, match="tukey_alpha"):
        _lib.check(lib.mach_beamform(ctypes.byref(_args(tukey_alpha=2.0)), None))


def test_empty_problem_is_a_no_op():
    lib = _lib.load()
    assert lib.mach_beamform(ctypes.byref(_args(n_vox=0)), None) == 0
    assert lib.mach_beamform(ctypes.byref(_args(n_frames=0)), None) == 0


@pytest.mark.parametrize("n_tx, n_rx", [(1, 0), (0, 4), (5, 0)])
def test_empty_sum_in_accumulate_mode_touches_nothing(n_tx, n_rx):
    """No transmits or no receive elements: validation lets the arrivals (and, with no elements, the
    coordinates and channel data) be NULL, so the call must not launch the kernel, which reads the
    arrivals of every voxel. In accumulate mode nothing is written either: MACH_OK with no device work
    (no GPU here, so any HIP call would fail with a status)."""
    lib = _lib.load()
    a = _args(n_tx=n_tx, n_rx=n_rx, accumulate=1)
    a.out = a.scan_coords_m = 256  # non-NULL placeholders, never dereferenced
    assert a.tx_wave_arrivals_s is None and a.tx_waves is None and a.channel_data is None
    assert lib.mach_beamform(ctypes.byref(a), None) == 0, _lib.last_error()
    assert _lib.last_error() == ""


def test_device_query_reports_status():
    lib = _lib.load()
    buf = ctypes.create_string_buffer(64)
    st = lib.mach_device_arch(buf, 64)
    assert st in (0, 4)
    if st == 0:
        assert buf.value.decode().startswith("gfx")
    else:
        assert "no HIP device" in _lib.last_error()


def test_batch_entry_points_report_errors_without_work():
    lib = _lib.load()
    h = ctypes.c_void_p()
    assert lib.mach_batch_create(0, 1, ctypes.byref(h)) == 1
    assert "max_requests" in _lib.last_error()
    assert lib.mach_batch_submit(None, None, 0, None) == 1
    assert lib.mach_batch_destroy(None) == 0
    st = lib.mach_batch_create(4, 1, ctypes.byref(h))
    if st == 0:  # a GPU is present
        assert lib.mach_batch_submit(h, None, 0, None) == 0
        assert lib.mach_batch_destroy(h) == 0
    else:
        assert st == 3 and "HIP error" in _lib.last_error()


def test_layout_hint_is_validated():
    lib = _lib.load()
    a = _args(layout=(3 << 4) | 1)  # 3 lanes per voxel is not a power of two
    a.out = 256  # non-NULL placeholder; validation fails before any device access
    a.channel_data = a.rx_coords_m = a.scan_coords_m = a.tx_wave_arrivals_s = 256
    assert lib.mach_beamform(ctypes.byref(a), None) == 1
    assert "layout" in _lib.last_error()


@pytest.mark.parametrize("layout, n_tx", [(0x9402, 2),   # lateral groups sum a single transmit
                                          (0xB402, 1),   # 8 voxels per lane is not compiled
                                          (0x9401, 1),   # 1 frame per lane is not compiled
                                          (0x9202, 1)])  # lateral groups take 64 lanes
def test_lateral_layout_codes_are_validated(layout, n_tx):
    lib = _lib.load()
    a = _args(layout=layout, n_tx=n_tx)
    a.out = 256  # non-NULL placeholder; validation fails before any device access
    a.channel_data = a.rx_coords_m = a.scan_coords_m = a.tx_wave_arrivals_s = 256
    assert lib.mach_beamform(ctypes.byref(a), None) == 1
    assert "lateral" in _lib.last_error()


def test_output_kind_and_peak_are_validated():
    lib = _lib.load()
    assert lib.mach_beamform(ctypes.byref(_args(output=7)), None) == 1
    assert "unsupported output kind" in _lib.last_error()
    assert lib.mach_beamform(ctypes.byref(_args(output=2, accumulate=1)), None) == 1  # envelope
    assert "envelope" in _lib.last_error()
    assert lib.mach_beamform(ctypes.byref(_args(output=0, peak=256)), None) == 1
    assert "peak" in _lib.last_error()
    # the power output needs `out` even without frames (a sum over nothing is 0)
    assert lib.mach_beamform(ctypes.byref(_args(output=1, accumulate=0, n_frames=0)), None) == 1
    assert "out is NULL" in _lib.last_error()


def test_to_db_validates_pointers():
    lib = _lib.load()
    assert lib.mach_to_db(None, 0, None, 0, None) == 0  # nothing to convert
    assert lib.mach_to_db(None, 4, None, 1, None) == 1
    assert "NULL" in _lib.last_error()


def test_arrivals_and_waves_are_exclusive():
    lib = _lib.load()
    a = _args(accumulate=0)
    a.out = a.channel_data = a.rx_coords_m = a.scan_coords_m = a.tx_wave_arrivals_s = 256  # never dereferenced
    a.tx_waves = 512
    assert lib.mach_beamform(ctypes.byref(a), None) == 1
    assert "not both" in _lib.last_error()
    assert lib.mach_tx_arrivals(None, 0, None, 5, None, None) == 0  # nothing to compute
    assert lib.mach_tx_arrivals(None, 2, None, 5, None, None) == 1


def test_noalloc_batch_rejects_before_device_work():
    """mach_batch_submit_noalloc: the size limit, NULL requests and invalid requests are refused before
    anything touches the device; an empty or all-empty submission is a no-op."""
    lib = _lib.load()
    arr = (_lib.BeamformArgs * 65)()
    assert lib.mach_batch_submit_noalloc(arr, 65, None) == 1
    assert "allocation-free maximum of 64" in _lib.last_error()
    assert lib.mach_batch_submit_noalloc(None, 0, None) == 0
    assert lib.mach_batch_submit_noalloc(None, 2, None) == 1
    assert "requests is NULL" in _lib.last_error()
    bad = (_lib.BeamformArgs * 2)(_args(n_vox=0), _args(tukey_alpha=3.0))
    assert lib.mach_batch_submit_noalloc(bad, 2, None) == 1
    assert _lib.last_error().startswith("request 1: Error: tukey_alpha")
    empty = (_lib.BeamformArgs * 2)(_args(n_vox=0), _args(n_rx=0, accumulate=1, out=256, scan_coords_m=256))
    assert lib.mach_batch_submit_noalloc(empty, 2, None) == 0, _lib.last_error()


def test_debug_library_exports_the_same_entry_points():
    """libmach_hip_debug.so (the bounds-checked build, MACH_AMD_DEBUG=1) is a drop-in for the release
    library: same exported functions."""
    if not os.path.exists(_lib.DEBUG_PATH):
        pytest.skip("debug build not present (python mach_amd/build.py --debug)")

    def exported(path):
        out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True)
        return {line.split()[-1] for line in out.stdout.splitlines() if line.split()[-1].startswith("mach_")}

    assert exported(_lib.DEBUG_PATH) == exported(_lib.RELEASE_PATH)
    assert set(declared_functions()) <= exported(_lib.DEBUG_PATH)


def test_timing_is_off_by_default():
    lib = _lib.load()
    k, c = ctypes.c_float(), ctypes.c_float()
    if os.environ.get("MACH_AMD_PROFILE", "0") in ("", "0"):
        assert lib.mach_last_timing(ctypes.byref(k), ctypes.byref(c)) == _lib.ERR_UNSUPPORTED
        assert k.value == -1.0 and c.value == -1.0
