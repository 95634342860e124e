"""PyMUST data helpers (reference tests/io/test_must.py): parameter extraction from a .mat file, the
local-file loader's integrity check, and the probe / grid helpers."""

import numpy as np
import pytest
import scipy.io

from mach_amd.io import must


def _write_mat(path):
    param = {"Nelements": 128.0, "pitch": 2.98e-4, "c": 1480.0, "fs": 20.832e6, "fc": 5.208e6, "t0": 0.0,
             "TXdelay": np.zeros((1, 128))}
    scipy.io.savemat(path, {"param": param, "RF": np.zeros((64, 128, 2), dtype=np.int16)})


def test_extract_pymust_params(tmp_path):
    path = str(tmp_path / "disk.mat")
    _write_mat(path)
    data = must.load_pymust_doppler_data(path, verify=False)
    assert "param" in data and "RF" in data and data["RF"].shape == (64, 128, 2)
    params = must.extract_pymust_params(data)
    assert params["Nelements"] == 128 and params["c"] == 1480.0 and params["fc"] == 5.208e6
    assert params["TXdelay"].shape == (1, 128)  # multi-element fields stay arrays
    assert params["fnumber"] == 1.0  # default when the file has none


def test_loader_checks_the_pinned_file_and_never_downloads(tmp_path):
    path = str(tmp_path / "disk.mat")
    _write_mat(path)
    with pytest.raises(ValueError, match="not the pinned"):
        must.load_pymust_doppler_data(path)
    with pytest.raises(FileNotFoundError):
        must.load_pymust_doppler_data(str(tmp_path / "missing.mat"))


def test_probe_and_grid_helpers_match_the_reference():
    pos = must.linear_probe_positions(10, 0.1)
    assert pos.shape == (10, 3) and pos.dtype == np.float32
    np.testing.assert_allclose(pos.mean(axis=0), np.zeros(3), atol=1e-7)
    grid = must.scan_grid(np.linspace(0, 1, 3), np.linspace(0, 1, 4), np.linspace(0, 1, 5))
    assert grid.shape == (60, 3)
