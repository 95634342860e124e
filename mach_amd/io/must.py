"""PyMUST data helpers (reference: ``src/mach/io/must.py``).

The reference downloads PyMUST's ``PWI_disk.mat`` rotating-disk recording -- the data of its published
benchmark -- into a cache (``must.py:12-28``). There is no network here: ``load_pymust_doppler_data``
reads a copy the caller already has, after checking it against the size and SHA-256 the reference
pins. ``extract_pymust_params`` (``must.py:31-68``), ``linear_probe_positions`` and ``scan_grid``
(``must.py:71-104``) behave as the reference's.
"""

from __future__ import annotations

import hashlib
import os

from ..probes import linear_probe_positions, scan_grid

__all__ = ["PWI_DISK_SIZE", "PWI_DISK_SHA256", "load_pymust_doppler_data", "extract_pymust_params",
           "linear_probe_positions", "scan_grid"]

# The published file the reference pins (must.py:21-27).
PWI_DISK_SIZE = 2_085_623
PWI_DISK_SHA256 = "c349dc1d677c561434fd0e4a74142c4a0a44b7e6ae0a42d446c078a528ef58c1"


def load_pymust_doppler_data(path: str, verify: bool = True) -> dict:
    """Load a local copy of PyMUST's ``PWI_disk.mat`` (RF data and acquisition parameters).

    Args:
        path: the .mat file.
        verify: check the file's size and SHA-256 against the ones the reference pins.

    Raises:
        FileNotFoundError: if ``path`` does not exist (nothing is downloaded).
        ValueError: if ``verify`` and the file is not the pinned one.
    """
    import scipy.io

    if not os.path.exists(path):
        raise FileNotFoundError(f"{path}: PyMUST's PWI_disk.mat is not available locally (this package does not "
                                "download data)")
    if verify:
        size = os.path.getsize(path)
        h = hashlib.sha256()
        with open(path, "rb") as fh:
            for chunk in iter(lambda: fh.read(1 << 20), b""):
                h.update(chunk)
        if size != PWI_DISK_SIZE or h.hexdigest() != PWI_DISK_SHA256:
            raise ValueError(f"{path} is not the pinned PWI_disk.mat (size {size}, sha256 {h.hexdigest()})")
    return scipy.io.loadmat(path, struct_as_record=False)


def extract_pymust_params(mat_data: dict) -> dict:
    """Acquisition parameters from the ``param`` struct of a PyMUST .mat file: fields of one element
    become scalars, others stay arrays, and ``fnumber`` defaults to 1.0 when absent (``must.py:31-68``)."""
    mat_param = mat_data["param"][0][0]
    param = {}
    for k, c in mat_param.__dict__.items():
        if k == "_fieldnames":
            continue
        if c.size == 1:
            c = c[0][0]
        param[k] = c
    if "fnumber" not in param:
        param["fnumber"] = 1.0
    return param
