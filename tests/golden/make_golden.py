"""Writes tests/golden/das_small.npz: small delay-and-sum cases with their FP64 oracle outputs.

The reference commits no numeric vectors for its kernel (its tests assert shapes only), so these
fixtures are produced by the pinned oracle (oracle/das64.py, cross-checked against oracle/das_c.py
here) from seeded inputs. They freeze the oracle's outputs so that a later change to the oracle or the
kernel that alters results is caught. Regenerate with ``python tests/golden/make_golden.py``.
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle.das64 import das64  # noqa: E402
from oracle.das_c import das_c  # noqa: E402


def cases():
    rng = np.random.default_rng(20261203)
    for interp in (0, 1, 2):
        for kind in ("rf", "iq", "iq16"):
            for alpha in (0.0, 0.5):
                T, E, S, F, V = 2, 5, 32, 2, 8
                D = rng.normal(size=(T, E, S, F)).astype(np.float32)
                if kind != "rf":
                    D = (D + 1j * rng.normal(size=D.shape)).astype(np.complex64)
                if kind == "iq16":
                    D = np.stack([D.real, D.imag], -1).astype(np.float16)
                rx = (rng.normal(size=(E, 3)) * 1e-3).astype(np.float32)
                vox = (rng.normal(size=(V, 3)) * 1e-3 + [0, 0, 2e-3]).astype(np.float32)
                arr = (rng.normal(size=(T, V)) * 1e-6 + 0.5e-6).astype(np.float32)
                params = np.array([0.8, -1e-7, 20e6, 1540.0, 5e6 if kind != "rf" else 0.0, alpha, interp])
                yield f"{kind}_i{interp}_a{int(alpha * 10)}", D, rx, vox, arr, params


def main():
    out = {}
    for name, D, rx, vox, arr, p in cases():
        k = dict(f_number=p[0], rx_start_s=p[1], sampling_freq_hz=p[2], sound_speed_m_s=p[3],
                 modulation_freq_hz=p[4], tukey_alpha=p[5], interp_type=int(p[6]))
        y = das64(D, rx, vox, arr, **k)
        np.testing.assert_allclose(das_c(D, rx, vox, arr, **k), y, rtol=1e-12, atol=1e-12)
        for key, val in (("D", D), ("rx", rx), ("vox", vox), ("arr", arr), ("params", p), ("y", y)):
            out[f"{name}__{key}"] = val
    path = os.path.join(HERE, "das_small.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes,", len(out) // 6, "cases")


if __name__ == "__main__":
    main()
