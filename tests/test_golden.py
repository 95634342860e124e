"""The oracle still reproduces the committed golden vectors (tests/golden/das_small.npz)."""

import os

import numpy as np
import pytest

from oracle.das64 import das64
from oracle.das_c import das_c

PATH = os.path.join(os.path.dirname(__file__), "golden", "das_small.npz")


def load_cases():
    z = np.load(PATH, allow_pickle=False)
    names = sorted({k.split("__")[0] for k in z.files})
    return [(n, {k.split("__")[1]: z[k] for k in z.files if k.startswith(n + "__")}) for n in names]


CASES = load_cases()


def params_kw(p):
    return dict(f_number=p[0], rx_start_s=p[1], sampling_freq_hz=p[2], sound_speed_m_s=p[3],
                modulation_freq_hz=p[4], tukey_alpha=p[5], interp_type=int(p[6]))


@pytest.mark.parametrize("name,case", CASES, ids=[c[0] for c in CASES])
def test_oracle_reproduces_golden(name, case):
    k = params_kw(case["params"])
    y = case["y"]
    assert np.abs(y).max() > 0
    np.testing.assert_allclose(das64(case["D"], case["rx"], case["vox"], case["arr"], **k), y, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(das_c(case["D"], case["rx"], case["vox"], case["arr"], **k), y, rtol=1e-12, atol=1e-12)
