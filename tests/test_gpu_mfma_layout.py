"""Pin the FP64 MFMA lane layout the exchange kernels rely on:
A[i][k] from lane i+16k, B[k][j] from lane j+16k, D[lane/16 + 4e][lane%16]."""
import numpy as np
import pytest

from chem_amd import _lib

pytestmark = pytest.mark.gpu


def test_mfma_f64_16x16x4_layout():
    rng = np.random.default_rng(0)
    A = rng.integers(-8, 8, size=(16, 4)).astype(np.float64)
    B = rng.integers(-8, 8, size=(4, 16)).astype(np.float64)
    a_lane = np.array([A[l % 16, l // 16] for l in range(64)])
    b_lane = np.array([B[l // 16, l % 16] for l in range(64)])
    out = np.zeros(256)
    _lib.check(_lib.lib().mchip_probe_mfma_f64(_lib.dptr(a_lane), _lib.dptr(b_lane), _lib.dptr(out)))
    D = A @ B
    got = np.zeros((16, 16))
    for l in range(64):
        for e in range(4):
            got[l // 16 + 4 * e, l % 16] = out[l * 4 + e]
    if not np.array_equal(got, D):
        # report the layout actually observed
        found = {}
        for l in range(64):
            for e in range(4):
                hits = np.argwhere(D == out[l * 4 + e])
                found[(l, e)] = [tuple(h) for h in hits[:3]]
        print("observed layout (lane, elem) -> candidates:", {k: v for k, v in list(found.items())[:24]})
    assert np.array_equal(got, D)
