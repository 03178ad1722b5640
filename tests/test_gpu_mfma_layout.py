"""Pin the FP64 MFMA lane layout the exchange kernels rely on:
A[i][k] from lane i+16k, B[k][j] from lane j+16k, D[lane/16 + 4e][lane%16]."""
import numpy as np
import pytest

from chem_amd import _lib

pytestmark = pytest.mark.gpu


class _Probe:
    """The test-only probe library tests/native/libmchip_probe.so (built with
    the package); the probes are not part of libmchip.so."""

    _h = None

    def __getattr__(self, name):
        import ctypes
        import os
        if _Probe._h is None:
            path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "native", "libmchip_probe.so")
            if not os.path.exists(path):
                raise _lib.MchipError("%s not built; run __graft_entry__.build()" % path)
            h = ctypes.CDLL(path)
            h.mchip_probe_last_error.restype = ctypes.c_char_p
            _Probe._h = h
        return getattr(_Probe._h, name)


_probe = _Probe()


def _check(rc):
    if rc != 0:
        raise _lib.MchipError(_probe.mchip_probe_last_error().decode())


def test_mfma_f64_16x16x4_layout():
    rng = np.random.default_rng(0)
    A = rng.integers(-8, 8, size=(16, 4)).astype(np.float64)
    B = rng.integers(-8, 8, size=(4, 16)).astype(np.float64)
    a_lane = np.array([A[l % 16, l // 16] for l in range(64)])
    b_lane = np.array([B[l // 16, l % 16] for l in range(64)])
    out = np.zeros(256)
    _check(_probe.mchip_probe_mfma_f64(_lib.dptr(a_lane), _lib.dptr(b_lane), _lib.dptr(out)))
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


def _observe(kind, per):
    rng = np.random.default_rng(1 + kind)
    K = 4 * per
    A = rng.integers(-4, 4, size=(16, K)).astype(np.float32)
    B = rng.integers(-4, 4, size=(K, 16)).astype(np.float32)
    # hypothesis: lane l supplies A[l%16][per*(l//16)+q], B[per*(l//16)+q][l%16]
    a_lane = np.zeros(64 * per, np.float32)
    b_lane = np.zeros(64 * per, np.float32)
    for l in range(64):
        for q in range(per):
            a_lane[l * per + q] = A[l % 16, per * (l // 16) + q]
            b_lane[l * per + q] = B[per * (l // 16) + q, l % 16]
    out = np.zeros(256, np.float32)
    _check(_probe.mchip_probe_mfma_lp(kind, _lib.fptr(a_lane), _lib.fptr(b_lane), _lib.fptr(out)))
    return A.astype(np.float64) @ B.astype(np.float64), out


@pytest.mark.parametrize("kind,per", [(0, 1), (1, 8)])
def test_mfma_lowprec_layout(kind, per):
    """f32 16x16x4 and bf16 16x16x32: D[4*(lane/16)+e][lane%16]."""
    D, out = _observe(kind, per)
    got = np.zeros((16, 16))
    for l in range(64):
        for e in range(4):
            got[4 * (l // 16) + e, l % 16] = out[l * 4 + e]
    if not np.array_equal(got, D):
        alt = np.zeros((16, 16))
        for l in range(64):
            for e in range(4):
                alt[l // 16 + 4 * e, l % 16] = out[l * 4 + e]
        print("kind %d: alternative layout D[lane/16+4e] matches: %s" % (kind, np.array_equal(alt, D)))
    assert np.array_equal(got, D)


def test_global_load_lds_placement():
    """global_load_lds_dwordx4 writes lane i's 16 bytes at LDS base + 16 i (the
    direct-to-LDS exchange kernels rely on it); nothing outside the 1 KB window
    is touched."""
    import ctypes
    rng = np.random.default_rng(3)
    src = rng.integers(1, 2 ** 31, size=(64, 4), dtype=np.uint32)
    perm = rng.permutation(64).astype(np.int32)
    out = np.zeros((128, 4), dtype=np.uint32)
    u32 = ctypes.POINTER(ctypes.c_uint32)
    _check(_probe.mchip_probe_lds_load(src.ctypes.data_as(u32), _lib.iptr(perm), out.ctypes.data_as(u32)))
    assert np.array_equal(out[:64], np.zeros((64, 4), np.uint32))
    assert np.array_equal(out[64:], src[perm])


def test_mfma_f32_rounding():
    """Rounding of v_mfma_f32_16x16x4_f32's four-term sum: every case lands
    within one ulp of the exact sum (the FP32 contractions' error model); the
    printed line says whether each case rounded to nearest-even."""
    A = np.zeros((16, 4), np.float64)
    B = np.zeros((4, 16), np.float64)
    B[:, 0] = 1.0
    cases = [
        (1.0, 1.25 * 2.0 ** -24, 0.0, 0.0),                  # above half an ulp: up
        (1.0, -0.75 * 2.0 ** -25, 0.0, 0.0),                 # below 1, under half an ulp: stays 1
        (1.0, 2.0 ** -24, 0.0, 0.0),                         # tie: even (1)
        (1.0 + 2.0 ** -23, 2.0 ** -24, 0.0, 0.0),            # tie: even (1 + 2^-22)
        (1.0, 2.0 ** -25, 2.0 ** -25, 2.0 ** -25),           # three sub-half-ulp terms: 1 + 2^-23 if fused
        (2.0 ** 20, -(2.0 ** 20), 2.0 ** -30, 0.0),          # cancellation keeps the small term
    ]
    for i, c in enumerate(cases):
        A[i, :] = c
    a_lane = np.zeros(64, np.float32)
    b_lane = np.zeros(64, np.float32)
    for l in range(64):
        a_lane[l] = A[l % 16, l // 16]
        b_lane[l] = B[l // 16, l % 16]
    assert np.array_equal(a_lane.astype(np.float64)[:64].reshape(4, 16).T, A.astype(np.float32))
    out = np.zeros(256, np.float32)
    _check(_probe.mchip_probe_mfma_lp(0, _lib.fptr(a_lane), _lib.fptr(b_lane), _lib.fptr(out)))
    got = np.zeros((16, 16))
    for l in range(64):
        for e in range(4):
            got[4 * (l // 16) + e, l % 16] = out[l * 4 + e]
    modes = []
    for i, c in enumerate(cases):
        exact = float(np.sum(np.array(c, np.float64)))
        rne = float(np.float32(exact))
        g = got[i, 0]
        ulp = float(np.spacing(np.float32(abs(exact)))) if exact != 0 else 2.0 ** -149
        modes.append("case %d: exact %.17g got %.17g rne %s" % (i, exact, g, g == rne))
        assert abs(g - exact) <= ulp
    print("\n".join(modes))
