"""GPU Cholesky factorisation and triangular inverse of the metric (the
library's own, blocked, block size 64) against numpy, at sizes that are not
multiples of the block, the non-positive-definite error, and the factor's
reproducibility while other host threads keep the device busy with FP64
exchange builds (the load under which rocsolver_dpotrf/dtrtri returned wrong
factors)."""
import threading

import numpy as np
import pytest

from chem_amd import desc, ints

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from chem_amd.engine import Context
    yield Context(0)


@pytest.mark.parametrize("n", [1, 63, 64, 65, 200, 1000])
def test_cholesky_and_inverse_vs_numpy(ctx, n):
    from chem_amd.engine import Metric
    print("case: metric factors n=%d" % n, flush=True)
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n))
    V = A @ A.T + n * np.eye(n)
    m = Metric(ctx, V)
    Linv, Vinv = m.get()
    L = np.linalg.cholesky(V)
    Linv_ref = np.linalg.inv(L)
    cond = np.linalg.cond(V)
    assert np.array_equal(np.triu(Linv, 1), np.zeros_like(Linv))
    assert np.abs(Linv - Linv_ref).max() < 50 * cond * 2.2e-16 * np.abs(Linv_ref).max()
    assert np.abs(Linv @ V @ Linv.T - np.eye(n)).max() < 1e-11
    assert np.abs(Vinv @ V - np.eye(n)).max() < 1e-11


def test_not_positive_definite(ctx):
    from chem_amd.engine import Metric
    print("case: indefinite metric", flush=True)
    rng = np.random.default_rng(1)
    A = rng.standard_normal((150, 150))
    V = A @ A.T + 150 * np.eye(150)
    V[100, 100] = -1.0
    with pytest.raises(RuntimeError, match="not positive definite"):
        Metric(ctx, V)


def test_factor_reproducible_under_concurrent_fp64_load(ctx):
    """Factor one 24-water metric 40 times while three other contexts run plain
    FP64 exchange builds: every factor is bit-for-bit the first one."""
    from chem_amd.engine import Context, DeviceB, Metric
    print("case: metric factor under concurrent FP64 exchange", flush=True)
    molm = desc.Molecule(desc.water_cluster(24, 7))
    V = ints.int2c2e(molm)
    mol = desc.Molecule(desc.water_cluster(8, 3))
    Vs = ints.int2c2e(mol)
    work = []
    for k in range(3):
        c = Context(0)
        dB = DeviceB.build_on_gpu(c, mol, 1e-10)
        met = Metric(c, Vs)
        C = np.ascontiguousarray(np.random.default_rng(k).standard_normal((mol.nbf, mol.nocc)))
        work.append((dB, met, C))
    stop = [False]
    builds = [0]

    def load(k):
        dB, met, C = work[k]
        while not stop[0]:
            dB.dfk_mo(C, met)
            builds[0] += 1

    ths = [threading.Thread(target=load, args=(k,)) for k in range(3)]
    for t in ths:
        t.start()
    try:
        ref = None
        for _ in range(40):
            m = Metric(ctx, V)
            Linv, Vinv = m.get()
            m.release()
            if ref is None:
                ref = (Linv, Vinv)
            else:
                assert np.array_equal(Linv, ref[0]) and np.array_equal(Vinv, ref[1])
    finally:
        stop[0] = True
        for t in ths:
            t.join()
    print("concurrent FP64 builds during the test: %d" % builds[0])
    assert builds[0] > 0
