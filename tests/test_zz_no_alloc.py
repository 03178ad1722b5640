"""The allocation-free DF-J entry point (mchip_dfj_dev_noalloc) against the
allocating one (mchip_dfj_dev), through the C ABI: bit for bit on two calls in
a row, the second on a smaller tensor, and a clear error for a tensor whose
workspace exceeds the fixed array.  The two entry points run the same kernels
in the same order on different workspaces, so their J must be identical."""
import numpy as np
import pytest

from chem_amd import _lib, desc, ints

pytestmark = pytest.mark.gpu


def _dfj_both(ctx, dB, m, P):
    from chem_amd.engine import DeviceArray
    L = _lib.lib()
    n = dB.nbf
    Pd = DeviceArray(ctx, (n, n), np.ascontiguousarray(P))
    J1 = DeviceArray(ctx, (n, n))
    J2 = DeviceArray(ctx, (n, n))
    _lib.check(L.mchip_dfj_dev(dB.h, m.h, Pd.ptr, 2.0, J1.ptr))
    _lib.check(L.mchip_dfj_dev_noalloc(dB.h, m.h, Pd.ptr, 2.0, J2.ptr))
    ctx.sync()
    a, b = J1.download(), J2.download()
    for x in (Pd, J1, J2):
        x.release()
    return a, b


def test_noalloc_dfj_matches_allocating_entry_point():
    from chem_amd.engine import Context, DeviceB, Metric
    ctx = Context(0)
    L = _lib.lib()
    rng = np.random.default_rng(77)
    for nwater in (3, 1):                       # the second call on a smaller tensor
        print("case: no-alloc DF-J, %d waters" % nwater, flush=True)
        mol = desc.Molecule(desc.water_cluster(nwater, 4))
        dB = DeviceB(ctx, ints.BTensorHost.build(mol, 1e-10))
        m = Metric(ctx, ints.int2c2e(mol))
        C = rng.standard_normal((mol.nbf, mol.nocc)) * 0.3
        J1, J2 = _dfj_both(ctx, dB, m, C @ C.T)
        assert np.abs(J1).max() > 0
        assert np.array_equal(J1, J2)
        dB.release()
        m.release()
    # a tensor whose DF-J workspace is above the fixed array: one 2200-function
    # block fully significant (4.84e6 pair elements, 155 MB of workspace)
    print("case: no-alloc DF-J, oversized workspace", flush=True)
    nbf, naux = 2200, 8
    dense = rng.standard_normal((naux, nbf, nbf)) * 1e-3
    dense = 0.5 * (dense + dense.transpose(0, 2, 1))
    Bh = ints.BTensorHost.from_dense(dense, [naux], [nbf], np.ones((1, 1, 1), dtype=np.uint8))
    dB = DeviceB(ctx, Bh)
    m = Metric(ctx, np.eye(naux) * 2.0)
    from chem_amd.engine import DeviceArray
    Pd = DeviceArray(ctx, (nbf, nbf), np.zeros((nbf, nbf)))
    Jd = DeviceArray(ctx, (nbf, nbf))
    rc = L.mchip_dfj_dev_noalloc(dB.h, m.h, Pd.ptr, 2.0, Jd.ptr)
    assert rc != 0
    assert b"more than the fixed" in L.mchip_last_error()
    # the allocating entry point still handles it, and the context stays usable
    _lib.check(L.mchip_dfj_dev(dB.h, m.h, Pd.ptr, 2.0, Jd.ptr))
    ctx.sync()
    assert np.all(Jd.download() == 0.0)
    for x in (Pd, Jd):
        x.release()
    dB.release()
    m.release()
    ctx.close()
