"""GPU three-centre integrals against the host engine: same screening mask
bit-exactly, values to 1e-12 relative (different FMA contraction only)."""
import numpy as np
import pytest

from chem_amd import desc, ints

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from chem_amd.engine import Context
    yield Context(0)


@pytest.mark.parametrize("nwater,dfbasis", [(1, "cc-pvdz-jkfit"), (3, "cc-pvdz-jkfit"), (2, "cc-pvdz-ri")])
def test_gpu_B_matches_host(ctx, nwater, dfbasis):
    from chem_amd.engine import DeviceB
    mol = desc.Molecule(desc.water_cluster(nwater, 21), dfbasis=dfbasis)
    print("case: GPU B n=%d %s" % (nwater, dfbasis), flush=True)
    Bh = ints.BTensorHost.build(mol, 1e-10)
    dB = DeviceB.build_on_gpu(ctx, mol, 1e-10)
    assert np.array_equal(dB.mask(), Bh.mask())
    assert dB.nelem == Bh.nelem
    ref = Bh.dense()
    got = dB.dense_rows(0, mol.naux)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print("relerr %.3e" % err)
    assert err < 1e-12
