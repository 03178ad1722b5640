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


@pytest.mark.parametrize("nwater,dfbasis", [(1, "cc-pvdz-jkfit"), (3, "cc-pvdz-jkfit"), (2, "cc-pvdz-ri")])
def test_gpu_1e_and_metric_match_host(ctx, nwater, dfbasis):
    """S, T, V and the Coulomb metric from the GPU class kernels against the
    host engine (same formulas; FMA contraction and the top-degree-only
    auxiliary Hermite expansion differ at rounding level)."""
    mol = desc.Molecule(desc.water_cluster(nwater, 21), dfbasis=dfbasis)
    print("case: GPU 1e/2c2e n=%d %s" % (nwater, dfbasis), flush=True)
    S, T, V = ints.int1e(mol)
    Sg, Tg, Vg = ints.int1e_gpu(ctx, mol)
    for name, a, b in (("S", S, Sg), ("T", T, Tg), ("V", V, Vg)):
        err = np.abs(a - b).max() / np.abs(a).max()
        print("%s relerr %.3e" % (name, err))
        assert err < 1e-12
        assert np.array_equal(b, b.T)
    M = ints.int2c2e(mol)
    Mg = ints.int2c2e_gpu(ctx, mol)
    err = np.abs(M - Mg).max() / np.abs(M).max()
    print("M relerr %.3e" % err)
    assert err < 1e-12
    assert np.array_equal(Mg, Mg.T)


@pytest.mark.parametrize("nwater,seed", [(1, 21), (3, 21), (8, 2)])
def test_gpu_schwarz_matches_host(ctx, nwater, seed):
    """AO Schwarz factors from the GPU class kernels against the host engine:
    values to 1e-12 relative, symmetric, and the screening mask built from
    them equal bit for bit to the library's (host) mask."""
    mol = desc.Molecule(desc.water_cluster(nwater, seed))
    print("case: GPU Schwarz n=%d seed=%d" % (nwater, seed), flush=True)
    Qh, Qx = ints.schwarz(mol)
    Qg = ints.schwarz_ao_gpu(ctx, mol)
    assert np.array_equal(Qg, Qg.T)
    err = np.abs(Qg - Qh).max() / np.abs(Qh).max()
    big = Qh > 1e-8 * Qh.max()
    rel = (np.abs(Qg - Qh)[big] / Qh[big]).max()
    print("relerr %.3e  elementwise %.3e" % (err, rel))
    assert err < 1e-12 and rel < 1e-11
    for eps in (1e-10, 1e-8, 1e-6):
        mask, _, _ = ints.schwarz_mask(mol, eps)
        mine = desc.schwarz_block_mask(desc.block_norms(Qg, mol.blocks), desc.block_norms(Qx, mol.dfblocks), eps)
        assert np.array_equal(mask, mine)
