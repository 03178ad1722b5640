"""GPU SOS-AO-MP2 against the oracle (same orbitals, Laplace points and metric
factors): energies to 1e-9 relative; per-point terms to 1e-9 relative."""
import numpy as np
import pytest

from chem_amd import desc, ints
from oracle import dfock, mp2 as omp2, scf as oscf

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from chem_amd.engine import Context
    yield Context(0)


@pytest.mark.parametrize("nwater", [1, 3])
def test_sos_mp2_vs_oracle(ctx, nwater):
    from chem_amd.engine import DeviceB, Metric
    from chem_amd.mp2 import sos_mp2, laplace_for
    mol = desc.Molecule(desc.water_cluster(nwater, 5), dfbasis="cc-pvdz-ri")
    print("case: MP2 water n=%d" % nwater, flush=True)
    S, T, V = ints.int1e(mol)
    Bh = ints.BTensorHost.build(mol, 1e-10)
    B = Bh.dense()
    M = ints.int2c2e(mol)
    Vinv0, _, invsqrt0 = dfock.metric_factors(M)
    jk = lambda P, C: (dfock.df_j(B, Vinv0, P), dfock.dfmo_k(B, invsqrt0, C))
    r = oscf.rhf(S, T, V, mol.nuclear_repulsion(), mol.nocc, jk, bsz=mol.b, thr=1e-9)
    dB = DeviceB(ctx, Bh)
    met = Metric(ctx, M)
    _, Vinv = met.get()
    w, a, _ = laplace_for(r["eps"], mol.nocc, 6)
    got = sos_mp2(mol, dB, met, S, r["C"], r["eps"], points=(w, a))
    E, terms = omp2.sos_mp2(B, Vinv, S, r["C"], r["eps"], mol.nocc, w, a, mol.b, Bmask=Bh.mask(),
                            return_terms=True)
    print("E_gpu %.12f E_oracle %.12f rel %.2e" % (got["energy"], E, abs(got["energy"] - E) / E))
    assert abs(got["energy"] - E) < 1e-9 * E
    assert np.abs(got["terms"] - np.array(terms)).max() < 1e-9 * E
