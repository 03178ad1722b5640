"""GPU restricted HF against the oracle SCF (same integrals, same metric
factors): total energies agree to 1e-10 Eh; the converged density satisfies
P S P = P and tr(P S) = nocc."""
import numpy as np
import pytest

from chem_amd import desc, ints
from oracle import dfock, scf as oscf

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from chem_amd.engine import Context
    yield Context(0)


@pytest.mark.parametrize("nwater,kmode", [(1, "dfmo"), (1, "dfmo_transformed"), (2, "dfmo")])
def test_rhf_energy_vs_oracle(ctx, nwater, kmode):
    from chem_amd.engine import DeviceB, Metric
    from chem_amd.scf import run_rhf
    mol = desc.Molecule(desc.water_cluster(nwater, 7))
    print("case: RHF water n=%d kmode=%s" % (nwater, kmode), flush=True)
    S, T, V = ints.int1e(mol)
    Bh = ints.BTensorHost.build(mol, 1e-10)
    M = ints.int2c2e(mol)
    dB = DeviceB(ctx, Bh)
    m = Metric(ctx, M)
    Linv, Vinv = m.get()
    if kmode == "dfmo_transformed":
        dB.transform(m)
    res = run_rhf(mol, dB, m, S, T + V, threshold=1e-8, kmode=kmode)
    assert res.converged
    Bd = Bh.dense()
    jk = lambda P, C: (dfock.df_j(Bd, Vinv, P, bsz=mol.b),
                       dfock.dfmo_k(Bd, Linv.T, C, bsz=mol.b, xsz=mol.x, osz=mol.oa()))
    ref = oscf.rhf(S, T, V, mol.nuclear_repulsion(), mol.nocc, jk, bsz=mol.b, thr=1e-8)
    print("E gpu %.12f oracle %.12f diff %.2e iters %d/%d" % (
        res.energy, ref["energy"], res.energy - ref["energy"], res.iterations, ref["iterations"]))
    assert abs(res.energy - ref["energy"]) < 1e-10
    C = res.C[:, :mol.nocc]
    P = C @ C.T
    assert np.abs(P @ S @ P - P).max() < 1e-8
    assert abs(np.trace(P @ S) - mol.nocc) < 1e-10
    assert np.abs(res.eps - ref["eps"]).max() < 1e-7
