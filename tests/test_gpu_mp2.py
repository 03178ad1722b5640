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


# FP32 operands and products (2^-24 rounding) with FP64-flushed accumulation,
# contracted in the orthonormal auxiliary basis (B~ = L^-1 B) so that the
# energy e = sum Z'^2 is well conditioned in the FP32 intermediates: the
# batched energy agrees with the FP64 path to ~1e-7 relative (2 and 4
# waters; ~2e-8 for 32 waters).  Bar: 1e-6 relative for the energy and every
# per-point term.
BATCH_TOL = 1e-6


@pytest.mark.parametrize("nwater,nlap", [(2, 8), (4, 3)])
def test_batched_mp2_vs_fp64(ctx, nwater, nlap):
    """All Laplace points in one grouped FP32 hipGraph against the FP64
    per-point path (same orbitals, points and metric), and against the oracle."""
    from chem_amd.engine import DeviceB, Metric
    from chem_amd.mp2 import sos_mp2, laplace_for
    from chem_amd.scf import run_rhf
    mol = desc.Molecule(desc.water_cluster(nwater, 9), dfbasis="cc-pvdz-ri")
    print("case: batched MP2 water n=%d, %d points" % (nwater, nlap), flush=True)
    S, T, V = ints.int1e(mol)
    Bh = ints.BTensorHost.build(mol, 1e-10)
    M = ints.int2c2e(mol)
    dB = DeviceB(ctx, Bh)
    met = Metric(ctx, M)
    # orbitals from the FP64 GPU SCF in the jkfit basis
    molj = desc.Molecule(desc.water_cluster(nwater, 9))
    dBj = DeviceB(ctx, ints.BTensorHost.build(molj, 1e-10))
    metj = Metric(ctx, ints.int2c2e(molj))
    r = run_rhf(molj, dBj, metj, S, T + V, threshold=1e-9)
    w, a, _ = laplace_for(r.eps, mol.nocc, nlap)
    ref = sos_mp2(mol, dB, met, S, r.C, r.eps, points=(w, a))
    got = sos_mp2(mol, dB, met, S, r.C, r.eps, points=(w, a), batched=True)
    _, Vinv = met.get()
    E, terms = omp2.sos_mp2(Bh.dense(), Vinv, S, r.C, r.eps, mol.nocc, w, a, mol.b, Bmask=Bh.mask(),
                            return_terms=True)
    rel = abs(got["energy"] - ref["energy"]) / ref["energy"]
    print("E_fp64 %.12f E_batched %.12f rel %.2e (oracle %.12f); graph %.4f s vs per-point %.4f s" % (
        ref["energy"], got["energy"], rel, E, got["t_z"], ref["t_z"]))
    assert abs(ref["energy"] - E) < 1e-9 * E
    assert rel < BATCH_TOL
    assert np.abs(got["terms"] - ref["terms"]).max() < BATCH_TOL * ref["energy"]
    # a second call reproduces the graph's result bit for bit
    again = sos_mp2(mol, dB, met, S, r.C, r.eps, points=(w, a), batched=True)
    assert again["energy"] == got["energy"]


def test_z_builder_vs_oracle(ctx):
    """Z_k(X,Y) of every Laplace point, element by element: the FP64 path
    against the oracle (1e-10 of max|Z_k|) and the grouped FP32 graph against
    the FP64 path."""
    from chem_amd.engine import DeviceB, Metric
    from chem_amd.mp2 import z_builder, laplace_for
    mol = desc.Molecule(desc.water_cluster(2, 9), dfbasis="cc-pvdz-ri")
    print("case: Z builder water n=2, 8 points", flush=True)
    S, T, V = ints.int1e(mol)
    Bh = ints.BTensorHost.build(mol, 1e-10)
    B = Bh.dense()
    M = ints.int2c2e(mol)
    molj = desc.Molecule(desc.water_cluster(2, 9))
    Bj = ints.BTensorHost.build(molj, 1e-10).dense()
    Vinvj, _, invsqrtj = dfock.metric_factors(ints.int2c2e(molj))
    jk = lambda P, C: (dfock.df_j(Bj, Vinvj, P), dfock.dfmo_k(Bj, invsqrtj, C))
    r = oscf.rhf(S, T, V, molj.nuclear_repulsion(), molj.nocc, jk, bsz=molj.b, thr=1e-9)
    dB = DeviceB(ctx, Bh)
    met = Metric(ctx, M)
    _, Vinv = met.get()
    w, a, _ = laplace_for(r["eps"], mol.nocc, 8)
    Z64, t64 = z_builder(mol, dB, met, S, r["C"], r["eps"], w, a)
    Z32, t32 = z_builder(mol, dB, met, S, r["C"], r["eps"], w, a, batched=True)
    E, terms, Zo = omp2.sos_mp2(B, Vinv, S, r["C"], r["eps"], mol.nocc, w, a, mol.b, Bmask=Bh.mask(),
                                return_terms=True, return_z=True)
    for k in range(len(w)):
        scale = np.abs(Zo[k]).max()
        e64 = np.abs(Z64[k] - Zo[k]).max() / scale
        e32 = np.abs(Z32[k] - Z64[k]).max() / scale
        asym = np.abs(Z32[k] - Z32[k].T).max() / scale
        print("point %d: max|Z| %.3e  fp64 vs oracle %.2e  fp32 graph vs fp64 %.2e  (asym %.1e)  "
              "term rel %.2e" % (k, scale, e64, e32, asym, abs(t32[k] - t64[k]) / E), flush=True)
        assert e64 < 1e-10
        # Z = L Z' L^T back from the orthonormal basis: ~1e-5 of max|Z| (the
        # energy is taken from Z' and is far more accurate, checked below)
        assert e32 < 5e-5
    assert np.abs(t64 - np.array(terms)).max() < 1e-9 * E
    assert np.abs(t32 - t64).max() < BATCH_TOL * E


@pytest.mark.parametrize("nwater", [1, 3])
def test_ll_full_and_llmp_mem_vs_oracle(ctx, nwater):
    """The other Z builders behind mp::zmethod: LL_FULL_Z (pivoted Cholesky of
    the virtual pseudo-density, B(X|ia) route) against the oracle's LL_FULL
    restatement to 1e-9 relative; LLMP_MEM_Z (the same contractions as
    LLMP_FULL) reproduces the LLMP_FULL energy bit for bit."""
    from chem_amd.engine import DeviceB, Metric
    from chem_amd.mp2 import sos_mp2, laplace_for
    mol = desc.Molecule(desc.water_cluster(nwater, 5), dfbasis="cc-pvdz-ri")
    print("case: LL_FULL / LLMP_MEM water n=%d" % nwater, flush=True)
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
    ll = sos_mp2(mol, dB, met, S, r["C"], r["eps"], points=(w, a), zmethod="ll_full")
    full = sos_mp2(mol, dB, met, S, r["C"], r["eps"], points=(w, a))
    mem = sos_mp2(mol, dB, met, S, r["C"], r["eps"], points=(w, a), zmethod="llmp_mem")
    E, terms = omp2.sos_mp2(B, Vinv, S, r["C"], r["eps"], mol.nocc, w, a, mol.b, Bmask=Bh.mask(),
                            return_terms=True, zmethod="ll_full")
    print("LL_FULL gpu %.12f oracle %.12f rel %.2e; LLMP_FULL %.12f" % (
        ll["energy"], E, abs(ll["energy"] - E) / E, full["energy"]))
    assert abs(ll["energy"] - E) < 1e-9 * E
    assert np.abs(ll["terms"] - np.array(terms)).max() < 1e-9 * E
    assert abs(ll["energy"] - full["energy"]) < 1e-8 * E
    assert mem["energy"] == full["energy"]


def test_z_compute_from_factors(ctx):
    """mp::Z from given factors (set_occ_coeff / set_vir_density /
    set_vir_coeff): LLMP and LL_FULL Z against the oracle's restatements, with
    random factors on two waters, to 1e-12 of max|Z|."""
    from chem_amd.engine import DeviceB
    from chem_amd.mp2 import z_compute
    mol = desc.Molecule(desc.water_cluster(2, 5), dfbasis="cc-pvdz-ri")
    print("case: Z from factors, 2 waters", flush=True)
    Bh = ints.BTensorHost.build(mol, 1e-10)
    B = Bh.dense()
    dB = DeviceB(ctx, Bh)
    rng = np.random.default_rng(11)
    n = mol.nbf
    Lo = rng.standard_normal((n, 6))
    Lv = rng.standard_normal((n, 9))
    Pv = Lv @ Lv.T
    pm = omp2.shellpair_mask(Bh.mask(), mol.b)
    Zl = z_compute(dB, Lo, Pvir=Pv)
    Zo = omp2.llmp_full_z(B, Lo, Pv, pm)
    assert np.abs(Zl - Zo).max() < 1e-12 * np.abs(Zo).max()
    Zf = z_compute(dB, Lo, Lvir=Lv, zmethod="ll_full")
    Zfo = omp2.ll_full_z(B, Lo, Lv)
    assert np.abs(Zf - Zfo).max() < 1e-12 * np.abs(Zfo).max()
