"""Parity at the workloads' full sizes (32 and 64 waters), where the smaller
tests cannot reach: occupied widths above 128, hundreds of atom blocks,
several thousand aux functions.

* 32 waters, cc-pVDZ-jkfit (768 basis functions, 3712 aux, 160 occupied):
  the GPU-built B's block mask equals the host engine's, bit for bit, on a
  sample of aux blocks; plain DF-J against the oracle's extended-precision J
  and plain DFMO-K (three occupied batches, as hfmod runs it) against the
  oracle's DFMO, both fed with B's dense rows.
* 64 waters (1536 / 7424 / 320): the chosen form's BF16x3 and FP32 exchanges
  against the FP64 exchange at the converged SCF orbitals (Frobenius norms),
  the FP64 exchange on the L^-1 B copy against the plain one, and |dE_SCF|
  between the chosen form and the plain FP64 SCF with both converged to 1e-8.
* 32 waters, cc-pVDZ-ri, 8 Laplace points: the grouped FP32 MP2 graph against
  the FP64 per-point energy, and rows of one Laplace point's Z against the
  oracle's LLMP_FULL restatement.
"""
import time

import numpy as np
import pytest

from chem_amd import desc, ints
from oracle import dfock, mp2 as omp2

pytestmark = pytest.mark.gpu

SEED = 1234            # the generator seed of the 32- and 64-water workloads
ROWS = 256             # aux rows per dense batch fed to the oracle


@pytest.fixture(scope="module")
def ctx():
    from chem_amd.engine import Context
    c = Context(0)
    yield c
    c.close()


def _relerr(a, b):
    return float(np.abs(a - b).max() / max(1e-300, np.abs(b).max()))


def _frob(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.timeout(900)
def test_32_water_plain_j_k_vs_oracle(ctx):
    from chem_amd.engine import DeviceB, Metric
    mol = desc.Molecule(desc.water_cluster(32, SEED))
    print("case: 32 waters nbf=%d naux=%d nocc=%d" % (mol.nbf, mol.naux, mol.nocc), flush=True)
    assert (mol.nbf, mol.naux, mol.nocc) == (768, 3712, 160)
    mask, _, _ = ints.schwarz_mask(mol, 1e-10)
    t = time.perf_counter()
    dB = DeviceB.build_on_gpu(ctx, mol, 1e-10, mask=mask)
    ctx.sync()
    print("GPU B built in %.2f s, %.3e elements kept" % (time.perf_counter() - t, dB.nnz_kept()), flush=True)
    gmask = dB.mask()
    # host engine on a sample of aux blocks (every other X block screened
    # out): its post-fill filter must drop exactly the blocks the GPU dropped
    xs = [0, 1, 17, 40, 71, 95]
    assert max(xs) < len(mol.dfblocks)
    sub = np.zeros_like(mask)
    sub[xs] = mask[xs]
    t = time.perf_counter()
    Bh = ints.BTensorHost.build(mol, 1e-10, mask=sub)
    print("host B on %d aux blocks in %.2f s" % (len(xs), time.perf_counter() - t), flush=True)
    hmask = Bh.mask()
    assert np.array_equal(hmask[xs], gmask[xs])
    assert np.all(gmask <= mask)                 # the filter only drops blocks
    xoff = np.concatenate([[0], np.cumsum(mol.x)])
    for xb in xs:
        x0, x1 = int(xoff[xb]), int(xoff[xb + 1])
        hg, gg = Bh.dense_rows(x0, x1), dB.dense_rows(x0, x1)
        assert np.abs(hg - gg).max() <= 1e-12 * np.abs(hg).max()
    del Bh

    m = Metric(ctx, ints.int2c2e(mol))
    Linv, Vinv = m.get()
    rng = np.random.default_rng(32)
    C = rng.standard_normal((mol.nbf, mol.nocc)) * 0.1
    P = C @ C.T
    rows = lambda x0, x1: dB.dense_rows(x0, x1)

    t = time.perf_counter()
    J = dB.dfj(P, m)
    tj = time.perf_counter() - t
    t = time.perf_counter()
    K = dB.dfk_mo(C, m, occ_nbatches=3)
    tk = time.perf_counter() - t
    print("GPU plain J %.3f s, K (3 occupied batches) %.3f s (host copies included)" % (tj, tk), flush=True)
    t = time.perf_counter()
    Jref = dfock.df_j_ext(rows, Vinv, P, x_rows=(mol.naux, ROWS))
    print("oracle J (extended precision) %.1f s" % (time.perf_counter() - t), flush=True)
    t = time.perf_counter()
    Kref = dfock.dfmo_k_rows(rows, mol.naux, ROWS, Linv, C)
    print("oracle K %.1f s" % (time.perf_counter() - t), flush=True)
    ej, ek = _relerr(J, Jref), _relerr(K, Kref)
    print("32 waters: relerr J %.2e K %.2e" % (ej, ek), flush=True)
    # compensated GPU sums against extended-precision oracle sums: a few ulps
    assert ej < 1e-13
    # K: same L^-1 on both sides, FP64 sums over 3712 x 768 terms
    assert ek < 1e-12
    assert np.array_equal(K, K.T)
    assert np.array_equal(J, J.T)
    # the batched K equals the unbatched one to rounding
    K1 = dB.dfk_mo(C, m)
    assert _relerr(K1, K) < 1e-13


@pytest.mark.timeout(900)
def test_64_water_chosen_form_vs_fp64(ctx):
    from chem_amd.engine import DeviceB, Metric
    from chem_amd.scf import run_rhf
    mol = desc.Molecule(desc.water_cluster(64, SEED))
    print("case: 64 waters nbf=%d naux=%d nocc=%d" % (mol.nbf, mol.naux, mol.nocc), flush=True)
    assert (mol.nbf, mol.naux, mol.nocc) == (1536, 7424, 320)
    S, T, V = ints.int1e(mol)
    dB = DeviceB.build_on_gpu(ctx, mol, 1e-10)
    m = Metric(ctx, ints.int2c2e(mol))
    t = time.perf_counter()
    ref = run_rhf(mol, dB, m, S, T + V, threshold=1e-8, kmode="dfmo")
    print("plain FP64 SCF: E %.10f, %d iterations, %.1f s" % (ref.energy, ref.iterations,
                                                             time.perf_counter() - t), flush=True)
    assert ref.converged
    C = np.ascontiguousarray(ref.C[:, :mol.nocc])
    K_plain = dB.dfk_mo(C, m)
    ctx.trim()
    dB.make_lowp(m, f32=True, bf16=True, f64=True)
    K64 = dB.dfk_mo_tilde64(C)
    K32 = dB.dfk_mo_f32(C)
    K16 = dB.dfk_mo_bf16(C)
    e64, e32, e16 = _frob(K64, K_plain), _frob(K32, K_plain), _frob(K16, K_plain)
    print("K Frobenius rel. vs plain FP64: FP64 copy %.2e  FP32 %.2e  BF16x3 %.2e (max abs %.2e)" % (
        e64, e32, e16, np.abs(K16 - K_plain).max()), flush=True)
    assert e64 < 1e-12
    assert e32 < 3e-7
    assert e16 < 5e-6
    t = time.perf_counter()
    chosen = run_rhf(mol, dB, m, S, T + V, threshold=1e-8, kmode="mixed_auto")
    print("chosen-form SCF: E %.10f, %d iterations, %.1f s" % (chosen.energy, chosen.iterations,
                                                              time.perf_counter() - t), flush=True)
    assert chosen.converged
    dE = chosen.energy - ref.energy
    nocc = mol.nocc
    deo = np.abs(chosen.eps[:nocc] - ref.eps[:nocc]).max()
    dev = np.abs(chosen.eps[nocc:] - ref.eps[nocc:]).max()
    print("64 waters, both at 1e-8: dE %.2e Eh; max |d eps| occupied %.2e virtual %.2e" % (dE, deo, dev),
          flush=True)
    assert abs(dE) < 1e-6
    assert deo < 1e-5 and dev < 1e-5


@pytest.mark.timeout(900)
def test_32_water_batched_mp2_and_z_rows(ctx):
    from chem_amd.engine import DeviceB, Metric
    from chem_amd.mp2 import laplace_for, sos_mp2, z_builder
    from chem_amd.scf import run_rhf
    atoms = desc.water_cluster(32, SEED)
    molj = desc.Molecule(atoms)
    mol = desc.Molecule(atoms, dfbasis="cc-pvdz-ri")
    print("case: 32-water SOS-MP2 nbf=%d naux(ri)=%d" % (mol.nbf, mol.naux), flush=True)
    S, T, V = ints.int1e(mol)
    dBj = DeviceB.build_on_gpu(ctx, molj, 1e-10)
    mj = Metric(ctx, ints.int2c2e(molj))
    r = run_rhf(molj, dBj, mj, S, T + V, threshold=1e-8)
    assert r.converged
    dBj.release()
    mj.release()
    ctx.trim()
    dB = DeviceB.build_on_gpu(ctx, mol, 1e-10)
    m = Metric(ctx, ints.int2c2e(mol))
    w, a, _ = laplace_for(r.eps, mol.nocc, 8)
    ref = sos_mp2(mol, dB, m, S, r.C, r.eps, points=(w, a))
    got = sos_mp2(mol, dB, m, S, r.C, r.eps, points=(w, a), batched=True)
    rel = abs(got["energy"] - ref["energy"]) / ref["energy"]
    print("E_fp64 %.12f E_graph %.12f rel %.2e; Z builders %.3f s per-point vs %.3f s graph" % (
        ref["energy"], got["energy"], rel, ref["t_z"], got["t_z"]), flush=True)
    assert rel < 1e-6
    assert np.abs(got["terms"] - ref["terms"]).max() < 1e-6 * ref["energy"]
    # one Laplace point's Z, rows sampled, against the oracle
    Z, terms = z_builder(mol, dB, m, S, r.C, r.eps, w[:1], a[:1])
    assert abs(terms[0] - ref["terms"][0]) < 1e-10 * ref["energy"]
    Lo, Pv, rank = omp2.point_factors(S, r.C, r.eps, mol.nocc, w[0], a[0], mol.b)
    pm = omp2.shellpair_mask(dB.mask(), mol.b)
    xs = [0, 1, 500, 1344, 2000, mol.naux - 1]
    t = time.perf_counter()
    Zo = omp2.llmp_z_rows(lambda x0, x1: dB.dense_rows(x0, x1), mol.naux, ROWS, Lo, Pv, pm, xs)
    ez = np.abs(Z[0][xs] - Zo).max() / np.abs(Zo).max()
    print("Z rows of point 0 vs oracle: rel %.2e (rank %d, oracle %.1f s)" % (ez, rank,
                                                                              time.perf_counter() - t), flush=True)
    assert ez < 1e-10
