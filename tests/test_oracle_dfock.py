"""Pins of the oracle's DF-J/K and RHF restatements (CPU).

* the four exchange formulations of the reference (DFMO, DFAO, DFMEM, DFLMO
  with occupied coefficients) agree to rounding;
* J and K are symmetric, K is negative semidefinite, 1/2 tr(P J) > 0;
* batching (nu, occupied, X batches) changes nothing beyond rounding;
* complete-auxiliary limit: s functions on one centre with the aux basis made
  of all pair products; DF is then exact and J/K equal the closed-form
  four-centre contractions;
* RHF invariants at convergence and DIIS independence of the energy.
"""
import math

import numpy as np
import pytest

from chem_amd import desc, ints
from oracle import dfock, scf as oscf


def _water_system(n=1, seed=0):
    mol = desc.Molecule(desc.water_cluster(n, seed))
    Bh = ints.BTensorHost.build(mol, 1e-10)
    return mol, Bh.dense(), ints.int2c2e(mol)


def test_exchange_formulations_agree():
    mol, B, V = _water_system(2, 3)
    Vinv, Linv, invsqrt = dfock.metric_factors(V)
    rng = np.random.default_rng(1)
    C = rng.standard_normal((mol.nbf, mol.nocc)) * 0.2
    P = C @ C.T
    k_mo = dfock.dfmo_k(B, invsqrt, C, bsz=mol.b, xsz=mol.x, osz=mol.oa(), occ_nbatches=3)
    k_ao = dfock.dfao_k(B, dfock.fitting_coefficients(B, Vinv), P, bsz=mol.b, xsz=mol.x)
    k_mem = dfock.dfmem_k(B, Vinv, P, bsz=mol.b, xsz=mol.x)
    k_lmo = dfock.dflmo_k_sym(B, Vinv, C)
    s = np.abs(k_mo).max()
    for k in (k_ao, k_mem, k_lmo):
        assert np.abs(k - k_mo).max() < 1e-10 * s
    assert np.abs(k_mo - k_mo.T).max() < 1e-12 * s
    assert np.linalg.eigvalsh(k_mo).max() < 1e-10 * s
    # batching does not matter
    k1 = dfock.dfmo_k(B, invsqrt, C, occ_nbatches=1, nbatches_b=1, nbatches_x=1)
    assert np.abs(k1 - k_mo).max() < 1e-11 * s


def test_coulomb_properties():
    mol, B, V = _water_system(1, 0)
    Vinv, _, _ = dfock.metric_factors(V)
    rng = np.random.default_rng(2)
    C = rng.standard_normal((mol.nbf, mol.nocc)) * 0.2
    P = C @ C.T
    J = dfock.df_j(B, Vinv, P, bsz=mol.b)
    J1 = dfock.df_j(B, Vinv, P, bsz=mol.b, nbatches_b=1)
    assert np.abs(J - J.T).max() < 1e-11 * np.abs(J).max()
    assert np.abs(J - J1).max() < 1e-11 * np.abs(J).max()
    assert 0.5 * np.sum(2 * P * J) > 0
    # dense einsum restatement
    g = np.einsum("xmn,mn->x", B, 2 * P)
    Jd = np.einsum("x,xmn->mn", Vinv @ g, B)
    assert np.abs(J - Jd).max() < 1e-10 * np.abs(J).max()


def test_complete_auxiliary_limit_matches_exact_jk():
    """One centre, s functions with exponents a_i; aux = s functions with
    exponents a_i + a_j (every product).  DF reproduces the exact integrals."""
    alphas = [0.3, 1.1, 4.0]
    aux = sorted({a + b for a in alphas for b in alphas})
    O = np.zeros(3)
    shells = [desc.Shell(0, O, 0, np.array([a]), np.array([1.0])) for a in alphas]
    dfshells = [desc.Shell(0, O, 0, np.array([g]), np.array([1.0])) for g in aux]

    class M:
        pass
    mol = M()
    mol.shells, mol.dfshells = shells, dfshells
    mol.blocks, mol.dfblocks = [[0, 1, 2]], [list(range(len(aux)))]
    mol.b, mol.x = [3], [len(aux)]
    mol.nbf, mol.naux = 3, len(aux)
    mol.atoms = [desc.Atom(1, 0, 0, 0)]
    mol.basis_handle = desc.BasisHandle(shells)
    mol.dfbasis_handle = desc.BasisHandle(dfshells)
    B = ints.BTensorHost.build(mol, 0.0).dense()
    V = ints.int2c2e(mol)
    Vinv, _, invsqrt = dfock.metric_factors(V)
    # exact (ij|kl) for normalised s Gaussians on one centre
    N = [(2 * a / math.pi) ** 0.75 for a in alphas]
    eri = np.zeros((3, 3, 3, 3))
    for i in range(3):
        for j in range(3):
            for k in range(3):
                for l in range(3):
                    p = alphas[i] + alphas[j]
                    q = alphas[k] + alphas[l]
                    eri[i, j, k, l] = N[i] * N[j] * N[k] * N[l] * 2 * math.pi ** 2.5 / (p * q * math.sqrt(p + q))
    C = np.array([[0.5], [0.3], [-0.2]])
    P = C @ C.T
    J_ex = np.einsum("ijkl,kl->ij", eri, 2 * P)
    K_ex = -np.einsum("ikjl,kl->ij", eri, P)
    J = dfock.df_j(B, Vinv, P)
    K = dfock.dfmo_k(B, invsqrt, C)
    assert np.abs(J - J_ex).max() < 1e-9 * np.abs(J_ex).max()
    assert np.abs(K - K_ex).max() < 1e-9 * np.abs(K_ex).max()


@pytest.fixture(scope="module")
def water_scf():
    mol, B, V = _water_system(1, 0)
    S, T, Vn = ints.int1e(mol)
    Vinv, _, invsqrt = dfock.metric_factors(V)
    jk = lambda P, C: (dfock.df_j(B, Vinv, P, bsz=mol.b), dfock.dfmo_k(B, invsqrt, C))
    return mol, S, T, Vn, jk


def test_rhf_invariants(water_scf):
    mol, S, T, Vn, jk = water_scf
    r = oscf.rhf(S, T, Vn, mol.nuclear_repulsion(), mol.nocc, jk, bsz=mol.b, thr=1e-9)
    P = r["P"]
    assert np.abs(P @ S @ P - P).max() < 1e-10
    assert abs(np.trace(P @ S) - mol.nocc) < 1e-10
    e = r["F"] @ P @ S - S @ P @ r["F"]
    assert np.sqrt((e * e).mean()) < 1e-8
    # DF-HF/cc-pVDZ water at this geometry sits within 1e-3 Eh of -76.0268
    assert abs(r["energy"] + 76.0268) < 1e-3
    r2 = oscf.rhf(S, T, Vn, mol.nuclear_repulsion(), mol.nocc, jk, bsz=mol.b, thr=1e-9, do_diis=False,
                  max_iter=400)
    assert abs(r["energy"] - r2["energy"]) < 1e-10


def test_orthogonalizer_and_filter():
    S = np.array([[1.0, 0.5], [0.5, 1.0]])
    X = oscf.orthogonalizer(S)
    assert np.abs(X @ S @ X - np.eye(2)).max() < 1e-14
    M = np.arange(16.0).reshape(4, 4) * 1e-12
    M[2:, 2:] = 1e-9
    f = oscf.block_filter(M, [2, 2], 1e-10)
    assert not f[:2, :2].any() and f[2:, 2:].any()


def _exact_round(fr):
    """Nearest double to a Fraction (float() of a Fraction rounds correctly)."""
    return float(fr)


def test_df_j_ext_matches_exact_rational():
    """The extended-precision DF-J equals exact rational arithmetic (rounded to
    FP64 at the same three points: gamma, V^-1 gamma, J) to within one ulp of
    each J element, on a tiny ill-conditioned case where plain FP64 sums in a
    different order disagree far beyond that."""
    from fractions import Fraction as F
    rng = np.random.default_rng(5)
    naux, nbf = 7, 4
    B = rng.standard_normal((naux, nbf, nbf))
    B = 0.5 * (B + B.transpose(0, 2, 1))
    A = rng.standard_normal((naux, naux))
    # nearly singular metric: V^-1 has entries of order 1e8
    w = np.array([1e-8, 1e-6, 1e-3, 1, 2, 3, 4])
    Q, _ = np.linalg.qr(A)
    V = (Q * w) @ Q.T
    Vinv = (Q / w) @ Q.T
    Vinv = 0.5 * (Vinv + Vinv.T)
    C = rng.standard_normal((nbf, 2))
    P = C @ C.T
    J = dfock.df_j_ext(B, Vinv, P)
    Pt = [[F(2 * P[m, n]) for n in range(nbf)] for m in range(nbf)]
    gamma = [_exact_round(sum(F(B[x, m, n]) * Pt[m][n] for m in range(nbf) for n in range(nbf)))
             for x in range(naux)]
    gq = [_exact_round(sum(F(gamma[x]) * F(Vinv[x, y]) for x in range(naux))) for y in range(naux)]
    Jx = np.array([[_exact_round(sum(F(gq[x]) * F(B[x, m, n]) for x in range(naux))) for n in range(nbf)]
                   for m in range(nbf)])
    ulp = np.spacing(np.abs(Jx))
    assert np.all(np.abs(J - Jx) <= ulp)
    # the plain FP64 restatement misses that bound here (the check is sharp)
    Jp = dfock.df_j(B, Vinv, P)
    assert (np.abs(Jp - Jx) > ulp).any()


def test_df_j_ext_row_batches():
    """The callable (X-row batch) form agrees with the dense one to one ulp
    (only the extended-precision association of the X sum differs), and with
    the plain FP64 restatement to its cond(V) bound."""
    mol, B, V = _water_system(1, 0)
    Vinv, _, _ = dfock.metric_factors(V)
    rng = np.random.default_rng(3)
    C = rng.standard_normal((mol.nbf, mol.nocc)) * 0.2
    P = C @ C.T
    J = dfock.df_j_ext(B, Vinv, P)
    Jb = dfock.df_j_ext(lambda x0, x1: B[x0:x1], Vinv, P, x_rows=(B.shape[0], 13))
    assert np.all(np.abs(J - Jb) <= np.spacing(np.abs(J)))
    Jp = dfock.df_j(B, Vinv, P, bsz=mol.b)
    assert np.abs(J - Jp).max() < 10 * np.linalg.cond(V) * 2.2e-16 * np.abs(J).max()


def test_dfmo_k_rows_matches_dfmo_k():
    """The X-row-batched DFMO restatement (for tensors too large to hold
    dense) equals the batched dense one to rounding."""
    mol, B, V = _water_system(2, 3)
    _, Linv, invsqrt = dfock.metric_factors(V)
    C = np.random.default_rng(4).standard_normal((mol.nbf, mol.nocc)) * 0.2
    K = dfock.dfmo_k(B, invsqrt, C, bsz=mol.b, xsz=mol.x, osz=mol.oa(), occ_nbatches=3)
    Kr = dfock.dfmo_k_rows(lambda x0, x1: B[x0:x1], B.shape[0], 17, Linv, C)
    assert np.abs(Kr - K).max() < 1e-13 * np.abs(K).max()


def test_dflmo_uv_pins():
    """The two-sided DFLMO restatement: with u = C and v = C^T it is the
    symmetric DFLMO (= DFMO) exchange; each element equals a direct triple sum
    (a transposed u or v, or V^-1 applied on the wrong side, breaks either);
    swapping u and v transposes K."""
    mol, B, V = _water_system(1, 2)
    Vinv, _, invsqrt = dfock.metric_factors(V)
    rng = np.random.default_rng(8)
    C = rng.standard_normal((mol.nbf, mol.nocc)) * 0.3
    Ksym = dfock.dflmo_k_sym(B, Vinv, C)
    Kuv = dfock.dflmo_k_uv(B, Vinv, C, C.T)
    assert np.abs(Kuv - Ksym).max() < 1e-12 * np.abs(Ksym).max()
    U = rng.standard_normal((mol.nbf, 3))
    W = rng.standard_normal((3, mol.nbf))
    K = dfock.dflmo_k_uv(B, Vinv, U, W)
    assert np.abs(K - K.T).max() > 1e-3 * np.abs(K).max()       # genuinely not symmetric
    for m_, n_ in ((0, 1), (5, 17), (23, 2)):
        bu = np.array([[B[x, m_, :] @ U[:, i] for i in range(3)] for x in range(B.shape[0])])
        bv = np.array([[B[x, n_, :] @ W[i, :] for i in range(3)] for x in range(B.shape[0])])
        direct = -sum(bu[x, i] * Vinv[x, y] * bv[y, i] for x in range(B.shape[0]) for y in range(B.shape[0])
                      for i in range(3))
        assert abs(K[m_, n_] - direct) < 1e-11 * np.abs(K).max()
    # (V^-1 now meets the other factor: the explicit inverse's cond(V) eps noise)
    Kt = dfock.dflmo_k_uv(B, Vinv, W.T, U.T)
    assert np.abs(Kt - K.T).max() < 10 * np.linalg.cond(V) * 2.2e-16 * np.abs(K).max()
