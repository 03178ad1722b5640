"""FP64 restatement of the reference's restricted Hartree-Fock loop.

Follows src/hf/hfscf.cpp:
  one_electron          :131-181  S, T, V, X = orthogonalizer(S), H = T + V,
                                  then block filter(eps) of S, T, V, X, H
  compute_guess "core"  src/hf/hfguess.cpp:127-141  F = H, diag_fock
  form_fock             :228-276  F = H + J + K (K is returned negated)
  compute_errmat        :298-324  e = F P S - S P F
  compute_scf_energy    :278-296  E = tr(H P) + tr(F P)  (RHF, P = P_alpha)
  loop                  :386-459  RMS(e) < thr and iter > 0 -> stop; DIIS; diag
  diag_fock             src/hf/hfdiagfock.cpp:12-211  F' = X^T F X, eig,
                                  C = X C', P = C_occ C_occ^T
  orthogonalizer        src/math/linalg/orthogonalizer.cpp:14-45 (threshold is
                                  an int = 0: eigenvalues < 0 -> 0, else 1/sqrt)
  diis_helper           src/math/solvers/diis.hpp:44-215 (start 1, min 2,
                                  max 10; when full the error vector of largest
                                  norm is dropped; B-matrix Lagrange solve)

  compute_guess "SAD"   src/hf/hfguess.cpp:142-452  per element an atomic
                        SCF (fractional, spin-averaged occupations,
                        src/desc/molecule.cpp:37-93; exact J and K; core
                        guess; no DIIS; thr 1e-6, hfmod.hpp:50-53), the
                        atomic alpha densities on the atoms' diagonal blocks,
                        the first exchange on the density's pivoted Cholesky
                        factor (hfguess.cpp:436-452)

Pins: at convergence P S P = P, tr(P S) = nocc, |F P S - S P F| < thr; the
energy does not depend on DIIS settings; a closed-shell energy from two
independent K formulations agrees.  SAD: the occupation rule on atoms whose
configurations are known (He, C, Na, Ar, ...); closed-shell atoms (He, Ne),
where fractional and integer occupation coincide, reproduce the published
RHF/cc-pVDZ energies; a one-function atom reproduces E = 2 f h + f^2 (ss|ss)
(f = occupation), which fixes the occupation weighting and the 2J - K
combination; tr(P S) = Z / 2.  The oracle's integrals (S, T, V, ERIs) are the
library's, which are pinned on their own by closed forms and derivative
relations.
"""
import bisect

import numpy as np


def block_filter(M, bsz, eps):
    """Drop blocks with Frobenius norm < eps (DBCSR filter, block level)."""
    if eps is None or eps <= 0:
        return M
    off = np.concatenate([[0], np.cumsum(bsz)])
    out = M.copy()
    for i in range(len(bsz)):
        for j in range(len(bsz)):
            blk = out[off[i]:off[i + 1], off[j]:off[j + 1]]
            if np.sqrt((blk * blk).sum()) < eps:
                blk[...] = 0.0
    return out


def orthogonalizer(S):
    """X = U diag(d) U^T, d = 0 for eigenvalues < 0 else 1/sqrt(eig)."""
    w, U = np.linalg.eigh(S)
    d = np.where(w < 0, 0.0, 1.0 / np.sqrt(np.where(w < 0, 1.0, w)))
    return (U * d) @ U.T


class DIIS:
    """diis.hpp:19-215."""

    def __init__(self, start=1, min_vecs=2, max_vecs=10):
        self.start, self.min, self.max = start, min_vecs, max_vecs
        self.delta, self.trial = [], []
        self.B = np.zeros((0, 0))
        self.coeffs = np.zeros(0)

    def compute_extrapolation_parameters(self, T, err, it):
        if it < self.start:
            return
        reduce = len(self.delta) >= self.max
        self.delta.append(err.copy())
        norms = [np.linalg.norm(e) for e in self.delta]
        max_pos = int(np.argmax(norms))
        if reduce:
            del self.delta[max_pos]
        self.trial.append(T.copy())
        if reduce:
            del self.trial[max_pos]
        nerr = len(self.delta)
        v = np.array([np.sum(self.delta[i] * self.delta[nerr - 1]) for i in range(nerr)])
        if reduce:
            Bn = np.delete(np.delete(self.B, max_pos, 0), max_pos, 1)
            B = np.zeros((nerr, nerr))
            B[:nerr - 1, :nerr - 1] = Bn[:nerr - 1, :nerr - 1]
        else:
            B = np.zeros((nerr, nerr))
            B[:nerr - 1, :nerr - 1] = self.B
        B[nerr - 1, :] = v
        B[:, nerr - 1] = v
        self.B = B
        Bs = np.zeros((nerr + 1, nerr + 1))
        Bs[:nerr, :nerr] = B
        Bs[nerr, :nerr] = -1
        Bs[:nerr, nerr] = -1
        rhs = np.zeros(nerr + 1)
        rhs[nerr] = -1
        x = np.linalg.lstsq(Bs, rhs, rcond=None)[0]
        self.coeffs = x[:nerr]

    def extrapolate(self, T, it):
        if it >= self.start and len(self.trial) >= self.min:
            return sum(c * t for c, t in zip(self.coeffs, self.trial))
        return T


def rhf(S, T, V, Enuc, nocc, jk, bsz=None, filter_eps=1e-10, thr=1e-8, max_iter=100,
        do_diis=True, diis_start=1, diis_min=2, diis_max=10, log=None, occ=None, guess_P=None):
    """Restricted HF with a J/K callback jk(P_alpha, C_occ) -> (J, K).

    occ: factors of the occupied columns (fractional occupation,
    hfdiagfock.cpp scales C_occ by them; P = C_occ diag(occ^2) C_occ^T).
    guess_P: the first density (SAD); the first callback then gets its
    pivoted Cholesky factor as C_occ (hfscf.cpp:389-396)."""
    nbf = S.shape[0]
    bsz = bsz or [nbf]
    X = orthogonalizer(S)
    H = T + V
    S = block_filter(S, bsz, filter_eps)
    X = block_filter(X, bsz, filter_eps)
    H = block_filter(H, bsz, filter_eps)

    def diag(F):
        Fp = X.T @ F @ X
        eps, Cp = np.linalg.eigh(Fp)
        C = X @ Cp
        return eps, C

    f = np.ones(nocc) if occ is None else np.asarray(occ, dtype=np.float64)

    def density(C):
        Co = C[:, :nocc] * f
        return Co @ Co.T, Co

    if guess_P is None:
        eps, C = diag(H)
        P, Co = density(C)
    else:
        from .mp2 import pivoted_cholesky
        eps, C = None, None
        P = np.array(guess_P, dtype=np.float64)
        Co, _ = pivoted_cholesky(P, filter_eps=filter_eps)
    diis = DIIS(diis_start, diis_min, diis_max)
    E = 0.0
    it = 0
    history = []
    while True:
        J, K = jk(P, Co)
        F = H + J + K
        e = F @ P @ S - S @ P @ F
        E = float(np.sum(H * P) + np.sum(F * P))
        rms = np.sqrt(np.sum(e * e) / (nbf * nbf))
        history.append((E + Enuc, rms))
        if log:
            log(it, E + Enuc, rms)
        if rms < thr and it > 0:
            break
        if it > max_iter:
            raise RuntimeError("HF did not converge.")
        if do_diis:
            diis.compute_extrapolation_parameters(F, e, it)
            F = diis.extrapolate(F, it)
        eps, C = diag(F)
        P, Co = density(C)
        it += 1
    return dict(energy=E + Enuc, e_elec=E, C=C, eps=eps, P=P, F=F, J=J, K=K,
                iterations=it, history=history)


CONF_ORB = (0, 2, 10, 18, 36, 54, 86, 118)   # electrons of the closed shells (hfguess.cpp:24)


def sad_occupations(Z):
    """molecule.cpp:37-93 (fractional, spin averaged, neutral): Z/2 alpha
    electrons; the orbitals below the last closed shell under Z hold one
    each, the next shell's orbitals share the rest.  Returns the factors of
    the occupied columns (square roots of the occupations)."""
    if not 1 <= Z <= 118:
        raise ValueError("atomic number out of range")
    i = bisect.bisect_left(CONF_ORB, Z)     # first closed-shell count >= Z
    if Z < CONF_ORB[i]:
        i -= 1
    if CONF_ORB[i] == Z:
        ncore, nact = 0, CONF_ORB[i] // 2
    else:
        ncore, nact = CONF_ORB[i] // 2, CONF_ORB[i + 1] // 2 - CONF_ORB[i] // 2
    occ = np.ones(ncore + nact)
    occ[ncore:] = np.sqrt((0.5 * Z - ncore) / nact)
    return occ


def exact_jk(eri):
    """Restricted exact J and K of an atom (J[P_alpha + P_beta] and the
    negated K[P_alpha]) from the full ERI tensor (ij|kl)."""
    def jk(P, C):
        J = np.einsum("ijkl,kl->ij", eri, P, optimize=True)
        K = np.einsum("ikjl,kl->ij", eri, P, optimize=True)
        return 2.0 * J, -K
    return jk


def atomic_scf(S, T, V, eri, Z, thr=1e-6, max_iter=100):
    """The atomic SCF of the SAD guess (hfguess.cpp:248-292): fractional
    occupation, exact J/K, core guess, no DIIS.  Returns rhf()'s dict."""
    occ = sad_occupations(Z)
    return rhf(S, T, V, 0.0, len(occ), exact_jk(eri), filter_eps=0.0, thr=thr, max_iter=max_iter,
               do_diis=False, occ=occ)


def sad_density(nbf, atoms, atomic_density):
    """hfguess.cpp:398-413: the alpha density of the molecule, block diagonal,
    atoms = [(Z, first function, number of functions)], atomic_density(Z) the
    element's atomic alpha density (computed once per element from its first
    atom, hfguess.cpp:146-160)."""
    P = np.zeros((nbf, nbf))
    cache = {}
    for Z, off, n in atoms:
        if Z not in cache:
            cache[Z] = atomic_density(Z)
        P[off:off + n, off:off + n] = cache[Z]
    return P
