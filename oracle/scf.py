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

Pins: at convergence P S P = P, tr(P S) = nocc, |F P S - S P F| < thr; the
energy does not depend on DIIS settings; a closed-shell energy from two
independent K formulations agrees.
"""
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
        do_diis=True, diis_start=1, diis_min=2, diis_max=10, log=None):
    """Restricted HF with a J/K callback jk(P_alpha, C_occ) -> (J, K)."""
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

    eps, C = diag(H)
    P = C[:, :nocc] @ C[:, :nocc].T
    diis = DIIS(diis_start, diis_min, diis_max)
    E = 0.0
    it = 0
    history = []
    while True:
        J, K = jk(P, C[:, :nocc])
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
        P = C[:, :nocc] @ C[:, :nocc].T
        it += 1
    return dict(energy=E + Enuc, e_elec=E, C=C, eps=eps, P=P, F=F, J=J, K=K,
                iterations=it, history=history)
