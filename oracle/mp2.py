"""FP64 restatement of the reference's batched CD-LT-SOS-RI-MP2.

Follows src/mp/mp_compute.cpp:27-407 and src/mp/z_full.cpp:69-397:
  * LLT of the overlap S = L_S L_S^T (mp_compute.cpp:167-178, _ORTHOGONALIZE)
  * per Laplace point k (mp_compute.cpp:232-382):
      C_o' = C_o diag(exp(+alpha eps_i / 2)),  C_v' = C_v diag(exp(-alpha eps_a / 2))
      P~_o = omega^(1/4) (L_S^-1 C_o')(L_S^-1 C_o')^T
      P_v  = omega^(1/4) C_v' C_v'^T
      L~   = pivoted incomplete Cholesky of P~_o (piv_cd.cpp:125-483, threshold
             N eps max_diag, columns reordered by LMO position :485-560)
      L_o  = L_S L~, then block filter(eps) of L_o and P_v
      Z    = LLMP_FULL_Z: b = L_o^T B; b2 = P_v b; b3 = L_o b2 on B's
             shell-pair pattern; Z_XY = sum b3(X|mn) B(Y|mn)
             (LL_FULL_Z: L_v = L_S CD(P~_v), b(X,i,a) = L_v^T b; Z = sum_ia b b)
      e_k  = tr(Z~ Z~) with Z~ = Z V^-1      (:361-381)
  * E = sum_k e_k (printed positive), scaled E = 1.3 E (:384-389).

Pins (tests/test_oracle_mp2.py): pivoted CD reconstructs P to its threshold
and its rank equals nocc; Laplace energy against the canonical DF-SOS sum
sum_{ijab} (ia|jb)^2 / (e_a + e_b - e_i - e_j) within the quadrature error.
The Laplace weights/exponents are an input here (they come from the minimax
solver, pinned separately against the reference's tabulated solutions).
"""
import numpy as np


def block_filter(M, rsz, csz, eps):
    """Zero blocks with Frobenius norm < eps."""
    ro = np.concatenate([[0], np.cumsum(rsz)])
    co = np.concatenate([[0], np.cumsum(csz)])
    out = M.copy()
    for i in range(len(rsz)):
        for j in range(len(csz)):
            blk = out[ro[i]:ro[i + 1], co[j]:co[j + 1]]
            if np.sqrt((blk * blk).sum()) < eps:
                blk[...] = 0.0
    return out


def split_range(n, split):
    if n <= 0:
        return []
    nb = n // split if n % split == 0 else n // split + 1
    out = [split] * nb
    if n % split:
        out[-1] = n % split
    return out


def pivoted_cholesky(A, eps=None, filter_eps=1e-10):
    """piv_cd.cpp:125-483: diagonal pivoting on the largest |diag|, stop when
    the pivot falls below N eps max|diag|; negative pivots above it throw.
    Returns L (N x rank) in the original row order with columns reordered by
    the LMO position rule of reorder_and_reduce (piv_cd.cpp:485-560)."""
    U = np.array(A, dtype=np.float64, copy=True)
    N = U.shape[0]
    perm = np.arange(N)
    maxd = np.abs(np.diag(U)).max() if N else 0.0
    thresh = eps if eps is not None else N * np.finfo(np.float64).eps * maxd
    L = np.zeros((N, N))
    rank = N
    for I in range(N):
        d = np.abs(np.diag(U)[I:])
        j = I + int(np.argmax(d))          # first index of the maximum
        if j != I:
            U[[I, j], :] = U[[j, I], :]
            U[:, [I, j]] = U[:, [j, I]]
            L[[I, j], :] = L[[j, I], :]
            perm[[I, j]] = perm[[j, I]]
        piv = U[I, I]
        if piv < 0 and abs(piv) > thresh:
            raise RuntimeError("Negative Pivot element. CD not possible.")
        if abs(piv) < thresh:
            rank = I
            break
        s = np.sqrt(piv)
        L[I, I] = s
        L[I + 1:, I] = U[I + 1:, I] / s
        U[I + 1:, I + 1:] -= np.outer(U[I + 1:, I], U[I + 1:, I]) / piv
    L = L[:, :rank]
    Lo = np.zeros_like(L)
    Lo[perm, :] = L                         # rows back to original order
    # reorder_and_reduce: sort columns by the midpoint of first/last row with value > eps
    pos = np.zeros(rank)
    for c in range(rank):
        rows = np.where(Lo[:, c] > filter_eps)[0]
        pos[c] = (N - 1) if len(rows) == 0 else 0.5 * (rows[0] + rows[-1])
    order = np.argsort(pos, kind="stable")
    return Lo[:, order], rank


def llmp_full_z(B, Lo, Pv, pair_mask=None):
    """z_full.cpp:69-397 (dense restatement; b3 restricted to B's pattern)."""
    b = np.einsum("mi,xmn->xin", Lo, B, optimize=True)            # mi, Xmn -> Xin
    b2 = np.einsum("Nn,xin->xiN", Pv, b, optimize=True)           # Nn, Xin -> XiN
    b3 = np.einsum("Mi,xiN->xMN", Lo, b2, optimize=True)          # Mi, XiN -> XMN
    if pair_mask is not None:
        b3 = b3 * pair_mask[None]
    naux = B.shape[0]
    return b3.reshape(naux, -1) @ B.reshape(naux, -1).T            # Mmn, Nmn -> MN


def ll_full_z(B, Lo, Lv):
    """LL_FULL_Z (z_full.cpp:428-692): b(X,i,N) = sum_m L_o(m,i) B(X|mN);
    b(X,i,a) = sum_N L_v(N,a) b(X,i,N); Z_XY = sum_ia b(X,i,a) b(Y,i,a)."""
    b = np.einsum("mi,xmn->xin", Lo, B, optimize=True)            # mi, Xmn -> Xin
    c = np.einsum("na,xin->xia", Lv, b, optimize=True)            # Na, XiN -> Xia
    naux = B.shape[0]
    c = c.reshape(naux, -1)
    return c @ c.T                                                # Mia, Nia -> MN


def shellpair_mask(Bmask, bsz):
    """get_shellpairs (z_common.cpp:7-60): (M,N) on if any X block is kept,
    expanded to functions."""
    pairs = Bmask.any(axis=0)
    off = np.concatenate([[0], np.cumsum(bsz)])
    nbf = off[-1]
    m = np.zeros((nbf, nbf))
    for i in range(len(bsz)):
        for j in range(len(bsz)):
            if pairs[i, j]:
                m[off[i]:off[i + 1], off[j]:off[j + 1]] = 1.0
    return m


def point_factors(S, C, eps, nocc, w, a, bsz, mo_split=5, filter_eps=1e-10):
    """The pseudo-density factors of one Laplace point (mp_compute.cpp:232-302):
    L_o = L_S CD(omega^(1/4) L_S^-1 C_o' C_o'^T L_S^-T) and P_v = omega^(1/4)
    C_v' C_v'^T, both block-filtered.  Returns (L_o, P_v, rank)."""
    Co, Cv = C[:, :nocc], C[:, nocc:]
    eo, ev = eps[:nocc], eps[nocc:]
    LS = np.linalg.cholesky(S)
    LSinv = np.linalg.solve(LS, np.eye(S.shape[0]))
    Coe = Co * np.exp(0.5 * eo * a)[None, :]
    Cve = Cv * np.exp(-0.5 * ev * a)[None, :]
    cortho = LSinv @ Coe
    Po_t = w ** 0.25 * cortho @ cortho.T
    Pv = w ** 0.25 * Cve @ Cve.T
    Lt, rank = pivoted_cholesky(Po_t, filter_eps=filter_eps)
    Lo = block_filter(LS @ Lt, bsz, split_range(rank, mo_split), filter_eps)
    Pv = block_filter(Pv, bsz, bsz, filter_eps)
    return Lo, Pv, rank


def llmp_z_rows(rows, naux, step, Lo, Pv, pair_mask, xs):
    """Rows xs of LLMP_FULL_Z's Z (z_full.cpp:69-397) with B supplied in X-row
    batches (rows(x0, x1) -> dense rows): b3(X) = pair_mask * (L_o L_o^T B(X)
    P_v) for X in xs (b = L_o^T B, b2 = P_v b, b3 = L_o b2 restricted to B's
    shell-pair pattern), then Z(X, Y) = sum_mn b3(X|mn) B(Y|mn) for every Y."""
    xs = list(xs)
    nbf = Lo.shape[0]
    LL = Lo @ Lo.T
    b3 = np.empty((len(xs), nbf * nbf))
    for k, x in enumerate(xs):
        b3[k] = (pair_mask * (LL @ rows(x, x + 1)[0] @ Pv.T)).reshape(-1)
    Z = np.empty((len(xs), naux))
    for y0 in range(0, naux, step):
        y1 = min(naux, y0 + step)
        Z[:, y0:y1] = b3 @ rows(y0, y1).reshape(y1 - y0, -1).T
    return Z


def sos_mp2(B, Vinv, S, C, eps, nocc, omega, alpha, bsz, mo_split=5, filter_eps=1e-10,
            Bmask=None, return_terms=False, return_z=False, zmethod="llmp_full"):
    """mpmod::compute.  Returns the positive sum E = sum_k e_k (unscaled);
    with return_terms also the e_k, with return_z also the Z_k of every point.
    zmethod "ll_full" factors the orthogonalised virtual pseudo-density too
    (mp_compute.cpp:309-336) and builds Z with ll_full_z."""
    pm = shellpair_mask(Bmask, bsz) if Bmask is not None else None
    terms, Zs = [], []
    for w, a in zip(omega, alpha):
        Lo, Pv, rank = point_factors(S, C, eps, nocc, w, a, bsz, mo_split, filter_eps)
        if zmethod == "ll_full":
            LS = np.linalg.cholesky(S)
            LSinv = np.linalg.solve(LS, np.eye(S.shape[0]))
            Cve = C[:, nocc:] * np.exp(-0.5 * eps[nocc:] * a)[None, :]
            cvo = LSinv @ Cve
            Lvt, rank_v = pivoted_cholesky(w ** 0.25 * cvo @ cvo.T, filter_eps=filter_eps)
            Lv = block_filter(LS @ Lvt, bsz, split_range(rank_v, mo_split), filter_eps)
            Z = ll_full_z(B, Lo, Lv)
        else:
            Z = llmp_full_z(B, Lo, Pv, pm)
        if return_z:
            Zs.append(Z)
        Zt = Z @ Vinv
        terms.append(float(np.sum(Zt * Zt.T)))
    E = float(np.sum(terms))
    out = (E,) + ((terms,) if return_terms else ()) + ((np.array(Zs),) if return_z else ())
    return out if len(out) > 1 else E


def canonical_sos(B, Vinv, C, eps, nocc):
    """sum_{ijab} (ia|jb)^2 / (e_a + e_b - e_i - e_j) with DF integrals."""
    Co, Cv = C[:, :nocc], C[:, nocc:]
    Bia = np.einsum("xmn,mi,na->xia", B, Co, Cv, optimize=True)
    naux = B.shape[0]
    Linv = np.linalg.cholesky(Vinv)                                 # Vinv = Linv Linv^T
    Q = np.einsum("xy,xia->yia", Linv, Bia, optimize=True)          # (ia|jb) = sum Q Q
    eo, ev = eps[:nocc], eps[nocc:]
    E = 0.0
    for i in range(nocc):
        iajb = np.einsum("ya,yjb->ajb", Q[:, i, :], Q, optimize=True)
        den = ev[:, None, None] + ev[None, None, :] - eo[i] - eo[None, :, None]
        E += float(np.sum(iajb * iajb / den))
    return E
