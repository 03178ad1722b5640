"""FP64 restatement of the reference's density-fitted J and K builders.

Follows src/fock/jkdf.cpp:
  DF_J::compute_J      jkdf.cpp:69-190   (P_tot -> gamma -> V^-1 gamma -> J)
  DFMO_K::compute_K    jkdf.cpp:221-462  (HT1 = B C, HT2 = L^-1 HT1, K = -HT2^T HT2)
  DFAO_K::compute_K    jkdf.cpp:510-648  (cbar = B P, K = -C_fit . cbar) with
  dfitting::compute    src/ints/fitting.cpp:44-161 (C_fit = V^-1 B)
  DFMEM_K::compute_K   jkdf.cpp:719-859  (DFAO with V^-1 B formed per batch)
  DFLMO_K (sym)        jkdf.cpp:890-1090 (K = -sum ht (V^-1 ht), ht = B C)
and the metric inversion of src/ints/aoloader.cpp:14-32 + src/math/linalg/LLT.cpp
(V = L L^T, invsqrt := L^-T, V^-1 = L^-T L^-1).

Batching (nu, occupied and X batches from make_blk_bounds) is reproduced; it
only changes the floating-point summation order.  DBCSR's on-the-fly filtering
of contraction outputs (filter(eps/nbatches) at jkdf.cpp:124,166,364,419) is a
third-party behaviour that nothing here can pin; the rule used throughout this
engine is: no filtering of contraction outputs (parity unpinned against DBCSR
for blocks below eps).

Pins (tests/test_oracle_dfock.py): dense einsum restatements on tiny inputs;
DFMO = DFAO = DFMEM = DFLMO to rounding; J and K symmetric; K negative
semidefinite; 1/2 tr(P J) > 0; complete-auxiliary limit against exact 4-centre
J/K for s-only systems with closed-form integrals.
"""
import numpy as np



def make_blk_bounds(blksizes, nbatches, blkmap=None):
    """Batch bounds in block indices (src/dbcsrx/dbcsr_btensor.hpp:35-95)."""
    nblks = len(blksizes)
    nbatches = min(nbatches, nblks)
    per = sum(blksizes) / nbatches
    out, cur, ibatch, first = [], 0, 1, 0
    for i in range(nblks):
        cur += blksizes[i]
        cc = blkmap[i] if blkmap is not None else i
        if i != nblks - 1:
            nc = blkmap[i + 1] if blkmap is not None else i + 1
        else:
            nc = -1 if blkmap is not None else i + 1
        if cur >= ibatch * per and nc != cc:
            out.append([first, i])
            first = i + 1
            ibatch += 1
    if not out:
        raise RuntimeError("Block bounds are zero.")
    return out


def metric_factors(V):
    """V = L L^T; returns (V^-1, L^-1, invsqrt = L^-T)  (aoloader.cpp:14-32)."""
    L = np.linalg.cholesky(V)
    Linv = np.linalg.solve(L, np.eye(V.shape[0]))
    Linv = np.tril(Linv)
    Vinv = Linv.T @ Linv
    return Vinv, Linv, Linv.T


def _elem_bounds(blksizes, blk_bounds):
    # make_blk_bounds closes its last batch only when nbatches * (nele /
    # nbatches) does not round above nele (dbcsr_btensor.hpp:46,77); the
    # reference then leaves the trailing blocks out of every batch.  As a
    # checker the oracle refuses such a split instead of dropping terms.
    flat = [b for b0, b1 in blk_bounds for b in range(b0, b1 + 1)]
    if flat != list(range(len(blksizes))):
        raise RuntimeError("batch bounds %s do not cover the %d blocks" % (blk_bounds, len(blksizes)))
    off = np.concatenate([[0], np.cumsum(blksizes)])
    return [(off[b0], off[b1 + 1]) for b0, b1 in blk_bounds]


def df_j(B, Vinv, P_alpha, P_beta=None, bsz=None, nbatches_b=5, blkmap=None):
    """DF_J::compute_J (jkdf.cpp:69-190).  B dense [naux, nbf, nbf]."""
    ptot = 2.0 * P_alpha if P_beta is None else P_alpha + P_beta
    nbf = B.shape[1]
    if bsz is None:
        bsz = [nbf]
    nub = _elem_bounds(bsz, make_blk_bounds(bsz, nbatches_b, blkmap))
    gamma = np.zeros(B.shape[0])
    for n0, n1 in nub:                        # jkdf.cpp:111-128  XMN, MN_ -> X_
        gamma += np.einsum("xmn,mn->x", B[:, :, n0:n1], ptot[:, n0:n1], optimize=True)
    gq = Vinv.T @ gamma                       # jkdf.cpp:137-141  X_, XY -> Y_
    J = np.zeros((nbf, nbf))
    for n0, n1 in nub:                        # jkdf.cpp:153-170  X_, XMN -> MN_
        J[:, n0:n1] += np.einsum("x,xmn->mn", gq, B[:, :, n0:n1], optimize=True)
    return J


def df_j_ext(B, Vinv, P_alpha, P_beta=None, x_rows=None):
    """DF_J::compute_J (jkdf.cpp:69-190) with every sum carried in x86 extended
    precision (numpy longdouble, 64-bit significand) and J rounded to FP64 once:
    gamma = B.P_tot (:111-128), gq(Y) = sum_X gamma(X) V^-1(X,Y) (:137-141),
    J = sum_X gq(X) B(X|..) (:153-170).  Summation order then no longer matters
    (the reference's batch loops only reorder these sums), so this is the
    order-free value the reference's FP64 sums scatter around; V^-1 amplifies
    that scatter by up to cond(V).

    B is a dense [naux, nbf, nbf] array, or a callable (x0, x1) -> the dense
    rows x0..x1-1 with x_rows giving (naux, rows per call), for tensors too
    large to hold dense."""
    ld = np.longdouble
    ptot = (2.0 * P_alpha if P_beta is None else P_alpha + P_beta).astype(ld)
    if callable(B):
        naux, step = x_rows
        rows = lambda x0, x1: B(x0, x1)
    else:
        naux, step = B.shape[0], max(1, B.shape[0])
        rows = lambda x0, x1: B[x0:x1]
    nbf = ptot.shape[0]
    gamma = np.zeros(naux, dtype=ld)
    for x0 in range(0, naux, step):
        x1 = min(naux, x0 + step)
        b = rows(x0, x1).reshape(x1 - x0, -1).astype(ld)
        gamma[x0:x1] = b @ ptot.reshape(-1)
    gq = gamma.astype(np.float64).astype(ld) @ np.asarray(Vinv, dtype=ld)
    gq = gq.astype(np.float64).astype(ld)
    J = np.zeros(nbf * nbf, dtype=ld)
    for x0 in range(0, naux, step):
        x1 = min(naux, x0 + step)
        b = rows(x0, x1).reshape(x1 - x0, -1).astype(ld)
        J += gq[x0:x1] @ b
    return J.astype(np.float64).reshape(nbf, nbf)


def dfmo_k(B, invsqrt, C_occ, bsz=None, xsz=None, osz=None, occ_nbatches=1,
           nbatches_b=5, nbatches_x=5, bmap=None, xmap=None):
    """DFMO_K::compute_K for one spin (jkdf.cpp:221-462); returns K (negated)."""
    naux, nbf, _ = B.shape
    nocc = C_occ.shape[1]
    bsz = bsz or [nbf]
    xsz = xsz or [naux]
    osz = osz or [nocc]
    nub = _elem_bounds(bsz, make_blk_bounds(bsz, nbatches_b, bmap))
    xb = _elem_bounds(xsz, make_blk_bounds(xsz, nbatches_x, xmap))
    ob = _elem_bounds(osz, make_blk_bounds(osz, occ_nbatches))
    K = np.zeros((nbf, nbf))
    for i0, i1 in ob:                                   # jkdf.cpp:334
        C = C_occ[:, i0:i1]
        HT1 = np.zeros((naux, i1 - i0, nbf))
        for n0, n1 in nub:                              # XMN, Mi -> XiN  (:361)
            HT1[:, :, n0:n1] += np.einsum("xmn,mi->xin", B[:, :, n0:n1], C, optimize=True)
        # XiN, XY -> YiN with XY = invsqrt (:384)
        HT2 = np.einsum("xin,xy->yin", HT1, invsqrt, optimize=True)
        for x0, x1 in xb:                               # XiM, XiN -> MN (:417)
            H = HT2[x0:x1].reshape(-1, nbf)
            K += H.T @ H
    return -K


def dfmo_k_rows(rows, naux, step, Linv, C_occ):
    """DFMO_K::compute_K (jkdf.cpp:221-462) for one spin with B supplied in
    X-row batches, for tensors too large to hold dense: HT1 = B C
    ("XMN, Mi -> XiN", :361), HT2 = L^-1 HT1 ("XiN, XY -> YiN" with
    invsqrt = L^-T, :384), K = -HT2^T HT2 ("XiM, XiN -> MN", :417).
    rows(x0, x1) returns the dense rows x0..x1-1 of B; Linv is the lower
    triangular L^-1 (row-major).  Returns K (negated)."""
    nbf, nocc = C_occ.shape
    HT1 = np.empty((naux, nocc, nbf))
    Ct = np.ascontiguousarray(C_occ.T)
    for x0 in range(0, naux, step):
        x1 = min(naux, x0 + step)
        HT1[x0:x1] = np.matmul(Ct, rows(x0, x1))          # sum_m C(m,i) B(x|m n)
    HT2 = (np.asarray(Linv) @ HT1.reshape(naux, -1)).reshape(naux * nocc, nbf)
    del HT1
    return -(HT2.T @ HT2)


def dflmo_k_uv(B, Vinv, U, V):
    """DFLMO_K non-symmetric branch (jkdf.cpp:1093-1331): htv = "Xmn, in -> Xmi"
    with v_mb = V (nocc x nbf), htu = "Xmn, ni -> Xmi" with u_bm = U (nbf x
    nocc), htvfit = "XY, Ymi -> Xmi" with V^-1, K = "Xmi, Xni -> mn" of htu and
    htvfit, negated (:1323).  Returns K (not symmetric)."""
    htu = np.einsum("xmn,ni->xmi", B, U, optimize=True)
    htv = np.einsum("xmn,in->xmi", B, V, optimize=True)
    htvfit = np.einsum("xy,ymi->xmi", Vinv, htv, optimize=True)
    return -np.einsum("xmi,xni->mn", htu, htvfit, optimize=True)


def fitting_coefficients(B, Vinv):
    """dfitting::compute: C_fit(X|mn) = sum_Y Vinv(X,Y) B(Y|mn) (fitting.cpp:124-128)."""
    naux, nbf, _ = B.shape
    return (Vinv @ B.reshape(naux, -1)).reshape(naux, nbf, nbf)


def dfao_k(B, Cfit, P, bsz=None, xsz=None, nbatches_b=5, nbatches_x=5, bmap=None, xmap=None):
    """DFAO_K::compute_K for one spin (jkdf.cpp:510-648); returns K (negated)."""
    naux, nbf, _ = B.shape
    bsz = bsz or [nbf]
    xsz = xsz or [naux]
    nub = _elem_bounds(bsz, make_blk_bounds(bsz, nbatches_b, bmap))
    xb = _elem_bounds(xsz, make_blk_bounds(xsz, nbatches_x, xmap))
    K = np.zeros((nbf, nbf))
    for x0, x1 in xb:
        for s0, s1 in nub:
            # XMN, NL -> XML with L restricted to the nu batch (:566)
            cbar = np.einsum("xmn,nl->xml", B[x0:x1], P[:, s0:s1], optimize=True)
            # XNS, XMS -> MN over X batch and S in the nu batch (:605)
            K += np.einsum("xns,xms->mn", Cfit[x0:x1, :, s0:s1], cbar, optimize=True)
    return -K


def dfmem_k(B, Vinv, P, bsz=None, xsz=None, nbatches_b=5, nbatches_x=5, bmap=None, xmap=None):
    """DFMEM_K (jkdf.cpp:719-859): C_fit rows formed per X batch as
    sum over Y batches of Vinv(X,Y) B(Y|..)  (XY,Ymr->Xmr; Xmr,rs->Xms; Xns,Xms->mn)."""
    naux, nbf, _ = B.shape
    bsz = bsz or [nbf]
    xsz = xsz or [naux]
    xb = _elem_bounds(xsz, make_blk_bounds(xsz, nbatches_x, xmap))
    nub = _elem_bounds(bsz, make_blk_bounds(bsz, nbatches_b, bmap))
    K = np.zeros((nbf, nbf))
    for x0, x1 in xb:
        cfit = np.zeros((x1 - x0, nbf, nbf))
        for y0, y1 in xb:
            cfit += np.einsum("xy,ymr->xmr", Vinv[x0:x1, y0:y1], B[y0:y1], optimize=True)
        for s0, s1 in nub:
            t = np.einsum("xmr,rs->xms", cfit, P[:, s0:s1], optimize=True)
            K += np.einsum("xns,xms->mn", B[x0:x1, :, s0:s1], t, optimize=True)
    return -K


def dflmo_k_sym(B, Vinv, C):
    """DFLMO_K symmetric branch (jkdf.cpp:890-1090): ht = B C over all columns of
    the given C, K = -sum ht (V^-1 ht)."""
    naux, nbf, _ = B.shape
    ht = np.einsum("xmn,mi->xin", B, C, optimize=True)
    vht = np.einsum("xy,yin->xin", Vinv, ht, optimize=True)
    K = np.einsum("xim,xin->mn", ht, vht, optimize=True)
    return -K
