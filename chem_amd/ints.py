"""AO integrals and the screened three-centre tensor (host side).

Mirrors the registry keys the reference's aoloader produces for the DF Fock
path (src/ints/aoloader.cpp:34-537): ovlp_bb, kin_bb, pot_bb, coul_xx,
scr_xbb (Schwarz screener) and coul_xbb (B).
"""
import ctypes

import numpy as np

from . import _lib
from .desc import block_norms, schwarz_block_mask


def int1e(mol):
    L = _lib.lib()
    n = mol.nbf
    S = np.zeros((n, n))
    T = np.zeros((n, n))
    V = np.zeros((n, n))
    Z, xyz = mol.charges()
    _lib.check(L.mchip_int1e(mol.basis_handle.h, len(mol.atoms), _lib.dptr(Z), _lib.dptr(xyz),
                             _lib.dptr(S), _lib.dptr(T), _lib.dptr(V)))
    return S, T, V


def int2c2e(mol):
    n = mol.naux
    M = np.zeros((n, n))
    _lib.check(_lib.lib().mchip_int2c2e(mol.dfbasis_handle.h, _lib.dptr(M)))
    return M


def int1e_gpu(ctx, mol):
    """S, T, V of int1e computed on the context's GPU."""
    n = mol.nbf
    S, T, V = np.zeros((n, n)), np.zeros((n, n)), np.zeros((n, n))
    Z, xyz = mol.charges()
    _lib.check(_lib.lib().mchip_int1e_gpu(ctx.h, mol.basis_handle.h, len(mol.atoms), _lib.dptr(Z),
                                          _lib.dptr(xyz), _lib.dptr(S), _lib.dptr(T), _lib.dptr(V)))
    return S, T, V


def int2c2e_gpu(ctx, mol):
    """The Coulomb metric of int2c2e computed on the context's GPU."""
    n = mol.naux
    M = np.zeros((n, n))
    _lib.check(_lib.lib().mchip_int2c2e_gpu(ctx.h, mol.dfbasis_handle.h, _lib.dptr(M)))
    return M


def schwarz_ao_gpu(ctx, mol):
    """AO Schwarz factors (nshell x nshell) computed on the context's GPU."""
    n = len(mol.shells)
    Q = np.zeros((n, n))
    _lib.check(_lib.lib().mchip_schwarz_ao_gpu(ctx.h, mol.basis_handle.h, _lib.dptr(Q)))
    return Q


def int4c2e(mol):
    """Exact four-centre ERIs (ij|kl) over the AO basis, nbf^4 row-major
    (host; the SAD guess's atomic SCFs use them)."""
    n = mol.nbf
    out = np.zeros((n, n, n, n))
    _lib.check(_lib.lib().mchip_int4c2e(mol.basis_handle.h, _lib.dptr(out)))
    return out


def schwarz(mol):
    Qao = np.zeros((len(mol.shells), len(mol.shells)))
    Qx = np.zeros(len(mol.dfshells))
    _lib.check(_lib.lib().mchip_schwarz(mol.basis_handle.h, mol.dfbasis_handle.h,
                                        _lib.dptr(Qao), _lib.dptr(Qx)))
    return Qao, Qx


def int3c2e_shell(mol, ia, ib, ic):
    la, lb, lc = mol.shells[ia].l, mol.shells[ib].l, mol.dfshells[ic].l
    out = np.zeros((2 * lc + 1, 2 * la + 1, 2 * lb + 1))
    _lib.check(_lib.lib().mchip_int3c2e_shell(mol.basis_handle.h, mol.dfbasis_handle.h,
                                              ia, ib, ic, _lib.dptr(out)))
    return out


def schwarz_mask(mol, eps):
    """Block mask of the Schwarz screener (screening.cpp:8-30), computed by the
    library (mchip_schwarz_mask, the same code the C++ layer uses).  Returns
    (mask, block norms of the AO pair factors, block norms of the aux
    factors)."""
    nbb, nxb = len(mol.blocks), len(mol.dfblocks)
    bsh = np.array([len(b) for b in mol.blocks], dtype=np.int32)
    xsh = np.array([len(b) for b in mol.dfblocks], dtype=np.int32)
    mask = np.zeros((nxb, nbb, nbb), dtype=np.uint8)
    nmn = np.zeros((nbb, nbb))
    nx = np.zeros(nxb)
    _lib.check(_lib.lib().mchip_schwarz_mask(mol.basis_handle.h, mol.dfbasis_handle.h, nbb, _lib.iptr(bsh), nxb,
                                             _lib.iptr(xsh), float(eps), _lib.u8ptr(mask), _lib.dptr(nmn),
                                             _lib.dptr(nx)))
    return mask, nmn, nx


class BTensorHost:
    """Host B(X|mu nu) in pair-slab layout (owner of an mchip_btensor)."""

    def __init__(self, handle, xsz, bsz):
        self.h = handle
        self.xsz = list(xsz)
        self.bsz = list(bsz)
        dims = (ctypes.c_int * 4)()
        ne = ctypes.c_int64()
        _lib.check(_lib.lib().mchip_btensor_info(handle, dims, ctypes.byref(ne)))
        self.naux, self.nbf, self.nxb, self.nbb = dims[0], dims[1], dims[2], dims[3]
        self.nelem = ne.value

    @classmethod
    def build(cls, mol, eps, mask=None):
        if mask is None:
            mask, _, _ = schwarz_mask(mol, eps)
        mask = np.ascontiguousarray(mask, dtype=np.uint8)
        bsh = np.array([len(b) for b in mol.blocks], dtype=np.int32)
        xsh = np.array([len(b) for b in mol.dfblocks], dtype=np.int32)
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().mchip_btensor_build(
            mol.basis_handle.h, mol.dfbasis_handle.h, len(bsh), _lib.iptr(bsh), len(xsh),
            _lib.iptr(xsh), _lib.u8ptr(mask), float(eps), ctypes.byref(h)))
        return cls(h, mol.x, mol.b)

    @classmethod
    def from_dense(cls, dense, xsz, bsz, mask):
        dense = np.ascontiguousarray(dense, dtype=np.float64)
        mask = np.ascontiguousarray(mask, dtype=np.uint8)
        b = np.array(bsz, dtype=np.int32)
        x = np.array(xsz, dtype=np.int32)
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().mchip_btensor_from_dense(len(b), _lib.iptr(b), len(x), _lib.iptr(x),
                                                       _lib.u8ptr(mask), _lib.dptr(dense),
                                                       ctypes.byref(h)))
        return cls(h, xsz, bsz)

    def mask(self):
        m = np.zeros((self.nxb, self.nbb, self.nbb), dtype=np.uint8)
        _lib.check(_lib.lib().mchip_btensor_mask(self.h, _lib.u8ptr(m)))
        return m

    def dense(self):
        out = np.zeros((self.naux, self.nbf, self.nbf))
        _lib.check(_lib.lib().mchip_btensor_to_dense(self.h, _lib.dptr(out)))
        return out

    def dense_rows(self, x0, x1):
        """Dense B(X|mu nu) for aux functions x0 <= X < x1."""
        out = np.zeros((x1 - x0, self.nbf, self.nbf))
        _lib.check(_lib.lib().mchip_btensor_dense_rows(self.h, x0, x1, _lib.dptr(out)))
        return out

    def nnz_kept(self):
        """Elements inside kept blocks (the algorithmic nnz of B)."""
        m = self.mask().astype(np.int64)
        return int(np.einsum("x,xmn,m,n->", np.array(self.xsz), m, np.array(self.bsz),
                             np.array(self.bsz)))

    def __del__(self):
        try:
            if getattr(self, "h", None):
                _lib.lib().mchip_btensor_destroy(self.h)
                self.h = None
        except Exception:
            pass
