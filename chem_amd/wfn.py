"""HDF5 wavefunction files in the reference's layout (libmchip_wfn.so).

The reference's driver writes, per job tag, the molecule (write_molecule,
src/desc/molecule.hpp:322-352) and the SCF result (write_hfwfn,
src/desc/wfn.hpp:266-289) into its output file, and its MP2 driver reads an
SCF result back (read_hfwfn, wfn.hpp:291-326).  The same files are written and
read here, so an SCF run on the GPU can feed the reference's MP2 driver and
the reference's SCF output can feed this package's MP2.
"""
import ctypes
import os
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _lib

_WFN = None
WFN_LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libmchip_wfn.so")


def lib():
    global _WFN
    if _WFN is None:
        if not os.path.exists(WFN_LIB):
            raise RuntimeError("libmchip_wfn.so is not built (chem_amd.build.build_wfn needs the HDF5 C library)")
        L = ctypes.CDLL(WFN_LIB)
        c_int, c_double, c_char_p = ctypes.c_int, ctypes.c_double, ctypes.c_char_p
        P = ctypes.POINTER(c_double)
        Pi = ctypes.POINTER(c_int)
        L.mchip_wfn_last_error.restype = c_char_p
        L.mchip_wfn_last_error.argtypes = []
        sigs = {
            "mchip_wfn_create": [c_char_p],
            "mchip_wfn_write_molecule": [c_char_p, c_char_p, c_char_p, c_int, c_int, c_int, c_int, P, P],
            "mchip_wfn_write_hfwfn": [c_char_p, c_char_p, c_int, c_int, c_int, P, P, P, P, c_int, c_int, P, P, P, P,
                                      c_double, c_double, c_double],
            "mchip_wfn_read_hfwfn": [c_char_p, c_char_p, c_int, c_int, c_int, c_int, c_int, P, P, P, P, P, P, P, P,
                                     Pi, P],
        }
        for name, args in sigs.items():
            f = getattr(L, name)
            f.restype = c_int
            f.argtypes = args
        _WFN = L
    return _WFN


def _check(rc):
    if rc != 0:
        raise RuntimeError(lib().mchip_wfn_last_error().decode())


def _p(a):
    return None if a is None else _lib.dptr(a)


def _b(s):
    return os.fspath(s).encode()


@dataclass
class HfWavefunction:
    """desc::hf_wavefunction (wfn.hpp:21-111): occupied / virtual coefficients
    (nbf x n) and orbital energies per spin (beta None when restricted), the
    electronic SCF energy, the nuclear repulsion and their sum."""
    c_bo_A: np.ndarray
    c_bv_A: np.ndarray
    eps_occ_A: np.ndarray
    eps_vir_A: np.ndarray
    scf_energy: float
    nuc_energy: float
    wfn_energy: float
    c_bo_B: Optional[np.ndarray] = None
    c_bv_B: Optional[np.ndarray] = None
    eps_occ_B: Optional[np.ndarray] = None
    eps_vir_B: Optional[np.ndarray] = None

    @classmethod
    def from_scf(cls, mol, res):
        """The wavefunction hfmod::compute returns (hfscf.cpp:475-552) from
        an SCF result: columns and energies split at nocc."""
        o = mol.nocc
        C = np.ascontiguousarray(res.C)
        enuc = mol.nuclear_repulsion()
        return cls(np.ascontiguousarray(C[:, :o]), np.ascontiguousarray(C[:, o:]), np.array(res.eps[:o]),
                   np.array(res.eps[o:]), res.energy - enuc, enuc, res.energy)

    def coefficients(self):
        """All MO columns, occupied first (the C the SCF returned)."""
        return np.hstack([self.c_bo_A, self.c_bv_A])

    def eps(self):
        return np.concatenate([self.eps_occ_A, self.eps_vir_A])


def create(path):
    """Create (truncate) an output file (data_handler, create_mode::truncate)."""
    _check(lib().mchip_wfn_create(_b(path)))


def write_molecule(path, tag, mol, name="molecule", mo_split=5):
    """write_molecule: name, charge, mult, mo_split and the atoms (Z, x, y, z
    in bohr) of a closed-shell molecule."""
    Z, xyz = mol.charges()
    charge = int(round(Z.sum())) - 2 * mol.nocc
    _check(lib().mchip_wfn_write_molecule(_b(path), tag.encode(), name.encode(), charge, 1, mo_split, len(Z),
                                          _lib.dptr(Z), _lib.dptr(xyz)))


def write_hfwfn(path, tag, wfn):
    nbf = wfn.c_bo_A.shape[0]
    arr = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)
    a = [arr(x) for x in (wfn.c_bo_A, wfn.c_bv_A, wfn.eps_occ_A, wfn.eps_vir_A)]
    b = [arr(x) for x in (wfn.c_bo_B, wfn.c_bv_B, wfn.eps_occ_B, wfn.eps_vir_B)]
    nob = 0 if b[0] is None else b[0].shape[1]
    nvb = 0 if b[1] is None else b[1].shape[1]
    _check(lib().mchip_wfn_write_hfwfn(_b(path), tag.encode(), nbf, a[0].shape[1], a[1].shape[1], *map(_p, a),
                                       nob, nvb, *map(_p, b), wfn.scf_energy, wfn.nuc_energy, wfn.wfn_energy))


def read_hfwfn(path, tag, nbf, nocc_a, nocc_b=None):
    """read_hfwfn with the molecule's dimensions (nbf functions, nocc_a /
    nocc_b occupied; nvir = nbf - nocc)."""
    nocc_b = nocc_a if nocc_b is None else nocc_b
    nva, nvb = nbf - nocc_a, nbf - nocc_b
    a = [np.zeros((nbf, nocc_a)), np.zeros((nbf, nva)), np.zeros(nocc_a), np.zeros(nva)]
    b = [np.zeros((nbf, nocc_b)), np.zeros((nbf, nvb)), np.zeros(nocc_b), np.zeros(nvb)]
    present = ctypes.c_int()
    en = np.zeros(3)
    _check(lib().mchip_wfn_read_hfwfn(_b(path), tag.encode(), nbf, nocc_a, nva, nocc_b, nvb, *map(_p, a),
                                      *map(_p, b), ctypes.byref(present), _lib.dptr(en)))
    beta = [b[i] if present.value & (1 << i) else None for i in range(4)]
    return HfWavefunction(a[0], a[1], a[2], a[3], float(en[0]), float(en[1]), float(en[2]), *beta)
