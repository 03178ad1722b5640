"""Restricted Hartree-Fock driver (host side of hf::hfmod, src/hf/hfscf.cpp).

Integrals and descriptors are prepared here; the SCF loop itself (J, K, Fock
assembly, DIIS, diagonalisation) runs on the GPU inside libmchip.so
(mchip_scf_run).
"""
import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _lib, ints
from .engine import DeviceB, Metric

KMODES = {"dfmo": 0, "dfmo_transformed": 1, "mixed_f32": 2, "mixed_bf16": 3, "mixed_auto": 4}


@dataclass
class ScfResult:
    energy: float
    iterations: int
    converged: bool
    C: np.ndarray
    eps: np.ndarray
    history_e: list = field(default_factory=list)
    history_rms: list = field(default_factory=list)
    t_j: list = field(default_factory=list)
    t_k: list = field(default_factory=list)
    energy_iter: float = 0.0
    t_final_k: float = 0.0


def sad_density(mol):
    """SAD guess density (alpha, nbf x nbf) of the molecule: atomic fractional-
    occupation SCFs with exact integrals on the host (mchip_sad_guess,
    hfguess.cpp:142-452)."""
    Z, xyz = mol.charges()
    P = np.zeros((mol.nbf, mol.nbf))
    _lib.check(_lib.lib().mchip_sad_guess(mol.basis_handle.h, len(Z), _lib.dptr(Z), _lib.dptr(xyz), _lib.dptr(P)))
    return P


def run_rhf(mol, dB, metric, S, H, threshold=1e-8, max_iter=100, do_diis=True, diis_start=1,
            diis_min=2, diis_max=10, filter_eps=1e-10, kmode="dfmo", occ_nbatches=1, guess="core"):
    """guess: "core" (diagonalised core Hamiltonian) or "sad" (superposition of
    atomic densities, the reference's default)."""
    n = mol.nbf
    if guess == "sad":
        Pg = np.ascontiguousarray(sad_density(mol))
        _lib.check(_lib.lib().mchip_ctx_set_guess(dB.ctx.h, n, _lib.dptr(Pg)))
    elif guess != "core":
        raise ValueError("guess must be core or sad")
    bsz = np.array(mol.b, dtype=np.int32)
    S = np.ascontiguousarray(S, dtype=np.float64)
    H = np.ascontiguousarray(H, dtype=np.float64)
    C = np.zeros((n, n))
    eps = np.zeros(n)
    mh = max_iter + 3
    he, hr, tj, tk = (np.zeros(mh) for _ in range(4))
    e = ctypes.c_double()
    it = ctypes.c_int()
    conv = ctypes.c_int()
    nh = ctypes.c_int()
    m = None if metric is None else metric.h
    _lib.check(_lib.lib().mchip_scf_run(
        dB.h, m, n, mol.nocc, len(bsz), _lib.iptr(bsz), _lib.dptr(S), _lib.dptr(H),
        mol.nuclear_repulsion(), threshold, max_iter, int(do_diis), diis_start, diis_min,
        diis_max, filter_eps, KMODES[kmode], occ_nbatches, ctypes.byref(e), ctypes.byref(it),
        ctypes.byref(conv), _lib.dptr(C), _lib.dptr(eps), mh, _lib.dptr(he), _lib.dptr(hr),
        _lib.dptr(tj), _lib.dptr(tk), ctypes.byref(nh)))
    k = nh.value
    ei = ctypes.c_double()
    tf = ctypes.c_double()
    _lib.check(_lib.lib().mchip_scf_last_extra(ctypes.byref(ei), ctypes.byref(tf)))
    return ScfResult(e.value, it.value, bool(conv.value), C, eps, list(he[:k]), list(hr[:k]),
                     list(tj[:k]), list(tk[:k]), ei.value, tf.value)
