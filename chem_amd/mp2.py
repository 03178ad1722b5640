"""SOS-AO-MP2 driver (host side of mp::mpmod, src/mp/mp_compute.cpp).

Laplace points from the minimax solver (chem_amd.laplace) on
[2(e_LUMO - e_HOMO), 2(e_max - e_min)] (mp_compute.cpp:64-69); the
pseudo-densities, pivoted Cholesky, Z builder and energy run on the GPU
(mchip_mp2_sos).
"""
import ctypes

import numpy as np

from . import _lib, laplace

C_OS = 1.3   # mpmod.hpp:26


def laplace_for(eps, nocc, nlap):
    ymin = 2.0 * (eps[nocc] - eps[nocc - 1])
    ymax = 2.0 * (eps[-1] - eps[0])
    w, a, err = laplace.laplace_points(nlap, ymin, ymax)
    return w, a, err


ZMETHODS = {"llmp_full": 0, "llmp_mem": 1, "ll_full": 2}   # mp::zmethod (z_builder.hpp)


def sos_mp2(mol, dB, metric, S, C, eps, nlap=5, mo_split=5, filter_eps=1e-10, points=None, batched=False,
            zmethod="llmp_full"):
    """Returns dict(energy (positive sum), scaled (1.3 x), terms, omega, alpha, t_z, t_prep).
    zmethod: "llmp_full" / "llmp_mem" (one GPU path) or "ll_full" (FP64 only).

    batched=False: FP64, one launch sequence per Laplace point.  batched=True:
    the contractions of all points as grouped FP32 launches (FP64
    accumulation) captured in one hipGraph; t_z is the graph's time."""
    n, nocc = mol.nbf, mol.nocc
    if points is None:
        w, a, _ = laplace_for(eps, nocc, nlap)
    else:
        w, a = points
    w = np.ascontiguousarray(w, dtype=np.float64)
    a = np.ascontiguousarray(a, dtype=np.float64)
    S = np.ascontiguousarray(S, dtype=np.float64)
    C = np.ascontiguousarray(C, dtype=np.float64)
    eps = np.ascontiguousarray(eps, dtype=np.float64)
    bsz = np.array(mol.b, dtype=np.int32)
    e = ctypes.c_double()
    tz = ctypes.c_double()
    tp = ctypes.c_double()
    terms = np.zeros(len(w))
    args = (dB.h, metric.h, n, nocc, _lib.dptr(S), _lib.dptr(C), _lib.dptr(eps), len(w), _lib.dptr(w),
            _lib.dptr(a), len(bsz), _lib.iptr(bsz), mo_split, filter_eps, ctypes.byref(e), _lib.dptr(terms),
            ctypes.byref(tz))
    zm = ZMETHODS[zmethod]
    if batched:
        if zm == 2:
            raise ValueError("the grouped FP32 graph implements the LLMP Z builders")
        _lib.check(_lib.lib().mchip_mp2_sos_batched(*args, ctypes.byref(tp)))
    elif zm == 0:
        _lib.check(_lib.lib().mchip_mp2_sos(*args))
    else:
        _lib.check(_lib.lib().mchip_mp2_sos_z(dB.h, metric.h, zm, *args[2:]))
    return dict(energy=e.value, scaled=C_OS * e.value, terms=terms, omega=w, alpha=a, t_z=tz.value,
                t_prep=tp.value)


def z_builder(mol, dB, metric, S, C, eps, omega, alpha, mo_split=5, filter_eps=1e-10, batched=False):
    """Z_k(X,Y) of every Laplace point (nlap x naux x naux) and the terms e_k:
    the tensor LLMP_FULL_Z hands to the energy loop (z_full.cpp:373-377,
    mp_compute.cpp:361-389).  batched selects the grouped FP32 graph."""
    w = np.ascontiguousarray(omega, dtype=np.float64)
    a = np.ascontiguousarray(alpha, dtype=np.float64)
    S = np.ascontiguousarray(S, dtype=np.float64)
    C = np.ascontiguousarray(C, dtype=np.float64)
    eps = np.ascontiguousarray(eps, dtype=np.float64)
    bsz = np.array(mol.b, dtype=np.int32)
    naux = dB.naux
    Z = np.zeros((len(w), naux, naux))
    terms = np.zeros(len(w))
    _lib.check(_lib.lib().mchip_mp2_z(dB.h, metric.h, mol.nbf, mol.nocc, _lib.dptr(S), _lib.dptr(C), _lib.dptr(eps),
                                      len(w), _lib.dptr(w), _lib.dptr(a), len(bsz), _lib.iptr(bsz), mo_split,
                                      filter_eps, int(bool(batched)), _lib.dptr(Z), _lib.dptr(terms)))
    return Z, terms


def z_compute(dB, Locc, Pvir=None, Lvir=None, zmethod="llmp_full"):
    """One Z (naux x naux) from given factors, FP64 (mp::Z::compute,
    z_builder.hpp:43-116): LLMP builders take the virtual pseudo-density Pvir,
    LL_FULL its Cholesky factor Lvir."""
    zm = ZMETHODS[zmethod]
    Lo = np.ascontiguousarray(Locc, dtype=np.float64)
    n, ro = Lo.shape
    Pv = None if Pvir is None else np.ascontiguousarray(Pvir, dtype=np.float64)
    Lv = None if Lvir is None else np.ascontiguousarray(Lvir, dtype=np.float64)
    rv = 0 if Lv is None else Lv.shape[1]
    Z = np.zeros((dB.naux, dB.naux))
    _lib.check(_lib.lib().mchip_z_compute(dB.h, zm, n, _lib.dptr(Lo), ro, _lib.dptr(Pv), _lib.dptr(Lv), rv,
                                          _lib.dptr(Z)))
    return Z
