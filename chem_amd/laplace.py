"""Laplace quadrature weights and exponents for SOS-AO-MP2.

Follows the reference's choice of starting values (lap_rddata.F90: the largest
tabulated range R0 <= R for the requested number of points, else the
smallest) and its scaling of [1, R] to [ymin, ymax] with
ymin = 2 (e_LUMO - e_HOMO), ymax = 2 (e_max - e_min) (mp_compute.cpp:64-69);
the Remez solve itself runs in libmchip.so (mchip_laplace_minimax).
"""
import ctypes
import json
import os

import numpy as np

from . import _lib

_TABLE = None


def table():
    global _TABLE
    if _TABLE is None:
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "laplace_init.json")
        with open(path) as f:
            _TABLE = json.load(f)["orders"]
    return _TABLE


def start_values(n, R):
    """lap_rddata.F90: largest tabulated R0 <= R, else the smallest R0."""
    entries = sorted(table()[str(n)], key=lambda e: e["R"])
    pick = None
    for e in entries:
        if e["R"] <= R:
            pick = e
    if pick is None:
        pick = entries[0]
    w0 = np.array([float(v) for v in pick["omega"]])
    a0 = np.array([float(v) for v in pick["alpha"]])
    return pick["R"], w0, a0, pick


def minimax_unit(n, R, maxit=60):
    """Best exponential sum for 1/x on [1, R]: (w, a, max error, levelling)."""
    R0, w0, a0, _ = start_values(n, R)
    w = np.zeros(n)
    a = np.zeros(n)
    err = ctypes.c_double()
    equi = ctypes.c_double()
    _lib.check(_lib.lib().mchip_laplace_minimax(n, float(R), float(R0), _lib.dptr(w0), _lib.dptr(a0),
                                                maxit, _lib.dptr(w), _lib.dptr(a), ctypes.byref(err),
                                                ctypes.byref(equi)))
    return w, a, err.value, equi.value


def laplace_points(n, ymin, ymax):
    """omega, alpha with 1/y ~ sum_k omega_k exp(-alpha_k y) on [ymin, ymax]."""
    w, a, err, _ = minimax_unit(n, ymax / ymin)
    return w / ymin, a / ymin, err / ymin
