"""DF-HF request loop on one GPU (Python side of mchip_server_*).

The loop itself runs in C++ inside libmchip.so (worker threads with their own
streams, HBM-budget admission in arrival order, B and densities resident in
HBM for the life of a request).  This module submits molecules and collects
results; it has no counterpart in the reference, which runs one molecule per
process (src/main.cpp).
"""
import ctypes
import os

import numpy as np

from . import _lib
from .scf import KMODES


class Server:
    def __init__(self, device=0, budget_bytes=200e9, nslots=4, cpu_threads=None):
        if cpu_threads is None:
            cpu_threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().mchip_server_create(device, float(budget_bytes), nslots, cpu_threads,
                                                  ctypes.byref(h)))
        self.h = h
        self._keep = {}

    def submit(self, mol, threshold=1e-6, kmode="mixed_auto", screen_eps=1e-10):
        """Queue one RHF request; returns its id."""
        Z, xyz = mol.charges()
        bsh = np.array([len(b) for b in mol.blocks], dtype=np.int32)
        xsh = np.array([len(b) for b in mol.dfblocks], dtype=np.int32)
        rid = ctypes.c_int()
        _lib.check(_lib.lib().mchip_server_submit(
            self.h, mol.basis_handle.h, mol.dfbasis_handle.h, len(Z), _lib.dptr(Z), _lib.dptr(xyz),
            len(bsh), _lib.iptr(bsh), len(xsh), _lib.iptr(xsh), mol.nocc, mol.nuclear_repulsion(),
            threshold, KMODES[kmode], screen_eps, ctypes.byref(rid)))
        return rid.value

    def result(self, rid, wait=True):
        """dict with energy, iterations, converged, t_submit/t_admit/t_done
        (server clock, s), latency, est_bytes, slot; None if not done."""
        done = ctypes.c_int()
        e = ctypes.c_double()
        it = ctypes.c_int()
        conv = ctypes.c_int()
        t = np.zeros(10)
        _lib.check(_lib.lib().mchip_server_result(self.h, rid, int(wait), ctypes.byref(done), ctypes.byref(e),
                                                  ctypes.byref(it), ctypes.byref(conv), _lib.dptr(t)))
        if not done.value:
            return None
        return dict(id=rid, energy=e.value, iterations=it.value, converged=bool(conv.value),
                    t_submit=t[0], t_admit=t[1], t_done=t[2], latency=t[2] - t[0],
                    est_bytes=t[3], slot=int(t[4]), energy_iter=t[5], t_host=t[6], t_build=t[7],
                    t_setup=t[8], t_scf=t[9])

    def stats(self):
        s = np.zeros(6)
        _lib.check(_lib.lib().mchip_server_stats(self.h, _lib.dptr(s)))
        return dict(bytes_in_use=s[0], peak_bytes=s[1], completed=int(s[2]), failed=int(s[3]),
                    max_live=int(s[4]), clock=s[5])

    def set_batching(self, on):
        """Cross-request batching of the BF16x3 exchange builds (on by default)."""
        _lib.check(_lib.lib().mchip_server_set_batching(self.h, int(bool(on))))

    def batch_stats(self):
        s = np.zeros(3)
        _lib.check(_lib.lib().mchip_server_batch_stats(self.h, _lib.dptr(s)))
        return dict(batches=int(s[0]), builds=int(s[1]), largest=int(s[2]))

    def clock(self):
        return self.stats()["clock"]

    def close(self):
        """Finish every submitted request, then stop the workers."""
        if getattr(self, "h", None):
            _lib.check(_lib.lib().mchip_server_destroy(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
