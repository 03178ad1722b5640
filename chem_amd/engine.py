"""Device-side objects: context, resident B tensor, Coulomb metric.

Thin owners of the C ABI handles (include/mchip.h); every compute call goes to
libmchip.so.
"""
import ctypes
import weakref

import numpy as np

from . import _lib


def device_count():
    n = ctypes.c_int()
    _lib.check(_lib.lib().mchip_device_count(ctypes.byref(n)))
    return n.value


class Context:
    """One GPU, one HIP stream, one rocBLAS handle.

    Objects created on a context (DeviceB, Metric) register here and are
    released before the context itself."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().mchip_ctx_create(device, ctypes.byref(h)))
        self.h = h
        self.device = device
        self._children = weakref.WeakSet()

    def _adopt(self, obj):
        self._children.add(obj)

    def sync(self):
        _lib.check(_lib.lib().mchip_ctx_sync(self.h))

    def set_timing(self, on):
        _lib.check(_lib.lib().mchip_ctx_set_timing(self.h, int(bool(on))))

    def set_graphs(self, on):
        """Replay repeated exchange and DF-J builds from captured hipGraphs
        (default on); off: direct launches."""
        _lib.check(_lib.lib().mchip_ctx_set_graphs(self.h, int(bool(on))))

    def set_escalation(self, rms=1e-5, stall=3):
        """When the escalation K mode switches BF16x3 -> FP32: at an error RMS
        <= rms (0: off) or after `stall` iterations without progress (0: off)."""
        _lib.check(_lib.lib().mchip_ctx_set_escalation(self.h, float(rms), int(stall)))

    def trim(self):
        """Release the work arrays kept for reuse by earlier builds."""
        _lib.check(_lib.lib().mchip_ctx_trim(self.h))

    def phase_times(self):
        """Accumulated seconds: K1, K2, K3, J, K_inc, split, -, -."""
        out = np.zeros(8)
        _lib.check(_lib.lib().mchip_ctx_phase_times(self.h, _lib.dptr(out), 8))
        return out

    @property
    def stream(self):
        s = ctypes.c_void_p()
        _lib.check(_lib.lib().mchip_ctx_stream(self.h, ctypes.byref(s)))
        return s.value

    def close(self):
        if getattr(self, "h", None):
            for ch in list(self._children):
                ch.release()
            _lib.lib().mchip_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceArray:
    """A float64 array in the context's GPU memory (for the _dev exchange entry
    points, which keep the occupied coefficients and K resident)."""

    def __init__(self, ctx, shape, host=None):
        self.ctx = ctx
        self.shape = tuple(shape)
        self.nbytes = int(np.prod(self.shape)) * 8
        p = ctypes.c_void_p()
        _lib.check(_lib.lib().mchip_dev_alloc(ctx.h, self.nbytes, ctypes.byref(p)))
        self.ptr = p
        if host is not None:
            self.upload(host)

    def upload(self, host):
        a = np.ascontiguousarray(host, dtype=np.float64)
        if a.shape != self.shape:
            raise ValueError("DeviceArray.upload: shape %s, expected %s" % (a.shape, self.shape))
        _lib.check(_lib.lib().mchip_dev_copy_in(self.ctx.h, self.ptr, a.ctypes.data_as(ctypes.c_void_p),
                                                self.nbytes))

    def download(self):
        a = np.empty(self.shape)
        _lib.check(_lib.lib().mchip_dev_copy_out(self.ctx.h, a.ctypes.data_as(ctypes.c_void_p), self.ptr,
                                                 self.nbytes))
        return a

    def release(self):
        if getattr(self, "ptr", None) is not None and self.ptr.value:
            _lib.lib().mchip_dev_free(self.ctx.h, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


class Metric:
    """V = L L^T, L^-1 and V^-1 resident on the GPU."""

    def __init__(self, ctx, V):
        V = np.ascontiguousarray(V, dtype=np.float64)
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().mchip_metric_create(ctx.h, V.shape[0], _lib.dptr(V), ctypes.byref(h)))
        self.h = h
        self.ctx = ctx
        self.naux = V.shape[0]
        ctx._adopt(self)

    def get(self):
        n = self.naux
        Linv_cm = np.zeros((n, n))
        Vinv = np.zeros((n, n))
        _lib.check(_lib.lib().mchip_metric_get(self.ctx.h, self.h, _lib.dptr(Linv_cm), _lib.dptr(Vinv)))
        # column-major storage -> row-major lower-triangular L^-1
        return Linv_cm.T.copy(), Vinv

    def release(self):
        if getattr(self, "h", None):
            _lib.lib().mchip_metric_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


class DeviceB:
    """B(X|mu nu) resident in HBM, pair-slab layout."""

    def __init__(self, ctx, host_b=None, _handle=None):
        if _handle is None:
            h = ctypes.c_void_p()
            _lib.check(_lib.lib().mchip_dbt_upload(ctx.h, host_b.h, ctypes.byref(h)))
        else:
            h = _handle
        self.h = h
        self.ctx = ctx
        dims = (ctypes.c_int * 4)()
        ne = ctypes.c_int64()
        _lib.check(_lib.lib().mchip_dbt_info(h, dims, ctypes.byref(ne)))
        self.naux, self.nbf, self.nxb, self.nbb = dims[0], dims[1], dims[2], dims[3]
        self.nelem = ne.value
        self.transformed = False
        if host_b is not None:
            self.xsz, self.bsz = list(host_b.xsz), list(host_b.bsz)
        ctx._adopt(self)

    @classmethod
    def build_on_gpu(cls, ctx, mol, eps, mask=None):
        """Three-centre integrals computed, screened and laid out on the GPU."""
        from . import ints
        if mask is None:
            mask, _, _ = ints.schwarz_mask(mol, eps)
        mask = np.ascontiguousarray(mask, dtype=np.uint8)
        bsh = np.array([len(b) for b in mol.blocks], dtype=np.int32)
        xsh = np.array([len(b) for b in mol.dfblocks], dtype=np.int32)
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().mchip_dbt_build_gpu(ctx.h, mol.basis_handle.h, mol.dfbasis_handle.h,
                                                  len(bsh), _lib.iptr(bsh), len(xsh), _lib.iptr(xsh),
                                                  _lib.u8ptr(mask), float(eps), ctypes.byref(h)))
        obj = cls(ctx, _handle=h)
        obj.xsz, obj.bsz = list(mol.x), list(mol.b)
        return obj

    def mask(self):
        m = np.zeros((self.nxb, self.nbb, self.nbb), dtype=np.uint8)
        _lib.check(_lib.lib().mchip_dbt_mask(self.h, _lib.u8ptr(m)))
        return m

    def graph_stats(self):
        """(captured graphs held by this tensor, builds that replayed one)."""
        n, r = ctypes.c_int64(), ctypes.c_int64()
        _lib.check(_lib.lib().mchip_dbt_graph_stats(self.h, ctypes.byref(n), ctypes.byref(r)))
        return n.value, r.value

    def nnz_kept(self):
        """Elements inside kept blocks (the algorithmic nnz of B)."""
        m = self.mask().astype(np.int64)
        return int(np.einsum("x,xmn,m,n->", np.array(self.xsz), m, np.array(self.bsz),
                             np.array(self.bsz)))

    def dense_rows(self, x0, x1):
        out = np.zeros((x1 - x0, self.nbf, self.nbf))
        _lib.check(_lib.lib().mchip_dbt_dense_rows(self.h, x0, x1, _lib.dptr(out)))
        return out

    def transform(self, metric):
        _lib.check(_lib.lib().mchip_dbt_transform(self.h, metric.h))
        self.transformed = True

    def make_f32(self):
        _lib.check(_lib.lib().mchip_dbt_make_f32(self.h))

    def dfk_mo_f32(self, C_occ, occ_nbatches=1):
        C = np.ascontiguousarray(C_occ, dtype=np.float64)
        K = np.zeros((self.nbf, self.nbf))
        _lib.check(_lib.lib().mchip_dfk_mo_f32(self.h, _lib.dptr(C), C.shape[1], occ_nbatches,
                                               _lib.dptr(K)))
        return K

    def make_lowp(self, metric, f32=True, bf16=True, f64=False):
        """FP32 / BF16 split copies of L^-1 B made from the untransformed B, which
        stays FP64 for J and the energy (reference formulas).  f64 also keeps the
        FP64 L^-1 B slabs, which the FP64 exchanges of the mixed modes then use."""
        _lib.check(_lib.lib().mchip_dbt_make_lowp(self.h, metric.h, int(f32), int(bf16), int(f64)))

    def dfk_mo_tilde64(self, C_occ, occ_nbatches=1):
        """K from the FP64 L^-1 B copy (make_lowp(..., f64=True))."""
        C = np.ascontiguousarray(C_occ, dtype=np.float64)
        K = np.zeros((self.nbf, self.nbf))
        _lib.check(_lib.lib().mchip_dfk_mo_tilde64(self.h, _lib.dptr(C), C.shape[1], occ_nbatches, _lib.dptr(K)))
        return K

    def make_bf16(self):
        """BF16 hi/lo split copy of the transformed B (chosen-form exchange)."""
        _lib.check(_lib.lib().mchip_dbt_make_bf16(self.h))

    def dfk_mo_bf16(self, C_occ, occ_nbatches=1):
        C = np.ascontiguousarray(C_occ, dtype=np.float64)
        K = np.zeros((self.nbf, self.nbf))
        _lib.check(_lib.lib().mchip_dfk_mo_bf16(self.h, _lib.dptr(C), C.shape[1], occ_nbatches,
                                                _lib.dptr(K)))
        return K

    def dfk_dev(self, form, C_dev, nocc, K_dev, metric=None, occ_nbatches=1):
        """K into the DeviceArray K_dev (nbf x nbf) from the DeviceArray C_dev
        (nbf x ldc, first nocc columns), without host copies or a
        synchronisation: form "fp64" (plain, with the metric unless B is
        transformed), "fp64_copy" (the FP64 copy of L^-1 B), "f32" or "bf16x3"."""
        # the kernels read nbf x ldc and write nbf x nbf through raw device
        # pointers: check the arrays on the host before any launch
        for name, a in (("C_dev", C_dev), ("K_dev", K_dev)):
            if getattr(a, "ptr", None) is None or not a.ptr.value:
                raise ValueError("dfk_dev: %s has been released" % name)
            if len(a.shape) != 2 or a.shape[0] != self.nbf:
                raise ValueError("dfk_dev: %s has shape %s, expected (%d, n)" % (name, a.shape, self.nbf))
        if K_dev.shape != (self.nbf, self.nbf):
            raise ValueError("dfk_dev: K_dev has shape %s, expected (%d, %d)" % (K_dev.shape, self.nbf, self.nbf))
        ldc = C_dev.shape[1]
        if not 0 <= nocc <= ldc:
            raise ValueError("dfk_dev: nocc %d outside 0..%d (columns of C_dev)" % (nocc, ldc))
        if form not in ("fp64", "fp64_copy", "f32", "bf16x3"):
            raise ValueError("dfk_dev: unknown form %r" % (form,))
        L = _lib.lib()
        if form == "fp64":
            m = None if metric is None else metric.h
            rc = L.mchip_dfk_mo_dev(self.h, m, C_dev.ptr, ldc, nocc, occ_nbatches, K_dev.ptr)
        else:
            fn = {"fp64_copy": L.mchip_dfk_mo_tilde64_dev, "f32": L.mchip_dfk_mo_f32_dev,
                  "bf16x3": L.mchip_dfk_mo_bf16_dev}[form]
            rc = fn(self.h, C_dev.ptr, ldc, nocc, occ_nbatches, K_dev.ptr)
        _lib.check(rc)

    def exchange_energy_dev(self, C_dev, nocc):
        """tr(K P) of the FP64 exchange on the FP64 copy of L^-1 B at the
        occupied columns of the DeviceArray C_dev (P = C C^T), without forming
        K (synchronises)."""
        if getattr(C_dev, "ptr", None) is None or not C_dev.ptr.value:
            raise ValueError("exchange_energy_dev: C_dev has been released")
        if len(C_dev.shape) != 2 or C_dev.shape[0] != self.nbf or not 0 <= nocc <= C_dev.shape[1]:
            raise ValueError("exchange_energy_dev: C_dev shape %s / nocc %d" % (C_dev.shape, nocc))
        e = ctypes.c_double()
        _lib.check(_lib.lib().mchip_dfk_energy_tilde64_dev(self.h, C_dev.ptr, C_dev.shape[1], nocc,
                                                            ctypes.byref(e)))
        return e.value

    def exchange_energy_metric_dev(self, metric, C_dev, nocc, occ_nbatches=1):
        """The same tr(K P) from the untransformed B and the metric (plain
        FP64 exchange formula) without forming K: -sum_X ||(L^-1 M)_X||^2,
        M_X = C^T B_X C (synchronises)."""
        if getattr(C_dev, "ptr", None) is None or not C_dev.ptr.value:
            raise ValueError("exchange_energy_metric_dev: C_dev has been released")
        if len(C_dev.shape) != 2 or C_dev.shape[0] != self.nbf or not 0 <= nocc <= C_dev.shape[1]:
            raise ValueError("exchange_energy_metric_dev: C_dev shape %s / nocc %d" % (C_dev.shape, nocc))
        e = ctypes.c_double()
        _lib.check(_lib.lib().mchip_dfk_energy_dev(self.h, metric.h, C_dev.ptr, C_dev.shape[1], nocc,
                                                    int(occ_nbatches), ctypes.byref(e)))
        return e.value

    @staticmethod
    def dfk_bf16_group(builds):
        """BF16x3 exchanges of several tensors in one grouped launch:
        builds = [(DeviceB, C_dev, nocc, K_dev)], all on one context; each
        K_dev equals the build alone (dfk_dev "bf16x3") bit for bit."""
        n = len(builds)
        if n == 0:
            return
        ctx = builds[0][0].ctx
        for dB, C_dev, nocc, K_dev in builds:
            if dB.ctx is not ctx:
                raise ValueError("dfk_bf16_group: the tensors must share one context")
            for name, a in (("C_dev", C_dev), ("K_dev", K_dev)):
                if getattr(a, "ptr", None) is None or not a.ptr.value:
                    raise ValueError("dfk_bf16_group: %s has been released" % name)
                if len(a.shape) != 2 or a.shape[0] != dB.nbf:
                    raise ValueError("dfk_bf16_group: %s has shape %s, expected (%d, n)" % (name, a.shape, dB.nbf))
            if K_dev.shape != (dB.nbf, dB.nbf):
                raise ValueError("dfk_bf16_group: K_dev has shape %s" % (K_dev.shape,))
            if not 0 < nocc <= C_dev.shape[1]:
                raise ValueError("dfk_bf16_group: nocc %d outside 1..%d" % (nocc, C_dev.shape[1]))
        vp = ctypes.c_void_p * n
        ts = vp(*[b[0].h for b in builds])
        cs = vp(*[b[1].ptr.value for b in builds])
        ks = vp(*[b[3].ptr.value for b in builds])
        ldc = np.array([b[1].shape[1] for b in builds], dtype=np.int32)
        nocc = np.array([b[2] for b in builds], dtype=np.int32)
        _lib.check(_lib.lib().mchip_dfk_mo_bf16_group_dev(n, ts, cs, _lib.iptr(ldc), _lib.iptr(nocc), ks))

    def dfk_ao(self, P, metric=None, form="fp64", occ_nbatches=1):
        """Exchange from a density (DFAO_K / DFMEM_K): pivoted Cholesky of P on
        the GPU, then the occupied-factor exchange.  form: fp64, f32, bf16."""
        P = np.ascontiguousarray(P, dtype=np.float64)
        K = np.zeros((self.nbf, self.nbf))
        m = None if metric is None else metric.h
        f = {"fp64": 0, "f32": 1, "bf16": 2}[form]
        _lib.check(_lib.lib().mchip_dfk_ao(self.h, m, _lib.dptr(P), f, occ_nbatches, _lib.dptr(K)))
        return K

    def dfk_lmo_uv(self, U, V, metric=None, occ_nbatches=1, fp64_copy=False):
        """Two-sided exchange of DFLMO_K (set_sym(false)): U nbf x nocc (left
        coefficients), V nocc x nbf (right coefficients, the reference's v_mb);
        K (negated, not symmetric)."""
        U = np.ascontiguousarray(U, dtype=np.float64)
        V = np.ascontiguousarray(V, dtype=np.float64)
        if U.shape[0] != self.nbf or V.shape != (U.shape[1], self.nbf):
            raise ValueError("U must be nbf x nocc and V nocc x nbf")
        K = np.zeros((self.nbf, self.nbf))
        m = None if metric is None else metric.h
        _lib.check(_lib.lib().mchip_dfk_lmo_uv(self.h, m, _lib.dptr(U), _lib.dptr(V), U.shape[1], occ_nbatches,
                                               int(bool(fp64_copy)), _lib.dptr(K)))
        return K

    def bf16_tile(self, ng):
        """(column tile, threads per workgroup) picked by the BF16 autotuner for
        ng occupied columns, or None before the first build with that ng."""
        code = ctypes.c_int()
        _lib.check(_lib.lib().mchip_dfk_bf16_tile(self.h, ng, ctypes.byref(code)))
        return None if code.value == 0 else (code.value // 10000, code.value % 10000)

    def bf16_plan(self, ng):
        """The autotuned BF16 exchange pipeline for ng occupied columns (None
        before the first build with that ng)."""
        plan = np.zeros(8, dtype=np.int32)
        ms = np.zeros(32, dtype=np.float32)
        nms = ctypes.c_int()
        _lib.check(_lib.lib().mchip_dfk_bf16_plan(self.h, ng, _lib.iptr(plan), _lib.fptr(ms), ctypes.byref(nms)))
        if not plan[7]:
            return None
        k = nms.value
        ns = self.bf16_shapes()[1]
        return dict(gemm_shape=int(plan[0]), bm=int(plan[1]), bn=int(plan[2]), threads=int(plan[3]),
                    direct_to_lds=bool(plan[4]), pipeline=int(plan[4]), stages=int(plan[5]),
                    syrk_variant=int(plan[6]),
                    gemm_ms=[float(v) for v in ms[:k - ns]], syrk_ms=[float(v) for v in ms[k - ns:k]])

    def f32_plan(self, ng):
        """The autotuned FP32 exchange pipeline for ng occupied columns (None
        before the first FP32 build with that ng)."""
        plan = np.zeros(8, dtype=np.int32)
        ms = np.zeros(32, dtype=np.float32)
        nms = ctypes.c_int()
        _lib.check(_lib.lib().mchip_dfk_f32_plan(self.h, ng, _lib.iptr(plan), _lib.fptr(ms), ctypes.byref(nms)))
        if not plan[7]:
            return None
        k = nms.value
        ns = self.f32_shapes()[1]
        return dict(gemm_shape=int(plan[0]), bm=int(plan[1]), bn=int(plan[2]), threads=int(plan[3]),
                    stages=int(plan[5]), syrk_stages=int(plan[6]) + 2,
                    gemm_ms=[float(v) for v in ms[:k - ns]], syrk_ms=[float(v) for v in ms[k - ns:k]])

    def f32_force(self, ng, gemm_shape, syrk_variant):
        """Use the given FP32 pipeline for ng columns instead of the autotuner's."""
        _lib.check(_lib.lib().mchip_dfk_f32_force(self.h, ng, gemm_shape, syrk_variant))

    @staticmethod
    def f32_shapes():
        """Number of FP32 slab-GEMM shapes and of FP32 SYRK variants."""
        ng, ns = ctypes.c_int(), ctypes.c_int()
        _lib.check(_lib.lib().mchip_dfk_f32_shapes(ctypes.byref(ng), ctypes.byref(ns)))
        return ng.value, ns.value

    def f64_plan(self, ng):
        """The autotuned FP64 exchange pieces for ng occupied columns: slab-GEMM
        shape of the FP64 L^-1 B copy route and SYRK variant (None if neither is
        tuned yet; -1 for the one not yet tuned)."""
        plan = np.zeros(8, dtype=np.int32)
        ms = np.zeros(32, dtype=np.float32)
        nms = ctypes.c_int()
        _lib.check(_lib.lib().mchip_dfk_f64_plan(self.h, ng, _lib.iptr(plan), _lib.fptr(ms), ctypes.byref(nms)))
        if not plan[7]:
            return None
        k = nms.value
        ns = self.f64_shapes()[1]
        return dict(gemm_shape=int(plan[0]), bm=int(plan[1]), bn=int(plan[2]), threads=int(plan[3]),
                    stages=int(plan[5]), syrk_variant=int(plan[6]),
                    gemm_ms=[float(v) for v in ms[:k - ns]], syrk_ms=[float(v) for v in ms[k - ns:k]])

    def f64_force(self, ng, gemm_shape, syrk_variant):
        """Use the given FP64 slab-GEMM shape and SYRK variant for ng columns."""
        _lib.check(_lib.lib().mchip_dfk_f64_force(self.h, ng, gemm_shape, syrk_variant))

    @staticmethod
    def f64_shapes():
        """Number of FP64 slab-GEMM shapes and of FP64 SYRK variants."""
        ng, ns = ctypes.c_int(), ctypes.c_int()
        _lib.check(_lib.lib().mchip_dfk_f64_shapes(ctypes.byref(ng), ctypes.byref(ns)))
        return ng.value, ns.value

    def bf16_force(self, ng, gemm_shape, syrk_variant):
        """Use the given BF16 pipeline for ng columns instead of the autotuner's."""
        _lib.check(_lib.lib().mchip_dfk_bf16_force(self.h, ng, gemm_shape, syrk_variant))

    @staticmethod
    def bf16_shapes():
        """[(gemm_shape, direct_to_lds)], number of SYRK variants."""
        ng, ns = ctypes.c_int(), ctypes.c_int()
        d = np.zeros(64, dtype=np.int32)
        _lib.check(_lib.lib().mchip_dfk_bf16_shapes(ctypes.byref(ng), ctypes.byref(ns), _lib.iptr(d)))
        return [(i, bool(d[i])) for i in range(ng.value)], ns.value

    def download(self):
        out = np.zeros(self.nelem)
        _lib.check(_lib.lib().mchip_dbt_download(self.h, _lib.dptr(out)))
        return out

    def dfj(self, P, metric=None, pscale=2.0):
        P = np.ascontiguousarray(P, dtype=np.float64)
        J = np.zeros_like(P)
        m = None if metric is None else metric.h
        _lib.check(_lib.lib().mchip_dfj(self.h, m, _lib.dptr(P), pscale, _lib.dptr(J)))
        return J

    def dfk_mo(self, C_occ, metric=None, occ_nbatches=1):
        C = np.ascontiguousarray(C_occ, dtype=np.float64)
        K = np.zeros((self.nbf, self.nbf))
        m = None if metric is None else metric.h
        _lib.check(_lib.lib().mchip_dfk_mo(self.h, m, _lib.dptr(C), C.shape[1], occ_nbatches,
                                           _lib.dptr(K)))
        return K

    def release(self):
        if getattr(self, "h", None):
            _lib.lib().mchip_dbt_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass
