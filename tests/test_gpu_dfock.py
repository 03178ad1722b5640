"""GPU parity of the plain FP64 DF-J / DFMO-K path against the oracle.

The contraction kernels are compared with the oracle fed the same metric
factors the GPU produced (rocSOLVER potrf/trtri), so the bar isolates the
kernels: max |dJ|, |dK| <= 1e-13 relative to max |J|, |K|.  The metric itself
is pinned separately (V V^-1 = I, L^-1 against LAPACK within cond(V)*eps).
Block masks are compared bit-exactly.
"""
import numpy as np
import pytest

from chem_amd import desc, ints
from oracle import dfock

pytestmark = pytest.mark.gpu

TOL = 1e-13
EPS = 2.220446049250313e-16


def _jtol(V):
    """J = B^T V^-1 (B P) amplifies summation-order differences by cond(V)."""
    return max(TOL, 10.0 * np.linalg.cond(V) * EPS)


def _relerr(a, b):
    return float(np.abs(a - b).max() / max(1e-300, np.abs(b).max()))


@pytest.fixture(scope="module")
def ctx():
    from chem_amd.engine import Context
    yield Context(0)


def _system(nwater, seed=7, eps=1e-10):
    atoms = desc.water_cluster(nwater, seed)
    mol = desc.Molecule(atoms)
    B = ints.BTensorHost.build(mol, eps)
    V = ints.int2c2e(mol)
    return mol, B, V


def _random_occ(nbf, nocc, seed):
    rng = np.random.default_rng(seed)
    C = rng.standard_normal((nbf, nocc)) * 0.3
    return C, C @ C.T


@pytest.mark.parametrize("nwater", [1, 3])
def test_dfj_dfk_plain_vs_oracle(ctx, nwater):
    from chem_amd.engine import DeviceB, Metric
    mol, Bh, V = _system(nwater)
    print("case: water cluster n=%d nbf=%d naux=%d" % (nwater, mol.nbf, mol.naux), flush=True)
    Bd = Bh.dense()
    dB = DeviceB(ctx, Bh)
    m = Metric(ctx, V)
    Linv, Vinv = m.get()
    C, P = _random_occ(mol.nbf, mol.nocc, 11 + nwater)
    Jref = dfock.df_j(Bd, Vinv, P, bsz=mol.b)
    Kref = dfock.dfmo_k(Bd, Linv.T, C, bsz=mol.b, xsz=mol.x, osz=mol.oa())
    J = dB.dfj(P, m)
    K = dB.dfk_mo(C, m)
    print("relerr J %.3e K %.3e cond %.3e" % (_relerr(J, Jref), _relerr(K, Kref), np.linalg.cond(V)))
    assert _relerr(J, Jref) < _jtol(V)
    assert _relerr(K, Kref) < _jtol(V)
    assert np.array_equal(K, K.T)
    # occupied batching does not change K beyond rounding
    K3 = dB.dfk_mo(C, m, occ_nbatches=3)
    assert _relerr(K3, K) < TOL


def test_metric_factors(ctx):
    from chem_amd.engine import Metric
    mol, _, V = _system(1)
    m = Metric(ctx, V)
    Linv, Vinv = m.get()
    Vinv_ref, Linv_ref, _ = dfock.metric_factors(V)
    cond = np.linalg.cond(V)
    assert np.array_equal(np.triu(Linv, 1), np.zeros_like(Linv))
    assert _relerr(Linv, Linv_ref) < 10 * cond * 2.2e-16
    assert np.abs(Linv @ V @ Linv.T - np.eye(V.shape[0])).max() < 1e-9
    assert np.abs(Vinv @ V - np.eye(V.shape[0])).max() < 1e-8


def test_transformed_B_matches_plain(ctx):
    from chem_amd.engine import DeviceB, Metric
    mol, Bh, V = _system(2)
    C, P = _random_occ(mol.nbf, mol.nocc, 5)
    m = Metric(ctx, V)
    dB = DeviceB(ctx, Bh)
    J0 = dB.dfj(P, m)
    K0 = dB.dfk_mo(C, m)
    dB.transform(m)
    J1 = dB.dfj(P)
    K1 = dB.dfk_mo(C)
    # same L^-1, different association: (L^-1 B) C against L^-1 (B C)
    print("relerr J %.3e K %.3e" % (_relerr(J1, J0), _relerr(K1, K0)))
    assert _relerr(J1, J0) < _jtol(V)
    assert _relerr(K1, K0) < _jtol(V)


@pytest.mark.parametrize("nocc", [1, 5, 17, 70])
def test_synthetic_sparse_B(ctx, nocc):
    """Random B with a random block mask, including empty pairs and slabs."""
    from chem_amd.engine import DeviceB, Metric
    rng = np.random.default_rng(100 + nocc)
    bsz = [14, 5, 5, 3, 14, 1, 7]
    xsz = [23, 70, 9, 40]
    nbf, naux = sum(bsz), sum(xsz)
    mask = (rng.random((len(xsz), len(bsz), len(bsz))) < 0.6).astype(np.uint8)
    mask = np.maximum(mask, mask.transpose(0, 2, 1))
    mask[:, 5, :] = 0          # slab with no significant pair
    mask[:, :, 5] = 0
    boff = np.concatenate([[0], np.cumsum(bsz)])
    xoff = np.concatenate([[0], np.cumsum(xsz)])
    dense = rng.standard_normal((naux, nbf, nbf))
    dense = 0.5 * (dense + dense.transpose(0, 2, 1))
    for x in range(len(xsz)):
        for a in range(len(bsz)):
            for b in range(len(bsz)):
                if not mask[x, a, b]:
                    dense[xoff[x]:xoff[x + 1], boff[a]:boff[a + 1], boff[b]:boff[b + 1]] = 0
    Bh = ints.BTensorHost.from_dense(dense, xsz, bsz, mask)
    assert np.array_equal(Bh.mask(), mask)
    assert np.array_equal(Bh.dense(), dense)
    A = rng.standard_normal((naux, naux))
    V = A @ A.T + naux * np.eye(naux)
    C = rng.standard_normal((nbf, nocc))
    P = C @ C.T
    dB = DeviceB(ctx, Bh)
    m = Metric(ctx, V)
    Linv, Vinv = m.get()
    J = dB.dfj(P, m)
    K = dB.dfk_mo(C, m)
    Jref = dfock.df_j(dense, Vinv, P)
    Kref = dfock.dfmo_k(dense, Linv.T, C)
    print("relerr J %.3e K %.3e" % (_relerr(J, Jref), _relerr(K, Kref)))
    assert _relerr(J, Jref) < _jtol(V)
    assert _relerr(K, Kref) < _jtol(V)
    assert np.array_equal(dB.download(), dB.download())


def test_fp32_exchange_accuracy(ctx):
    """Chosen-form FP32 exchange against the FP64 one on the same B~."""
    from chem_amd.engine import DeviceB, Metric
    mol, Bh, V = _system(3)
    C, P = _random_occ(mol.nbf, mol.nocc, 9)
    m = Metric(ctx, V)
    dB = DeviceB(ctx, Bh)
    dB.transform(m)
    dB.make_f32()
    K64 = dB.dfk_mo(C)
    K32 = dB.dfk_mo_f32(C)
    K32b = dB.dfk_mo_f32(C, occ_nbatches=2)
    err = _relerr(K32, K64)
    print("FP32 K relerr %.3e" % err)
    assert err < 2e-6
    assert np.array_equal(K32, K32.T)
    assert _relerr(K32b, K32) < 2e-6
    # deterministic: the same call gives the same bits
    assert np.array_equal(K32, dB.dfk_mo_f32(C))


@pytest.mark.parametrize("nocc", [15, 70, 333])
def test_f32_every_pipeline(ctx, nocc):
    """Every FP32 slab-GEMM shape with every SYRK variant (direct-to-LDS, 2-4
    stages) on a block-sparse B with ragged blocks, an empty slab, several
    128-row SYRK tiles and partial row / column tiles: all give bitwise the
    same exchange (only the tiling of the output differs, not the order of any
    sum) and agree with FP64 to 2e-6 relative."""
    from chem_amd.engine import DeviceB, Metric
    rng = np.random.default_rng(700 + nocc)
    bsz = [14, 5, 5, 3, 14, 1, 7, 9, 30, 40, 50]
    xsz = [23, 70, 9, 40, 17]
    nbf, naux = sum(bsz), sum(xsz)
    mask = (rng.random((len(xsz), len(bsz), len(bsz))) < 0.6).astype(np.uint8)
    mask = np.maximum(mask, mask.transpose(0, 2, 1))
    mask[:, 5, :] = 0
    mask[:, :, 5] = 0
    boff = np.concatenate([[0], np.cumsum(bsz)])
    xoff = np.concatenate([[0], np.cumsum(xsz)])
    dense = rng.standard_normal((naux, nbf, nbf))
    dense = 0.5 * (dense + dense.transpose(0, 2, 1))
    for x in range(len(xsz)):
        for a in range(len(bsz)):
            for b in range(len(bsz)):
                if not mask[x, a, b]:
                    dense[xoff[x]:xoff[x + 1], boff[a]:boff[a + 1], boff[b]:boff[b + 1]] = 0
    Bh = ints.BTensorHost.from_dense(dense, xsz, bsz, mask)
    A = rng.standard_normal((naux, naux))
    V = A @ A.T + naux * np.eye(naux)
    C = rng.standard_normal((nbf, nocc))
    dB = DeviceB(ctx, Bh)
    m = Metric(ctx, V)
    dB.make_lowp(m, f32=True, bf16=False)
    K64 = dB.dfk_mo(C, m)
    ngemm, nsyrk = dB.f32_shapes()
    Kfirst = None
    for shape in range(ngemm):
        for v in range(nsyrk):
            print("case: nocc %d FP32 gemm shape %d syrk %d" % (nocc, shape, v), flush=True)
            dB.f32_force(nocc, shape, v)
            K = dB.dfk_mo_f32(C)
            err = _relerr(K, K64)
            print("  relerr %.3e" % err)
            assert err < 2e-6
            assert np.array_equal(K, K.T)
            if Kfirst is None:
                Kfirst = K
            assert np.array_equal(K, Kfirst)
    # the autotuner picks one of them
    dB2 = DeviceB(ctx, Bh)
    dB2.make_lowp(m, f32=True, bf16=False)
    assert np.array_equal(dB2.dfk_mo_f32(C), Kfirst)
    plan = dB2.f32_plan(nocc)
    assert plan is not None and 0 <= plan["gemm_shape"] < ngemm
    assert len(plan["gemm_ms"]) == ngemm and len(plan["syrk_ms"]) == nsyrk


@pytest.mark.parametrize("nocc", [15, 70, 333])
def test_f64_every_pipeline(ctx, nocc):
    """Every FP64 slab-GEMM shape on the padded FP64 copy of L^-1 B with every
    FP64 SYRK variant (direct-to-LDS, 2-4 stages), on a block-sparse B with
    ragged blocks, an empty slab, several 128-row SYRK tiles and partial row /
    column tiles: all give bitwise the same exchange, equal to the plain DFMO
    exchange (L^-1 applied per build, same SYRK) to 1e-12 relative, also in
    occupied batches whose width is not a multiple of the 8-column step."""
    from chem_amd.engine import DeviceB, Metric
    rng = np.random.default_rng(900 + nocc)
    bsz = [14, 5, 5, 3, 14, 1, 7, 9, 30, 40, 50]
    xsz = [23, 70, 9, 40, 17]
    nbf, naux = sum(bsz), sum(xsz)
    mask = (rng.random((len(xsz), len(bsz), len(bsz))) < 0.6).astype(np.uint8)
    mask = np.maximum(mask, mask.transpose(0, 2, 1))
    mask[:, 5, :] = 0
    mask[:, :, 5] = 0
    boff = np.concatenate([[0], np.cumsum(bsz)])
    xoff = np.concatenate([[0], np.cumsum(xsz)])
    dense = rng.standard_normal((naux, nbf, nbf))
    dense = 0.5 * (dense + dense.transpose(0, 2, 1))
    for x in range(len(xsz)):
        for a in range(len(bsz)):
            for b in range(len(bsz)):
                if not mask[x, a, b]:
                    dense[xoff[x]:xoff[x + 1], boff[a]:boff[a + 1], boff[b]:boff[b + 1]] = 0
    Bh = ints.BTensorHost.from_dense(dense, xsz, bsz, mask)
    A = rng.standard_normal((naux, naux))
    V = A @ A.T + naux * np.eye(naux)
    C = rng.standard_normal((nbf, nocc))
    dB = DeviceB(ctx, Bh)
    m = Metric(ctx, V)
    Kplain = dB.dfk_mo(C, m)
    Kplain3 = dB.dfk_mo(C, m, occ_nbatches=3)
    Linv, _ = m.get()
    Kref = dfock.dfmo_k(dense, Linv.T, C)
    assert _relerr(Kplain, Kref) < _jtol(V)
    assert np.abs(Kplain3 - Kplain).max() < 1e-12 * np.abs(Kplain).max()
    dB.make_lowp(m, f32=False, bf16=False, f64=True)
    ngemm, nsyrk = dB.f64_shapes()
    Kfirst = None
    for shape in range(ngemm):
        for v in range(nsyrk):
            print("case: nocc %d FP64 gemm shape %d syrk %d" % (nocc, shape, v), flush=True)
            dB.f64_force(nocc, shape, v)
            K = dB.dfk_mo_tilde64(C)
            assert np.array_equal(K, K.T)
            if Kfirst is None:
                Kfirst = K
                rel = np.abs(K - Kplain).max() / np.abs(Kplain).max()
                print("  rel vs plain %.3e" % rel)
                assert rel < 1e-12
            assert np.array_equal(K, Kfirst)
    dB2 = DeviceB(ctx, Bh)
    dB2.make_lowp(m, f32=False, bf16=False, f64=True)
    assert np.array_equal(dB2.dfk_mo_tilde64(C), Kfirst)
    plan = dB2.f64_plan(nocc)
    assert plan is not None and 0 <= plan["gemm_shape"] < ngemm and 0 <= plan["syrk_variant"] < nsyrk
    K2 = dB2.dfk_mo_tilde64(C, occ_nbatches=2)
    assert np.abs(K2 - Kfirst).max() < 1e-12 * np.abs(Kfirst).max()
    # released work arrays are allocated again by the next build
    ctx.trim()
    assert np.array_equal(dB2.dfk_mo_tilde64(C), Kfirst)
    assert np.array_equal(dB.dfk_mo(C, m), Kplain)


@pytest.mark.parametrize("symmetric", [True, False])
def test_lowp_copies_canonical_half_and_full(ctx, symmetric):
    """The L^-1 B copies (FP64, FP32, BF16 hi/lo) made from the canonical
    half of a symmetric tensor with mirrored blocks (ragged blocks, an empty
    slab), and over every column when one kept pair lacks its transpose: the
    exchanges on the copies equal the plain DFMO exchange to each form's
    tolerance."""
    from chem_amd.engine import DeviceB, Metric
    rng = np.random.default_rng(4242)
    bsz = [14, 5, 5, 3, 14, 1, 7, 9, 30, 12]
    xsz = [23, 70, 9, 40, 17]
    nbf, naux = sum(bsz), sum(xsz)
    mask = (rng.random((len(xsz), len(bsz), len(bsz))) < 0.6).astype(np.uint8)
    mask = np.maximum(mask, mask.transpose(0, 2, 1))
    mask[:, 5, :] = 0
    mask[:, :, 5] = 0
    if not symmetric:
        mask[:, 2, 7] = 1
        mask[:, 7, 2] = 0
        mask[1, 7, 2] = 0
    boff = np.concatenate([[0], np.cumsum(bsz)])
    xoff = np.concatenate([[0], np.cumsum(xsz)])
    dense = rng.standard_normal((naux, nbf, nbf))
    dense = 0.5 * (dense + dense.transpose(0, 2, 1))
    for x in range(len(xsz)):
        for a in range(len(bsz)):
            for b in range(len(bsz)):
                if not mask[x, a, b]:
                    dense[xoff[x]:xoff[x + 1], boff[a]:boff[a + 1], boff[b]:boff[b + 1]] = 0
    print("case: low-precision copies, symmetric mask %s" % symmetric, flush=True)
    assert np.array_equal(mask, mask.transpose(0, 2, 1)) == symmetric
    assert np.array_equal(dense, dense.transpose(0, 2, 1)) == symmetric
    Bh = ints.BTensorHost.from_dense(dense, xsz, bsz, mask)
    A = rng.standard_normal((naux, naux))
    m = Metric(ctx, A @ A.T + naux * np.eye(naux))
    C = rng.standard_normal((nbf, 37)) * 0.3
    dB = DeviceB(ctx, Bh)
    Kplain = dB.dfk_mo(C, m)
    dB.make_lowp(m, f32=True, bf16=True, f64=True)
    for name, K, tol in (("fp64_copy", dB.dfk_mo_tilde64(C), 1e-12), ("f32", dB.dfk_mo_f32(C), 2e-6),
                         ("bf16x3", dB.dfk_mo_bf16(C), BF16_TOL)):
        rel = _relerr(K, Kplain)
        print("  %s relerr vs plain %.2e" % (name, rel))
        assert rel < tol


def test_device_resident_exchange_matches_host(ctx):
    """The _dev entry points (occupied coefficients and K in device arrays,
    a row stride larger than the occupied count) give bitwise the host-pointer
    results for every form."""
    from chem_amd.engine import DeviceArray, DeviceB, Metric
    mol, Bh, V = _system(3)
    C, _ = _random_occ(mol.nbf, mol.nocc, 13)
    m = Metric(ctx, V)
    dB = DeviceB(ctx, Bh)
    dB.make_lowp(m, f32=True, bf16=True, f64=True)
    ld = mol.nocc + 3
    Cwide = np.zeros((mol.nbf, ld))
    Cwide[:, :mol.nocc] = C
    Cd = DeviceArray(ctx, Cwide.shape, Cwide)
    Kd = DeviceArray(ctx, (mol.nbf, mol.nbf))
    host = {"fp64": lambda: dB.dfk_mo(C, m), "fp64_copy": lambda: dB.dfk_mo_tilde64(C),
            "f32": lambda: dB.dfk_mo_f32(C), "bf16x3": lambda: dB.dfk_mo_bf16(C)}
    for form, fn in host.items():
        print("case: device-resident %s" % form, flush=True)
        dB.dfk_dev(form, Cd, mol.nocc, Kd, metric=m)
        ctx.sync()
        assert np.array_equal(Kd.download(), fn()), form
    Cd.release()
    Kd.release()


def test_device_resident_exchange_rejects_bad_arguments(ctx):
    """The _dev exchange paths check what they can before any launch: array
    shapes and released arrays in the Python layer, the occupied count against
    the row stride in the C entry points (called directly here)."""
    import ctypes
    from chem_amd import _lib
    from chem_amd.engine import DeviceArray, DeviceB, Metric
    mol, Bh, V = _system(1)
    m = Metric(ctx, V)
    dB = DeviceB(ctx, Bh)
    dB.make_lowp(m, f32=True, bf16=True, f64=True)
    n, no = mol.nbf, mol.nocc
    Cd = DeviceArray(ctx, (n, no), np.zeros((n, no)))
    Kd = DeviceArray(ctx, (n, n))
    Kbad = DeviceArray(ctx, (n, n - 1))
    Cbad = DeviceArray(ctx, (n + 1, no))
    for form in ("fp64", "fp64_copy", "f32", "bf16x3"):
        print("case: bad arguments, %s" % form, flush=True)
        with pytest.raises(ValueError):
            dB.dfk_dev(form, Cd, no, Kbad, metric=m)
        with pytest.raises(ValueError):
            dB.dfk_dev(form, Cbad, no, Kd, metric=m)
        with pytest.raises(ValueError):
            dB.dfk_dev(form, Cd, no + 1, Kd, metric=m)
    with pytest.raises(ValueError):
        dB.dfk_dev("fp16", Cd, no, Kd, metric=m)
    L = _lib.lib()
    for name in ("mchip_dfk_mo_f32_dev", "mchip_dfk_mo_tilde64_dev", "mchip_dfk_mo_bf16_dev"):
        rc = getattr(L, name)(dB.h, Cd.ptr, no, no + 1, 1, Kd.ptr)
        assert rc != 0, name
        assert b"outside 0..ldc" in L.mchip_last_error(), name
    rc = L.mchip_dfk_mo_dev(dB.h, m.h, Cd.ptr, no, -1, 1, Kd.ptr)
    assert rc != 0 and b"outside 0..ldc" in L.mchip_last_error()
    # the checks left the context usable: a valid call still works
    dB.dfk_dev("bf16x3", Cd, no, Kd)
    ctx.sync()
    assert np.all(Kd.download() == 0.0)   # C = 0 gives K = 0
    Cd.release()
    with pytest.raises(ValueError):
        dB.dfk_dev("f32", Cd, no, Kd)
    for a in (Kd, Kbad, Cbad):
        a.release()


# BF16x3 carries every operand with 16 significant bits (hi + lo, 2^-17
# relative rounding) and accumulates in FP32 over chunks of <= 256 aux
# functions: K agrees with the FP64 exchange on the same B~ to a few 1e-6
# relative; the bar is 5e-5 of max |K|.
BF16_TOL = 5e-5


def test_bf16_exchange_accuracy(ctx):
    """Chosen-form BF16 split exchange against the FP64 one on the same B~."""
    from chem_amd.engine import DeviceB, Metric
    mol, Bh, V = _system(3)
    C, P = _random_occ(mol.nbf, mol.nocc, 9)
    m = Metric(ctx, V)
    dB = DeviceB(ctx, Bh)
    dB.transform(m)
    dB.make_bf16()
    K64 = dB.dfk_mo(C)
    K16 = dB.dfk_mo_bf16(C)
    err = _relerr(K16, K64)
    print("BF16x3 K relerr %.3e (column tile, threads) %s" % (err, dB.bf16_tile(mol.nocc)))
    assert err < BF16_TOL
    assert np.array_equal(K16, K16.T)
    bn, nthr = dB.bf16_tile(mol.nocc)
    assert bn in (64, 128, 160, 192, 256, 320) and nthr in (256, 512)
    plan = dB.bf16_plan(mol.nocc)
    print("BF16 plan", plan)
    assert plan is not None and plan["bn"] == bn
    # occupied batches (ragged: 15 = 8 + 7) and determinism
    K16b = dB.dfk_mo_bf16(C, occ_nbatches=2)
    assert _relerr(K16b, K64) < BF16_TOL
    assert np.array_equal(K16, dB.dfk_mo_bf16(C))


@pytest.mark.parametrize("nocc", [1, 17, 70])
def test_bf16_synthetic_sparse_B(ctx, nocc):
    """BF16 exchange on a random block-sparse B with an empty slab and ragged
    block sizes (row padding and zero tiles) against the FP64 exchange."""
    from chem_amd.engine import DeviceB, Metric
    rng = np.random.default_rng(300 + nocc)
    bsz = [14, 5, 5, 3, 14, 1, 7]
    xsz = [23, 70, 9, 40]
    nbf, naux = sum(bsz), sum(xsz)
    mask = (rng.random((len(xsz), len(bsz), len(bsz))) < 0.6).astype(np.uint8)
    mask = np.maximum(mask, mask.transpose(0, 2, 1))
    mask[:, 5, :] = 0
    mask[:, :, 5] = 0
    boff = np.concatenate([[0], np.cumsum(bsz)])
    xoff = np.concatenate([[0], np.cumsum(xsz)])
    dense = rng.standard_normal((naux, nbf, nbf))
    dense = 0.5 * (dense + dense.transpose(0, 2, 1))
    for x in range(len(xsz)):
        for a in range(len(bsz)):
            for b in range(len(bsz)):
                if not mask[x, a, b]:
                    dense[xoff[x]:xoff[x + 1], boff[a]:boff[a + 1], boff[b]:boff[b + 1]] = 0
    Bh = ints.BTensorHost.from_dense(dense, xsz, bsz, mask)
    A = rng.standard_normal((naux, naux))
    V = A @ A.T + naux * np.eye(naux)
    C = rng.standard_normal((nbf, nocc))
    dB = DeviceB(ctx, Bh)
    m = Metric(ctx, V)
    dB.transform(m)
    dB.make_bf16()
    K64 = dB.dfk_mo(C)
    K16 = dB.dfk_mo_bf16(C)
    err = _relerr(K16, K64)
    print("nocc %d BF16x3 K relerr %.3e" % (nocc, err))
    assert err < BF16_TOL
    assert np.array_equal(K16, K16.T)


@pytest.mark.parametrize("nwater", [1, 3])
def test_dfao_from_density_vs_oracle(ctx, nwater):
    """DFAO_K / DFMEM_K from a density: the GPU factorises P (pivoted Cholesky)
    and runs the occupied-factor exchange; compared with the oracle's fitted-
    coefficient DFAO and batched DFMEM contractions on the same metric factors.
    The two routes differ by how V^-1 enters (L^-1 on both sides vs V^-1 on
    one), so the bar is the metric-conditioning bound 10 cond(V) eps."""
    from chem_amd.engine import DeviceB, Metric
    mol, Bh, V = _system(nwater)
    Bd = Bh.dense()
    dB = DeviceB(ctx, Bh)
    m = Metric(ctx, V)
    Linv, Vinv = m.get()
    C, P = _random_occ(mol.nbf, mol.nocc, 21 + nwater)
    K = dB.dfk_ao(P, m)
    Kao = dfock.dfao_k(Bd, dfock.fitting_coefficients(Bd, Vinv), P, bsz=mol.b, xsz=mol.x)
    Kmem = dfock.dfmem_k(Bd, Vinv, P, bsz=mol.b, xsz=mol.x)
    Kmo = dB.dfk_mo(C, m)
    e_ao, e_mem, e_mo = _relerr(K, Kao), _relerr(K, Kmem), _relerr(K, Kmo)
    print("DFAO from P: relerr vs oracle DFAO %.2e DFMEM %.2e, vs GPU DFMO %.2e" % (e_ao, e_mem, e_mo))
    assert e_ao < _jtol(V) and e_mem < _jtol(V)
    assert e_mo < 1e-12
    # the BF16 form through the same route, on the transformed tensor
    dB.transform(m)
    dB.make_bf16()
    K16 = dB.dfk_ao(P, form="bf16")
    assert _relerr(K16, Kmo) < BF16_TOL


@pytest.mark.parametrize("nwater", [1, 3])
def test_dfao_indefinite_density_vs_oracle(ctx, nwater):
    """DFAO_K from an indefinite density (a difference of two densities, as in
    a response or a density update): the pivoted Cholesky stops at a negative
    pivot, and the exchange runs on the density's eigenvector factors,
    K(P) = K(C+) - K(C-).  Compared with the oracle's fitted-coefficient DFAO
    (which takes any P) and with the difference of the two GPU DFMO builds."""
    from chem_amd.engine import DeviceB, Metric
    mol, Bh, V = _system(nwater)
    Bd = Bh.dense()
    dB = DeviceB(ctx, Bh)
    m = Metric(ctx, V)
    Linv, Vinv = m.get()
    C1, P1 = _random_occ(mol.nbf, mol.nocc, 31 + nwater)
    C2, P2 = _random_occ(mol.nbf, mol.nocc - 1, 41 + nwater)
    P = P1 - 0.5 * P2
    w = np.linalg.eigvalsh(P)
    assert w.min() < -1e-3 * w.max()          # really indefinite
    print("case: indefinite density, %d waters" % nwater, flush=True)
    K = dB.dfk_ao(P, m)
    Kao = dfock.dfao_k(Bd, dfock.fitting_coefficients(Bd, Vinv), P, bsz=mol.b, xsz=mol.x)
    Kdiff = dB.dfk_mo(C1, m) - 0.5 * dB.dfk_mo(C2, m)
    e_ao, e_diff = _relerr(K, Kao), _relerr(K, Kdiff)
    print("indefinite DFAO: relerr vs oracle DFAO %.2e, vs DFMO difference %.2e" % (e_ao, e_diff))
    assert e_ao < _jtol(V)
    assert e_diff < 1e-11


def test_pivoted_cholesky_rank_and_factor(ctx):
    """The device pivoted Cholesky on a rank-deficient PSD matrix across several
    panels: the rank found is the matrix rank and L L^T reproduces it (through
    the DFAO route: K(P) equals K of the generating factor)."""
    from chem_amd.engine import DeviceB, Metric
    mol, Bh, V = _system(3)
    dB = DeviceB(ctx, Bh)
    m = Metric(ctx, V)
    rng = np.random.default_rng(5)
    for r in (1, 31, 32, 33, 70):
        G = rng.standard_normal((mol.nbf, min(r, mol.nbf)))
        P = G @ G.T
        print("case: rank %d of %d" % (r, mol.nbf), flush=True)
        K = dB.dfk_ao(P, m)
        Kg = dB.dfk_mo(np.ascontiguousarray(G), m)
        assert _relerr(K, Kg) < 1e-11


def test_lowp_copies_from_untransformed_B(ctx):
    """The low-precision copies of L^-1 B made slab by slab from the untransformed
    B give bitwise the same exchange as the copies of the in-place transformed
    tensor (same rocBLAS triangular multiply per slab), and B itself stays the
    untransformed FP64 tensor (plain J and K unchanged)."""
    from chem_amd.engine import DeviceB, Metric
    mol, Bh, V = _system(3)
    C, P = _random_occ(mol.nbf, mol.nocc, 5)
    m = Metric(ctx, V)
    a = DeviceB(ctx, Bh)
    J0, K0 = a.dfj(P, m), a.dfk_mo(C, m)
    a.make_lowp(m)
    assert np.array_equal(a.dfj(P, m), J0) and np.array_equal(a.dfk_mo(C, m), K0)
    b = DeviceB(ctx, Bh)
    b.transform(m)
    b.make_f32()
    b.make_bf16()
    assert np.array_equal(a.dfk_mo_f32(C), b.dfk_mo_f32(C))
    assert np.array_equal(a.dfk_mo_bf16(C), b.dfk_mo_bf16(C))
    assert _relerr(a.dfk_mo_bf16(C), K0) < BF16_TOL


@pytest.mark.parametrize("nocc", [15, 70])
def test_bf16_every_pipeline(ctx, nocc):
    """Every BF16 slab-GEMM shape with every matching SYRK variant (register-
    staged with FP32 W; direct-to-LDS pipelined with bf16 hi/lo W) gives the
    same exchange as FP64 to the BF16 bar, on a block-sparse B with ragged
    blocks, an empty slab and partial row / column tiles."""
    from chem_amd.engine import DeviceB, Metric
    rng = np.random.default_rng(500 + nocc)
    bsz = [14, 5, 5, 3, 14, 1, 7, 9]
    xsz = [23, 70, 9, 40, 17]
    nbf, naux = sum(bsz), sum(xsz)
    mask = (rng.random((len(xsz), len(bsz), len(bsz))) < 0.6).astype(np.uint8)
    mask = np.maximum(mask, mask.transpose(0, 2, 1))
    mask[:, 5, :] = 0
    mask[:, :, 5] = 0
    boff = np.concatenate([[0], np.cumsum(bsz)])
    xoff = np.concatenate([[0], np.cumsum(xsz)])
    dense = rng.standard_normal((naux, nbf, nbf))
    dense = 0.5 * (dense + dense.transpose(0, 2, 1))
    for x in range(len(xsz)):
        for a in range(len(bsz)):
            for b in range(len(bsz)):
                if not mask[x, a, b]:
                    dense[xoff[x]:xoff[x + 1], boff[a]:boff[a + 1], boff[b]:boff[b + 1]] = 0
    Bh = ints.BTensorHost.from_dense(dense, xsz, bsz, mask)
    A = rng.standard_normal((naux, naux))
    V = A @ A.T + naux * np.eye(naux)
    C = rng.standard_normal((nbf, nocc))
    dB = DeviceB(ctx, Bh)
    m = Metric(ctx, V)
    dB.make_lowp(m)
    K64 = dB.dfk_mo(C, m)
    shapes, nsyrk = dB.bf16_shapes()
    for shape, direct in shapes:
        for v in (range(1, nsyrk) if direct else [0]):
            print("case: nocc %d gemm shape %d syrk %d" % (nocc, shape, v), flush=True)
            dB.bf16_force(nocc, shape, v)
            K = dB.dfk_mo_bf16(C)
            err = _relerr(K, K64)
            print("  relerr %.3e" % err)
            assert err < BF16_TOL
            assert np.array_equal(K, K.T)


def _synthetic_B(rng, bsz, xsz, density=0.6):
    nbf, naux = sum(bsz), sum(xsz)
    mask = (rng.random((len(xsz), len(bsz), len(bsz))) < density).astype(np.uint8)
    mask = np.maximum(mask, mask.transpose(0, 2, 1))
    boff = np.concatenate([[0], np.cumsum(bsz)])
    xoff = np.concatenate([[0], np.cumsum(xsz)])
    dense = rng.standard_normal((naux, nbf, nbf))
    dense = 0.5 * (dense + dense.transpose(0, 2, 1))
    for x in range(len(xsz)):
        for a in range(len(bsz)):
            for b in range(len(bsz)):
                if not mask[x, a, b]:
                    dense[xoff[x]:xoff[x + 1], boff[a]:boff[a + 1], boff[b]:boff[b + 1]] = 0
    A = rng.standard_normal((naux, naux))
    V = A @ A.T + naux * np.eye(naux)
    return ints.BTensorHost.from_dense(dense, xsz, bsz, mask), dense, V


@pytest.mark.parametrize("nocc,nbatch", [(129, 2), (193, 3), (97, 2)])
def test_batched_exchange_crosses_tile_edge(ctx, nocc, nbatch):
    """Occupied batches whose last batch is narrower than the others by enough
    that its column tiles end before the work array's padded width (129 = 65 +
    64: BN 64 tiles of the last batch stop at column 64 of a 72/80/96-wide W).
    Every form must zero those columns itself, or the previous batch's
    orbitals are counted twice.  Each batched K equals its unbatched K."""
    from chem_amd.engine import DeviceB, Metric
    rng = np.random.default_rng(1300 + nocc)
    bsz = [14, 5, 5, 14, 5, 5, 14, 5, 5, 14, 5, 5, 14, 5, 5, 14, 5, 5, 14, 5, 5, 14, 5, 5, 14, 5]
    xsz = [40, 17, 17, 40, 17]
    Bh, dense, V = _synthetic_B(rng, bsz, xsz)
    nbf = sum(bsz)
    assert nbf >= nocc
    C = rng.standard_normal((nbf, nocc)) * 0.3
    P = C @ C.T
    dB = DeviceB(ctx, Bh)
    m = Metric(ctx, V)
    dB.make_lowp(m, f32=True, bf16=True, f64=True)
    Kplain = dB.dfk_mo(C, m)
    Linv, _ = m.get()
    Kref = dfock.dfmo_k(dense, Linv.T, C)
    assert _relerr(Kplain, Kref) < _jtol(V)
    forms = {"plain": (lambda nb: dB.dfk_mo(C, m, occ_nbatches=nb), 1e-12),
             "fp64_copy": (lambda nb: dB.dfk_mo_tilde64(C, occ_nbatches=nb), 1e-12),
             "f32": (lambda nb: dB.dfk_mo_f32(C, occ_nbatches=nb), 2e-6),
             "bf16x3": (lambda nb: dB.dfk_mo_bf16(C, occ_nbatches=nb), BF16_TOL),
             "ao_fp64": (lambda nb: dB.dfk_ao(P, m, occ_nbatches=nb), 1e-10)}
    for name, (fn, tol) in forms.items():
        print("case: batched exchange nocc %d nbatch %d form %s" % (nocc, nbatch, name), flush=True)
        K1 = fn(1)
        Kb = fn(nbatch)
        e1, eb, ebb = _relerr(K1, Kref), _relerr(Kb, Kref), _relerr(Kb, K1)
        print("  relerr unbatched %.2e batched %.2e batched vs unbatched %.2e" % (e1, eb, ebb))
        assert e1 < tol and eb < tol and ebb < tol


@pytest.mark.parametrize("nbatch", [1, 3])
def test_two_sided_exchange_vs_oracle(ctx, nbatch):
    """DFLMO_K's two-sided exchange (set_sym(false), jkdf.cpp:1093-1331) with
    left coefficients u (nbf x nocc) and right coefficients v (nocc x nbf):
    against the oracle's restatement on a 3-water system and a ragged sparse
    synthetic B, through the plain route (L^-1 per build) and the FP64 copy of
    L^-1 B, in one and three occupied batches.  With u = C and v = C^T it is
    the symmetric DFMO exchange."""
    from chem_amd.engine import DeviceB, Metric
    rng = np.random.default_rng(40 + nbatch)
    cases = []
    mol, Bh, V = _system(3)
    cases.append(("3 waters", Bh, Bh.dense(), V, mol.nbf))
    Bs, dense, Vs = _synthetic_B(rng, [14, 5, 5, 3, 14, 1, 7, 9, 30], [23, 70, 9, 40, 17])
    cases.append(("synthetic", Bs, dense, Vs, sum([14, 5, 5, 3, 14, 1, 7, 9, 30])))
    for name, Bh_, dense_, V_, nbf in cases:
        print("case: two-sided exchange %s, %d batches" % (name, nbatch), flush=True)
        nocc = 13
        U = rng.standard_normal((nbf, nocc)) * 0.3
        W = rng.standard_normal((nocc, nbf)) * 0.3
        dB = DeviceB(ctx, Bh_)
        m = Metric(ctx, V_)
        _, Vinv = m.get()
        Kref = dfock.dflmo_k_uv(dense_, Vinv, U, W)
        Kp = dB.dfk_lmo_uv(U, W, metric=m, occ_nbatches=nbatch)
        dB.make_lowp(m, f32=False, bf16=False, f64=True)
        Kc = dB.dfk_lmo_uv(U, W, occ_nbatches=nbatch, fp64_copy=True)
        ep, ec = _relerr(Kp, Kref), _relerr(Kc, Kref)
        print("  relerr plain %.2e fp64 copy %.2e (asym %.2e)" % (ep, ec, np.abs(Kref - Kref.T).max()))
        assert ep < _jtol(V_) and ec < _jtol(V_)
        assert _relerr(Kc, Kp) < 1e-12
        # u = C, v = C^T: the symmetric exchange
        Ks = dB.dfk_lmo_uv(U, np.ascontiguousarray(U.T), metric=m, occ_nbatches=nbatch)
        assert _relerr(Ks, dB.dfk_mo(U, m)) < 1e-12


def test_grouped_bf16_exchange_matches_alone(ctx):
    """Several tensors' BF16x3 exchanges in one grouped slab-GEMM launch and
    one grouped SYRK launch (the request loop's cross-request batch) give each
    tensor's K of the build alone bit for bit, for both grouped slab-GEMM
    shapes (more / fewer than 160 occupied columns) and both SYRK tiles (above
    / below 512 basis functions), with the FP64 exchange as the sanity bar."""
    from chem_amd import desc
    from chem_amd.engine import DeviceArray, DeviceB, Metric
    cases = [(2, 10), (3, 15), (10, 200), (22, 110), (22, 200)]   # (waters, occupied columns)
    builds, alone, ref = [], [], []
    for k, (nw, ncol) in enumerate(cases):
        mol = desc.Molecule(desc.water_cluster(nw, 40 + k))
        print("case: grouped exchange member %d: %d waters, %d columns, nbf %d" % (k, nw, ncol, mol.nbf), flush=True)
        dB = DeviceB.build_on_gpu(ctx, mol, 1e-10)
        m = Metric(ctx, ints.int2c2e(mol))
        dB.make_lowp(m, f32=False, bf16=True, f64=True)
        C, _ = _random_occ(mol.nbf, ncol, 70 + k)
        Cd = DeviceArray(ctx, C.shape, C)
        Ka = DeviceArray(ctx, (mol.nbf, mol.nbf))
        dB.dfk_dev("bf16x3", Cd, ncol, Ka)
        ctx.sync()
        alone.append(Ka.download())
        ref.append(dB.dfk_mo_tilde64(C))
        Kg = DeviceArray(ctx, (mol.nbf, mol.nbf))
        builds.append((dB, Cd, ncol, Kg))
        Ka.release()
    print("case: grouped launch of %d builds" % len(builds), flush=True)
    DeviceB.dfk_bf16_group(builds)
    ctx.sync()
    for k, (dB, Cd, ncol, Kg) in enumerate(builds):
        K = Kg.download()
        assert np.array_equal(K, alone[k]), k
        assert _relerr(K, ref[k]) < BF16_TOL, k
    # a second grouped launch with a subset reuses the tables and buffers
    sub = [builds[3], builds[0]]
    for b in sub:
        b[3].upload(np.zeros((b[0].nbf, b[0].nbf)))
    DeviceB.dfk_bf16_group(sub)
    ctx.sync()
    assert np.array_equal(sub[0][3].download(), alone[3])
    assert np.array_equal(sub[1][3].download(), alone[0])
    for b in builds:
        b[1].release()
        b[3].release()


@pytest.mark.parametrize("nw,ncol", [(1, 5), (6, 129), (12, 300)])
def test_exchange_energy_without_k(ctx, nw, ncol):
    """tr(K P) of the FP64 exchange on the FP64 copy, computed as
    -sum_X ||W_X^T C||^2 without forming K (the final energy of the chosen
    SCF), equals the trace of the full FP64 exchange and the oracle's to 1e-12
    relative, across one, two and three 128-wide occupied tiles."""
    from chem_amd import desc
    from chem_amd.engine import DeviceArray, DeviceB, Metric
    mol = desc.Molecule(desc.water_cluster(nw, 60 + nw))
    print("case: exchange energy, %d waters, %d columns" % (nw, ncol), flush=True)
    Bh = ints.BTensorHost.build(mol, 1e-10)
    dB = DeviceB(ctx, Bh)
    m = Metric(ctx, ints.int2c2e(mol))
    dB.make_lowp(m, f32=False, bf16=False, f64=True)
    C, P = _random_occ(mol.nbf, ncol, 90 + nw)
    Cd = DeviceArray(ctx, C.shape, C)
    ek = dB.exchange_energy_dev(Cd, ncol)
    full = float(np.sum(dB.dfk_mo_tilde64(C) * P))
    print("energy-only %.15e full trace %.15e rel %.2e" % (ek, full, abs(ek - full) / abs(full)))
    assert ek < 0
    assert abs(ek - full) <= 1e-12 * abs(full)
    if nw == 1:
        Linv, _ = m.get()
        ref = float(np.sum(dfock.dfmo_k(Bh.dense(), Linv.T, C) * P))
        assert abs(ek - ref) <= 1e-12 * abs(ref)
    Cd.release()


@pytest.mark.parametrize("nw,ncol,nbatch", [(1, 5, 1), (6, 129, 1), (6, 129, 2), (12, 300, 3)])
def test_exchange_energy_metric_without_k(ctx, nw, ncol, nbatch):
    """tr(K P) of the plain FP64 exchange (untransformed B, metric step)
    computed as -sum_X ||(L^-1 M)_X||^2, M_X = C^T B_X C, without forming K
    (the final energy of the chosen SCF when no FP64 copy is held), equals the
    trace of the full plain exchange to 1e-12 relative, in one to three
    occupied batches; at one water also the oracle's."""
    from chem_amd import desc
    from chem_amd.engine import DeviceArray, DeviceB, Metric
    mol = desc.Molecule(desc.water_cluster(nw, 60 + nw))
    print("case: metric exchange energy, %d waters, %d columns, %d batches" % (nw, ncol, nbatch), flush=True)
    Bh = ints.BTensorHost.build(mol, 1e-10)
    dB = DeviceB(ctx, Bh)
    m = Metric(ctx, ints.int2c2e(mol))
    C, P = _random_occ(mol.nbf, ncol, 90 + nw)
    Cd = DeviceArray(ctx, C.shape, C)
    ek = dB.exchange_energy_metric_dev(m, Cd, ncol, nbatch)
    full = float(np.sum(dB.dfk_mo(C, m) * P))
    print("energy-only %.15e full trace %.15e rel %.2e" % (ek, full, abs(ek - full) / abs(full)))
    assert ek < 0
    assert abs(ek - full) <= 1e-12 * abs(full)
    if nw == 1:
        Linv, _ = m.get()
        ref = float(np.sum(dfock.dfmo_k(Bh.dense(), Linv.T, C) * P))
        assert abs(ek - ref) <= 1e-12 * abs(ref)
    Cd.release()


@pytest.mark.parametrize("nbatch", [1, 3])
def test_graph_replay_matches_direct(ctx, nbatch):
    """Repeated builds with the same device operands replay a captured hipGraph
    of their launch chain (captured on the second call, replayed from the
    third): every replay gives the bits of the direct launches, for the
    BF16x3, FP32 and FP64-copy exchanges (one and three occupied batches) and
    DF-J, with the operands' contents changed between calls."""
    from chem_amd import _lib
    from chem_amd.engine import DeviceArray, DeviceB, Metric
    mol, Bh, V = _system(4)
    nbf, nocc = mol.nbf, mol.nocc
    m = Metric(ctx, V)
    dB = DeviceB(ctx, Bh)
    dB.make_lowp(m, f32=True, bf16=True, f64=True)
    Cs = [_random_occ(nbf, nocc, 100 + i)[0] for i in range(5)]
    Cd = DeviceArray(ctx, (nbf, nocc))
    Kd = DeviceArray(ctx, (nbf, nbf))
    Pd = DeviceArray(ctx, (nbf, nbf))

    def run(form, C):
        if form == "dfj":
            Pd.upload(C @ C.T)
            _lib.check(_lib.lib().mchip_dfj_dev(dB.h, m.h, Pd.ptr, 2.0, Kd.ptr))
        else:
            Cd.upload(C)
            dB.dfk_dev(form, Cd, nocc, Kd, occ_nbatches=nbatch)
        ctx.sync()
        return Kd.download()

    try:
        for form in ("bf16x3", "f32", "fp64_copy", "dfj"):
            print("case: graph replay %s nbatch=%d" % (form, nbatch), flush=True)
            ctx.set_graphs(False)
            ref = [run(form, C) for C in Cs]
            ctx.set_graphs(True)
            n0, r0 = dB.graph_stats()
            for i, C in enumerate(Cs):
                assert np.array_equal(run(form, C), ref[i]), (form, i)
            n1, r1 = dB.graph_stats()
            # first call direct (the key's first run), second captured, then replays
            assert n1 == n0 + 1 and r1 - r0 == len(Cs) - 1, (form, n0, n1, r0, r1)
    finally:
        ctx.set_graphs(True)
        for a in (Cd, Kd, Pd):
            a.release()
