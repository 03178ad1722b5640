"""GPU tests of the DF-HF request loop: each result is the same alone and in a
mix of concurrent requests, admission respects the HBM budget, and a served
FP64 request reproduces the direct FP64 SCF."""
import time

import numpy as np
import pytest

from chem_amd import desc, ints

pytestmark = pytest.mark.gpu


def _mol(n, seed):
    return desc.Molecule(desc.water_cluster(n, seed))


def test_alone_vs_mix():
    """A request's energy is identical (to 1e-12 Eh; the kernels reduce in a
    fixed order, so bitwise in practice) whether it runs alone or while other
    requests share the GPU."""
    from chem_amd.serve import Server
    target = _mol(3, 5)
    others = [_mol(4, 6), _mol(2, 7), _mol(5, 8), _mol(1, 9)]
    with Server(budget_bytes=50e9, nslots=3) as srv:
        print("case: serve alone, 3 waters mixed_auto", flush=True)
        r0 = srv.result(srv.submit(target, threshold=1e-7, kmode="mixed_auto"))
        print("case: serve mix of 6 requests", flush=True)
        ids = [srv.submit(m, threshold=1e-7, kmode="mixed_auto") for m in others[:2]]
        tid = srv.submit(target, threshold=1e-7, kmode="mixed_auto")
        ids += [srv.submit(m, threshold=1e-7, kmode="mixed_auto") for m in others[2:]]
        tid2 = srv.submit(target, threshold=1e-7, kmode="mixed_auto")
        rs = [srv.result(i) for i in ids]
        r1, r2 = srv.result(tid), srv.result(tid2)
        st = srv.stats()
    print("alone %.12f mix %.12f %.12f  max live %d" % (r0["energy"], r1["energy"], r2["energy"], st["max_live"]))
    assert r0["converged"] and r1["converged"] and r2["converged"]
    assert all(r["converged"] for r in rs)
    assert abs(r1["energy"] - r0["energy"]) <= 1e-12
    assert abs(r2["energy"] - r0["energy"]) <= 1e-12
    assert r1["iterations"] == r0["iterations"]
    assert st["completed"] == 7 and st["failed"] == 0
    assert st["max_live"] >= 2          # the mix did overlap on the device


def test_cross_request_batching():
    """With several requests live, their BF16x3 exchange builds run in shared
    grouped launches (batches of two or more builds), and every energy equals
    the one served with batching off, bit for bit."""
    from chem_amd.serve import Server
    mols = [_mol(n, 30 + n) for n in (2, 4, 3, 6, 5, 2)]
    res = {}
    for on in (False, True):
        with Server(budget_bytes=60e9, nslots=4) as srv:
            srv.set_batching(on)
            print("case: serve 6 requests, batching %s" % on, flush=True)
            ids = [srv.submit(m, threshold=1e-7, kmode="mixed_auto") for m in mols]
            res[on] = [srv.result(i) for i in ids]
            bs = srv.batch_stats()
            st = srv.stats()
        print("batching %s: %s, max live %d" % (on, bs, st["max_live"]))
        assert st["completed"] == len(mols) and st["failed"] == 0
        if on:
            assert bs["builds"] > 0 and bs["largest"] >= 2
        else:
            assert bs["builds"] == 0
    for a, b in zip(res[False], res[True]):
        assert a["converged"] and b["converged"]
        assert a["energy"] == b["energy"] and a["iterations"] == b["iterations"]


def test_served_fp64_matches_direct_scf():
    """kmode 0 through the loop (integrals on the GPU, mask in C++) equals the
    direct device SCF driven from Python on the same molecule and the same
    (GPU) one-electron integrals and metric."""
    from chem_amd.engine import Context, DeviceB, Metric
    from chem_amd.scf import run_rhf
    from chem_amd.serve import Server
    mol = _mol(2, 3)
    print("case: served FP64 vs direct, 2 waters", flush=True)
    with Server(budget_bytes=20e9, nslots=1) as srv:
        r = srv.result(srv.submit(mol, threshold=1e-8, kmode="dfmo"))
    ctx = Context(0)
    S, T, V = ints.int1e_gpu(ctx, mol)
    dB = DeviceB.build_on_gpu(ctx, mol, 1e-10)
    met = Metric(ctx, ints.int2c2e_gpu(ctx, mol))
    d = run_rhf(mol, dB, met, S, T + V, threshold=1e-8, kmode="dfmo")
    print("served %.12f direct %.12f" % (r["energy"], d.energy))
    assert r["converged"] and d.converged
    assert abs(r["energy"] - d.energy) < 1e-10
    ctx.close()


def test_admission_budget():
    """With a budget that holds one request at a time, requests are admitted
    one by one in arrival order and all complete; the peak stays in budget."""
    from chem_amd.serve import Server
    mols = [_mol(2, s) for s in (21, 22, 23)]
    with Server(budget_bytes=20e9, nslots=3) as probe:
        est = probe.result(probe.submit(mols[0], kmode="mixed_auto"))["est_bytes"]
    budget = 1.5 * est
    print("case: admission, budget %.3g bytes (estimate %.3g)" % (budget, est), flush=True)
    with Server(budget_bytes=budget, nslots=3) as srv:
        ids = [srv.submit(m, kmode="mixed_auto") for m in mols]
        rs = [srv.result(i) for i in ids]
        st = srv.stats()
    assert all(r["converged"] for r in rs)
    assert st["max_live"] == 1 and st["peak_bytes"] <= budget
    admits = [r["t_admit"] for r in rs]
    assert admits == sorted(admits)
    for a, b in zip(rs, rs[1:]):      # each admitted only after the previous one left
        assert b["t_admit"] >= a["t_done"] - 1e-3


def test_oversized_request_fails_cleanly():
    """A request whose estimate exceeds the whole budget fails with an error
    and does not block later requests."""
    from chem_amd import _lib
    from chem_amd.serve import Server
    with Server(budget_bytes=1e6, nslots=1) as srv:
        rid = srv.submit(_mol(1, 1), kmode="dfmo")
        with pytest.raises(_lib.MchipError, match="budget"):
            srv.result(rid)
        st = srv.stats()
    assert st["failed"] == 1
