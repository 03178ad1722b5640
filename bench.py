"""Benchmark: DF-K seconds per SCF iteration on one MI355X.

Workload (default --config c2): seeded 64-water cluster, cc-pVDZ / cc-pVDZ-jkfit,
atomic blocks, Schwarz screening at 1e-10.

Plain path: the reference's DFMO algorithm in FP64 (K1 = B.C, K2 = L^-1 per
build, K3 = SYRK).  Chosen form: copies of B~ = L^-1 B made once from B (which
stays FP64 for J and the energy): a BF16 hi/lo split, an FP32 copy and an FP64
copy.  --form auto (default) runs the library's precision escalation: BF16x3
builds until the SCF error falls below --escalate (or stops improving), then
FP32 builds, then FP64 ones if FP32 stalls; every build runs on MFMA.  J and
the FP64 exchange energy tr(K P) (on the FP64 copy, -sum_X ||W_X^T C||^2
without forming K) at the converged orbitals give the reported energy with
the reference's formulas.

A step is one SCF iteration's DF-K work in the chosen form.  The timed region
replays the chosen SCF's sequence of builds (same forms, same count; steps
defaults to its iteration count) plus the final FP64 exchange energy, so
value = (sum of the step builds + final energy) / steps.  The line also carries the
plain FP64 seconds per build, |dE| between the two SCFs, the roofline of the
dominant kernel (K1 of the BF16x3 form) from HIP events on the library stream,
and the oracle timed on the host.

--config c3 times the batched SOS-MP2 graph and --config c4 the request stream;
each prints its own JSON line.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c0|c1|c2|c3|c4]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

CONFIGS = {
    "c0": dict(nwater=1, seed=0),
    "c1": dict(nwater=32, seed=1234),
    "c2": dict(nwater=64, seed=1234),
    # batched SOS-AO-MP2: 32 waters, cc-pVDZ / cc-pVDZ-ri, 8 Laplace points
    "c3": dict(nwater=32, seed=1234, nlap=8),
    # serving stream: Poisson arrivals at 4/s, clusters of U{4..64} waters
    "c4": dict(requests=200, rate=4.0, nmin=4, nmax=64, seed=2024, budget=200e9, thresh=1e-6),
    # a short stream of the c4 kind run after the default c2 line, so that the
    # default run also reports served latencies (not the c4 workload itself)
    "c4_sample": dict(requests=24, rate=4.0, nmin=4, nmax=16, seed=2024, budget=200e9, thresh=1e-6),
}
MFMA_F64_PEAK = 78.6e12
MFMA_F32_PEAK = 157.3e12
MFMA_BF16_PEAK = 2.5e15
HBM_PEAK = 8.0e12


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def cpu_baseline(mol, dB, Linv, C_occ, budget_s=20.0):
    """The oracle's DFMO-K (numpy/OpenBLAS, all host threads) timed on a bounded
    sample and scaled to one full K build:
      leg 1: oracle.dfock.dfmo_k on the first r aux rows of B with the r x r
             corner of L^-1 (K1 and K3 are linear in naux), scaled by naux / r;
      leg 2: the metric step K2 = L^-1 HT1 (naux^2 nbf nocc flop) timed as the
             full naux x naux triangular multiply on s of the nbf*nocc columns,
             scaled by nbf*nocc / s.
    Leg 1 also holds the r x r metric corner, so the sum overcounts K2 by r/naux."""
    from scipy.linalg.blas import dtrmm
    from oracle import dfock
    naux, nbf, nocc = mol.naux, mol.nbf, C_occ.shape[1]
    rows = min(naux, 32)
    while True:
        Bs = dB.dense_rows(0, rows)
        Ls = np.ascontiguousarray(Linv[:rows, :rows].T)
        t0 = time.perf_counter()
        dfock.dfmo_k(Bs, Ls, C_occ)
        dt1 = time.perf_counter() - t0
        if dt1 > budget_s / 4 or rows == naux:
            break
        rows = min(naux, rows * 2)
    del Bs
    ncol = nbf * nocc
    s = min(ncol, 256)
    L = np.asfortranarray(Linv)
    while True:
        X = np.asfortranarray(np.random.default_rng(0).standard_normal((naux, s)))
        t0 = time.perf_counter()
        dtrmm(1.0, L, X, side=0, lower=1)
        dt2 = time.perf_counter() - t0
        if dt2 > budget_s / 4 or s == ncol:
            break
        s = min(ncol, s * 2)
    est = dt1 * naux / rows + dt2 * ncol / s
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return dict(value=est, unit="s/iter", cores=cores, kind="port", cpu=cpu_model(),
                sample="oracle DFMO-K on aux rows 0..%d of %d (%.2f s, x naux/rows) + L^-1 "
                       "triangular multiply on %d of %d HT1 columns (%.2f s, x cols/sample)" % (
                           rows, naux, dt1, s, ncol, dt2))


def measured_traffic():
    """HBM bytes per launch of the exchange kernels at 64 waters, from the
    rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs summarised under profiles/
    (lines `traffic_bytes_per_launch <kernel> <bytes>`, calibrated against a
    copy kernel of known size)."""
    out = {}
    pdir = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles")
    if not os.path.isdir(pdir):
        return out
    for f in sorted(os.listdir(pdir)):
        if not f.endswith(".md"):
            continue
        with open(os.path.join(pdir, f)) as fh:
            for line in fh:
                parts = line.split()
                if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch":
                    out[parts[1]] = float(parts[2])
    return out


def measured_mfma_util():
    """MFMA busy fraction per kernel from the rocprofv3 --pmc runs summarised
    under profiles/ (lines `mfma_util <kernel> <fraction>`)."""
    out = {}
    pdir = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles")
    if not os.path.isdir(pdir):
        return out
    for f in sorted(os.listdir(pdir)):
        if not f.endswith(".md"):
            continue
        with open(os.path.join(pdir, f)) as fh:
            for line in fh:
                parts = line.split()
                if len(parts) == 3 and parts[0] == "mfma_util":
                    out[parts[1]] = float(parts[2])
    return out


def exact_kernels(form, plan):
    """Template names (as the profiler prints them, spaces removed) of the slab
    GEMM and SYRK an autotuned exchange plan runs: the keys of the PMC lines."""
    if not plan:
        return None, None
    wn = plan["threads"] // 128
    tm, tn = plan["bm"] // 32, plan["bn"] // (16 * wn)
    if form == "bf16x3":
        base = {0: "k_slab_gemm_bf16x3", 1: "k_slab_gemm_bf16x3_dl", 2: "k_slab_gemm_bf16x3_dla"}[plan["pipeline"]]
        k1 = "%s<2,%d,%d,%d,%d>" % (base, wn, tm, tn, plan["stages"])
        v = plan["syrk_variant"]
        k3 = "k_seg_syrk_bf16x3<128>" if v == 0 else (
            "k_seg_syrk_bf16x3_dl<2,256,8,%s>" % ("true" if v == 4 else "false") if v >= 4
            else "k_seg_syrk_bf16x3_dl<%d,128,4,true>" % (v + 1))
    else:
        k1 = "k_slab_gemm_%s_dl<2,%d,%d,%d,%d>" % (form, wn, tm, tn, plan["stages"])
        k3 = "k_seg_syrk_%s_dl<%d>" % (form, (plan["syrk_stages"] if form == "f32" else plan["syrk_variant"] + 2))
    return k1, k3


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def serve_stream(args, log_progress=True, cfg=None, ncheck=None):
    """The DF-HF request stream: seeded Poisson arrivals, water clusters of
    uniformly random size, the chosen form (BF16x3 escalating to FP32, FP64
    final energy) per request, admission against the HBM budget.  Returns the
    latency percentiles, throughput and the energy check against the plain
    FP64 path on a sample of the requests.  cfg: the stream (default c4)."""
    from chem_amd import desc
    from chem_amd.serve import Server
    if cfg is None:
        cfg = CONFIGS["c4"]
        nreq = args.requests or cfg["requests"]
    else:
        nreq = cfg["requests"]
    ncheck = args.serve_check if ncheck is None else ncheck
    rng = np.random.default_rng(cfg["seed"])
    arrivals = np.cumsum(rng.exponential(1.0 / cfg["rate"], nreq))
    sizes = rng.integers(cfg["nmin"], cfg["nmax"] + 1, nreq)
    mols = {}
    results = []
    with Server(budget_bytes=cfg["budget"], nslots=args.slots) as srv:
        srv.set_batching(not args.no_batch)
        t0 = time.perf_counter()
        ids = []
        for i in range(nreq):
            wait = t0 + arrivals[i] - time.perf_counter()
            if wait > 0:
                time.sleep(wait)
            mol = desc.Molecule(desc.water_cluster(int(sizes[i]), 5000 + i))
            mols[i] = mol
            ids.append(srv.submit(mol, threshold=cfg["thresh"], kmode="mixed_auto"))
        failures = []
        for i, rid in enumerate(ids):
            try:
                r = srv.result(rid)
            except Exception as e:          # a failed request is reported, not fatal
                failures.append((rid, int(sizes[i]), str(e)))
                log("request %d (%d waters) FAILED: %s" % (rid, sizes[i], e))
                results.append(None)
                continue
            r["nwater"] = int(sizes[i])
            results.append(r)
            if log_progress:
                log("request %d (%d waters): E=%.8f iters=%d latency %.2f s (admitted after %.2f s, slot %d; "
                    "host %.2f build %.2f setup %.2f scf %.2f s)" % (
                        rid, sizes[i], r["energy"], r["iterations"], r["latency"], r["t_admit"] - r["t_submit"],
                        r["slot"], r["t_host"], r["t_build"], r["t_setup"], r["t_scf"]))
        st = srv.stats()
        bst = srv.batch_stats()
    ok = [r for r in results if r is not None]
    lat = np.array([r["latency"] for r in ok])
    span = max(r["t_done"] for r in ok) - min(r["t_submit"] for r in ok)
    fock_builds = sum(r["iterations"] + 1 for r in ok)
    # energy check: plain FP64 path (kmode dfmo, the reference formulas) on a sample
    sample = [i for i in sorted(set(int(i) for i in np.linspace(0, nreq - 1, min(ncheck, nreq)).round()))
              if results[i] is not None] if ncheck > 0 else []
    # (both paths also at 1e-8: at the served threshold two FP64 SCFs that take
    # different iteration paths already differ by up to ~1e-6 Eh, so the tight
    # pair isolates the error of the chosen form's exchange)
    dEs, dEt = [], []
    with Server(budget_bytes=cfg["budget"], nslots=1) as ref:
        for i in sample:
            r64 = ref.result(ref.submit(mols[i], threshold=cfg["thresh"], kmode="dfmo"))
            dEs.append(abs(results[i]["energy"] - r64["energy"]))
            r64t = ref.result(ref.submit(mols[i], threshold=1e-8, kmode="dfmo"))
            rmt = ref.result(ref.submit(mols[i], threshold=1e-8, kmode="mixed_auto"))
            dEt.append(abs(rmt["energy"] - r64t["energy"]))
            if log_progress:
                log("check request %d (%d waters): |dE| vs plain FP64 %.2e at %.0e, %.2e with both at 1e-8" % (
                    i, sizes[i], dEs[-1], cfg["thresh"], dEt[-1]))
    return dict(requests=nreq, rate_per_s=cfg["rate"], sizes="U{%d..%d} waters" % (cfg["nmin"], cfg["nmax"]),
                budget_gb=cfg["budget"] / 1e9, slots=args.slots, p50_s=float(np.percentile(lat, 50)),
                p99_s=float(np.percentile(lat, 99)), mean_s=float(lat.mean()), max_s=float(lat.max()),
                scf_iterations_per_s=fock_builds / span, requests_per_s=len(ok) / span, span_s=span,
                max_live=st["max_live"], peak_hbm_gb=st["peak_bytes"] / 1e9, failed=st["failed"],
                failures=failures[:5], all_converged=all(r["converged"] for r in ok),
                max_abs_dE_vs_fp64=float(max(dEs)) if dEs else None, dE_sample=len(dEs),
                max_abs_dE_vs_fp64_both_1e8=float(max(dEt)) if dEt else None,
                cross_request_batching=not args.no_batch, exchange_batches=bst["batches"],
                batched_builds=bst["builds"], largest_batch=bst["largest"],
                mean_phase_s={k: float(np.mean([r[k] for r in ok])) for k in ("t_host", "t_build", "t_setup",
                                                                              "t_scf")})


def mp2_batched_bench(args):
    """SOS-AO-MP2 Z builders on the 32-water cluster: FP64 per-point launches
    (the reference's loop over Laplace points) against all 8 points as grouped
    FP32 launches (FP64 accumulation) in one hipGraph.  Orbitals from the FP64
    GPU SCF (cc-pVDZ-jkfit, 1e-8)."""
    from chem_amd import desc, ints
    from chem_amd.engine import Context, DeviceB, Metric
    from chem_amd.mp2 import sos_mp2, laplace_for
    from chem_amd.scf import run_rhf
    cfg = CONFIGS["c3"]
    atoms = desc.water_cluster(cfg["nwater"], cfg["seed"])
    molj = desc.Molecule(atoms)
    mol = desc.Molecule(atoms, dfbasis="cc-pvdz-ri")
    ctx = Context(0)
    S, T, V = ints.int1e(molj)
    dBj = DeviceB.build_on_gpu(ctx, molj, 1e-10)
    metj = Metric(ctx, ints.int2c2e(molj))
    r = run_rhf(molj, dBj, metj, S, T + V, threshold=1e-8)
    log("SCF: E=%.10f iters=%d" % (r.energy, r.iterations))
    dBj.release()
    metj.release()
    dB = DeviceB.build_on_gpu(ctx, mol, 1e-10)
    met = Metric(ctx, ints.int2c2e(mol))
    nnz = dB.nnz_kept()
    w, a, err = laplace_for(r.eps, mol.nocc, cfg["nlap"])
    t0 = time.perf_counter()
    ref = sos_mp2(mol, dB, met, S, r.C, r.eps, points=(w, a))
    t_ref = time.perf_counter() - t0
    log("FP64 per point: E=%.12f Z builders %.3f s (total %.2f s)" % (ref["energy"], ref["t_z"], t_ref))
    t0 = time.perf_counter()
    got = sos_mp2(mol, dB, met, S, r.C, r.eps, points=(w, a), batched=True)
    t_b = time.perf_counter() - t0
    log("batched graph: E=%.12f graph %.3f s (set-up %.2f s, total %.2f s)" % (
        got["energy"], got["t_z"], got["t_prep"], t_b))
    # algorithmic work per point: 3 * 2 nnz(B) nocc + 2 naux nnz(B)
    flops = cfg["nlap"] * (6.0 * nnz * mol.nocc + 2.0 * mol.naux * nnz)
    return dict(nwater=cfg["nwater"], nbf=mol.nbf, naux=mol.naux, nocc=mol.nocc, nlap=cfg["nlap"], nnz_B=nnz,
                laplace_err=float(err), E_fp64=ref["energy"], E_batched=got["energy"],
                dE=got["energy"] - ref["energy"], rel_err=abs(got["energy"] - ref["energy"]) / ref["energy"],
                max_term_err=float(np.abs(got["terms"] - ref["terms"]).max()),
                t_z_fp64_per_point=ref["t_z"], t_z_batched_graph=got["t_z"], t_batched_setup=got["t_prep"],
                speedup_z=ref["t_z"] / got["t_z"], graph_tflops=flops / got["t_z"] / 1e12)


def kernels_only(args):
    """The chosen form's exchange builds alone, for profiling: B (its GPU build
    timed), metric and the copies of L^-1 B of the 64-water cluster, then
    `steps` BF16x3, FP32 and FP64-copy builds each at orthonormalised random
    occupied coefficients (the kernels' work does not depend on the orbitals).
    Everything in this mode runs on the library's own kernels (no vendor BLAS
    call)."""
    from chem_amd import desc, ints
    from chem_amd.engine import Context, DeviceB, Metric
    cfg = CONFIGS[args.config if args.config in ("c1", "c2") else "c2"]
    mol = desc.Molecule(desc.water_cluster(cfg["nwater"], cfg["seed"]), reorder=not args.no_reorder)
    ctx = Context(0)
    t = time.perf_counter()
    dB = DeviceB.build_on_gpu(ctx, mol, 1e-10)
    ctx.sync()
    out = {"b_build_gpu_s": time.perf_counter() - t}
    log("GPU build of B %.3f s" % out["b_build_gpu_s"])
    met = Metric(ctx, ints.int2c2e(mol))
    dB.make_lowp(met, f64=True)
    S, _, _ = ints.int1e(mol)
    rng = np.random.default_rng(7)
    C = rng.standard_normal((mol.nbf, mol.nocc))
    w, U = np.linalg.eigh(C.T @ S @ C)
    C = np.ascontiguousarray(C @ U / np.sqrt(w))
    steps = args.steps or 3
    for name, fn in (("bf16x3", dB.dfk_mo_bf16), ("f32", dB.dfk_mo_f32), ("fp64_copy", dB.dfk_mo_tilde64)):
        fn(C)
        ctx.sync()
        t = time.perf_counter()
        for _ in range(steps):
            fn(C)
        ctx.sync()
        out[name + "_build_s"] = (time.perf_counter() - t) / steps
        log("%s build %.4f s" % (name, out[name + "_build_s"]))
    K64 = dB.dfk_mo_tilde64(C)
    for name, fn in (("bf16x3", dB.dfk_mo_bf16), ("f32", dB.dfk_mo_f32)):
        out[name + "_frob_rel_vs_fp64"] = float(np.linalg.norm(fn(C) - K64) / np.linalg.norm(K64))
    out["bf16_plan"] = dB.bf16_plan(mol.nocc)
    out["f32_plan"] = dB.f32_plan(mol.nocc)
    out["f64_plan"] = dB.f64_plan(mol.nocc)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0, help="0: the chosen SCF's iteration count")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c2", choices=sorted(CONFIGS))
    ap.add_argument("--scf-thresh", type=float, default=1e-6,
                    help="convergence of the chosen-form SCF whose build mix is timed")
    ap.add_argument("--ref-thresh", type=float, default=1e-8,
                    help="convergence of the plain FP64 reference SCF (and of the chosen-form accuracy rerun)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--kernels-only", action="store_true",
                    help="profiling mode: exchange builds only (no SCF), one JSON line of timings")
    ap.add_argument("--escalate", type=float, default=1e-5,
                    help="auto form: error RMS at which BF16x3 builds give way to FP32 (0: off)")
    ap.add_argument("--stall", type=int, default=3,
                    help="auto form: iterations without progress before BF16x3 gives way to FP32 (0: off)")
    ap.add_argument("--requests", type=int, default=0, help="c4: number of requests (0: 200)")
    ap.add_argument("--no-serve-sample", action="store_true",
                    help="c2: skip the short served stream (24 requests of 4-16 waters) after the DF-K line")
    ap.add_argument("--slots", type=int, default=4, help="c4: request workers")
    ap.add_argument("--no-batch", action="store_true",
                    help="c4: no cross-request batching of the exchange builds (each on its worker's stream)")
    ap.add_argument("--serve-check", type=int, default=4,
                    help="c4: requests re-run on the plain FP64 path for the energy check")
    ap.add_argument("--no-reorder", action="store_true",
                    help="keep the generator's atom order (the reference's \"reorder\": false)")
    ap.add_argument("--form", default="auto", choices=["auto", "bf16", "f32"],
                    help="precision of the chosen-form exchange builds (auto: BF16x3 until the "
                         "SCF error reaches --escalate or stalls, then FP32, then FP64 if FP32 stalls)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only (the DF-K path runs on a single MI355X); "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    from chem_amd import desc, ints
    from chem_amd.engine import Context, DeviceArray, DeviceB, Metric
    from chem_amd.scf import run_rhf

    if args.kernels_only:
        print(json.dumps(kernels_only(args)), flush=True)
        return

    if args.config == "c3":
        mb = mp2_batched_bench(args)
        out = {
            "metric": "SOS-AO-MP2 Z builders, 8 Laplace points in one grouped hipGraph (FP32, FP64 accumulation)",
            "value": mb["t_z_batched_graph"], "unit": "s", "n_gpus": 1, "steps": 1, "warmup": 0,
            "ms_per_step": mb["t_z_batched_graph"] * 1e3, "higher_is_better": False, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32 contractions, f64 accumulation and energies",
            "data": "synthetic seeded water cluster, real integrals",
            "config": {"workload": "SOS-AO-MP2 (c3): %d waters cc-pVDZ/cc-pVDZ-ri, %d Laplace points" % (
                mb["nwater"], mb["nlap"])},
            "mp2": mb,
        }
        print(json.dumps(out), flush=True)
        return

    if args.config == "c4":
        sv = serve_stream(args)
        out = {
            "metric": "DF-HF serving: p99 time-to-converged-energy",
            "value": sv["p99_s"], "unit": "s", "n_gpus": 1, "steps": sv["requests"], "warmup": 0,
            "ms_per_step": sv["mean_s"] * 1e3, "higher_is_better": False, "scaling": "weak",
            "vs_baseline": None, "dtype": "bf16x3 split (K builds) + f64 (J, final K, SCF)",
            "data": "synthetic seeded water clusters, real integrals",
            "config": {"workload": "request stream (c4): Poisson %.0f/s, %s, cc-pVDZ/cc-pVDZ-jkfit, "
                                   "conv %.0e, HBM budget %.0f GB" % (
                                       sv["rate_per_s"], sv["sizes"], CONFIGS["c4"]["thresh"], sv["budget_gb"])},
            "serving": sv,
        }
        print(json.dumps(out), flush=True)
        return

    cfg = CONFIGS[args.config]
    t0 = time.perf_counter()
    mol = desc.Molecule(desc.water_cluster(cfg["nwater"], cfg["seed"]), reorder=not args.no_reorder)
    log("molecule: %d waters nbf=%d naux=%d nocc=%d" % (cfg["nwater"], mol.nbf, mol.naux, mol.nocc))
    S, T, V = ints.int1e(mol)
    M = ints.int2c2e(mol)
    ctx = Context(0)
    tb = time.perf_counter()
    dB = DeviceB.build_on_gpu(ctx, mol, 1e-10)
    ctx.sync()
    t_build = time.perf_counter() - tb
    nnz = dB.nnz_kept()
    log("B: stored %.3e kept %.3e elements (%.1f%% of dense) built on the GPU in %.1f s" % (
        dB.nelem, nnz, 100.0 * nnz / (mol.naux * mol.nbf ** 2), t_build))
    met = Metric(ctx, M)

    # plain FP64 path: DFMO with the metric applied per build (reference
    # algorithm), converged to --ref-thresh: the energy and orbital-energy reference
    t1 = time.perf_counter()
    ref = run_rhf(mol, dB, met, S, T + V, threshold=args.ref_thresh, kmode="dfmo")
    log("plain FP64 SCF (thr %.0e): E=%.10f iters=%d (%.1f s) K s/iter median %.4f" % (
        args.ref_thresh, ref.energy, ref.iterations, time.perf_counter() - t1,
        float(np.median(ref.t_k[1:] or ref.t_k))))
    C_occ = np.ascontiguousarray(ref.C[:, :mol.nocc])

    def time_builds(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        ctx.sync()
        t = time.perf_counter()
        for _ in range(steps):
            fn()
        ctx.sync()
        return (time.perf_counter() - t) / steps

    # plain FP64 build with the occupied coefficients and K resident in HBM
    C0d = DeviceArray(ctx, C_occ.shape, C_occ)
    K0d = DeviceArray(ctx, (mol.nbf, mol.nbf))
    plain_s = time_builds(lambda: dB.dfk_dev("fp64", C0d, mol.nocc, K0d, metric=met), 2, 1)
    C0d.release()
    K0d.release()

    K_plain = dB.dfk_mo(C_occ, met)
    cpu = None
    if not args.no_cpu_baseline:
        Linv, _ = met.get()
        cpu = cpu_baseline(mol, dB, Linv, C_occ)
        log("cpu baseline: %.2f s/iter (%s)" % (cpu["value"], cpu["sample"]))

    # the plain exchange's two naux x nbf x nocc work arrays are not needed from
    # here on; at 64 waters B and its three copies leave no room for them
    ctx.trim()
    # chosen form: low-precision copies of B~ = L^-1 B made once from B, which
    # stays FP64 for J, the final exchange and the energy (reference formulas)
    form = args.form
    tc = time.perf_counter()
    # (and the FP64 L^-1 B slabs: the FP64 exchanges of the chosen SCF contract
    # (L^-1 B) C instead of applying L^-1 to B C in every build)
    dB.make_lowp(met, f32=form in ("auto", "f32"), bf16=form in ("auto", "bf16"), f64=True)
    ctx.sync()
    t_copies = time.perf_counter() - tc
    kmode = {"auto": "mixed_auto", "bf16": "mixed_bf16", "f32": "mixed_f32"}[form]
    ctx.set_escalation(args.escalate, args.stall)
    t2 = time.perf_counter()
    chosen = run_rhf(mol, dB, met, S, T + V, threshold=args.scf_thresh, kmode=kmode)
    dE = chosen.energy - ref.energy
    nbuild = len(chosen.t_k)
    if form == "auto":
        # the library's escalation rule (select_k_builders, K mode 4): stage 0
        # BF16x3, 1 FP32, 2 FP64; a stage ends at the RMS trigger (stage 0) or
        # after args.stall iterations (below the stage's arming RMS) without the
        # RMS falling below 0.8 x its best
        stage_of, stage, best, flat = [], 0, float("inf"), 0
        for i in range(nbuild):
            prev = chosen.history_rms[i - 1] if i > 0 else float("inf")
            if stage < 2 and np.isfinite(prev):
                nxt = stage == 0 and args.escalate > 0 and prev <= args.escalate
                if prev >= (1e-4, 1e-6)[stage]:     # stall counting is armed below these
                    flat = 0
                elif prev < 0.8 * best:
                    flat = 0
                elif args.stall > 0:
                    flat += 1
                    nxt = nxt or flat >= args.stall
                best = min(best, prev)
                if nxt:
                    stage, best, flat = stage + 1, float("inf"), 0
            stage_of.append(stage)
    else:
        stage_of = [0 if form == "bf16" else 1] * nbuild
    scf_mix = [stage_of.count(k) for k in range(3)]
    log("chosen SCF (%s, thr %.0e): E=%.10f (low-precision-K iterate %.10f) iters=%d (%d BF16 + %d FP32 + "
        "%d FP64 builds) dE vs plain %.3e (%.1f s)" % (form, args.scf_thresh, chosen.energy, chosen.energy_iter,
                                                      chosen.iterations, scf_mix[0], scf_mix[1], scf_mix[2], dE,
                                                      time.perf_counter() - t2))
    # the same chosen form converged as tightly as the reference: its own
    # error, free of the convergence error of a 1e-6 SCF
    chosen_tight = run_rhf(mol, dB, met, S, T + V, threshold=args.ref_thresh, kmode=kmode)
    dE_tight = chosen_tight.energy - ref.energy
    log("chosen SCF at thr %.0e: E=%.10f iters=%d dE vs plain %.3e" % (
        args.ref_thresh, chosen_tight.energy, chosen_tight.iterations, dE_tight))

    # the timed steps replay the chosen SCF's build mix: all of its builds in
    # order when steps is its build count (the default); otherwise each stage
    # gets its share of the steps (largest remainders), in stage order
    steps = args.steps or max(1, nbuild)
    if steps == nbuild:
        seq = list(stage_of)
    else:
        share = [steps * c / nbuild for c in scf_mix]
        cnt = [int(x) for x in share]
        for k in sorted(range(3), key=lambda k: share[k] - cnt[k], reverse=True)[:steps - sum(cnt)]:
            cnt[k] += 1
        seq = [k for k in range(3) for _ in range(cnt[k])]
    nb = [seq.count(k) for k in range(3)]
    n16 = nb[0]
    Cc = np.ascontiguousarray(chosen.C[:, :mol.nocc])
    # the occupied coefficients and K stay in HBM for the timed builds (as in
    # the library's SCF loop): no host copies inside the timed region
    Cd = DeviceArray(ctx, Cc.shape, Cc)
    Kd = DeviceArray(ctx, (mol.nbf, mol.nbf))

    def dev_build(kind):
        return lambda: dB.dfk_dev(kind, Cd, mol.nocc, Kd)
    k16 = dev_build("bf16x3") if form in ("auto", "bf16") else None
    k32 = dev_build("f32") if form in ("auto", "f32") else None
    k64 = dev_build("fp64_copy")

    def kfin():   # the final FP64 exchange energy tr(K P) of the reported energy
        return dB.exchange_energy_dev(Cd, mol.nocc)
    for _ in range(args.warmup):
        for k in (k16, k32, k64, kfin):
            if k:
                k()
    ctx.sync()
    t = time.perf_counter()
    for st in seq:
        (k16, k32, k64)[st]()
    ek = kfin()
    ctx.sync()
    step_s = (time.perf_counter() - t) / steps
    final_s = time_builds(kfin, 2, 0)
    # the energy-only evaluation against the trace of the full FP64 exchange
    k64()
    ctx.sync()
    ek_full = float(np.sum(Kd.download() * (Cc @ Cc.T)))
    log("final exchange energy %.12f, tr(K P) of the full FP64 exchange %.12f (rel %.1e), %.4f s" % (
        ek, ek_full, abs(ek - ek_full) / abs(ek_full), final_s))
    t16 = time_builds(k16, 3, 0) if k16 else None
    t32 = time_builds(k32, 3, 0) if k32 else None

    # roofline of the dominant kernel from HIP events on the library stream:
    # K1 (the B~ C slab GEMM) of the form that runs most of the builds
    def phases(fn, nrep=3):
        ctx.set_timing(True)
        for _ in range(nrep):
            fn()
        ph = ctx.phase_times() / nrep
        ctx.set_timing(False)
        return ph
    k1_flops = 2.0 * nnz * mol.nocc
    k3_flops = float(mol.naux) * mol.nocc * mol.nbf ** 2   # symmetric half of 2*naux*nocc*nbf^2
    kernels = {}
    if k16:
        ph = phases(k16)
        kernels["k_slab_gemm_bf16x3"] = dict(s=ph[0], tflops=k1_flops / ph[0] / 1e12,
                                             builds=n16, peak=MFMA_BF16_PEAK / 1e12)
        kernels["k_seg_syrk_bf16x3"] = dict(s=ph[2], tflops=k3_flops / ph[2] / 1e12,
                                            builds=n16, peak=MFMA_BF16_PEAK / 1e12)
        kernels["bf16_plan"] = dB.bf16_plan(mol.nocc)
    if k32:
        ph = phases(k32)
        kernels["k_slab_gemm_f32_dl"] = dict(s=ph[0], tflops=k1_flops / ph[0] / 1e12,
                                             builds=nb[1], peak=MFMA_F32_PEAK / 1e12)
        kernels["k_seg_syrk_f32_dl"] = dict(s=ph[2], tflops=k3_flops / ph[2] / 1e12,
                                            builds=nb[1], peak=MFMA_F32_PEAK / 1e12)
        kernels["f32_plan"] = dB.f32_plan(mol.nocc)
    ph = phases(k64, nrep=2)
    kernels["k_slab_gemm_f64_dl"] = dict(s=ph[0], tflops=k1_flops / ph[0] / 1e12, builds=1 + nb[2],
                                         peak=MFMA_F64_PEAK / 1e12)
    kernels["k_seg_syrk_f64_dl"] = dict(s=ph[2], tflops=k3_flops / ph[2] / 1e12, builds=1 + nb[2],
                                        peak=MFMA_F64_PEAK / 1e12)
    kernels["f64_plan"] = dB.f64_plan(mol.nocc)
    dom = max((k for k in kernels if isinstance(kernels[k], dict) and "s" in kernels[k]),
              key=lambda k: kernels[k]["s"] * kernels[k]["builds"])
    d = kernels[dom]
    # the exact kernel instances the autotuned plans run, for the PMC figures
    exact = {}
    for key, kform, pname in (("k_slab_gemm_bf16x3", "bf16x3", "bf16_plan"), ("k_seg_syrk_bf16x3", "bf16x3", "bf16_plan"),
                             ("k_slab_gemm_f32_dl", "f32", "f32_plan"), ("k_seg_syrk_f32_dl", "f32", "f32_plan"),
                             ("k_slab_gemm_f64_dl", "f64", "f64_plan"), ("k_seg_syrk_f64_dl", "f64", "f64_plan")):
        if key in kernels and kernels.get(pname):
            k1, k3 = exact_kernels(kform, kernels[pname])
            exact[key] = k1 if key.startswith("k_slab") else k3
            kernels[key]["instance"] = exact[key]
    traffic = measured_traffic().get(exact.get(dom, dom))
    roof = dict(bound="mfma", achieved=d["tflops"], peak=d["peak"], unit="TFLOP/s",
                frac=d["tflops"] / d["peak"], traffic=traffic, kernel=dom, kernel_s=d["s"],
                kernels=kernels)
    if traffic:
        # HBM bytes per launch from the committed PMC runs, over this run's launch time
        roof["hbm_gbs"] = traffic / d["s"] / 1e9
        roof["hbm_frac"] = roof["hbm_gbs"] / 8000.0
    util = measured_mfma_util().get(exact.get(dom, dom))
    if util is not None:
        roof["mfma_util_counter"] = util

    # accuracy of each low-precision K at the FP64 orbitals
    kerr = {}
    for name, fn in (("bf16x3", dB.dfk_mo_bf16 if k16 else None), ("f32", dB.dfk_mo_f32 if k32 else None)):
        if fn is None:
            continue
        dK = fn(C_occ) - K_plain
        kerr[name] = dict(max_abs=float(np.abs(dK).max()),
                          frob_rel=float(np.linalg.norm(dK) / np.linalg.norm(K_plain)))
        log("%s K error vs FP64: max %.3e, Frobenius rel %.3e" % (
            name, kerr[name]["max_abs"], kerr[name]["frob_rel"]))
    # orbital energies of the final FP64 Fock matrix at the chosen SCF's
    # converged density (F = H + J + K with the reference's J and an FP64 K),
    # against the plain SCF's, occupied and virtual separately
    from scipy.linalg import eigh
    no = mol.nocc

    def final_eps(res):
        Co = np.ascontiguousarray(res.C[:, :no])
        F = T + V + dB.dfj(Co @ Co.T, met) + dB.dfk_mo_tilde64(Co)
        return eigh(F, S, eigvals_only=True)
    Cp = np.ascontiguousarray(ref.C[:, :no])
    eps_ref = eigh(T + V + dB.dfj(Cp @ Cp.T, met) + dB.dfk_mo_tilde64(Cp), S, eigvals_only=True)
    deps = {}
    for tag, res in (("scf_thresh", chosen), ("both_tight", chosen_tight)):
        e = final_eps(res)
        deps[tag] = {"occupied": float(np.abs(e[:no] - eps_ref[:no]).max()),
                     "virtual": float(np.abs(e[no:] - eps_ref[no:]).max())}
        log("orbital energies of the final FP64 Fock (chosen SCF %s) vs the plain SCF's: max |d eps| "
            "occupied %.3e virtual %.3e" % (tag, deps[tag]["occupied"], deps[tag]["virtual"]))

    out = {
        "metric": "DF-K s/iter, %d H2O cc-pVDZ, chosen form vs FP64" % cfg["nwater"],
        "value": step_s,
        "unit": "s/iter",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": step_s * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"auto": "bf16x3 split then f32", "bf16": "bf16x3 split",
                  "f32": "f32"}[form] + " (K builds, f32 accumulation) + f64 (final exchange energy, J, SCF)",
        "data": "synthetic seeded water cluster, real integrals",
        "config": {"workload": "DF-K build, %d-water cluster cc-pVDZ/cc-pVDZ-jkfit (%s)" % (
            cfg["nwater"], args.config), "nbf": mol.nbf, "naux": mol.naux, "nocc": mol.nocc,
            "nnz_B": nnz, "screen_eps": 1e-10, "rcm_reorder": not args.no_reorder, "scf_thresh": args.scf_thresh},
        "plain_fp64_s_per_iter": plain_s,
        "speedup_vs_plain": plain_s / step_s,
        "chosen_form": form,
        "builds_timed": {"bf16x3": nb[0], "f32": nb[1], "fp64": nb[2], "fp64_final_energy": 1},
        "builds_scf": {"bf16x3": scf_mix[0], "f32": scf_mix[1], "fp64": scf_mix[2], "fp64_final_energy": 1},
        "bf16x3_build_s": t16,
        "f32_build_s": t32,
        "final_fp64_energy_s": final_s,
        "final_energy_vs_full_trace_rel": abs(ek - ek_full) / abs(ek_full),
        "lowp_copies_s": t_copies,
        "dE_scf": dE,
        "dE_scf_note": "chosen SCF at %.0e vs plain FP64 SCF at %.0e" % (args.scf_thresh, args.ref_thresh),
        "dE_scf_both_tight": dE_tight,
        "dK_vs_fp64": kerr,
        "max_deps_final_fock": deps,
        "b_build_gpu_s": t_build,
        "scf_iters": {"plain": ref.iterations, "chosen": chosen.iterations, "chosen_tight": chosen_tight.iterations},
        "E_fp64": ref.energy,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if not args.no_serve_sample and args.config == "c2":
        # the c2 tensors hold most of HBM: release them, then a short served
        # stream (reported beside the DF-K value, not as it)
        for obj in (Cd, Kd, dB, met):
            obj.release()
        ctx.close()
        sv = serve_stream(args, log_progress=False, cfg=CONFIGS["c4_sample"], ncheck=1)
        sc = CONFIGS["c4_sample"]
        out["serving_sample"] = {
            "workload": "request stream: %d requests, Poisson %.0f/s, U{%d..%d} waters, conv %.0e, %d workers "
                        "(a short stream of the c4 kind; c4 itself: --config c4)" % (
                            sc["requests"], sc["rate"], sc["nmin"], sc["nmax"], sc["thresh"], sv["slots"]),
            "p50_s": sv["p50_s"], "p99_s": sv["p99_s"], "mean_s": sv["mean_s"],
            "requests_per_s": sv["requests_per_s"], "scf_iterations_per_s": sv["scf_iterations_per_s"],
            "all_converged": sv["all_converged"], "failed": sv["failed"],
            "max_abs_dE_vs_fp64": sv["max_abs_dE_vs_fp64"],
            "max_abs_dE_vs_fp64_both_1e8": sv["max_abs_dE_vs_fp64_both_1e8"],
            "mean_phase_s": sv["mean_phase_s"]}
        log("serving sample: p50 %.2f s p99 %.2f s, %.2f requests/s" % (sv["p50_s"], sv["p99_s"],
                                                                        sv["requests_per_s"]))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
