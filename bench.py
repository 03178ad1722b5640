"""Benchmark: DF-K seconds per SCF iteration on one MI355X.

Workload (default --config c2): seeded 64-water cluster, cc-pVDZ / cc-pVDZ-jkfit,
atomic blocks, Schwarz screening at 1e-10.

Plain path: the reference's DFMO algorithm in FP64 (K1 = B.C, K2 = L^-1 per
build, K3 = SYRK).  Chosen form: B~ = L^-1 B made once and stored as a BF16
hi/lo split (--form bf16, default) or in FP32 (--form f32); every SCF
iteration builds K on MFMA with FP32 accumulation and FP64 reduction, and one
FP64 K at the converged orbitals gives the reported energy.

A step is one SCF iteration's DF-K work in the chosen form.  The timed region
runs `steps` FP32 builds plus the one final FP64 build at the converged
orbitals, so value = (steps * t_fp32 + t_fp64_final) / steps; steps defaults
to the chosen SCF's iteration count.  The line also carries the plain FP64
seconds per build, |dE| between the two SCFs, the roofline of the dominant
kernel (K1, the B.C slab GEMM) from HIP events on the library stream, and the
oracle timed on the host.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c0|c1|c2]
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
    return dict(value=est, unit="s/iter", cores=cores, kind="port",
                sample="oracle DFMO-K on aux rows 0..%d of %d (%.2f s, x naux/rows) + L^-1 "
                       "triangular multiply on %d of %d HT1 columns (%.2f s, x cols/sample)" % (
                           rows, naux, dt1, s, ncol, dt2))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0, help="0: the chosen SCF's iteration count")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c2", choices=sorted(CONFIGS))
    ap.add_argument("--scf-thresh", type=float, default=1e-6)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--form", default="bf16", choices=["bf16", "f32"],
                    help="precision of the chosen-form exchange builds")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (the DF-K path runs on a single "
              "MI355X); --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    from chem_amd import desc, ints
    from chem_amd.engine import Context, DeviceB, Metric
    from chem_amd.scf import run_rhf

    cfg = CONFIGS[args.config]
    t0 = time.perf_counter()
    mol = desc.Molecule(desc.water_cluster(cfg["nwater"], cfg["seed"]))
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

    # plain FP64 path: DFMO with the metric applied per build (reference algorithm)
    t1 = time.perf_counter()
    ref = run_rhf(mol, dB, met, S, T + V, threshold=args.scf_thresh, kmode="dfmo")
    log("plain FP64 SCF: E=%.10f iters=%d (%.1f s) K s/iter median %.4f" % (
        ref.energy, ref.iterations, time.perf_counter() - t1, float(np.median(ref.t_k[1:] or ref.t_k))))
    C_occ = np.ascontiguousarray(ref.C[:, :mol.nocc])

    # time the plain K builds at the converged orbitals
    def time_builds(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        ctx.sync()
        t = time.perf_counter()
        for _ in range(steps):
            fn()
        ctx.sync()
        return (time.perf_counter() - t) / steps

    plain_s = time_builds(lambda: dB.dfk_mo(C_occ, met), 2, 1)

    K_plain = dB.dfk_mo(C_occ, met)
    cpu = None
    if not args.no_cpu_baseline:
        Linv, _ = met.get()
        cpu = cpu_baseline(mol, dB, Linv, C_occ)
        log("cpu baseline: %.2f s/iter (%s)" % (cpu["value"], cpu["sample"]))

    # chosen form: B~ = L^-1 B once (FP64), low-precision copy, mixed SCF
    dB.transform(met)
    form = args.form
    if form == "bf16":
        dB.make_bf16()
        kfun, kmode, bpeak = dB.dfk_mo_bf16, "mixed_bf16", MFMA_BF16_PEAK
    else:
        dB.make_f32()
        kfun, kmode, bpeak = dB.dfk_mo_f32, "mixed_f32", MFMA_F32_PEAK
    t2 = time.perf_counter()
    chosen = run_rhf(mol, dB, None, S, T + V, threshold=args.scf_thresh, kmode=kmode)
    dE = chosen.energy - ref.energy
    log("chosen SCF (%s): E=%.10f (low-precision-K iterate %.10f) iters=%d dE=%.3e (%.1f s)" % (
        form, chosen.energy, chosen.energy_iter, chosen.iterations, dE, time.perf_counter() - t2))
    steps = args.steps or max(1, chosen.iterations)
    Cc = np.ascontiguousarray(chosen.C[:, :mol.nocc])
    for _ in range(args.warmup):
        kfun(Cc)
        dB.dfk_mo(Cc)
    ctx.sync()
    t = time.perf_counter()
    for _ in range(steps):
        kfun(Cc)
    dB.dfk_mo(Cc)
    ctx.sync()
    step_s = (time.perf_counter() - t) / steps
    t64 = time_builds(lambda: dB.dfk_mo(Cc), 2, 0)
    tlow = time_builds(lambda: kfun(Cc), 3, 0)

    # roofline of the dominant kernel (K1 = B~ C slab GEMM) from HIP events on the library stream
    ctx.set_timing(True)
    nrep = 3
    for _ in range(nrep):
        kfun(Cc)
    ph = ctx.phase_times() / nrep
    ctx.set_timing(False)
    k1_flops = 2.0 * nnz * mol.nocc
    k3_flops = float(mol.naux) * mol.nocc * mol.nbf ** 2   # symmetric half of 2*naux*nocc*nbf^2
    achieved = k1_flops / ph[0]
    roof = dict(bound="mfma", achieved=achieved / 1e12, peak=bpeak / 1e12, unit="TFLOP/s",
                frac=achieved / bpeak, traffic=None,
                kernel="k_slab_gemm_bf16x3" if form == "bf16" else "k_slab_gemm<float>",
                kernel_s=ph[0], k3_s=ph[2], k3_achieved=k3_flops / ph[2] / 1e12,
                split_s=ph[5])
    if form == "bf16":
        # three BF16 products per useful multiply-add: issued MFMA work is 3x the algorithmic flops
        roof["mfma_issued_frac"] = 3.0 * achieved / bpeak
        roof["k3_mfma_issued_frac"] = 3.0 * k3_flops / ph[2] / bpeak
        roof["column_tile"] = dB.bf16_tile(mol.nocc)

    # accuracy of the chosen-form K at the FP64 orbitals, and of the orbital energies
    K_low = kfun(C_occ)
    dK = K_low - K_plain
    kerr = dict(max_abs=float(np.abs(dK).max()),
                frob_rel=float(np.linalg.norm(dK) / np.linalg.norm(K_plain)))
    deps = float(np.abs(chosen.eps - ref.eps).max())
    log("K error vs FP64: max %.3e, Frobenius rel %.3e; max |d eps| %.3e" % (
        kerr["max_abs"], kerr["frob_rel"], deps))

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
        "dtype": ("bf16x3 split, f32 accumulation" if form == "bf16" else "f32") +
                 " (K builds) + f64 (final K, J, SCF)",
        "data": "synthetic seeded water cluster, real integrals",
        "config": {"workload": "DF-K build, %d-water cluster cc-pVDZ/cc-pVDZ-jkfit (%s)" % (
            cfg["nwater"], args.config), "nbf": mol.nbf, "naux": mol.naux, "nocc": mol.nocc,
            "nnz_B": nnz, "screen_eps": 1e-10},
        "plain_fp64_s_per_iter": plain_s,
        "speedup_vs_plain": plain_s / step_s,
        "chosen_form": form,
        "chosen_build_s": tlow,
        "transformed_fp64_build_s": t64,
        "dE_scf": dE,
        "dK_vs_fp64": kerr,
        "max_deps_orbital": deps,
        "b_build_gpu_s": t_build,
        "scf_iters": {"plain": ref.iterations, "chosen": chosen.iterations},
        "E_fp64": ref.energy,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
