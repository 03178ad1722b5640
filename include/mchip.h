/*
 * mchip.h -- C ABI of the MI355X-native DF Fock build and SOS-AO-MP2 engine.
 *
 * Every entry point returns MCHIP_OK (0) on success and a nonzero status on
 * failure; mchip_last_error() then returns a thread-local message.  The C++
 * builder layer (include/mchip_fock.hpp) rethrows such failures as
 * std::runtime_error, which is how the reference reports every error
 * (src/main.cpp:115-147).
 *
 * Matrices are dense row-major double arrays unless stated otherwise.
 * Pointers named *_dev are device (HBM) pointers; all other pointers are host.
 * Device entry points are asynchronous on the context's stream unless they
 * return host data.
 */
#ifndef MCHIP_H
#define MCHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MCHIP_OK 0
#define MCHIP_ERR 1

const char* mchip_last_error(void);

/* ------------------------------------------------------------------------
 * Basis sets and AO integrals (host).
 * Replaces desc::cluster_basis + ints::aofactory (src/desc/basis.cpp:278-360,
 * src/ints/aofactory.cpp:273-310 normalisation, :380-430 libcint setup).
 * ---------------------------------------------------------------------- */
typedef struct mchip_basis_s* mchip_basis;

/* One shell per entry; exps/coefs are concatenated per shell (nprim[i]
 * primitives each, raw contraction coefficients as read from the basis file),
 * centers is nshell x 3 in bohr, atom gives the owning atom of each shell.
 * Coefficients are normalised as aofactory.cpp:273-310 does. */
int mchip_basis_create(int nshell, const int* l, const int* nprim,
                       const double* exps, const double* coefs,
                       const double* centers, const int* atom,
                       mchip_basis* out);
int mchip_basis_destroy(mchip_basis b);
int mchip_basis_nbf(mchip_basis b, int* nbf);
/* Normalised contraction coefficients, same layout as coefs on input. */
int mchip_basis_norm_coefs(mchip_basis b, double* coefs);

/* Boys function F_0..F_mmax(T) (test hook for the integral engine). */
int mchip_boys(int mmax, double T, double* F);

/* S, T, V (nbf x nbf); Z/xyz describe natom point charges (bohr).
 * aofactory::ao_overlap / ao_kinetic / ao_nuclear. */
int mchip_int1e(mchip_basis b, int natom, const double* Z, const double* xyz,
                double* S, double* T, double* V);
/* Coulomb metric (X|Y), naux x naux.  aofactory::ao_2c2e(coulomb). */
int mchip_int2c2e(mchip_basis aux, double* V);
/* Schwarz factors: Qao nshell_ao^2 (sqrt sum|(ab|ab)|), Qaux nshell_aux.
 * aofactory::ao_schwarz / ao_3cschwarz (integrals.cpp:511-689). */
int mchip_schwarz(mchip_basis ao, mchip_basis aux, double* Qao, double* Qaux);
/* The Schwarz block mask of schwarz_screener (screening.cpp:8-30) for the
 * given atomic blocking (shells per block, contiguous): Frobenius block norms
 * of the shell factors, norm_mn nbb x nbb and norm_x nxb (either may be NULL),
 * and mask[X][M][N] = (float)(norm_x[X] * norm_mn[M][N]) > eps, nxb*nbb*nbb
 * bytes.  The one implementation the host and GPU tensor builds use. */
int mchip_schwarz_mask(mchip_basis ao, mchip_basis aux, int nbb, const int* ao_blk_shells, int nxb,
                       const int* aux_blk_shells, double eps, uint8_t* mask, double* norm_mn, double* norm_x);
/* One shell triple (c|ab), out[m_c][m_a][m_b] (test hook). */
int mchip_int3c2e_shell(mchip_basis ao, mchip_basis aux, int ia, int ib, int ic,
                        double* out);

/* Four-centre (mu nu|lambda sigma) of a small basis, row-major nbf^4 (the
 * reference's int2e_sph through libcint, used by its exact J/K). */
int mchip_int4c2e(mchip_basis b, double* out);

/* SAD guess (hfmod::compute_guess, src/hf/hfguess.cpp:142-452): per element,
 * a restricted SCF of the atom alone in its shells of the molecule's basis
 * with fractional, spin-averaged occupations (src/desc/molecule.cpp:37-93),
 * exact four-centre J/K, core guess, no DIIS, RMS(FPS - SPF) < 1e-6; the
 * atomic alpha densities on the diagonal blocks.  The shells of each atom
 * must be consecutive, in atom order.  P_alpha: nbf x nbf row-major. */
int mchip_sad_guess(mchip_basis ao, int natom, const double* Z, const double* xyz, double* P_alpha);
/* The atomic SCF alone (test hook): atom_basis holds one atom's shells, Z its
 * charge at xyz; P (nbf^2) its alpha density, energy the electronic energy
 * tr(HP) + tr(FP), iterations the SCF count, nocc / occ (capacity nbf) the
 * occupied orbitals and their coefficient factors (sqrt of the occupation). */
int mchip_sad_atom(mchip_basis atom_basis, double Z, const double* xyz, double threshold, double* P,
                   double* energy, int* iterations, int* nocc, double* occ);

/* ------------------------------------------------------------------------
 * Three-centre tensor B(X|mu nu), block-sparse (host container).
 * Replaces dbcsr::btensor<3> core mode (src/dbcsrx/dbcsr_btensor.hpp) and the
 * screened fill of aoloader::compute (src/ints/aoloader.cpp:341-377).
 * ---------------------------------------------------------------------- */
typedef struct mchip_btensor_s* mchip_btensor;

/* Blocks are contiguous shell ranges: ao_blk_shells[i] shells in AO block i.
 * schwarz_mask is nxb*nbb*nbb bytes (1 = compute the block triple).  After the
 * fill, blocks with Frobenius norm < filter_eps are dropped. */
int mchip_btensor_build(mchip_basis ao, mchip_basis aux,
                        int nbb, const int* ao_blk_shells,
                        int nxb, const int* aux_blk_shells,
                        const uint8_t* schwarz_mask, double filter_eps,
                        mchip_btensor* out);
/* Build from a dense naux x nbf x nbf array (tests, synthetic inputs). */
int mchip_btensor_from_dense(int nbb, const int* bsz, int nxb, const int* xsz,
                             const uint8_t* mask, const double* dense,
                             mchip_btensor* out);
int mchip_btensor_destroy(mchip_btensor t);
/* dims[0..3] = naux, nbf, nxb, nbb; *nelem = stored elements. */
int mchip_btensor_info(mchip_btensor t, int* dims, int64_t* nelem);
/* Final block mask after filtering, nxb*nbb*nbb bytes. */
int mchip_btensor_mask(mchip_btensor t, uint8_t* mask);
/* Dense copy naux x nbf x nbf (zeros outside the mask). */
int mchip_btensor_to_dense(mchip_btensor t, double* dense);
/* Dense copy of aux rows x0 <= X < x1 ((x1-x0) x nbf x nbf). */
int mchip_btensor_dense_rows(mchip_btensor t, int x0, int x1, double* dense);

/* ------------------------------------------------------------------------
 * Laplace quadrature for SOS-AO-MP2: minimax exponential sum
 * 1/x ~ sum_k w_k exp(-a_k x) on [1, R] by Remez iteration, started from a
 * tabulated solution (w0, a0) at range R0 and continued geometrically to R.
 * Replaces math::minimax / c_laplace_minimax (src/math/laplace/minimax.hpp:
 * 67-88).  err = max |1/x - s(x)|; equi = relative spread of |e| over the
 * alternation points (0 when fully levelled).
 * ---------------------------------------------------------------------- */
int mchip_laplace_minimax(int n, double R, double R0, const double* w0, const double* a0,
                          int maxit, double* w, double* a, double* err, double* equi);

/* ------------------------------------------------------------------------
 * Device context (one HIP stream + rocBLAS handle on one GPU).
 * Replaces the reference's world/cart (src/megalochem.hpp:13-80): one
 * process drives one GPU.
 * ---------------------------------------------------------------------- */
typedef struct mchip_ctx_s* mchip_ctx;
typedef struct mchip_dbt_s* mchip_dbt;        /* device-resident B */
typedef struct mchip_metric_s* mchip_metric;  /* Coulomb metric factors */

int mchip_device_count(int* n);
int mchip_ctx_create(int device, mchip_ctx* out);
int mchip_ctx_destroy(mchip_ctx c);
int mchip_ctx_sync(mchip_ctx c);
/* The context's hipStream_t (for callers that enqueue their own work). */
int mchip_ctx_stream(mchip_ctx c, void** stream);

/* Per-phase kernel timing with HIP events on the context's stream (adds a
 * synchronisation per phase; for measurement only).  Enabling resets the
 * accumulators.  Slots: 0 K1 (B.C), 1 K2 (metric), 2 K3 (SYRK),
 * 3 J total, 4 incremental K, 5 operand splitting. */
int mchip_ctx_set_timing(mchip_ctx c, int on);

/* Graph replay of repeated builds (default on): the launch chain of an
 * exchange build (gather -> slab GEMM -> SYRK -> chunk reduction, over every
 * occupied batch) of the BF16x3, FP32 and FP64-copy forms, and of a DF-J
 * build, is captured into a hipGraph on its second call with the same
 * operand pointers, widths and tuned shapes, and replayed from the third on
 * (one graph launch per build).  A replay gives the same bits as the direct
 * launches.  The reference builds through DBCSR calls (jkdf.cpp:334-422) and
 * has no counterpart; this is the context's switch (0 = direct launches),
 * e.g. for timing both. */
int mchip_ctx_set_graphs(mchip_ctx c, int on);
/* Captured graphs held by tensor t and how many builds replayed one. */
int mchip_dbt_graph_stats(mchip_dbt t, int64_t* ngraphs, int64_t* replays);
/* When the escalation K mode (4) switches from the BF16x3 to the FP32
 * exchange: once the previous iteration's error RMS is <= rms (0 turns this
 * trigger off; default 1e-5), or after `stall` consecutive iterations below an
 * error RMS of 1e-4 whose RMS stayed above 0.8 x the lowest so far (0 turns it
 * off; default 3).  FP32 gives way to FP64 after such a stall below 1e-6.  No
 * counterpart in the reference. */
int mchip_ctx_set_escalation(mchip_ctx c, double rms, int stall);
/* First density of the next mchip_scf_run on this context: the alpha density
 * of a SAD guess (mchip_sad_guess; nbf x nbf row-major, copied), used by that
 * one run, whose first exchange then contracts every column of its pivoted
 * Cholesky factor (the SAD iteration, hfscf.cpp:389-396, jkdf.cpp:243-246).
 * NULL clears it (core guess, hfguess.cpp:127-141). */
int mchip_ctx_set_guess(mchip_ctx c, int nbf, const double* P_alpha);
/* Release the context's work arrays (they grow to the largest exchange run on
 * the context and are kept for reuse; the plain FP64 exchange holds two of
 * naux * nbf * nocc doubles).  The next build allocates what it needs again.
 * No counterpart in the reference. */
int mchip_ctx_trim(mchip_ctx c);
/* Device arrays on the context's GPU for callers that keep the occupied
 * coefficients and K resident between builds (the _dev entry points):
 * allocate, free (after the context's pending work), and synchronous copies
 * in and out on the context's stream.  No counterpart in the reference (its
 * matrices live in DBCSR on the host). */
int mchip_dev_alloc(mchip_ctx c, size_t bytes, void** p);
int mchip_dev_free(mchip_ctx c, void* p);
int mchip_dev_copy_in(mchip_ctx c, void* dst, const void* src, size_t bytes);
int mchip_dev_copy_out(mchip_ctx c, void* dst, const void* src, size_t bytes);
int mchip_ctx_phase_times(mchip_ctx c, double* out, int n);

/* Copy a host B to HBM (FP64).  dbcsr::btensor decompress_init/decompress
 * (dbcsr_btensor.hpp:1133-1348) become no-ops: B stays resident. */
int mchip_dbt_upload(mchip_ctx c, mchip_btensor host, mchip_dbt* out);
int mchip_dbt_destroy(mchip_dbt t);
/* Compute B on the GPU (same screening, filter and layout as
 * mchip_btensor_build; AO shells up to d, uncontracted aux basis). */
int mchip_dbt_build_gpu(mchip_ctx c, mchip_basis ao, mchip_basis aux, int nbb,
                        const int* ao_blk_shells, int nxb, const int* aux_blk_shells,
                        const uint8_t* schwarz_mask, double filter_eps, mchip_dbt* out);
/* The same S, T, V and (X|Y) computed on the context's GPU (AO shells up to
 * d, uncontracted auxiliary shells up to f), into host arrays; they replace
 * the same aofactory calls (aofactory.cpp:380-430) on the device path. */
int mchip_int1e_gpu(mchip_ctx c, mchip_basis b, int natom, const double* Z, const double* xyz,
                    double* S, double* T, double* V);
int mchip_int2c2e_gpu(mchip_ctx c, mchip_basis aux, double* V);
/* AO Schwarz factors Q[a][b] = sqrt(sum_s |(ab|ab)_s|), nshell x nshell, on
 * the context's GPU (the same factors as mchip_schwarz's Qao; the reference's
 * Schwarz screener, integrals.cpp:511-689 and screening.cpp:8-30). */
int mchip_schwarz_ao_gpu(mchip_ctx c, mchip_basis ao, double* Qao);
/* dims[0..3] = naux, nbf, nxb, nbb; *nelem = stored elements. */
int mchip_dbt_info(mchip_dbt t, int* dims, int64_t* nelem);
int mchip_dbt_mask(mchip_dbt t, uint8_t* mask);
/* Dense host copy of aux rows x0 <= X < x1 of the device B. */
int mchip_dbt_dense_rows(mchip_dbt t, int x0, int x1, double* dense);
/* Copy the device slabs back (same layout as the host container). */
int mchip_dbt_download(mchip_dbt t, double* host);

/* Coulomb metric from host V (naux x naux): V = L L^T, L^-1, V^-1 on the GPU.
 * Replaces aoloader::invert / math::LLT (aoloader.cpp:14-32, LLT.cpp:8-109). */
int mchip_metric_create(mchip_ctx c, int naux, const double* V, mchip_metric* out);
int mchip_metric_destroy(mchip_metric m);
/* Host copies: Linv column-major lower triangular, Vinv symmetric. */
int mchip_metric_get(mchip_ctx c, mchip_metric m, double* Linv, double* Vinv);
/* One-time B <- L^-1 B.  Afterwards J/K calls take m = NULL. */
int mchip_dbt_transform(mchip_dbt t, mchip_metric m);

/* DF-J: J = B^T V^-1 (B . (pscale*P))  -- DF_J::compute_J (jkdf.cpp:69-190);
 * pscale = 2 for a closed-shell alpha density (jkdf.cpp:83-90). */
int mchip_dfj_dev(mchip_dbt t, mchip_metric m, const double* P_dev, double pscale, double* J_dev);
int mchip_dfj(mchip_dbt t, mchip_metric m, const double* P, double pscale, double* J);
/* mchip_dfj_dev without any device allocation: no counterpart in the
 * reference; for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Its workspace is a
 * fixed 128 MiB device array per GPU, held until the process ends; a tensor
 * whose DF-J workspace is larger is rejected with an error (about 83 MB at 64
 * waters, cc-pVDZ-jkfit).  Same kernels, same J bit for bit as mchip_dfj_dev;
 * every call on one GPU shares the array, so calls that overlap on one GPU
 * (two host threads, or two streams) race. */
int mchip_dfj_dev_noalloc(mchip_dbt t, mchip_metric m, const double* P_dev, double pscale, double* J_dev);

/* DFMO-K: K = -(L^-1 B C)^T (L^-1 B C) for C = nbf x nocc (row-major, leading
 * dimension ldc), returned negated as the reference does
 * (DFMO_K::compute_K, jkdf.cpp:221-462, scale at :436-438).
 * The _dev exchange entry points take device pointers (C: nbf x ldc, K:
 * nbf x nbf, the caller's to size) and reject nocc outside 0..ldc. */
int mchip_dfk_mo_dev(mchip_dbt t, mchip_metric m, const double* C_dev, int ldc, int nocc,
                     int occ_nbatches, double* K_dev);
int mchip_dfk_mo(mchip_dbt t, mchip_metric m, const double* C, int nocc, int occ_nbatches,
                 double* K);

/* Chosen form, FP32: make the FP32 copy of the transformed B once, then
 * K = -(B~ C)^T (B~ C) with FP32 MFMA and FP64 split-K reduction. */
int mchip_dbt_make_f32(mchip_dbt t);
int mchip_dfk_mo_f32_dev(mchip_dbt t, const double* C_dev, int ldc, int nocc, int occ_nbatches,
                         double* K_dev);
int mchip_dfk_mo_f32(mchip_dbt t, const double* C, int nocc, int occ_nbatches, double* K);

/* FP32 and/or BF16 split copies of B~ = L^-1 B made slab by slab from the
 * untransformed B, which stays FP64 in place: J and the reported energy keep
 * the reference's formulas (DF_J with explicit V^-1, jkdf.cpp:69-190), and the
 * per-iteration exchange of the chosen form (mchip_dfk_mo_f32 / _bf16, SCF
 * kmodes 2-4) reads the copies.  f64 also keeps the FP64 slabs of L^-1 B: the
 * FP64 exchanges of kmodes 2-4 (the final one, and the last escalation stage)
 * then contract (L^-1 B) C instead of applying L^-1 to B C per build
 * (jkdf.cpp:361-417 with the L^-1 product associated the other way). */
int mchip_dbt_make_lowp(mchip_dbt t, mchip_metric m, int f32, int bf16, int f64);
/* K from the FP64 L^-1 B copy (host C nbf x nocc row-major, K nbf x nbf). */
int mchip_dfk_mo_tilde64(mchip_dbt t, const double* C, int nocc, int occ_nbatches, double* K);

/* BF16 split-precision copy of the transformed B (hi = bf16(x), lo = bf16(x - hi),
 * rows padded to a multiple of 32), made once per geometry.  Chosen form of the
 * B-contractions of DFMO_K::compute_K (src/fock/jkdf.cpp:221-462). */
int mchip_dbt_make_bf16(mchip_dbt t);
/* K = -sum_X (B~_X C)^T (B~_X C) with both contractions as three BF16 MFMA
 * products (lo*hi + hi*lo + hi*hi), FP32 accumulation within chunks of at most
 * 256 aux functions, FP64 across chunks.  Replaces the per-iteration
 * DFMO_K::compute_K (jkdf.cpp:221-462) in the chosen form. */
/* The FP64 exchange on the FP64 copy of L^-1 B with device-resident C
 * (row-major nbf x ldc, first nocc columns) and K (nbf x nbf); no
 * synchronisation (the context's stream). */
int mchip_dfk_mo_tilde64_dev(mchip_dbt t, const double* C_dev, int ldc, int nocc, int occ_nbatches,
                             double* K_dev);
/* tr(K P) of the FP64 exchange on the FP64 copy of L^-1 B at the device
 * coefficients C_dev (row-major nbf x ldc, first nocc columns; P = C C^T),
 * computed as -sum_X ||(L^-1 B C)_X^T C||^2 without forming K: the exchange
 * part of E = tr(H P) + tr(F P) (hfscf.cpp:278-296), which is all the
 * reported energy needs of the final exchange.  Synchronises; *ek on the host. */
int mchip_dfk_energy_tilde64_dev(mchip_dbt t, const double* C_dev, int ldc, int nocc, double* ek);
/* The same tr(K P) from the untransformed B and the metric (the plain FP64
 * exchange's formula, jkdf.cpp:221-462), without K and without the nbf-sized
 * metric step: -sum_X ||(L^-1 M)_X||^2 with M_X = C^T B_X C (nocc x nocc),
 * in occ_nbatches batches of the first occupied index.  Synchronises. */
int mchip_dfk_energy_dev(mchip_dbt t, mchip_metric m, const double* C_dev, int ldc, int nocc, int occ_nbatches,
                         double* ek);
int mchip_dfk_mo_bf16_dev(mchip_dbt t, const double* C_dev, int ldc, int nocc, int occ_nbatches,
                          double* K_dev);
/* The BF16x3 exchange of n tensors on one context in one slab-GEMM launch
 * and one SYRK launch over all their tiles (one occupied batch each; tensor j
 * with device C_dev[j], row-major nbf_j x ldc[j], first nocc[j] > 0 columns,
 * into K_dev[j]).  Each K_dev[j] equals what mchip_dfk_mo_bf16_dev gives for
 * that tensor alone, bit for bit.  The reference builds one molecule per
 * process; this is the contraction the request loop batches across live
 * requests.  No synchronisation (the context's stream). */
int mchip_dfk_mo_bf16_group_dev(int n, const mchip_dbt* ts, const double* const* C_dev, const int* ldc,
                                const int* nocc, double* const* K_dev);
int mchip_dfk_mo_bf16(mchip_dbt t, const double* C, int nocc, int occ_nbatches, double* K);
/* Exchange from a density matrix P (nbf x nbf, symmetric PSD): DFAO_K::compute_K
 * (src/fock/jkdf.cpp:510-648) and DFMEM_K::compute_K (jkdf.cpp:719-859),
 * K = -sum B(X|m l) V^-1(X,Y) P(l,s) B(Y|s n), returned negated like the
 * reference.  P is factorised on the GPU (pivoted Cholesky, P = L L^T) and the
 * occupied-factor exchange runs with C := L.  form 0 = FP64 (metric needed unless
 * B is transformed), 1 = FP32, 2 = BF16x3 (transformed B with its copy). */
int mchip_dfk_ao_dev(mchip_dbt t, mchip_metric m, const double* P_dev, int form, int occ_nbatches,
                     double* K_dev);
int mchip_dfk_ao(mchip_dbt t, mchip_metric m, const double* P, int form, int occ_nbatches, double* K);
/* The two-sided exchange of DFLMO_K (src/fock/jkdf.cpp:1093-1331, taken when
 * the builder has set_sym(false), as the ADC drivers use it):
 * K(m,n) = -sum_{X,Y,i} (B u)(X|m i) V^-1(X,Y) (B v^T)(Y|n i), u nbf x nocc
 * (the reference's u_bm, set_coeff_left_*) and v nocc x nbf (v_mb,
 * set_coeff_right_*), both row-major; K nbf x nbf row-major, returned negated
 * like the reference and not symmetric.  fp64_copy = 1: on the FP64 copy of
 * L^-1 B (mchip_dbt_make_lowp with f64); 0: L^-1 applied to each
 * half-transformed tensor (m needed unless B is transformed).  FP64. */
int mchip_dfk_lmo_uv(mchip_dbt t, mchip_metric m, const double* U, const double* V, int nocc, int occ_nbatches,
                     int fp64_copy, double* K);
/* Launch shape picked by the BF16 slab-GEMM autotuner for ng occupied columns,
 * as column_tile * 10000 + threads_per_workgroup (0 before the first build with
 * that ng).  No counterpart in the reference. */
int mchip_dfk_bf16_tile(mchip_dbt t, int ng, int* bn);
/* The full autotuned BF16 pipeline for ng occupied columns: plan[8] = {GEMM
 * shape index, row tile, column tile, threads, direct-to-LDS (0 register-
 * staged, 1 one pipeline for both operands, 2 the B~ rows in a pipeline of
 * their own), LDS buffers or pipeline stages, SYRK variant (0 register-staged,
 * 1-3 direct-to-LDS 128-row tiles with 2-4 stages, 4 direct-to-LDS 256-row
 * tiles), 1 if tuned}; ms[] (capacity 32) = the timings of every
 * GEMM shape then every SYRK variant (-1 = not tried), *nms entries.  No
 * counterpart in the reference. */
int mchip_dfk_bf16_plan(mchip_dbt t, int ng, int* plan, float* ms, int* nms);
/* The autotuned FP32 exchange pipeline (direct-to-LDS FP32 MFMA slab GEMM and
 * SYRK on the padded FP32 copy) for ng occupied columns, chosen on the first
 * FP32 build with that ng: plan[8] = {GEMM shape index, row tile, column tile,
 * threads, 1, pipeline stages, SYRK variant (0-2: 2-4 stages), 1 if tuned};
 * ms[] (capacity 32) = the timings of every GEMM shape then every SYRK variant
 * (-1 = not tried), *nms entries.  No counterpart in the reference. */
int mchip_dfk_f32_plan(mchip_dbt t, int ng, int* plan, float* ms, int* nms);
/* Fix the FP32 pipeline for ng columns instead of tuning it; shapes: the
 * number of GEMM shapes and SYRK variants.  No counterpart in the reference. */
int mchip_dfk_f32_force(mchip_dbt t, int ng, int gemm_shape, int syrk_variant);
int mchip_dfk_f32_shapes(int* ngemm, int* nsyrk);
/* The autotuned FP64 exchange pieces for ng occupied columns: the slab GEMM
 * on the padded FP64 copy of L^-1 B (mchip_dfk_mo_tilde64) and the FP64 SYRK
 * every FP64 exchange ends with; same fields as mchip_dfk_f32_plan, with
 * plan[0] (GEMM shape) and plan[6] (SYRK variant, 0-2: 2-4 stages) -1 until
 * tuned.  force fixes both; shapes gives their counts.  No counterpart in the
 * reference. */
int mchip_dfk_f64_plan(mchip_dbt t, int ng, int* plan, float* ms, int* nms);
int mchip_dfk_f64_force(mchip_dbt t, int ng, int gemm_shape, int syrk_variant);
int mchip_dfk_f64_shapes(int* ngemm, int* nsyrk);
/* Fix the BF16 pipeline for ng columns instead of tuning it (direct-to-LDS GEMM
 * shapes pair with SYRK variants 1-4, the others with 0); shapes: count of GEMM
 * shapes and SYRK variants and, per GEMM shape, whether it is direct-to-LDS
 * (direct has room for ngemm entries).  No counterpart in the reference. */
int mchip_dfk_bf16_force(mchip_dbt t, int ng, int gemm_shape, int syrk_variant);
int mchip_dfk_bf16_shapes(int* ngemm, int* nsyrk, int* direct);

/* ------------------------------------------------------------------------
 * Restricted Hartree-Fock on the GPU (hf::hfmod::compute, src/hf/hfscf.cpp:
 * 326-552): core guess, F = H + J + K, e = FPS - SPF, E = tr(HP) + tr(FP),
 * DIIS (diis.hpp), diag via X^T F X.  S and H = T + V are host row-major.
 * kmode: 0 = DFMO with the metric applied per iteration (reference algorithm),
 *        1 = DFMO on the transformed B (mchip_dbt_transform),
 *        2 = mixed: FP32 exchange in every iteration, one FP64 exchange at
 *            the converged orbitals for the reported energy,
 *        3 = mixed with the BF16x3 exchange in every iteration (needs
 *            mchip_dbt_make_bf16), FP64 exchange for the reported energy,
 *        4 = precision escalation: BF16x3 until the trigger set by
 *            mchip_ctx_set_escalation fires (by default: error RMS <= 1e-5,
 *            or three iterations without progress), then FP32 until it stalls in
 *            the same way, then FP64; FP64 for the reported energy (needs the
 *            BF16 copy; without the FP32 copy BF16x3 gives way to FP64).
 * Outputs C (nbf x nbf row-major), orbital energies and per-iteration
 * history (energy, RMS error, J and K seconds) up to max_hist entries.
 * ---------------------------------------------------------------------- */
int mchip_scf_run(mchip_dbt t, mchip_metric m, int nbf, int nocc, int nbb, const int* bsz,
                  const double* S, const double* H, double enuc, double threshold,
                  int max_iter, int do_diis, int diis_start, int diis_min, int diis_max,
                  double filter_eps, int kmode, int occ_nbatches, double* energy,
                  int* iterations, int* converged, double* C, double* eps, int max_hist,
                  double* hist_e, double* hist_rms, double* t_j, double* t_k, int* nhist);
/* Of the last mchip_scf_run on this thread: the energy with the iteration's
 * own K and the seconds of the final FP64 K build (mixed mode). */
int mchip_scf_last_extra(double* energy_iter, double* t_final_k);

/* ------------------------------------------------------------------------
 * Request loop (DF-HF serving).  No counterpart in the reference, which runs
 * one molecule per process; each request runs hf::hfmod::compute with the DF
 * builders (src/hf/hfscf.cpp:326-552).
 *
 * nslots worker threads (one stream and rocBLAS handle each) take requests
 * in arrival order, prepare the host side (one-electron integrals, Coulomb
 * metric, Schwarz block mask, HBM estimate), wait until the estimate fits
 * budget_bytes next to the resident requests (admission in arrival order),
 * then build B on the GPU and run the SCF with everything resident.  Work of
 * different requests overlaps on the device.  The one-electron integrals and
 * the metric are computed on the GPU, the Schwarz factors on the host
 * (cpu_threads split over the workers).
 * kmode as for mchip_scf_run (0-4); screen_eps is the block threshold.
 * Requests get no FP32 copy, and kmode 4 here escalates on a stall only (no
 * RMS trigger): BF16x3 until the SCF converges or stalls, then FP64.
 * result: times[10] = {submit, admit, done (server clock, s), HBM estimate
 * (bytes), worker slot, energy of the last low-precision iteration, seconds of
 * host preparation, B build, metric and copies, SCF}; wait = 1 blocks.
 * A failed request reports its error through the return status.
 * stats = {bytes in use, peak bytes, completed, failed, max live requests,
 * server clock}.
 * ---------------------------------------------------------------------- */
typedef struct mchip_server_s* mchip_server;
int mchip_server_create(int device, double budget_bytes, int nslots, int cpu_threads, mchip_server* out);
int mchip_server_destroy(mchip_server s);   /* finishes every submitted request */
int mchip_server_submit(mchip_server s, mchip_basis ao, mchip_basis aux, int natom, const double* Z,
                        const double* xyz, int nbb, const int* ao_blk_shells, int nxb,
                        const int* aux_blk_shells, int nocc, double enuc, double threshold, int kmode,
                        double screen_eps, int* id);
int mchip_server_result(mchip_server s, int id, int wait, int* done, double* energy, int* iterations,
                        int* converged, double* times);
int mchip_server_stats(mchip_server s, double* stats);
/* Cross-request batching of the BF16x3 exchange builds (on by default; for
 * requests that start after the call): the builds that live requests hand in
 * while a batch runs form the next batch, launched as one grouped slab GEMM
 * and one grouped SYRK (mchip_dfk_mo_bf16_group_dev).  A request's result is
 * the same bits alone and in any batch.  batch_stats[3] = {batches, builds,
 * largest batch}. */
int mchip_server_set_batching(mchip_server s, int enabled);
int mchip_server_batch_stats(mchip_server s, double* batch_stats);

/* ------------------------------------------------------------------------
 * SOS-AO-MP2 on the GPU: mp::mpmod::compute with LLMP_FULL_Z
 * (src/mp/mp_compute.cpp:27-407, src/mp/z_full.cpp:69-397).  t must hold the
 * untransformed B(X|mu nu) (RI auxiliary basis), m its Coulomb metric.
 * S, C (nbf x nbf row-major, occupied columns first) and eps are host data;
 * omega/alpha are the nlap Laplace weights/exponents (mchip_laplace_minimax
 * scaled to [ymin, ymax]).  energy = sum_k e_k (positive, unscaled, as
 * printed at mp_compute.cpp:384-389); terms[k] = e_k; t_z = seconds in the
 * Z builder (M2-M5).
 * ---------------------------------------------------------------------- */
int mchip_mp2_sos(mchip_dbt t, mchip_metric m, int nbf, int nocc, const double* S, const double* C,
                  const double* eps, int nlap, const double* omega, const double* alpha, int nbb,
                  const int* bsz, int mo_split, double filter_eps, double* energy, double* terms,
                  double* t_z);
/* The same energy with the contractions of all Laplace points (b = B L_o, P_v b,
 * b3 on B's pattern, Z = sum b3 B) as grouped launches captured in one hipGraph,
 * FP32 with FP64 accumulation; pseudo-densities, pivoted Cholesky and the
 * energies stay FP64.  t_z = seconds of the graph launch, t_prep = seconds of
 * the FP64 per-point set-up and autotuning.  Chosen form of LLMP_FULL_Z over
 * the Laplace loop of mp_compute.cpp:232-389. */
int mchip_mp2_sos_batched(mchip_dbt t, mchip_metric m, int nbf, int nocc, const double* S, const double* C,
                          const double* eps, int nlap, const double* omega, const double* alpha, int nbb,
                          const int* bsz, int mo_split, double filter_eps, double* energy, double* terms,
                          double* t_z, double* t_prep);
/* mchip_mp2_sos with a choice of Z builder (mp::zmethod, z_builder.hpp):
 * 0 LLMP_FULL, 1 LLMP_MEM (the same contractions; one GPU path), 2 LL_FULL
 * (z_full.cpp:428-692: L_v = L_S CD(P~_v) from the orthogonalised virtual
 * pseudo-density, mp_compute.cpp:309-336; b(X,i,a) = L_v^T b(X,:,i);
 * Z = sum_ia b(X,i,a) b(Y,i,a)).  FP64. */
int mchip_mp2_sos_z(mchip_dbt t, mchip_metric m, int zmethod, int nbf, int nocc, const double* S, const double* C,
                    const double* eps, int nlap, const double* omega, const double* alpha, int nbb, const int* bsz,
                    int mo_split, double filter_eps, double* energy, double* terms, double* t_z);
/* One Z from given factors (mp::Z: set_occ_coeff / set_vir_density /
 * set_vir_coeff, compute, zmat; z_builder.hpp:43-116).  Locc: nbf x rank_o
 * row-major; Pvir: nbf x nbf (zmethod 0/1, ignored for 2); Lvir: nbf x rank_v
 * row-major (zmethod 2, ignored otherwise).  Z: naux x naux row-major, full
 * symmetric, FP64.  t holds the untransformed B. */
int mchip_z_compute(mchip_dbt t, int zmethod, int nbf, const double* Locc, int rank_o, const double* Pvir,
                    const double* Lvir, int rank_v, double* Z);
/* The Z builder alone: Z_k(X,Y) = sum_{mu nu} b3_k(X|mu nu) B(Y|mu nu) of every
 * Laplace point k into Z (nlap x naux x naux doubles, row-major; zero for a
 * point whose occupied pseudo-density has rank 0), with terms[k] = e_k as
 * above.  batched = 0: the FP64 per-point path of mchip_mp2_sos; 1: the
 * grouped FP32 graph of mchip_mp2_sos_batched.  The tensor LLMP_FULL_Z
 * produces per point (z_full.cpp:373-377, consumed at mp_compute.cpp:361-389). */
int mchip_mp2_z(mchip_dbt t, mchip_metric m, int nbf, int nocc, const double* S, const double* C,
                const double* eps, int nlap, const double* omega, const double* alpha, int nbb, const int* bsz,
                int mo_split, double filter_eps, int batched, double* Z, double* terms);

#ifdef __cplusplus
}
#endif

#endif /* MCHIP_H */
