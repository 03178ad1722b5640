// Device drivers of the DF Fock build.
#pragma once

#include "dev.hpp"

namespace mchip {

void ctx_init(Ctx& c, int device);
void ctx_free(Ctx& c);
void dbt_upload(Ctx& c, const BLayout& L, const double* host_data, DevBT& t);
void metric_create(Ctx& c, int naux, const double* V_host, Metric& m);
void transform_B(Ctx& c, DevBT& t, const Metric& m);
void linv_apply(Ctx& c, const Metric& m, const double* src, int64_t R, double* dst);
// FP32 slab GEMM on an FP32 copy of the slabs (data), column tile bn (64 or
// 128; 0 picks by ng).  f64_sums: FP64 running sums of the FP32 MFMA
// accumulators, flushed every 16 products.
void slab_gemm_f32(Ctx& c, DevBT& t, const float* data, const float* G, int ldg, int ng, float* out, int ldo,
                   int bn, bool f64_sums = false);
void slab_gemm(Ctx& c, DevBT& t, const double* G, int ldg, int ng, double* out, int ldo,
               double alpha);
// J from density P (device, nbf x nbf); the total density is pscale * P.
void dfj(Ctx& c, DevBT& t, const Metric* m, const double* P, double* J, double pscale);
// The same with the caller's workspace ws (dfj_workspace_bytes(t) bytes,
// 8-byte aligned; every element is written before it is read).
size_t dfj_workspace_bytes(const DevBT& t);
void dfj_ws(Ctx& c, DevBT& t, const Metric* m, const double* P, double* J, double pscale, double* ws);
// K = -sum ... from occupied coefficients C (device, row-major nbf x ldc).
// tilde64: FP64 slabs of L^-1 B (t.data64t) to contract instead of applying
// L^-1 to B C per build (the same K; K1 and K3 only).
void dfk_mo(Ctx& c, DevBT& t, const Metric* m, const double* C, int ldc, int nocc, double* K,
            int occ_nbatches, const double* tilde64 = nullptr);

// FP32 copy of the transformed B and the FP32-MFMA exchange (FP64 output).
void dbt_make_f32(Ctx& c, DevBT& t);
void dfk_mo_f32(Ctx& c, DevBT& t, const double* C, int ldc, int nocc, double* K, int occ_nbatches);
void dbt_make_bf16(Ctx& c, DevBT& t);
// f64: also keep the FP64 L^-1 B slabs (t.data64t) for the FP64 exchange.
void dbt_make_lowp(Ctx& c, DevBT& t, const Metric& m, bool f32, bool bf16, bool f64 = false);
// FP32 copy of L^-1 B (made slab by slab in FP64 from the untransformed B)
// into out (L.total floats, slab layout); t is left unchanged.
void tilde_to_f32(Ctx& c, DevBT& t, const Metric& m, float* out);
void dbt_make_f32_from(Ctx& c, DevBT& t, const float* b32);
int64_t dbt_padded_cols(const DevBT& t);
void slab_gemm_f32_f64sum(Ctx& c, DevBT& t, const float* C, int ldc, int ng, float* G, float* out, int ldo,
                          int depth);
void dfk_mo_bf16(Ctx& c, DevBT& t, const double* C, int ldc, int nocc, double* K, int occ_nbatches);
// tr(K P) of the FP64 exchange on the FP64 copy of L^-1 B at the occupied C
// (device, row-major nbf x ldc, first nocc columns; P = C C^T), computed as
// -sum_X ||W_X^T C||^2 without forming K (host result, synchronises).
double dfk_energy_tilde64(Ctx& c, DevBT& t, const double* C, int ldc, int nocc);
// tr(K P) of the FP64 exchange on the untransformed B with the metric, without
// K: -sum_X ||(L^-1 M)_X||^2 with M_X = C^T B_X C, in occupied batches of the
// first index.  Host result; synchronises the stream.
double dfk_energy_metric(Ctx& c, DevBT& t, const Metric& m, const double* C, int ldc, int nocc, int occ_nbatches);
// Grouped BF16x3 exchange over several tensors (one occupied batch each): the
// coefficients of one build are gathered on its owner's stream c.stream
// (prepare), then every build runs in one slab-GEMM and one SYRK launch on
// stream s (launch), K[j] = -sum W^T W of build j, the same bits as
// dfk_mo_bf16 gives.  gb keeps the launch tables between calls.
struct GroupBufs {
  DevBuf reqs, tiles[2], items[2];
};
void dfk_bf16_group_prepare(Ctx& c, DevBT& t, const double* C, int ldc, int nocc, ExGroupReq& r);
void dfk_bf16_group_launch(hipStream_t s, const std::vector<DevBT*>& ts, const std::vector<ExGroupReq>& reqs,
                           const std::vector<double*>& K, GroupBufs& gb);
int dfk_bf16_tuned_bn(const DevBT& t, int ng);
void dfk_bf16_plan(const DevBT& t, int ng, int* plan, float* ms, int* nms);
// Tuned FP32 exchange pipeline for ng (same fields as dfk_bf16_plan; plan[4] = 1, plan[5] = stages).
void dfk_f32_plan(const DevBT& t, int ng, int* plan, float* ms, int* nms);
void dfk_f32_force(DevBT& t, int ng, int k1, int k3);
int dfk_f32_num_shapes(int* nsyrk);
// Tuned FP64 pipeline (same fields; plan[0] / plan[6] = -1 until tuned).
void dfk_f64_plan(const DevBT& t, int ng, int* plan, float* ms, int* nms);
void dfk_f64_force(DevBT& t, int ng, int k1, int k3);
int dfk_f64_num_shapes(int* nsyrk);
void dfk_bf16_force(DevBT& t, int ng, int k1, int k3);
int dfk_bf16_num_shapes(int* nsyrk);
int dfk_bf16_shape_dl(int k1);
void dfk_ao(Ctx& c, DevBT& t, const Metric* m, const double* P_dev, double* K, int occ_nbatches, int form);
// Two-sided exchange (DFLMO_K, set_sym(false)): K = -(L^-1 B U)^T (L^-1 B Vt)
// per aux function, summed; U, Vt device nbf x ld (row-major), K not symmetric.
void dfk_uv(Ctx& c, DevBT& t, const Metric* m, const double* U, int ldu, const double* Vt, int ldv, int nocc,
            double* K, int occ_nbatches, bool fp64_copy);
const Tile* slab_tiles(Ctx& c, DevBT& t, int ng, int* ntiles, int BN = 0);
// out (slab layout of B) = rows (X, nu in N) of A (lda, klen) times gathered rows of G
void slab_out_gemm(Ctx& c, DevBT& t, const double* A, int lda, int klen, const double* G, int ldg,
                   double* out);
// Z = sum over slabs of A_slab B_slab^T (naux x naux, symmetric by construction)
void slab_syrk(Ctx& c, DevBT& t, const double* A, const double* B, double* Z, DevBuf& partial);

}  // namespace mchip
