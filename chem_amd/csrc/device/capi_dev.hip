// C ABI: device context, device-resident B, metric, DF-J and DF-K.
#include <cstring>
#include <memory>
#include <mutex>
#include <string>

#include "../../../include/mchip.h"
#include "../capi_common.hpp"
#include "dfock.hpp"
#include "ints12_dev.hpp"
#include "ints_dev.hpp"

using namespace mchip;

struct mchip_btensor_s;   // host container (capi_host.cpp)
namespace mchip {
const Basis& basis_of(mchip_basis b);
const BLayout& btensor_layout(mchip_btensor t);
const double* btensor_data(mchip_btensor t);
}  // namespace mchip

struct mchip_ctx_s {
  Ctx c;
  std::vector<double> guess_P;   // first density of the next SCF on this context (empty: core guess)
};
struct mchip_dbt_s {
  DevBT t;
  mchip_ctx_s* ctx;
  int device;
};
struct mchip_metric_s {
  Metric m;
  int device;
};

static void need_device(const mchip_ctx_s* c)
{
  if (!c) throw std::runtime_error("null context");
  HIP_CHECK(hipSetDevice(c->c.device));
}

extern "C" int mchip_device_count(int* n)
{
  MCHIP_TRY
  int k = 0;
  hipError_t e = hipGetDeviceCount(&k);
  *n = (e == hipSuccess) ? k : 0;
  MCHIP_CATCH
}

// Workspace of the allocation-free DF-J (mchip_dfj_dev_noalloc): a fixed
// device array, one copy per GPU for the life of the process.  Host code
// only ever uses its device address, looked up per device when a context is
// created on it (with that device current); the array's name is not a
// device address in host code.
static constexpr size_t kNoallocBytes = (size_t)128 << 20;
static __device__ __align__(256) double scratch[kNoallocBytes / sizeof(double)];
static std::mutex g_noalloc_mu;
static void* g_noalloc_ptr[64] = {};

// Device address of this GPU's copy of the array (the current device is
// `device`).
static double* noalloc_scratch(int device)
{
  if (device < 0 || device >= 64) throw std::runtime_error("device index out of range for the fixed workspace");
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (!g_noalloc_ptr[device]) {
    void* p = nullptr;
    HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw std::runtime_error("fixed workspace: no device address");
    g_noalloc_ptr[device] = p;
  }
  return static_cast<double*>(g_noalloc_ptr[device]);
}

extern "C" int mchip_ctx_create(int device, mchip_ctx* out)
{
  MCHIP_TRY
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
    throw std::runtime_error("no HIP device available");
  if (device < 0 || device >= n) throw std::runtime_error("device index out of range");
  auto c = std::make_unique<mchip_ctx_s>();
  ctx_init(c->c, device);   // makes `device` current
  (void)noalloc_scratch(device);
  *out = c.release();
  MCHIP_CATCH
}

extern "C" int mchip_ctx_set_guess(mchip_ctx c, int nbf, const double* P_alpha)
{
  MCHIP_TRY
  if (!c) throw std::runtime_error("null context");
  if (!P_alpha || nbf <= 0) {
    c->guess_P.clear();
  }
  else {
    c->guess_P.assign(P_alpha, P_alpha + (size_t)nbf * nbf);
  }
  MCHIP_CATCH
}

extern "C" int mchip_ctx_destroy(mchip_ctx c)
{
  MCHIP_TRY
  if (c) {
    (void)hipSetDevice(c->c.device);
    ctx_free(c->c);
    delete c;
  }
  MCHIP_CATCH
}

extern "C" int mchip_ctx_sync(mchip_ctx c)
{
  MCHIP_TRY
  need_device(c);
  HIP_CHECK(hipStreamSynchronize(c->c.stream));
  MCHIP_CATCH
}

extern "C" int mchip_ctx_stream(mchip_ctx c, void** stream)
{
  MCHIP_TRY
  *stream = (void*)c->c.stream;
  MCHIP_CATCH
}

extern "C" int mchip_ctx_set_escalation(mchip_ctx c, double rms, int stall)
{
  MCHIP_TRY
  if (!(rms >= 0.0)) throw std::runtime_error("mchip_ctx_set_escalation: the RMS threshold must be >= 0");
  if (stall < 0) throw std::runtime_error("mchip_ctx_set_escalation: the stall count must be >= 0");
  c->c.escalate_rms = rms;
  c->c.escalate_stall = stall;
  MCHIP_CATCH
}

extern "C" int mchip_dev_alloc(mchip_ctx c, size_t bytes, void** p)
{
  MCHIP_TRY
  need_device(c);
  if (!p) throw std::runtime_error("mchip_dev_alloc: null output pointer");
  *p = nullptr;
  if (bytes) HIP_CHECK(hipMalloc(p, bytes));
  MCHIP_CATCH
}

extern "C" int mchip_dev_free(mchip_ctx c, void* p)
{
  MCHIP_TRY
  need_device(c);
  HIP_CHECK(hipStreamSynchronize(c->c.stream));
  if (p) HIP_CHECK(hipFree(p));
  MCHIP_CATCH
}

extern "C" int mchip_dev_copy_in(mchip_ctx c, void* dst, const void* src, size_t bytes)
{
  MCHIP_TRY
  need_device(c);
  if (bytes) h2d(dst, src, bytes, c->c.stream);
  MCHIP_CATCH
}

extern "C" int mchip_dev_copy_out(mchip_ctx c, void* dst, const void* src, size_t bytes)
{
  MCHIP_TRY
  need_device(c);
  if (bytes) d2h(dst, src, bytes, c->c.stream);
  MCHIP_CATCH
}

extern "C" int mchip_ctx_trim(mchip_ctx c)
{
  MCHIP_TRY
  HIP_CHECK(hipStreamSynchronize(c->c.stream));
  c->c.scratch_a.release();
  c->c.scratch_b.release();
  c->c.scratch_c.release();
  c->c.scratch_partial.release();
  MCHIP_CATCH
}

extern "C" int mchip_ctx_set_timing(mchip_ctx c, int on)
{
  MCHIP_TRY
  c->c.timing = on != 0;
  for (double& v : c->c.phase) v = 0.0;
  MCHIP_CATCH
}

extern "C" int mchip_ctx_set_graphs(mchip_ctx c, int on)
{
  MCHIP_TRY
  c->c.graphs = on != 0;
  MCHIP_CATCH
}

extern "C" int mchip_dbt_graph_stats(mchip_dbt t, int64_t* ngraphs, int64_t* replays)
{
  MCHIP_TRY
  if (!t) throw std::runtime_error("mchip_dbt_graph_stats: null tensor");
  if (ngraphs) *ngraphs = (int64_t)t->t.graphs.exec.size();
  if (replays) *replays = t->t.graphs.replays;
  MCHIP_CATCH
}

extern "C" int mchip_ctx_phase_times(mchip_ctx c, double* out, int n)
{
  MCHIP_TRY
  for (int i = 0; i < n && i < 8; ++i) out[i] = c->c.phase[i];
  MCHIP_CATCH
}

extern "C" int mchip_dbt_upload(mchip_ctx c, mchip_btensor host, mchip_dbt* out)
{
  MCHIP_TRY
  need_device(c);
  auto d = std::make_unique<mchip_dbt_s>();
  d->ctx = c;
  d->device = c->c.device;
  dbt_upload(c->c, btensor_layout(host), btensor_data(host), d->t);
  *out = d.release();
  MCHIP_CATCH
}

extern "C" int mchip_dbt_build_gpu(mchip_ctx c, mchip_basis ao, mchip_basis aux, int nbb,
                                   const int* ao_blk_shells, int nxb, const int* aux_blk_shells,
                                   const uint8_t* schwarz_mask, double filter_eps, mchip_dbt* out)
{
  MCHIP_TRY
  need_device(c);
  std::vector<int> bsh(ao_blk_shells, ao_blk_shells + nbb), xsh(aux_blk_shells, aux_blk_shells + nxb);
  int sb = 0, sx = 0;
  for (int v : bsh) sb += v;
  for (int v : xsh) sx += v;
  if (sb != (int)basis_of(ao).shells.size() || sx != (int)basis_of(aux).shells.size())
    throw std::runtime_error("mchip_dbt_build_gpu: block shell counts do not cover the basis");
  std::vector<uint8_t> msk(schwarz_mask, schwarz_mask + (size_t)nxb * nbb * nbb);
  auto d = std::make_unique<mchip_dbt_s>();
  d->ctx = c;
  d->device = c->c.device;
  build_B_device(c->c, basis_of(ao), basis_of(aux), bsh, xsh, msk, filter_eps, d->t);
  *out = d.release();
  MCHIP_CATCH
}

extern "C" int mchip_int1e_gpu(mchip_ctx c, mchip_basis b, int natom, const double* Z, const double* xyz,
                               double* S, double* T, double* V)
{
  MCHIP_TRY
  need_device(c);
  if (natom < 0) throw std::runtime_error("mchip_int1e_gpu: negative charge count");
  std::vector<Charge> ch(natom);
  for (int i = 0; i < natom; ++i) {
    ch[i].Z = Z[i];
    for (int d = 0; d < 3; ++d) ch[i].xyz[d] = xyz[3 * i + d];
  }
  ints_1e_device(c->c, basis_of(b), ch, S, T, V);
  MCHIP_CATCH
}

extern "C" int mchip_int2c2e_gpu(mchip_ctx c, mchip_basis aux, double* V)
{
  MCHIP_TRY
  need_device(c);
  ints_2c2e_device(c->c, basis_of(aux), V);
  MCHIP_CATCH
}

extern "C" int mchip_schwarz_ao_gpu(mchip_ctx c, mchip_basis ao, double* Qao)
{
  MCHIP_TRY
  need_device(c);
  schwarz_ao_device(c->c, basis_of(ao), Qao);
  MCHIP_CATCH
}

extern "C" int mchip_dbt_info(mchip_dbt t, int* dims, int64_t* nelem)
{
  MCHIP_TRY
  dims[0] = t->t.L.naux;
  dims[1] = t->t.L.nbf;
  dims[2] = t->t.L.nxb;
  dims[3] = t->t.L.nbb;
  *nelem = t->t.L.total;
  MCHIP_CATCH
}

extern "C" int mchip_dbt_mask(mchip_dbt t, uint8_t* mask)
{
  MCHIP_TRY
  std::memcpy(mask, t->t.L.mask.data(), t->t.L.mask.size());
  MCHIP_CATCH
}

extern "C" int mchip_dbt_dense_rows(mchip_dbt t, int x0, int x1, double* dense)
{
  MCHIP_TRY
  need_device(t->ctx);
  const BLayout& L = t->t.L;
  if (x0 < 0 || x1 > L.naux || x0 > x1) throw std::runtime_error("mchip_dbt_dense_rows: bad range");
  const int nbf = L.nbf, nr = x1 - x0;
  std::memset(dense, 0, sizeof(double) * (size_t)nr * nbf * nbf);
  HIP_CHECK(hipStreamSynchronize(t->ctx->c.stream));
  std::vector<double> rows;
  for (int n = 0; n < L.nbb; ++n) {
    const int64_t R = (int64_t)L.bsz[n] * L.ncol[n];
    if (R == 0 || nr == 0) continue;
    rows.resize((size_t)nr * R);
    HIP_CHECK(hipMemcpy(rows.data(), t->t.data64.as<double>() + L.slab_off[n] + (int64_t)x0 * R,
                        rows.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int p = L.pair_ptr[n]; p < L.pair_ptr[n + 1]; ++p) {
      const int mb = L.pair_m[p], col = L.pair_col[p];
      for (int xi = 0; xi < nr; ++xi)
        for (int nu = 0; nu < L.bsz[n]; ++nu)
          for (int mu = 0; mu < L.bsz[mb]; ++mu)
            dense[((int64_t)xi * nbf + L.boff[mb] + mu) * nbf + L.boff[n] + nu] =
                rows[((int64_t)xi * L.bsz[n] + nu) * L.ncol[n] + col + mu];
    }
  }
  MCHIP_CATCH
}

extern "C" int mchip_dbt_destroy(mchip_dbt t)
{
  MCHIP_TRY
  if (t) {
    // the owning context may already be gone: synchronise the device instead
    (void)hipSetDevice(t->device);
    (void)hipDeviceSynchronize();
    delete t;
  }
  MCHIP_CATCH
}

extern "C" int mchip_dbt_download(mchip_dbt t, double* host)
{
  MCHIP_TRY
  need_device(t->ctx);
  HIP_CHECK(hipStreamSynchronize(t->ctx->c.stream));
  if (t->t.L.total)
    HIP_CHECK(hipMemcpy(host, t->t.data64.p, t->t.data64.bytes, hipMemcpyDeviceToHost));
  MCHIP_CATCH
}

extern "C" int mchip_metric_create(mchip_ctx c, int naux, const double* V, mchip_metric* out)
{
  MCHIP_TRY
  need_device(c);
  auto m = std::make_unique<mchip_metric_s>();
  m->device = c->c.device;
  metric_create(c->c, naux, V, m->m);
  *out = m.release();
  MCHIP_CATCH
}

extern "C" int mchip_metric_destroy(mchip_metric m)
{
  MCHIP_TRY
  if (m) {
    (void)hipSetDevice(m->device);
    (void)hipDeviceSynchronize();
    delete m;
  }
  MCHIP_CATCH
}

extern "C" int mchip_metric_get(mchip_ctx c, mchip_metric m, double* Linv, double* Vinv)
{
  MCHIP_TRY
  need_device(c);
  const size_t bytes = (size_t)m->m.naux * m->m.naux * sizeof(double);
  HIP_CHECK(hipStreamSynchronize(c->c.stream));
  if (Linv) HIP_CHECK(hipMemcpy(Linv, m->m.Linv.p, bytes, hipMemcpyDeviceToHost));
  if (Vinv) HIP_CHECK(hipMemcpy(Vinv, m->m.Vinv.p, bytes, hipMemcpyDeviceToHost));
  MCHIP_CATCH
}

extern "C" int mchip_dbt_transform(mchip_dbt t, mchip_metric m)
{
  MCHIP_TRY
  need_device(t->ctx);
  transform_B(t->ctx->c, t->t, m->m);
  MCHIP_CATCH
}

extern "C" int mchip_dfj_dev(mchip_dbt t, mchip_metric m, const double* P_dev, double pscale,
                             double* J_dev)
{
  MCHIP_TRY
  need_device(t->ctx);
  dfj(t->ctx->c, t->t, m ? &m->m : nullptr, P_dev, J_dev, pscale);
  MCHIP_CATCH
}

extern "C" int mchip_dfj_dev_noalloc(mchip_dbt t, mchip_metric m, const double* P_dev, double pscale,
                                     double* J_dev)
{
  MCHIP_TRY
  need_device(t->ctx);
  const size_t need = dfj_workspace_bytes(t->t);
  if (need > kNoallocBytes)
    throw std::runtime_error("mchip_dfj_dev_noalloc: this tensor's DF-J needs " + std::to_string(need) +
                             " bytes of workspace, more than the fixed " + std::to_string(kNoallocBytes) +
                             " (use mchip_dfj_dev)");
  dfj_ws(t->ctx->c, t->t, m ? &m->m : nullptr, P_dev, J_dev, pscale, noalloc_scratch(t->ctx->c.device));
  MCHIP_CATCH
}

// The _dev entry points read C through a raw device pointer: the occupied
// count must fit the row stride, the one size they can check.
static void check_occ(const char* fn, int ldc, int nocc)
{
  if (nocc < 0 || nocc > ldc)
    throw std::runtime_error(std::string(fn) + ": nocc " + std::to_string(nocc) + " outside 0..ldc (" +
                             std::to_string(ldc) + ")");
}

extern "C" int mchip_dfk_mo_dev(mchip_dbt t, mchip_metric m, const double* C_dev, int ldc, int nocc,
                                int occ_nbatches, double* K_dev)
{
  MCHIP_TRY
  need_device(t->ctx);
  check_occ("mchip_dfk_mo_dev", ldc, nocc);
  dfk_mo(t->ctx->c, t->t, m ? &m->m : nullptr, C_dev, ldc, nocc, K_dev, occ_nbatches);
  MCHIP_CATCH
}

// Host-pointer conveniences (copy in, compute, copy out, synchronous).
extern "C" int mchip_dfj(mchip_dbt t, mchip_metric m, const double* P, double pscale, double* J)
{
  MCHIP_TRY
  need_device(t->ctx);
  Ctx& c = t->ctx->c;
  const int nbf = t->t.L.nbf;
  const size_t bytes = (size_t)nbf * nbf * sizeof(double);
  DevBuf dP(bytes), dJ(bytes);
  h2d(dP.p, P, bytes, c.stream);
  dfj(c, t->t, m ? &m->m : nullptr, dP.as<double>(), dJ.as<double>(), pscale);
  d2h(J, dJ.p, bytes, c.stream);
  MCHIP_CATCH
}

extern "C" int mchip_dfk_mo(mchip_dbt t, mchip_metric m, const double* C, int nocc,
                            int occ_nbatches, double* K)
{
  MCHIP_TRY
  need_device(t->ctx);
  Ctx& c = t->ctx->c;
  const int nbf = t->t.L.nbf;
  const size_t kb = (size_t)nbf * nbf * sizeof(double);
  const size_t cb = (size_t)nbf * std::max(nocc, 1) * sizeof(double);
  DevBuf dC(cb), dK(kb);
  if (nocc > 0) h2d(dC.p, C, (size_t)nbf * nocc * sizeof(double), c.stream);
  dfk_mo(c, t->t, m ? &m->m : nullptr, dC.as<double>(), nocc, nocc, dK.as<double>(), occ_nbatches);
  d2h(K, dK.p, kb, c.stream);
  MCHIP_CATCH
}

extern "C" int mchip_dfk_lmo_uv(mchip_dbt t, mchip_metric m, const double* U, const double* V, int nocc,
                                int occ_nbatches, int fp64_copy, double* K)
{
  MCHIP_TRY
  need_device(t->ctx);
  if (nocc < 0) throw std::runtime_error("two-sided exchange: nocc < 0");
  Ctx& c = t->ctx->c;
  const int nbf = t->t.L.nbf;
  const size_t kb = (size_t)nbf * nbf * sizeof(double);
  const size_t cb = (size_t)nbf * std::max(nocc, 1) * sizeof(double);
  // v arrives as nocc x nbf (the reference's v_mb); the contraction wants v^T
  std::vector<double> vt((size_t)nbf * std::max(nocc, 1));
  for (int i = 0; i < nocc; ++i)
    for (int n = 0; n < nbf; ++n) vt[(size_t)n * nocc + i] = V[(size_t)i * nbf + n];
  DevBuf dU(cb), dVt(cb), dK(kb);
  if (nocc > 0) {
    h2d(dU.p, U, (size_t)nbf * nocc * sizeof(double), c.stream);
    h2d(dVt.p, vt.data(), (size_t)nbf * nocc * sizeof(double), c.stream);
  }
  dfk_uv(c, t->t, m ? &m->m : nullptr, dU.as<double>(), nocc, dVt.as<double>(), nocc, nocc, dK.as<double>(),
         occ_nbatches, fp64_copy != 0);
  d2h(K, dK.p, kb, c.stream);
  MCHIP_CATCH
}

#include "scf.hpp"

static thread_local double last_scf_extra[2] = {0.0, 0.0};

extern "C" int mchip_scf_last_extra(double* energy_iter, double* t_final_k)
{
  MCHIP_TRY
  *energy_iter = last_scf_extra[0];
  *t_final_k = last_scf_extra[1];
  MCHIP_CATCH
}

extern "C" int mchip_dbt_make_f32(mchip_dbt t)
{
  MCHIP_TRY
  need_device(t->ctx);
  dbt_make_f32(t->ctx->c, t->t);
  MCHIP_CATCH
}

extern "C" int mchip_dfk_mo_f32_dev(mchip_dbt t, const double* C_dev, int ldc, int nocc,
                                    int occ_nbatches, double* K_dev)
{
  MCHIP_TRY
  need_device(t->ctx);
  check_occ("mchip_dfk_mo_f32_dev", ldc, nocc);
  dfk_mo_f32(t->ctx->c, t->t, C_dev, ldc, nocc, K_dev, occ_nbatches);
  MCHIP_CATCH
}

extern "C" int mchip_dfk_mo_f32(mchip_dbt t, const double* C, int nocc, int occ_nbatches, double* K)
{
  MCHIP_TRY
  need_device(t->ctx);
  Ctx& c = t->ctx->c;
  const int nbf = t->t.L.nbf;
  const size_t kb = (size_t)nbf * nbf * sizeof(double);
  DevBuf dC((size_t)nbf * std::max(nocc, 1) * sizeof(double)), dK(kb);
  if (nocc > 0) h2d(dC.p, C, (size_t)nbf * nocc * sizeof(double), c.stream);
  dfk_mo_f32(c, t->t, dC.as<double>(), nocc, nocc, dK.as<double>(), occ_nbatches);
  d2h(K, dK.p, kb, c.stream);
  MCHIP_CATCH
}

extern "C" int mchip_dbt_make_bf16(mchip_dbt t)
{
  MCHIP_TRY
  need_device(t->ctx);
  dbt_make_bf16(t->ctx->c, t->t);
  MCHIP_CATCH
}

extern "C" int mchip_dbt_make_lowp(mchip_dbt t, mchip_metric m, int f32, int bf16, int f64)
{
  MCHIP_TRY
  need_device(t->ctx);
  if (!m) throw std::runtime_error("mchip_dbt_make_lowp: the metric is required");
  dbt_make_lowp(t->ctx->c, t->t, m->m, f32 != 0, bf16 != 0, f64 != 0);
  MCHIP_CATCH
}

extern "C" int mchip_dfk_mo_tilde64(mchip_dbt t, const double* C, int nocc, int occ_nbatches, double* K)
{
  MCHIP_TRY
  need_device(t->ctx);
  if (!t->t.data64t.p) throw std::runtime_error("mchip_dfk_mo_tilde64: no FP64 L^-1 B copy (mchip_dbt_make_lowp f64)");
  Ctx& c = t->ctx->c;
  const int n = t->t.L.nbf;
  DevBuf dC((size_t)n * std::max(nocc, 1) * sizeof(double)), dK((size_t)n * n * sizeof(double));
  if (nocc) h2d(dC.p, C, (size_t)n * nocc * sizeof(double), c.stream);
  dfk_mo(c, t->t, nullptr, dC.as<double>(), nocc, nocc, dK.as<double>(), occ_nbatches, t->t.data64t.as<double>());
  d2h(K, dK.p, (size_t)n * n * sizeof(double), c.stream);
  MCHIP_CATCH
}

extern "C" int mchip_dfk_mo_tilde64_dev(mchip_dbt t, const double* C_dev, int ldc, int nocc, int occ_nbatches,
                                        double* K_dev)
{
  MCHIP_TRY
  need_device(t->ctx);
  check_occ("mchip_dfk_mo_tilde64_dev", ldc, nocc);
  if (!t->t.data64t.p) throw std::runtime_error("mchip_dfk_mo_tilde64_dev: no FP64 L^-1 B copy (mchip_dbt_make_lowp f64)");
  dfk_mo(t->ctx->c, t->t, nullptr, C_dev, ldc, nocc, K_dev, occ_nbatches, t->t.data64t.as<double>());
  MCHIP_CATCH
}

extern "C" int mchip_dfk_mo_bf16_dev(mchip_dbt t, const double* C_dev, int ldc, int nocc,
                                     int occ_nbatches, double* K_dev)
{
  MCHIP_TRY
  need_device(t->ctx);
  check_occ("mchip_dfk_mo_bf16_dev", ldc, nocc);
  dfk_mo_bf16(t->ctx->c, t->t, C_dev, ldc, nocc, K_dev, occ_nbatches);
  MCHIP_CATCH
}

extern "C" int mchip_dfk_energy_tilde64_dev(mchip_dbt t, const double* C_dev, int ldc, int nocc, double* ek)
{
  MCHIP_TRY
  need_device(t->ctx);
  check_occ("mchip_dfk_energy_tilde64_dev", ldc, nocc);
  if (!ek) throw std::runtime_error("mchip_dfk_energy_tilde64_dev: null result pointer");
  *ek = dfk_energy_tilde64(t->ctx->c, t->t, C_dev, ldc, nocc);
  MCHIP_CATCH
}

extern "C" int mchip_dfk_energy_dev(mchip_dbt t, mchip_metric m, const double* C_dev, int ldc, int nocc,
                                    int occ_nbatches, double* ek)
{
  MCHIP_TRY
  need_device(t->ctx);
  check_occ("mchip_dfk_energy_dev", ldc, nocc);
  if (!ek) throw std::runtime_error("mchip_dfk_energy_dev: null result pointer");
  if (!m) throw std::runtime_error("mchip_dfk_energy_dev: the metric is required");
  *ek = dfk_energy_metric(t->ctx->c, t->t, m->m, C_dev, ldc, nocc, occ_nbatches);
  MCHIP_CATCH
}

extern "C" int mchip_dfk_mo_bf16_group_dev(int n, const mchip_dbt* ts, const double* const* C_dev, const int* ldc,
                                           const int* nocc, double* const* K_dev)
{
  MCHIP_TRY
  if (n <= 0) return MCHIP_OK;
  if (!ts || !C_dev || !ldc || !nocc || !K_dev) throw std::runtime_error("mchip_dfk_mo_bf16_group_dev: null argument");
  mchip_ctx ctx = ts[0]->ctx;
  need_device(ctx);
  Ctx& c = ctx->c;
  std::vector<DevBT*> tv(n);
  std::vector<ExGroupReq> reqs(n);
  std::vector<double*> K(n);
  for (int j = 0; j < n; ++j) {
    if (!ts[j] || ts[j]->ctx != ctx) throw std::runtime_error("mchip_dfk_mo_bf16_group_dev: the tensors must share one context");
    for (int i = 0; i < j; ++i)
      if (ts[i] == ts[j]) throw std::runtime_error("mchip_dfk_mo_bf16_group_dev: a tensor appears twice");
    check_occ("mchip_dfk_mo_bf16_group_dev", ldc[j], nocc[j]);
    if (nocc[j] <= 0) throw std::runtime_error("mchip_dfk_mo_bf16_group_dev: every build needs nocc > 0");
    tv[j] = &ts[j]->t;
    K[j] = K_dev[j];
    dfk_bf16_group_prepare(c, *tv[j], C_dev[j], ldc[j], nocc[j], reqs[j]);
  }
  static thread_local GroupBufs gb;
  dfk_bf16_group_launch(c.stream, tv, reqs, K, gb);
  MCHIP_CATCH
}

extern "C" int mchip_dfk_mo_bf16(mchip_dbt t, const double* C, int nocc, int occ_nbatches, double* K)
{
  MCHIP_TRY
  need_device(t->ctx);
  Ctx& c = t->ctx->c;
  const int nbf = t->t.L.nbf;
  const size_t kb = (size_t)nbf * nbf * sizeof(double);
  DevBuf dC((size_t)nbf * std::max(nocc, 1) * sizeof(double)), dK(kb);
  if (nocc > 0) h2d(dC.p, C, (size_t)nbf * nocc * sizeof(double), c.stream);
  dfk_mo_bf16(c, t->t, dC.as<double>(), nocc, nocc, dK.as<double>(), occ_nbatches);
  d2h(K, dK.p, kb, c.stream);
  MCHIP_CATCH
}

extern "C" int mchip_dfk_ao_dev(mchip_dbt t, mchip_metric m, const double* P_dev, int form,
                                int occ_nbatches, double* K_dev)
{
  MCHIP_TRY
  need_device(t->ctx);
  if (!t->t.transformed && !m) throw std::runtime_error("mchip_dfk_ao: the metric is required for an untransformed B");
  dfk_ao(t->ctx->c, t->t, m ? &m->m : nullptr, P_dev, K_dev, occ_nbatches, form);
  MCHIP_CATCH
}

extern "C" int mchip_dfk_ao(mchip_dbt t, mchip_metric m, const double* P, int form, int occ_nbatches,
                            double* K)
{
  MCHIP_TRY
  need_device(t->ctx);
  if (!t->t.transformed && !m) throw std::runtime_error("mchip_dfk_ao: the metric is required for an untransformed B");
  Ctx& c = t->ctx->c;
  const int nbf = t->t.L.nbf;
  const size_t kb = (size_t)nbf * nbf * sizeof(double);
  DevBuf dP(kb), dK(kb);
  h2d(dP.p, P, kb, c.stream);
  dfk_ao(c, t->t, m ? &m->m : nullptr, dP.as<double>(), dK.as<double>(), occ_nbatches, form);
  d2h(K, dK.p, kb, c.stream);
  MCHIP_CATCH
}

extern "C" int mchip_dfk_bf16_tile(mchip_dbt t, int ng, int* bn)
{
  MCHIP_TRY
  *bn = dfk_bf16_tuned_bn(t->t, ng);
  MCHIP_CATCH
}

extern "C" int mchip_dfk_bf16_force(mchip_dbt t, int ng, int gemm_shape, int syrk_variant)
{
  MCHIP_TRY
  dfk_bf16_force(t->t, ng, gemm_shape, syrk_variant);
  MCHIP_CATCH
}

extern "C" int mchip_dfk_bf16_shapes(int* ngemm, int* nsyrk, int* direct)
{
  MCHIP_TRY
  *ngemm = dfk_bf16_num_shapes(nsyrk);
  if (direct)
    for (int i = 0; i < *ngemm; ++i) direct[i] = dfk_bf16_shape_dl(i);
  MCHIP_CATCH
}

extern "C" int mchip_dfk_bf16_plan(mchip_dbt t, int ng, int* plan, float* ms, int* nms)
{
  MCHIP_TRY
  dfk_bf16_plan(t->t, ng, plan, ms, nms);
  MCHIP_CATCH
}

extern "C" int mchip_dfk_f32_plan(mchip_dbt t, int ng, int* plan, float* ms, int* nms)
{
  MCHIP_TRY
  dfk_f32_plan(t->t, ng, plan, ms, nms);
  MCHIP_CATCH
}

extern "C" int mchip_dfk_f32_force(mchip_dbt t, int ng, int gemm_shape, int syrk_variant)
{
  MCHIP_TRY
  dfk_f32_force(t->t, ng, gemm_shape, syrk_variant);
  MCHIP_CATCH
}

extern "C" int mchip_dfk_f32_shapes(int* ngemm, int* nsyrk)
{
  MCHIP_TRY
  *ngemm = dfk_f32_num_shapes(nsyrk);
  MCHIP_CATCH
}

extern "C" int mchip_dfk_f64_plan(mchip_dbt t, int ng, int* plan, float* ms, int* nms)
{
  MCHIP_TRY
  dfk_f64_plan(t->t, ng, plan, ms, nms);
  MCHIP_CATCH
}

extern "C" int mchip_dfk_f64_force(mchip_dbt t, int ng, int gemm_shape, int syrk_variant)
{
  MCHIP_TRY
  dfk_f64_force(t->t, ng, gemm_shape, syrk_variant);
  MCHIP_CATCH
}

extern "C" int mchip_dfk_f64_shapes(int* ngemm, int* nsyrk)
{
  MCHIP_TRY
  *ngemm = dfk_f64_num_shapes(nsyrk);
  MCHIP_CATCH
}

extern "C" int mchip_scf_run(mchip_dbt t, mchip_metric m, int nbf, int nocc, int nbb,
                             const int* bsz, const double* S, const double* H, double enuc,
                             double threshold, int max_iter, int do_diis, int diis_start,
                             int diis_min, int diis_max, double filter_eps, int kmode,
                             int occ_nbatches, double* energy, int* iterations, int* converged,
                             double* C, double* eps, int max_hist, double* hist_e,
                             double* hist_rms, double* t_j, double* t_k, int* nhist)
{
  MCHIP_TRY
  need_device(t->ctx);
  Ctx& c = t->ctx->c;
  if (nbf != t->t.L.nbf) throw std::runtime_error("mchip_scf_run: nbf does not match B");
  if (nocc < 0 || nocc > nbf) throw std::runtime_error("mchip_scf_run: bad nocc");
  ScfInput in;
  in.nbf = nbf;
  in.nocc = nocc;
  in.bsz.assign(bsz, bsz + nbb);
  in.S = S;
  in.H = H;
  in.enuc = enuc;
  in.threshold = threshold;
  in.max_iter = max_iter;
  in.do_diis = do_diis != 0;
  in.diis_start = diis_start;
  in.diis_min = diis_min;
  in.diis_max = diis_max;
  in.filter_eps = filter_eps;
  if (!t->ctx->guess_P.empty()) {
    if ((int64_t)t->ctx->guess_P.size() != (int64_t)nbf * nbf)
      throw std::runtime_error("mchip_scf_run: the guess density set on the context is not nbf x nbf");
    in.guess_P.swap(t->ctx->guess_P);
  }
  const Metric* met = m ? &m->m : nullptr;
  const Metric* jmet = t->t.transformed ? nullptr : met;
  KBuildFn kb, kfinal;
  EnergyFn kenergy;
  select_k_builders(c, t->t, met, kmode, nocc, occ_nbatches, threshold, kb, kfinal, 0, &kenergy);
  ScfOutput out;
  scf_run(c, t->t, jmet, in, out, kb, kfinal, kenergy);
  last_scf_extra[0] = out.energy_iter;
  last_scf_extra[1] = out.t_final_k;
  *energy = out.energy;
  *iterations = out.iterations;
  *converged = out.converged ? 1 : 0;
  if (C) std::memcpy(C, out.C.data(), out.C.size() * sizeof(double));
  if (eps) std::memcpy(eps, out.eps.data(), out.eps.size() * sizeof(double));
  const int nh = std::min<int>(max_hist, (int)out.history_e.size());
  for (int i = 0; i < nh; ++i) {
    if (hist_e) hist_e[i] = out.history_e[i];
    if (hist_rms) hist_rms[i] = out.history_rms[i];
    if (t_j) t_j[i] = out.t_j[i];
    if (t_k) t_k[i] = out.t_k[i];
  }
  *nhist = nh;
  MCHIP_CATCH
}

#include "mp2.hpp"

static int mp2_entry(mchip_dbt t, mchip_metric m, int nbf, int nocc, const double* S, const double* C,
                     const double* eps, int nlap, const double* omega, const double* alpha, int nbb,
                     const int* bsz, int mo_split, double filter_eps, double* energy, double* terms, double* t_z,
                     double* t_prep, bool batched, double* Z = nullptr, int zmethod = 0)
{
  MCHIP_TRY
  need_device(t->ctx);
  if (!m) throw std::runtime_error("mchip_mp2_sos: the metric is required");
  if (nlap <= 0) throw std::runtime_error("mchip_mp2_sos: at least one Laplace point is required");
  Mp2Input in;
  in.nbf = nbf;
  in.nocc = nocc;
  in.S = S;
  in.C = C;
  in.eps = eps;
  in.omega.assign(omega, omega + nlap);
  in.alpha.assign(alpha, alpha + nlap);
  in.bsz.assign(bsz, bsz + nbb);
  in.mo_split = mo_split;
  in.filter_eps = filter_eps;
  in.zmethod = zmethod;
  std::vector<double> tv;
  double tz = 0.0, tp = 0.0;
  if (batched)
    *energy = mp2_sos_batched(t->ctx->c, t->t, m->m, in, tv, &tz, &tp, Z);
  else
    *energy = mp2_sos(t->ctx->c, t->t, m->m, in, tv, &tz, Z);
  if (terms)
    for (int k = 0; k < nlap && k < (int)tv.size(); ++k) terms[k] = tv[k];
  if (t_z) *t_z = tz;
  if (t_prep) *t_prep = tp;
  MCHIP_CATCH
}

extern "C" int mchip_mp2_sos(mchip_dbt t, mchip_metric m, int nbf, int nocc, const double* S,
                             const double* C, const double* eps, int nlap, const double* omega,
                             const double* alpha, int nbb, const int* bsz, int mo_split,
                             double filter_eps, double* energy, double* terms, double* t_z)
{
  return mp2_entry(t, m, nbf, nocc, S, C, eps, nlap, omega, alpha, nbb, bsz, mo_split, filter_eps, energy, terms,
                   t_z, nullptr, false);
}

extern "C" int mchip_mp2_sos_batched(mchip_dbt t, mchip_metric m, int nbf, int nocc, const double* S,
                                     const double* C, const double* eps, int nlap, const double* omega,
                                     const double* alpha, int nbb, const int* bsz, int mo_split,
                                     double filter_eps, double* energy, double* terms, double* t_z,
                                     double* t_prep)
{
  return mp2_entry(t, m, nbf, nocc, S, C, eps, nlap, omega, alpha, nbb, bsz, mo_split, filter_eps, energy, terms,
                   t_z, t_prep, true);
}

extern "C" int mchip_mp2_sos_z(mchip_dbt t, mchip_metric m, int zmethod, int nbf, int nocc, const double* S,
                               const double* C, const double* eps, int nlap, const double* omega,
                               const double* alpha, int nbb, const int* bsz, int mo_split, double filter_eps,
                               double* energy, double* terms, double* t_z)
{
  return mp2_entry(t, m, nbf, nocc, S, C, eps, nlap, omega, alpha, nbb, bsz, mo_split, filter_eps, energy, terms,
                   t_z, nullptr, false, nullptr, zmethod);
}

extern "C" int mchip_z_compute(mchip_dbt t, int zmethod, int nbf, const double* Locc, int rank_o,
                               const double* Pvir, const double* Lvir, int rank_v, double* Z)
{
  MCHIP_TRY
  need_device(t->ctx);
  Ctx& c = t->ctx->c;
  DevBT& bt = t->t;
  if (bt.transformed) throw std::runtime_error("mchip_z_compute: needs the untransformed B");
  if (nbf != bt.L.nbf) throw std::runtime_error("mchip_z_compute: nbf does not match B");
  if (zmethod < 0 || zmethod > 2) throw std::runtime_error("mchip_z_compute: unknown Z method");
  if (rank_o < 0 || rank_v < 0) throw std::runtime_error("mchip_z_compute: negative rank");
  if (rank_o > 0 && !Locc) throw std::runtime_error("mchip_z_compute: L_occ is required");
  if (zmethod == 2 && rank_v > 0 && !Lvir) throw std::runtime_error("mchip_z_compute: LL_FULL needs L_vir");
  if (zmethod != 2 && !Pvir) throw std::runtime_error("mchip_z_compute: LLMP needs P_vir");
  const int naux = bt.L.naux;
  const size_t nn = (size_t)nbf * nbf;
  DevBuf dLo((size_t)nbf * std::max(rank_o, 1) * sizeof(double)), dPv, dLv,
      dZ((size_t)naux * naux * sizeof(double));
  if (rank_o) h2d(dLo.p, Locc, (size_t)nbf * rank_o * sizeof(double), c.stream);
  if (zmethod != 2) {
    dPv.alloc(nn * sizeof(double));
    h2d(dPv.p, Pvir, nn * sizeof(double), c.stream);
  }
  else if (rank_v) {
    dLv.alloc((size_t)nbf * rank_v * sizeof(double));
    h2d(dLv.p, Lvir, (size_t)nbf * rank_v * sizeof(double), c.stream);
  }
  z_point(c, bt, zmethod, dLo.as<double>(), rank_o, dPv.p ? dPv.as<double>() : nullptr,
          dLv.p ? dLv.as<double>() : nullptr, rank_v, dZ.as<double>());
  d2h(Z, dZ.p, (size_t)naux * naux * sizeof(double), c.stream);
  MCHIP_CATCH
}

extern "C" int mchip_mp2_z(mchip_dbt t, mchip_metric m, int nbf, int nocc, const double* S, const double* C,
                           const double* eps, int nlap, const double* omega, const double* alpha, int nbb,
                           const int* bsz, int mo_split, double filter_eps, int batched, double* Z,
                           double* terms)
{
  if (!Z) {
    mchip::set_error("mchip_mp2_z: Z must not be null");
    return MCHIP_ERR;
  }
  double e = 0.0;
  return mp2_entry(t, m, nbf, nocc, S, C, eps, nlap, omega, alpha, nbb, bsz, mo_split, filter_eps, &e, terms,
                   nullptr, nullptr, batched != 0, Z);
}

// ---------------------------------------------------------------------------
// Request loop
// ---------------------------------------------------------------------------
#include "serve.hpp"

struct mchip_server_s {
  std::unique_ptr<Server> s;
};

extern "C" int mchip_server_create(int device, double budget_bytes, int nslots, int cpu_threads,
                                   mchip_server* out)
{
  MCHIP_TRY
  if (budget_bytes <= 0) throw std::runtime_error("mchip_server_create: budget must be positive");
  auto s = std::make_unique<mchip_server_s>();
  s->s = std::make_unique<Server>(device, (size_t)budget_bytes, nslots, std::max(1, cpu_threads));
  *out = s.release();
  MCHIP_CATCH
}

extern "C" int mchip_server_destroy(mchip_server s)
{
  MCHIP_TRY
  delete s;
  MCHIP_CATCH
}

extern "C" int mchip_server_submit(mchip_server s, mchip_basis ao, mchip_basis aux, int natom, const double* Z,
                                   const double* xyz, int nbb, const int* ao_blk_shells, int nxb,
                                   const int* aux_blk_shells, int nocc, double enuc, double threshold,
                                   int kmode, double screen_eps, int* id)
{
  MCHIP_TRY
  if (kmode < 0 || kmode > 4) throw std::runtime_error("mchip_server_submit: unknown K mode");
  auto j = std::make_unique<SrvJob>();
  j->ao = basis_of(ao);
  j->aux = basis_of(aux);
  j->charges.resize(natom);
  for (int i = 0; i < natom; ++i) {
    j->charges[i].Z = Z[i];
    for (int d = 0; d < 3; ++d) j->charges[i].xyz[d] = xyz[3 * i + d];
  }
  j->ao_blk.assign(ao_blk_shells, ao_blk_shells + nbb);
  j->aux_blk.assign(aux_blk_shells, aux_blk_shells + nxb);
  int sb = 0, sx = 0;
  for (int v : j->ao_blk) sb += v;
  for (int v : j->aux_blk) sx += v;
  if (sb != (int)j->ao.shells.size() || sx != (int)j->aux.shells.size())
    throw std::runtime_error("mchip_server_submit: block shell counts do not cover the basis");
  if (nocc < 0 || nocc > j->ao.nbf) throw std::runtime_error("mchip_server_submit: bad nocc");
  j->nocc = nocc;
  j->enuc = enuc;
  j->threshold = threshold;
  j->kmode = kmode;
  j->eps = screen_eps;
  *id = s->s->submit(std::move(j));
  MCHIP_CATCH
}

extern "C" int mchip_server_result(mchip_server s, int id, int wait, int* done, double* energy, int* iterations,
                                   int* converged, double* times)
{
  MCHIP_TRY
  SrvJob j;
  const bool d = s->s->result(id, wait != 0, j);
  *done = d ? 1 : 0;
  if (d) {
    if (j.failed) throw std::runtime_error("request " + std::to_string(id) + " failed: " + j.error);
    *energy = j.energy;
    *iterations = j.iterations;
    *converged = j.converged ? 1 : 0;
    if (times) {
      times[0] = j.t_submit;
      times[1] = j.t_admit;
      times[2] = j.t_done;
      times[3] = (double)j.est_bytes;
      times[4] = (double)j.slot;
      times[5] = j.energy_iter;
      times[6] = j.t_host;
      times[7] = j.t_build;
      times[8] = j.t_setup;
      times[9] = j.t_scf;
    }
  }
  MCHIP_CATCH
}

extern "C" int mchip_server_set_batching(mchip_server s, int enabled)
{
  MCHIP_TRY
  s->s->set_batching(enabled != 0);
  MCHIP_CATCH
}

extern "C" int mchip_server_batch_stats(mchip_server s, double* stats)
{
  MCHIP_TRY
  s->s->batch_stats(stats);
  MCHIP_CATCH
}

extern "C" int mchip_server_stats(mchip_server s, double* stats)
{
  MCHIP_TRY
  s->s->stats(stats);
  stats[5] = s->s->clock();
  MCHIP_CATCH
}
