// Device drivers of the DF Fock build: B upload, Coulomb metric, DF-J and
// the DFMO exchange, FP64.
//
//   metric_create   aoloader::invert + math::LLT (aoloader.cpp:14-32,
//                   LLT.cpp:8-109): V = L L^T, L^-1, V^-1 = L^-T L^-1 on the
//                   library's own blocked FP64 kernels (linalg.hip).
//   dfj             DF_J::compute_J (jkdf.cpp:69-190).
//   dfk_mo          DFMO_K::compute_K (jkdf.cpp:221-462), K returned negated.
//   transform_B     B~ = L^-1 B (one-time), after which K = -sum_X (B~_X C)^T
//                   (B~_X C) and J = B~^T (B~ . P) need no metric per call.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <type_traits>

#include <rocsolver/rocsolver.h>

#include "dfock.hpp"
#include "ints_dev.hpp"
#include "kernels.hpp"
#include "linalg.hpp"
#include "mp2.hpp"

namespace mchip {

// Runs the launch chain `launch` of one build on c.stream.  The first run of a
// key is direct (it may tune shapes and upload tile tables, which synchronise
// the stream); the second is captured into a hipGraph, instantiated and
// launched; later runs replay that graph.  Every kernel argument of the chain
// derives from the key, and the graph reads the operands' current contents,
// so a replay computes what a direct run would, bit for bit.  Per-phase timing
// and a context with graphs off run every chain directly.
template <class F>
static void run_chain(Ctx& c, DevBT& t, const ChainKey& key, F&& launch)
{
  GraphCache& g = t.graphs;
  if (!c.graphs || c.timing) {
    launch();
    return;
  }
  auto it = g.exec.find(key);
  if (it != g.exec.end()) {
    HIP_CHECK(hipGraphLaunch(it->second, c.stream));
    ++g.replays;
    return;
  }
  if (g.seen.size() >= 1024) g.seen.clear();   // bounded: callers with ever new buffers
  if (g.seen.insert(key).second) {
    launch();
    return;
  }
  if (g.exec.size() >= 32) {   // bounded: drop every graph once no launch of them is in flight
    HIP_CHECK(hipStreamSynchronize(c.stream));
    for (auto& kv : g.exec) (void)hipGraphExecDestroy(kv.second);
    g.exec.clear();
  }
  hipGraph_t graph = nullptr;
  HIP_CHECK(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
  try {
    launch();
  }
  catch (...) {
    (void)hipStreamEndCapture(c.stream, &graph);
    if (graph) (void)hipGraphDestroy(graph);
    throw;
  }
  HIP_CHECK(hipStreamEndCapture(c.stream, &graph));
  hipGraphExec_t exec = nullptr;
  const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  HIP_CHECK(e);
  g.exec[key] = exec;
  HIP_CHECK(hipGraphLaunch(exec, c.stream));
  ++g.replays;
}

static int64_t pkey(const void* p) { return (int64_t)reinterpret_cast<uintptr_t>(p); }

// The tuned shapes of a plan for width ng (-1 while untuned).
static int64_t plan_key(const std::map<int, ExchangePlan>& plans, int ng)
{
  auto it = plans.find(ng);
  if (it == plans.end() || it->second.k1 < 0 || it->second.k3 < 0) return -1;
  return (int64_t)it->second.k1 * 1000 + it->second.k3;
}

__global__ void k_zero_upper_cm(double* A, int n)
{
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n * n) return;
  const int j = (int)(idx / n), i = (int)(idx % n);   // column-major (i,j)
  if (i < j) A[idx] = 0.0;
}

void ctx_init(Ctx& c, int device)
{
  c.device = device;
  HIP_CHECK(hipSetDevice(device));
  HIP_CHECK(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
  RB_CHECK(rocblas_create_handle(&c.blas));
  RB_CHECK(rocblas_set_stream(c.blas, c.stream));
  c.info.alloc(sizeof(int) * 4);
  HIP_CHECK(hipEventCreate(&c.ev0));
  HIP_CHECK(hipEventCreate(&c.ev1));
  HIP_CHECK(hipEventCreate(&c.tev[0]));
  HIP_CHECK(hipEventCreate(&c.tev[1]));
}

void ctx_free(Ctx& c)
{
  if (c.stream) (void)hipStreamSynchronize(c.stream);
  c.scratch_a.release();
  c.scratch_b.release();
  c.scratch_c.release();
  c.scratch_partial.release();
  c.syrk_tiles.clear();
  c.info.release();
  if (c.stream) (void)hipStreamSynchronize(c.stream);   // stream-ordered frees land first
  if (c.blas) rocblas_destroy_handle(c.blas);
  if (c.stream) (void)hipStreamDestroy(c.stream);
  if (c.ev0) (void)hipEventDestroy(c.ev0);
  if (c.ev1) (void)hipEventDestroy(c.ev1);
  for (auto& e : c.tev)
    if (e) (void)hipEventDestroy(e);
  c.blas = nullptr;
  c.stream = nullptr;
}

static void dbt_index(Ctx& c, const BLayout& L, DevBT& t);

void dbt_upload(Ctx& c, const BLayout& L, const double* host_data, DevBT& t)
{
  HIP_CHECK(hipSetDevice(c.device));
  t.L = L;
  t.data64.alloc((size_t)L.total * sizeof(double));
  if (L.total)
    h2d(t.data64.p, host_data, t.data64.bytes, c.stream);
  dbt_index(c, L, t);
}

void dbt_adopt(Ctx& c, const BLayout& L, DevBuf&& data, DevBT& t)
{
  HIP_CHECK(hipSetDevice(c.device));
  t.L = L;
  t.data64 = std::move(data);
  dbt_index(c, L, t);
}

static void dbt_index(Ctx& c, const BLayout& L, DevBT& t)
{
  // column -> AO function map per slab
  t.h_colmap_off.assign(L.nbb + 1, 0);
  t.h_colmap.clear();
  for (int n = 0; n < L.nbb; ++n) {
    for (int p = L.pair_ptr[n]; p < L.pair_ptr[n + 1]; ++p)
      for (int mu = 0; mu < L.bsz[L.pair_m[p]]; ++mu) t.h_colmap.push_back(L.boff[L.pair_m[p]] + mu);
    t.h_colmap_off[n + 1] = (int)t.h_colmap.size();
  }
  t.colmap_total = (int64_t)t.h_colmap.size();
  t.h_pg_off.assign(L.nbb + 1, 0);
  t.maxR = 0;
  for (int n = 0; n < L.nbb; ++n) {
    int64_t R = (int64_t)L.bsz[n] * L.ncol[n];
    t.h_pg_off[n + 1] = t.h_pg_off[n] + R;
    t.maxR = std::max<int64_t>(t.maxR, R);
  }
  t.pg_total = t.h_pg_off[L.nbb];
  // canonical half of each slab (columns of blocks M <= N) for DF-J, when
  // every significant pair (M, N) has its transpose (N, M)
  {
    std::vector<int2> cdn(std::max(L.nbb, 1), make_int2(0, 0));
    t.maxRc = 0;
    t.sym_mask = true;
    for (int x = 0; x < L.nxb && t.sym_mask; ++x)
      for (int m = 0; m < L.nbb && t.sym_mask; ++m)
        for (int n = 0; n < m; ++n)
          if (L.kept(x, m, n) != L.kept(x, n, m)) {
            t.sym_mask = false;
            break;
          }
    for (int n = 0; n < L.nbb; ++n) {
      int cd = L.ncol[n], ce = L.ncol[n];
      for (int p = L.pair_ptr[n]; p < L.pair_ptr[n + 1]; ++p) {
        const int m = L.pair_m[p];
        if (m < n) continue;
        cd = L.pair_col[p];
        ce = m == n ? L.pair_col[p] + L.bsz[m] : L.pair_col[p];
        break;
      }
      cdn[n] = make_int2(cd, ce);
      t.maxRc = std::max<int64_t>(t.maxRc, (int64_t)L.bsz[n] * ce);
    }
    t.cdn = upload_vec(cdn, c.stream);
  }
  std::vector<int64_t> so(L.slab_off.begin(), L.slab_off.end());
  t.slab_off = upload_vec(so, c.stream);
  t.ncol = upload_vec(L.ncol, c.stream);
  t.bsz = upload_vec(L.bsz, c.stream);
  std::vector<int> bo(L.boff.begin(), L.boff.end());
  t.boff = upload_vec(bo, c.stream);
  t.colmap = upload_vec(t.h_colmap, c.stream);
  t.colmap_off = upload_vec(t.h_colmap_off, c.stream);
  t.pg_off = upload_vec(t.h_pg_off, c.stream);
  t.transformed = false;
  HIP_CHECK(hipStreamSynchronize(c.stream));
}

// Column tile of the grouped slab GEMM for ng output columns.
static int slab_bn(int ng) { return ng <= 64 ? 64 : 128; }

// Tile table of the grouped slab GEMM for ng output columns (column tile BN,
// 0 = by ng).
const Tile* slab_tiles(Ctx& c, DevBT& t, int ng, int* ntiles, int BN)
{
  if (BN == 0) BN = slab_bn(ng);
  auto key = std::make_pair(BN, ng);
  auto it = t.tiles.find(key);
  if (it == t.tiles.end()) {
    std::vector<Tile> v;
    for (int n = 0; n < t.L.nbb; ++n) {
      const int64_t rows = (int64_t)t.L.naux * t.L.bsz[n];
      for (int64_t m0 = 0; m0 < rows; m0 += SG_BM)
        for (int n0 = 0; n0 < ng; n0 += BN) v.push_back(Tile{n, (int)m0, n0, 0});
    }
    t.ntiles[key] = (int)v.size();
    t.tiles[key] = upload_vec(v, c.stream);
    HIP_CHECK(hipStreamSynchronize(c.stream));   // v is pageable and dies here
    it = t.tiles.find(key);
  }
  *ntiles = t.ntiles[key];
  return it->second.as<Tile>();
}

template <class T, int FL = 0>
static void slab_gemm_t(Ctx& c, DevBT& t, const T* data, const T* G, int ldg, int ng, T* out, int ldo,
                        int bn = 0)
{
  if (bn == 0) bn = slab_bn(ng);
  int nt = 0;
  const Tile* tiles = slab_tiles(c, t, ng, &nt, bn);
  if (nt == 0) return;
  if (bn == 64)
    hipLaunchKernelGGL((k_slab_gemm<T, SG_BM, 64, FL>), dim3(nt), dim3(256), 0, c.stream, data,
                       t.slab_off.as<int64_t>(), t.ncol.as<int>(), t.bsz.as<int>(), t.boff.as<int>(),
                       t.colmap.as<int>(), t.colmap_off.as<int>(), tiles, t.L.naux, t.L.nbf, G, ldg,
                       ng, out, ldo);
  else
    hipLaunchKernelGGL((k_slab_gemm<T, SG_BM, 128, FL>), dim3(nt), dim3(256), 0, c.stream, data,
                       t.slab_off.as<int64_t>(), t.ncol.as<int>(), t.bsz.as<int>(), t.boff.as<int>(),
                       t.colmap.as<int>(), t.colmap_off.as<int>(), tiles, t.L.naux, t.L.nbf, G, ldg,
                       ng, out, ldo);
  HIP_CHECK(hipGetLastError());
}

void slab_gemm_f32(Ctx& c, DevBT& t, const float* data, const float* G, int ldg, int ng, float* out, int ldo,
                   int bn, bool f64_sums)
{
  if (bn != 0 && bn != 64 && bn != 128) throw std::runtime_error("slab_gemm_f32: column tile must be 64 or 128");
  if (f64_sums)
    slab_gemm_t<float, 1>(c, t, data, G, ldg, ng, out, ldo, bn);
  else
    slab_gemm_t<float>(c, t, data, G, ldg, ng, out, ldo, bn);
}

void slab_gemm(Ctx& c, DevBT& t, const double* G, int ldg, int ng, double* out, int ldo,
               double alpha)
{
  if (alpha != 1.0) throw std::runtime_error("slab_gemm: only alpha = 1 is supported");
  slab_gemm_t<double>(c, t, t.data64.as<double>(), G, ldg, ng, out, ldo);
}


void metric_create(Ctx& c, int naux, const double* V_host, Metric& m)
{
  HIP_CHECK(hipSetDevice(c.device));
  m.naux = naux;
  const size_t bytes = (size_t)naux * naux * sizeof(double);
  m.Linv.alloc(bytes);
  m.Vinv.alloc(bytes);
  double* A = m.Linv.as<double>();
  // upload, read back and compare before factorising (cheap next to the
  // factorisation, and the factor feeds every later step)
  std::vector<double> back((size_t)naux * naux);
  h2d(m.Linv.p, V_host, bytes, c.stream);
  d2h(back.data(), m.Linv.p, bytes, c.stream);
  if (std::memcmp(back.data(), V_host, bytes) != 0)
    throw std::runtime_error("metric upload does not reproduce the host matrix");
  const int h_info = chol_lower(c, A, naux, naux);
  if (h_info != 0)
    throw std::runtime_error("metric Cholesky failed (matrix not positive definite at pivot " +
                             std::to_string(h_info) + " of " + std::to_string(naux) + ")");
  hipLaunchKernelGGL(k_zero_upper_cm, dim3(nblk((int64_t)naux * naux, 256)), dim3(256), 0, c.stream, A, naux);
  trtri_lower(c, A, naux, naux);
  hipLaunchKernelGGL(k_zero_upper_cm, dim3(nblk((int64_t)naux * naux, 256)), dim3(256), 0, c.stream, A, naux);
  // V^-1 = L^-T L^-1: V^-1(i, j) = sum_{k >= max(i, j)} L^-1(k, i) L^-1(k, j), both
  // operands read down the columns of L^-1; exactly symmetric (the (i, j) and
  // (j, i) tiles sum the same products in the same order)
  dgemm_rm(c.stream, true, true, naux, naux, naux, A, naux, A, naux, m.Vinv.as<double>(), naux, 1.0, 0.0,
           DG_TRI_A_UPPER | DG_TRI_B_LOWER);
  HIP_CHECK(hipStreamSynchronize(c.stream));
}

// dst[x][r] = sum_{y <= x} L^-1(x, y) src[y][r] for naux rows x of R elements
// (L^-1 lower triangular, stored column-major): the rows of L^-1 B for a
// slab of R = bsz * ncol elements per aux function, or HT2 = L^-1 HT1 of the
// exchange (DFMO_K K2, jkdf.cpp:384).  Asynchronous.
void linv_apply(Ctx& c, const Metric& m, const double* src, int64_t R, double* dst)
{
  // dst = L^-1 src over x (src row-major naux x R): the column-major R x naux
  // view times (L^-1)^T, out of place, on rocBLAS's triangular multiply
  const double one = 1.0;
  RB_CHECK(rocblas_dtrmm_64(c.blas, rocblas_side_right, rocblas_fill_lower, rocblas_operation_transpose,
                            rocblas_diagonal_non_unit, R, m.naux, &one, m.Linv.as<double>(), m.naux, src, R, dst, R));
}

void transform_B(Ctx& c, DevBT& t, const Metric& m)
{
  if (t.transformed) throw std::runtime_error("B is already transformed");
  if (m.naux != t.L.naux) throw std::runtime_error("metric and B have different aux dimensions");
  int64_t maxslab = 0;
  for (int n = 0; n < t.L.nbb; ++n) maxslab = std::max<int64_t>(maxslab, t.L.slab_off[n + 1] - t.L.slab_off[n]);
  c.scratch_c.ensure((size_t)maxslab * sizeof(double));
  for (int n = 0; n < t.L.nbb; ++n) {
    const int64_t R = (int64_t)t.L.bsz[n] * t.L.ncol[n];
    if (R == 0) continue;
    double* slab = t.data64.as<double>() + t.L.slab_off[n];
    const size_t bytes = (size_t)R * t.L.naux * sizeof(double);
    HIP_CHECK(hipMemcpyAsync(c.scratch_c.p, slab, bytes, hipMemcpyDeviceToDevice, c.stream));
    // slab <- L^-1 slab over x (out of place, from the copy)
    linv_apply(c, m, c.scratch_c.as<double>(), R, slab);
  }
  HIP_CHECK(hipStreamSynchronize(c.stream));
  t.transformed = true;
}

// DF-J's chunking of the aux index for the column pass (k_j_cols): about
// 2e6 elements of the pair-gathered slabs per chunk, at most 64 chunks.
static void dfj_chunks(const DevBT& t, int* nch_out, int* xchunk_out)
{
  const int naux = t.L.naux;
  const int64_t pg = t.pg_total;
  int nch = (int)std::max<int64_t>(1, std::min<int64_t>(64, (2000000 + pg - 1) / std::max<int64_t>(pg, 1)));
  nch = std::max(1, std::min(nch, naux));
  const int xchunk = (naux + nch - 1) / nch;
  *nch_out = std::max(1, (naux + xchunk - 1) / std::max(1, xchunk));
  *xchunk_out = xchunk;
}

size_t dfj_workspace_bytes(const DevBT& t)
{
  int nch = 1, xchunk = 1;
  dfj_chunks(t, &nch, &xchunk);
  const size_t n_pg = (size_t)t.pg_total, n_gp = (size_t)t.L.nbb * t.L.naux, n_jp = (size_t)nch * t.pg_total;
  return (n_pg + 2 * n_gp + 2 * (size_t)t.L.naux + 2 * n_jp + n_pg) * sizeof(double);
}

void dfj(Ctx& c, DevBT& t, const Metric* m, const double* P, double* J, double pscale)
{
  c.scratch_c.ensure(dfj_workspace_bytes(t) + 256);
  double* ws = c.scratch_c.as<double>();
  int64_t ps = 0;
  std::memcpy(&ps, &pscale, sizeof(ps));
  const ChainKey key{4, pkey(P), pkey(J), ps, pkey(m ? m->Vinv.p : nullptr), pkey(ws), pkey(t.data64.p),
                     t.transformed ? 1 : 0, 0, 0, 0, 0, 0, 0};
  run_chain(c, t, key, [&] { dfj_ws(c, t, m, P, J, pscale, ws); });
}

void dfj_ws(Ctx& c, DevBT& t, const Metric* m, const double* P, double* J, double pscale, double* ws)
{
  PhaseTimer timer(c, PH_J);
  const BLayout& L = t.L;
  if (!t.transformed && !m) throw std::runtime_error("DF-J needs the metric for an untransformed B");
  const int naux = L.naux, nbf = L.nbf, nbb = L.nbb;
  const int64_t pg = t.pg_total;
  int nch = 1, xchunk = 1;
  dfj_chunks(t, &nch, &xchunk);
  // workspace: Pg | gpart, gpart_c | g | gq | Jpart, Jpart_c | Jg; every sum is
  // compensated (value + rounding-error pair), so J carries no summation-order
  // noise for V^-1 to amplify
  const size_t n_pg = (size_t)pg, n_gp = (size_t)nbb * naux, n_jp = (size_t)nch * pg;
  double* Pg = ws;
  double* gpart = Pg + n_pg;
  double* gpart_c = gpart + n_gp;
  double* g = gpart_c + n_gp;
  double* gq = g + naux;
  double* Jpart = gq + naux;
  double* Jpart_c = Jpart + n_jp;
  double* Jg = Jpart_c + n_jp;
  const unsigned gx = nblk(std::max<int64_t>(t.maxR, 1), 256);
  if (pg > 0) {
    hipLaunchKernelGGL(k_pair_gather, dim3(gx, nbb), dim3(256), 0, c.stream, P, nbf, t.ncol.as<int>(),
                       t.bsz.as<int>(), t.boff.as<int>(), t.colmap.as<int>(), t.colmap_off.as<int>(),
                       t.pg_off.as<int64_t>(), Pg);
    if (t.sym_mask)
      hipLaunchKernelGGL(k_j_gamma_sym, dim3(nblk(naux, J_ROWS), nbb), dim3(256), 0, c.stream,
                         t.data64.as<double>(), t.slab_off.as<int64_t>(), t.ncol.as<int>(), t.bsz.as<int>(),
                         t.cdn.as<int2>(), t.pg_off.as<int64_t>(), Pg, naux, gpart, gpart_c);
    else
      hipLaunchKernelGGL(k_j_gamma, dim3(nblk(naux, J_ROWS), nbb), dim3(256), 0, c.stream,
                         t.data64.as<double>(), t.slab_off.as<int64_t>(), t.ncol.as<int>(),
                         t.bsz.as<int>(), t.pg_off.as<int64_t>(), Pg, naux, gpart, gpart_c);
  }
  else {
    hipLaunchKernelGGL(k_fill, dim3(gs_blocks(2 * n_gp)), dim3(256), 0, c.stream, gpart, (int64_t)(2 * n_gp), 0.0);
  }
  hipLaunchKernelGGL(k_reduce_slabs, dim3(nblk(naux, 256)), dim3(256), 0, c.stream, gpart, gpart_c, nbb, naux,
                     pscale, g);
  if (t.transformed) {
    HIP_CHECK(hipMemcpyAsync(gq, g, naux * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
  }
  else {
    // gq = V^-1 g (V^-1 symmetric: its rows are its columns)
    hipLaunchKernelGGL(k_symv_dot2, dim3(nblk(naux, 4)), dim3(256), 0, c.stream, m->Vinv.as<double>(), g, naux, gq);
  }
  hipLaunchKernelGGL(k_fill, dim3(gs_blocks((int64_t)nbf * nbf)), dim3(256), 0, c.stream, J,
                     (int64_t)nbf * nbf, 0.0);
  if (pg > 0 && t.sym_mask) {
    // canonical half: the entries of Jpart past each slab's canonical columns
    // are never written; the reduction carries them along unread
    const unsigned gxc = nblk(std::max<int64_t>(t.maxRc, 1), 256);
    hipLaunchKernelGGL(k_j_cols_sym, dim3(gxc, nbb, nch), dim3(256), 0, c.stream, t.data64.as<double>(),
                       t.slab_off.as<int64_t>(), t.ncol.as<int>(), t.bsz.as<int>(), t.cdn.as<int2>(),
                       t.pg_off.as<int64_t>(), gq, naux, xchunk, pg, Jpart, Jpart_c);
    hipLaunchKernelGGL(k_reduce_chunks_dd, dim3(gs_blocks(pg)), dim3(256), 0, c.stream, Jpart, Jpart_c, nch, pg,
                       Jg);
    hipLaunchKernelGGL(k_pair_scatter_sym, dim3(gxc, nbb), dim3(256), 0, c.stream, Jg, nbf, t.ncol.as<int>(),
                       t.bsz.as<int>(), t.boff.as<int>(), t.cdn.as<int2>(), t.colmap.as<int>(),
                       t.colmap_off.as<int>(), t.pg_off.as<int64_t>(), J);
  }
  else if (pg > 0) {
    hipLaunchKernelGGL(k_j_cols, dim3(gx, nbb, nch), dim3(256), 0, c.stream, t.data64.as<double>(),
                       t.slab_off.as<int64_t>(), t.ncol.as<int>(), t.bsz.as<int>(), t.pg_off.as<int64_t>(),
                       gq, naux, xchunk, pg, Jpart, Jpart_c);
    hipLaunchKernelGGL(k_reduce_chunks_dd, dim3(gs_blocks(pg)), dim3(256), 0, c.stream, Jpart, Jpart_c, nch, pg,
                       Jg);
    hipLaunchKernelGGL(k_pair_scatter, dim3(gx, nbb), dim3(256), 0, c.stream, Jg, nbf, t.ncol.as<int>(),
                       t.bsz.as<int>(), t.boff.as<int>(), t.colmap.as<int>(), t.colmap_off.as<int>(),
                       t.pg_off.as<int64_t>(), J);
  }
  HIP_CHECK(hipGetLastError());
}

// FP64 direct-to-LDS exchange pieces (defined with the other precision forms below)
static void dfk_mo_tilde64_dl(Ctx& c, DevBT& t, const double* C, int ldc, int nocc, double* K, int occ_nbatches);
static void syrk_f64_exchange(Ctx& c, DevBT& t, const double* W, int ng, int ldw, int nb, int b, double* K);

void dfk_mo(Ctx& c, DevBT& t, const Metric* m, const double* C, int ldc, int nocc, double* K,
            int occ_nbatches, const double* tilde64)
{
  const BLayout& L = t.L;
  // tilde64: the padded FP64 copy of L^-1 B: K1 on it, no K2
  if (tilde64) {
    if (tilde64 != t.data64t.as<double>()) throw std::runtime_error("DFMO-K: tilde64 must be the tensor's FP64 copy");
    dfk_mo_tilde64_dl(c, t, C, ldc, nocc, K, occ_nbatches);
    return;
  }
  const bool pre = t.transformed;
  const double* bdata = t.data64.as<double>();
  if (!pre && !m) throw std::runtime_error("DFMO-K needs the metric for an untransformed B");
  const int naux = L.naux, nbf = L.nbf;
  if (nocc <= 0) {
    hipLaunchKernelGGL(k_fill, dim3(gs_blocks((int64_t)nbf * nbf)), dim3(256), 0, c.stream, K,
                       (int64_t)nbf * nbf, 0.0);
    return;
  }
  occ_nbatches = std::max(1, std::min(occ_nbatches, nocc));
  const int per = (nocc + occ_nbatches - 1) / occ_nbatches;
  const int nb = (nocc + per - 1) / per;
  // HT rows padded to whole 64-byte steps of the SYRK, zero past the batch width
  const int ldw = (per + 7) / 8 * 8;
  const size_t ht = (size_t)naux * nbf * ldw;
  c.scratch_a.ensure(ht * sizeof(double));
  if (!pre) c.scratch_b.ensure(ht * sizeof(double));
  for (int b = 0; b < nb; ++b) {
    const int i0 = b * per;
    const int w = std::min(per, nocc - i0);
    double* HT1 = c.scratch_a.as<double>();
    // K1: HT1[x][nu][i] = sum_mu B(x|mu nu) C(mu, i0+i)
    {
      PhaseTimer tm(c, PH_K1);
      if (ldw != w)
        hipLaunchKernelGGL(k_fill, dim3(gs_blocks((int64_t)ht)), dim3(256), 0, c.stream, HT1, (int64_t)ht, 0.0);
      slab_gemm_t<double>(c, t, bdata, C + i0, ldc, w, HT1, ldw);
    }
    const double* HT = HT1;
    if (!pre) {
      PhaseTimer tm(c, PH_K2);
      // K2: HT2 = L^-1 HT1 over x; column-major (W x naux) view times (L^-1)^T.
      // The plain path's metric step stays on rocBLAS's triangular multiply:
      // the vendor kernel (0.67 of the FP64 MFMA peak at 64 waters) is the
      // faster of the two, and this path is the FP64 reference the chosen
      // form is measured against.
      double* HT2 = c.scratch_b.as<double>();
      const int64_t W = (int64_t)nbf * ldw;
      const double one = 1.0;
      RB_CHECK(rocblas_dtrmm_64(c.blas, rocblas_side_right, rocblas_fill_lower,
                                rocblas_operation_transpose, rocblas_diagonal_non_unit, W, naux, &one,
                                m->Linv.as<double>(), naux, HT1, W, HT2, W));
      HT = HT2;
    }
    // K3: K = -sum_{x,i} HT[x][m][i] HT[x][n][i]
    PhaseTimer tm(c, PH_K3);
    syrk_f64_exchange(c, t, HT, w, ldw, nb, b, K);
  }
}

}  // namespace mchip

namespace mchip {

void slab_out_gemm(Ctx& c, DevBT& t, const double* A, int lda, int klen, const double* G, int ldg,
                   double* out)
{
  const BLayout& L = t.L;
  std::vector<Tile> v;
  for (int n = 0; n < L.nbb; ++n) {
    const int64_t rows = (int64_t)L.naux * L.bsz[n];
    for (int64_t m0 = 0; m0 < rows; m0 += SG_BM)
      for (int c0 = 0; c0 < L.ncol[n]; c0 += 128) v.push_back(Tile{n, (int)m0, c0, 0});
  }
  if (v.empty()) return;
  DevBuf tiles = upload_vec(v, c.stream);
  hipLaunchKernelGGL((k_slab_out_gemm<double, SG_BM, 128>), dim3((unsigned)v.size()), dim3(256), 0, c.stream,
                     A, lda, klen, t.slab_off.as<int64_t>(), t.ncol.as<int>(), t.bsz.as<int>(),
                     t.boff.as<int>(), t.colmap.as<int>(), t.colmap_off.as<int>(), tiles.as<Tile>(), L.naux,
                     L.nbf, G, ldg, out, nullptr, nullptr, (int64_t)0);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(c.stream));
}

void slab_syrk(Ctx& c, DevBT& t, const double* A, const double* B, double* Z, DevBuf& partial)
{
  const BLayout& L = t.L;
  const int n = L.naux;
  const int nt = (n + SG_BM - 1) / SG_BM;
  std::vector<int2> tl;
  for (int i = 0; i < nt; ++i)
    for (int j = i; j < nt; ++j) tl.push_back(int2{i, j});
  const int ntile = (int)tl.size();
  // chunks of slabs with roughly equal contraction length
  std::vector<int64_t> off(L.nbb), len(L.nbb);
  int64_t tot = 0;
  for (int s = 0; s < L.nbb; ++s) {
    off[s] = L.slab_off[s];
    len[s] = (int64_t)L.bsz[s] * L.ncol[s];
    tot += len[s];
  }
  const int want = std::max(1, std::min(L.nbb, 2048 / std::max(1, ntile)));
  std::vector<int> cs(1, 0);
  int64_t acc = 0;
  for (int s = 0; s < L.nbb; ++s) {
    acc += len[s];
    if ((int)cs.size() < want && acc * want >= tot * (int64_t)cs.size() && s + 1 < L.nbb) cs.push_back(s + 1);
  }
  cs.push_back(L.nbb);
  const int nch = (int)cs.size() - 1;
  partial.ensure((size_t)nch * n * n * sizeof(double));
  DevBuf d_tl = upload_vec(tl, c.stream), d_off = upload_vec(off, c.stream), d_len = upload_vec(len, c.stream),
         d_cs = upload_vec(cs, c.stream);
  hipLaunchKernelGGL((k_slabseg_gemm<double, SG_BM, SG_BM>), dim3(ntile * nch), dim3(256), 0, c.stream, A, B,
                     d_off.as<int64_t>(), d_len.as<int64_t>(), d_cs.as<int>(), n, d_tl.as<int2>(), ntile,
                     partial.as<double>());
  hipLaunchKernelGGL(k_reduce_sym<double>, dim3(nblk((int64_t)n * n, 256)), dim3(256), 0, c.stream, partial.as<double>(),
                     nch, n, 0, 1.0, Z);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(c.stream));
}

}  // namespace mchip

namespace mchip {

// ---------------------------------------------------------------------------
// BF16 split-precision exchange on the transformed B~
// ---------------------------------------------------------------------------

// Padded slab layout shared by the BF16 and FP32 copies of B~ (computed once
// per tensor): rows of slab N padded to ldp[N], a multiple of 32 elements, at
// element offset off16[N]; gathered coefficients of slab N at ldp_pre[N] * ng.
static void pad_layout(Ctx& c, DevBT& t)
{
  const BLayout& L = t.L;
  if (t.h_ldp_pre.size() == (size_t)L.nbb + 1) return;
  t.h_ldp.assign(L.nbb, 0);
  t.h_off16.assign(L.nbb + 1, 0);
  t.h_ldp_pre.assign(L.nbb + 1, 0);
  for (int n = 0; n < L.nbb; ++n) {
    t.h_ldp[n] = (L.ncol[n] + 31) / 32 * 32;
    t.h_off16[n + 1] = t.h_off16[n] + (int64_t)L.naux * L.bsz[n] * t.h_ldp[n];
    t.h_ldp_pre[n + 1] = t.h_ldp_pre[n] + t.h_ldp[n];
  }
  t.ldp = upload_vec(t.h_ldp, c.stream);
  t.off16 = upload_vec(t.h_off16, c.stream);
  t.ldp_pre = upload_vec(t.h_ldp_pre, c.stream);
}

static int64_t padded_elems(const DevBT& t) { return t.h_off16.empty() ? 0 : t.h_off16.back(); }

static void bf16_layout(Ctx& c, DevBT& t)
{
  pad_layout(c, t);
  const int64_t tot = padded_elems(t);
  t.data_hi.alloc((size_t)std::max<int64_t>(tot, 1) * sizeof(uint16_t));
  t.data_lo.alloc((size_t)std::max<int64_t>(tot, 1) * sizeof(uint16_t));
  t.bf16_plan.clear();
}

// The padded copies of B~ = L^-1 B from the canonical half of a symmetric
// tensor: per slab N, its columns of blocks M <= N (the leading ce columns,
// blocks in ascending order) packed into rows of length ce, transformed by
// L^-1 over x, written into slab N's copies, and the columns of each block
// M < N mirrored into slab M's columns for block N (B~(X|mu nu) = B~(X|nu mu)).
// Every column of every slab is written once: the non-canonical columns of
// slab M are the mirrors of the slabs of their (larger) blocks.  About half the
// triangular-multiply flops of the full transform.
static void tilde_slabs_sym(Ctx& c, DevBT& t, const Metric& m, float* out32, bool bf16, double* out64)
{
  const BLayout& L = t.L;
  const int naux = L.naux;
  // per slab: canonical width ce and the mirror segments of its blocks M < N
  std::vector<int> ce(L.nbb, 0), seg_off(L.nbb + 1, 0);
  std::vector<int4> seg;
  int64_t maxc = 1;
  for (int n = 0; n < L.nbb; ++n) {
    int e = L.ncol[n];
    for (int p = L.pair_ptr[n]; p < L.pair_ptr[n + 1]; ++p) {
      const int mb = L.pair_m[p];
      if (mb > n) {
        e = L.pair_col[p];
        break;
      }
      if (mb == n) {
        e = L.pair_col[p] + L.bsz[n];
        break;
      }
      // block mb < n: its mirror is slab mb's column block n
      int q = -1;
      for (int pp = L.pair_ptr[mb]; pp < L.pair_ptr[mb + 1]; ++pp)
        if (L.pair_m[pp] == n) q = L.pair_col[pp];
      if (q < 0) throw std::runtime_error("L^-1 B copies: the block mask is not symmetric");
      seg.push_back(make_int4(mb, L.pair_col[p], q, 0));
    }
    ce[n] = e;
    seg_off[n + 1] = (int)seg.size();
    maxc = std::max<int64_t>(maxc, (int64_t)naux * L.bsz[n] * e);
  }
  DevBuf packed((size_t)maxc * sizeof(double)), T((size_t)maxc * sizeof(double));
  DevBuf d_seg = upload_vec(seg.empty() ? std::vector<int4>(1) : seg, c.stream);
  uint16_t* hi = bf16 ? t.data_hi.as<uint16_t>() : nullptr;
  uint16_t* lo = bf16 ? t.data_lo.as<uint16_t>() : nullptr;
  for (int n = 0; n < L.nbb; ++n) {
    const int K = L.ncol[n], bN = L.bsz[n], e = ce[n];
    const int64_t rows = (int64_t)naux * bN;
    if (rows == 0 || K == 0) continue;
    if (e > 0) {
      const double* slab = t.data64.as<double>() + L.slab_off[n];
      HIP_CHECK(hipMemcpy2DAsync(packed.p, (size_t)e * sizeof(double), slab, (size_t)K * sizeof(double),
                                 (size_t)e * sizeof(double), (size_t)rows, hipMemcpyDeviceToDevice, c.stream));
      linv_apply(c, m, packed.as<double>(), (int64_t)bN * e, T.as<double>());
    }
    hipLaunchKernelGGL(k_tilde_canon, dim3(gs_blocks(rows * t.h_ldp[n])), dim3(256), 0, c.stream, T.as<double>(), n,
                       e, t.ncol.as<int>(), t.bsz.as<int>(), naux, t.ldp.as<int>(), t.off16.as<int64_t>(), out64,
                       out32, hi, lo);
    const int ns = seg_off[n + 1] - seg_off[n];
    if (ns > 0) {
      int64_t big = 0;
      for (int s = seg_off[n]; s < seg_off[n + 1]; ++s) big = std::max<int64_t>(big, (int64_t)L.bsz[seg[s].x] * bN);
      const unsigned gx = (unsigned)std::min<int64_t>(std::max<int64_t>(1, (int64_t)naux * big / 256), 256);
      hipLaunchKernelGGL(k_tilde_mirror, dim3(gx, ns), dim3(256), 0, c.stream, T.as<double>(), n, e,
                         d_seg.as<int4>() + seg_off[n], t.bsz.as<int>(), naux, t.ldp.as<int>(),
                         t.off16.as<int64_t>(), out64, out32, hi, lo);
    }
    HIP_CHECK(hipGetLastError());
  }
  HIP_CHECK(hipStreamSynchronize(c.stream));
}

// Low-precision copies of B~ = L^-1 B made slab by slab from the untransformed
// B, which stays in data64: J, the reported energy and every FP64 exchange
// keep the reference's formula (explicit V^-1 in J, L^-1 applied per build in
// DFMO), while the per-iteration exchange reads the FP32 and/or BF16 copies.
// Per slab: B_slab (L^-1)^T in FP64 (the slab's rows of L^-1 B) into a
// temporary, then its padded FP64 copy into out64 if not null, its FP32 copy
// into out32 if not null (in the padded layout if pad32, else in the slab
// layout of data64) and its bf16 hi/lo split into t's padded planes if bf16.
static void tilde_slabs(Ctx& c, DevBT& t, const Metric& m, float* out32, bool pad32, bool bf16,
                        double* out64 = nullptr)
{
  const BLayout& L = t.L;
  if (t.sym_mask && (pad32 || !out32)) {
    tilde_slabs_sym(c, t, m, out32, bf16, out64);
    return;
  }
  int64_t maxslab = 0;
  for (int n = 0; n < L.nbb; ++n) maxslab = std::max<int64_t>(maxslab, L.slab_off[n + 1] - L.slab_off[n]);
  DevBuf tmp((size_t)std::max<int64_t>(maxslab, 1) * sizeof(double));
  for (int n = 0; n < L.nbb; ++n) {
    const int64_t R = (int64_t)L.bsz[n] * L.ncol[n];
    if (R == 0) continue;
    const double* slab = t.data64.as<double>() + L.slab_off[n];
    double* dst = tmp.as<double>();
    linv_apply(c, m, slab, R, dst);
    const int64_t ne = R * L.naux;
    if (out64)
      hipLaunchKernelGGL(k_pad_f64_slab, dim3(gs_blocks((int64_t)L.naux * L.bsz[n] * t.h_ldp[n])), dim3(256), 0,
                         c.stream, dst, n, t.ncol.as<int>(), t.bsz.as<int>(), L.naux, t.ldp.as<int>(),
                         t.off16.as<int64_t>(), out64);
    if (out32 && pad32)
      hipLaunchKernelGGL(k_pad_f32_slab, dim3(gs_blocks((int64_t)L.naux * L.bsz[n] * t.h_ldp[n])), dim3(256), 0,
                         c.stream, dst, n, t.ncol.as<int>(), t.bsz.as<int>(), L.naux, t.ldp.as<int>(),
                         t.off16.as<int64_t>(), out32);
    else if (out32)
      hipLaunchKernelGGL(k_to_f32, dim3(gs_blocks(ne)), dim3(256), 0, c.stream, dst,
                         out32 + L.slab_off[n], ne);
    if (bf16)
      hipLaunchKernelGGL(k_split_slab, dim3(gs_blocks(ne)), dim3(256), 0, c.stream, dst, n,
                         t.ncol.as<int>(), t.bsz.as<int>(), L.naux, t.ldp.as<int>(), t.off16.as<int64_t>(),
                         t.data_hi.as<uint16_t>(), t.data_lo.as<uint16_t>());
    HIP_CHECK(hipGetLastError());
  }
  HIP_CHECK(hipStreamSynchronize(c.stream));
}

void dbt_make_lowp(Ctx& c, DevBT& t, const Metric& m, bool f32, bool bf16, bool f64)
{
  if (t.transformed) throw std::runtime_error("B is transformed: use the copies of the transformed tensor");
  if (m.naux != t.L.naux) throw std::runtime_error("metric and B have different aux dimensions");
  if (f32 || bf16 || f64) pad_layout(c, t);
  if (f32) t.data32.alloc((size_t)std::max<int64_t>(padded_elems(t), 1) * sizeof(float));
  if (bf16) bf16_layout(c, t);
  if (f64) t.data64t.alloc((size_t)std::max<int64_t>(padded_elems(t), 1) * sizeof(double));
  tilde_slabs(c, t, m, f32 ? t.data32.as<float>() : nullptr, true, bf16, f64 ? t.data64t.as<double>() : nullptr);
  if (f32) t.f32_plan.clear();
  t.f64_plan.clear();
  t.lowp_tilde = true;
}

void tilde_to_f32(Ctx& c, DevBT& t, const Metric& m, float* out)
{
  if (t.transformed) throw std::runtime_error("B is transformed: the FP32 copy is made from the untransformed B");
  if (m.naux != t.L.naux) throw std::runtime_error("metric and B have different aux dimensions");
  tilde_slabs(c, t, m, out, false, false);
}

void dbt_make_bf16(Ctx& c, DevBT& t)
{
  if (!t.transformed) throw std::runtime_error("the BF16 copy is made from the transformed B");
  const BLayout& L = t.L;
  bf16_layout(c, t);
  if (L.nbb > 0 && t.h_off16[L.nbb] > 0)
    hipLaunchKernelGGL(k_split_slabs, dim3(2048, L.nbb), dim3(256), 0, c.stream, t.data64.as<double>(),
                       t.slab_off.as<int64_t>(), t.ncol.as<int>(), t.bsz.as<int>(), L.naux,
                       t.ldp.as<int>(), t.off16.as<int64_t>(), t.data_hi.as<uint16_t>(),
                       t.data_lo.as<uint16_t>());
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(c.stream));
  t.lowp_tilde = true;
}

// Zero an exchange work array of the given byte count (a multiple of 8): the
// slab GEMMs write W only up to the column-tile edge past their width.
static void zero_w(Ctx& c, void* W, size_t bytes)
{
  const int64_t n = (int64_t)(bytes / sizeof(double));
  hipLaunchKernelGGL(k_fill, dim3(gs_blocks(n)), dim3(256), 0, c.stream, static_cast<double*>(W), n, 0.0);
  HIP_CHECK(hipGetLastError());
}

// Tile table of the direct-to-LDS slab GEMMs (row tile BM, column tile BN),
// laid out for the XCD mapping of the kernels (xcd_tile): the table is cut
// into 8 equal segments and segment x runs on XCD x.  Whole slabs are dealt to
// the segments heaviest first, each to the segment with the least work so far
// (a tile of slab N costs its ldp[N] / 32 contraction steps plus one step for
// its loads and stores), and the segments are padded with empty tiles (slab
// -1) to equal length.  The work per XCD is then balanced whatever the atom
// order, and a slab's tiles share one XCD's L2 for their coefficient rows.
static const Tile* bf16_tiles(Ctx& c, DevBT& t, int BM, int BN, int ng, int* ntiles)
{
  auto key = std::make_pair(-(BN * 1000 + BM), ng);
  auto it = t.tiles.find(key);
  if (it == t.tiles.end()) {
    const int nbb = t.L.nbb;
    std::vector<int64_t> ntl(nbb);
    std::vector<double> work(nbb);
    const int ncb = (ng + BN - 1) / BN;
    for (int n = 0; n < nbb; ++n) {
      const int64_t rows = (int64_t)t.L.naux * t.L.bsz[n];
      ntl[n] = (rows + BM - 1) / BM * ncb;
      const int ldp = t.h_ldp.size() == (size_t)nbb ? t.h_ldp[n] : (t.L.ncol[n] + 31) / 32 * 32;
      work[n] = (double)ntl[n] * (ldp / 32 + 1);
    }
    std::vector<int> order(nbb);
    for (int n = 0; n < nbb; ++n) order[n] = n;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return work[a] > work[b]; });
    std::vector<std::vector<Tile>> seg(8);
    double load[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int n : order) {
      int best = 0;
      for (int x = 1; x < 8; ++x)
        if (load[x] < load[best]) best = x;
      load[best] += work[n];
      // slabs without significant pairs still get tiles: they write the zero rows of W
      const int64_t rows = (int64_t)t.L.naux * t.L.bsz[n];
      for (int64_t m0 = 0; m0 < rows; m0 += BM)
        for (int n0 = 0; n0 < ng; n0 += BN) seg[best].push_back(Tile{n, (int)m0, n0, 0});
    }
    size_t len = 0;
    for (auto& v : seg) len = std::max(len, v.size());
    std::vector<Tile> v;
    v.reserve(8 * len);
    for (auto& sv : seg) {
      v.insert(v.end(), sv.begin(), sv.end());
      v.resize((v.size() + len - sv.size()), Tile{-1, 0, 0, 0});
    }
    if (v.size() >= ((size_t)1 << 31)) throw std::runtime_error("slab GEMM: too many tiles");
    t.ntiles[key] = (int)v.size();
    t.tiles[key] = upload_vec(v, c.stream);
    HIP_CHECK(hipStreamSynchronize(c.stream));
    it = t.tiles.find(key);
  }
  *ntiles = t.ntiles[key];
  return it->second.as<Tile>();
}

// Launch shapes of the BF16 slab GEMM: waves (WM x WN), MFMA tiles per wave
// (TM x TN), BM = 16 WM TM rows, BN = 16 WN TN columns.  DL = 0: register-
// staged loads, NB LDS buffers, FP32 W; DL = 1: direct-to-LDS loads, NB
// pipeline stages, W written as bf16 hi/lo; DL = 2: as 1 with the B~ rows in
// an NB-stage pipeline of their own and the gathered coefficients in two.
// The list keeps the shapes that were fastest for some width on MI355X.
struct Bf16Shape {
  int WM, WN, TM, TN, NB, DL;
  int bm() const { return 16 * WM * TM; }
  int bn() const { return 16 * WN * TN; }
  int threads() const { return 64 * WM * WN; }
};
static const Bf16Shape kBf16Shapes[] = {
    {2, 2, 4, 2, 2, 0}, {2, 4, 4, 2, 2, 0}, {2, 4, 4, 2, 4, 1}, {2, 4, 4, 3, 3, 1}, {2, 4, 4, 5, 2, 1},
    {2, 2, 8, 5, 3, 1}, {2, 2, 4, 5, 2, 1}, {2, 4, 4, 5, 4, 2}, {2, 4, 4, 5, 3, 2}, {2, 4, 4, 3, 4, 2},
    {2, 2, 4, 5, 2, 2}, {2, 4, 6, 5, 3, 2}, {2, 4, 6, 5, 2, 2}};
static constexpr int kNumBf16Shapes = sizeof(kBf16Shapes) / sizeof(kBf16Shapes[0]);
// SYRK variants: 0 = register-staged on FP32 W; 1..3 = direct-to-LDS on bf16
// hi/lo W with 2, 3, 4 stages
// 0 staged 128, 1-3 direct-to-LDS 128 (2-4 stages), 4 direct 256, 5 direct 256
// with the diagonal tiles loading both sides (every tile of a chunk equally long)
static constexpr int kNumSyrkVariants = 6;

static void slab_gemm_bf16(Ctx& c, DevBT& t, int shape, const uint16_t* ghi, const uint16_t* glo, int ng,
                           float* W, uint16_t* Whi, uint16_t* Wlo, int ldw)
{
  const Bf16Shape sh = kBf16Shapes[shape];
  int nt = 0;
  const Tile* tiles = bf16_tiles(c, t, sh.bm(), sh.bn(), ng, &nt);
  if (nt == 0) return;
#define MCHIP_BF16_SLAB(WM, WN, TM, TN, NB)                                                               \
  hipLaunchKernelGGL((k_slab_gemm_bf16x3<WM, WN, TM, TN, NB>), dim3(nt), dim3(64 * WM * WN), 0, c.stream, \
                     t.data_hi.as<uint16_t>(), t.data_lo.as<uint16_t>(), t.off16.as<int64_t>(),           \
                     t.ldp.as<int>(), t.ldp_pre.as<int64_t>(), t.bsz.as<int>(), t.boff.as<int>(), tiles,  \
                     t.L.naux, t.L.nbf, ghi, glo, ng, W, ldw)
#define MCHIP_BF16_SLAB_DL(WM, WN, TM, TN, S)                                                             \
  hipLaunchKernelGGL((k_slab_gemm_bf16x3_dl<WM, WN, TM, TN, S>), dim3(nt), dim3(64 * WM * WN), 0,         \
                     c.stream, t.data_hi.as<uint16_t>(), t.data_lo.as<uint16_t>(), t.off16.as<int64_t>(), \
                     t.ldp.as<int>(), t.ldp_pre.as<int64_t>(), t.bsz.as<int>(), t.boff.as<int>(), tiles,  \
                     t.L.naux, t.L.nbf, ghi, glo, ng, nullptr, Whi, Wlo, ldw)
#define MCHIP_BF16_SLAB_DLA(WM, WN, TM, TN, SA)                                                           \
  hipLaunchKernelGGL((k_slab_gemm_bf16x3_dla<WM, WN, TM, TN, SA>), dim3(nt), dim3(64 * WM * WN), 0,       \
                     c.stream, t.data_hi.as<uint16_t>(), t.data_lo.as<uint16_t>(), t.off16.as<int64_t>(), \
                     t.ldp.as<int>(), t.ldp_pre.as<int64_t>(), t.bsz.as<int>(), t.boff.as<int>(), tiles,  \
                     t.L.naux, t.L.nbf, ghi, glo, ng, Whi, Wlo, ldw)
  switch (shape) {
    case 0: MCHIP_BF16_SLAB(2, 2, 4, 2, 2); break;
    case 1: MCHIP_BF16_SLAB(2, 4, 4, 2, 2); break;
    case 2: MCHIP_BF16_SLAB_DL(2, 4, 4, 2, 4); break;
    case 3: MCHIP_BF16_SLAB_DL(2, 4, 4, 3, 3); break;
    case 4: MCHIP_BF16_SLAB_DL(2, 4, 4, 5, 2); break;
    case 5: MCHIP_BF16_SLAB_DL(2, 2, 8, 5, 3); break;
    case 6: MCHIP_BF16_SLAB_DL(2, 2, 4, 5, 2); break;
    case 7: MCHIP_BF16_SLAB_DLA(2, 4, 4, 5, 4); break;
    case 8: MCHIP_BF16_SLAB_DLA(2, 4, 4, 5, 3); break;
    case 9: MCHIP_BF16_SLAB_DLA(2, 4, 4, 3, 4); break;
    case 10: MCHIP_BF16_SLAB_DLA(2, 2, 4, 5, 2); break;
    case 11: MCHIP_BF16_SLAB_DLA(2, 4, 6, 5, 3); break;
    case 12: MCHIP_BF16_SLAB_DLA(2, 4, 6, 5, 2); break;
    default: throw std::runtime_error("BF16 slab GEMM: unknown launch shape");
  }
#undef MCHIP_BF16_SLAB
#undef MCHIP_BF16_SLAB_DL
#undef MCHIP_BF16_SLAB_DLA
  HIP_CHECK(hipGetLastError());
}

// Upper-triangle tile table of an n x n SYRK with BM x BM tiles (cached in t).
static const int2* syrk_tiles(Ctx& c, DevBT& t, int n, int BM, int* ntile)
{
  const auto key = std::make_pair(-1000000 - BM, n);
  const int nt = (n + BM - 1) / BM;
  *ntile = nt * (nt + 1) / 2;
  if (!t.tiles.count(key)) {
    std::vector<int2> tl;
    for (int i = 0; i < nt; ++i)
      for (int j = i; j < nt; ++j) tl.push_back(int2{i, j});
    t.tiles[key] = upload_vec(tl, c.stream);
    HIP_CHECK(hipStreamSynchronize(c.stream));
  }
  return t.tiles[key].as<int2>();
}

// tl/ntile: the 128 x 128 upper tiles; the 256-tile variant uses its own table.
static void seg_syrk_bf16(Ctx& c, DevBT& t, int variant, const float* W, const uint16_t* Whi, const uint16_t* Wlo,
                          int naux, int nbf, int ldw, const int2* tl, int ntile, int nch, int spc, float* partial)
{
  const dim3 grid((unsigned)(ntile * nch));
  const dim3 grid8((unsigned)((ntile * nch + 7) / 8 * 8));   // direct-to-LDS kernels: XCD mapping
  switch (variant) {
    case 0:
      hipLaunchKernelGGL((k_seg_syrk_bf16x3<SG_BM>), grid, dim3(256), 0, c.stream, W, naux, (int64_t)nbf * ldw, ldw,
                         nbf, tl, ntile, spc, partial);
      break;
    case 1:
      hipLaunchKernelGGL((k_seg_syrk_bf16x3_dl<2, 128, 4, true>), grid8, dim3(256), 0, c.stream, Whi, Wlo, naux,
                         (int64_t)nbf * ldw, ldw, nbf, tl, ntile, spc, partial);
      break;
    case 2:
      hipLaunchKernelGGL((k_seg_syrk_bf16x3_dl<3, 128, 4, true>), grid8, dim3(256), 0, c.stream, Whi, Wlo, naux,
                         (int64_t)nbf * ldw, ldw, nbf, tl, ntile, spc, partial);
      break;
    case 3:
      hipLaunchKernelGGL((k_seg_syrk_bf16x3_dl<4, 128, 4, true>), grid8, dim3(256), 0, c.stream, Whi, Wlo, naux,
                         (int64_t)nbf * ldw, ldw, nbf, tl, ntile, spc, partial);
      break;
    case 4: {
      int nt2 = 0;
      const int2* tl2 = syrk_tiles(c, t, nbf, 256, &nt2);
      hipLaunchKernelGGL((k_seg_syrk_bf16x3_dl<2, 256, 8, true>), dim3((unsigned)((nt2 * nch + 7) / 8 * 8)), dim3(512), 0,
                         c.stream, Whi,
                         Wlo, naux, (int64_t)nbf * ldw, ldw, nbf, tl2, nt2, spc, partial);
      break;
    }
    case 5: {
      int nt2 = 0;
      const int2* tl2 = syrk_tiles(c, t, nbf, 256, &nt2);
      hipLaunchKernelGGL((k_seg_syrk_bf16x3_dl<2, 256, 8, false>), dim3((unsigned)((nt2 * nch + 7) / 8 * 8)),
                         dim3(512), 0, c.stream, Whi, Wlo, naux, (int64_t)nbf * ldw, ldw, nbf, tl2, nt2, spc, partial);
      break;
    }
    default: throw std::runtime_error("BF16 SYRK: unknown variant");
  }
  HIP_CHECK(hipGetLastError());
}

// Autotuner of the BF16 exchange for ng occupied columns.  Every slab-GEMM
// shape whose column tile is not mostly empty and every SYRK variant is timed
// once on the live operands (after one warm launch); the pipeline (register-
// staged with FP32 W, or direct-to-LDS with bf16 hi/lo W) with the smallest
// sum of its best GEMM and best SYRK is kept per ng, with the timings.
template <class F>
static float time_launch(Ctx& c, F&& f)
{
  f();
  HIP_CHECK(hipEventRecord(c.ev0, c.stream));
  f();
  HIP_CHECK(hipEventRecord(c.ev1, c.stream));
  HIP_CHECK(hipEventSynchronize(c.ev1));
  float ms = 0.f;
  HIP_CHECK(hipEventElapsedTime(&ms, c.ev0, c.ev1));
  return ms;
}

static const ExchangePlan& tune_bf16(Ctx& c, DevBT& t, const uint16_t* ghi, const uint16_t* glo, int ng, float* W,
                                 uint16_t* Whi, uint16_t* Wlo, int ldw, const int2* tl, int ntile, int nch, int spc,
                                 float* partial)
{
  auto it = t.bf16_plan.find(ng);
  if (it != t.bf16_plan.end()) return it->second;
  ExchangePlan p;
  p.k1_ms.assign(kNumBf16Shapes, -1.f);
  p.k3_ms.assign(kNumSyrkVariants, -1.f);
  float best1[2] = {1e30f, 1e30f};
  int arg1[2] = {-1, -1};
  for (int v = 0; v < kNumBf16Shapes; ++v) {
    const int bn = kBf16Shapes[v].bn();
    if (bn > 64 && (ng + bn - 1) / bn * bn - ng >= bn / 2) continue;   // mostly empty tiles
    const float ms = time_launch(c, [&] { slab_gemm_bf16(c, t, v, ghi, glo, ng, W, Whi, Wlo, ldw); });
    p.k1_ms[v] = ms;
    const int dl = kBf16Shapes[v].DL ? 1 : 0;   // W as bf16 hi/lo: pairs with the direct-to-LDS SYRKs
    if (ms < best1[dl]) {
      best1[dl] = ms;
      arg1[dl] = v;
    }
  }
  float best3[2] = {1e30f, 1e30f};
  int arg3[2] = {-1, -1};
  for (int v = 0; v < kNumSyrkVariants; ++v) {
    const float ms =
        time_launch(c, [&] { seg_syrk_bf16(c, t, v, W, Whi, Wlo, t.L.naux, t.L.nbf, ldw, tl, ntile, nch, spc, partial); });
    p.k3_ms[v] = ms;
    const int dl = v > 0 ? 1 : 0;
    if (ms < best3[dl]) {
      best3[dl] = ms;
      arg3[dl] = v;
    }
  }
  const int fam = (arg1[1] >= 0 && best1[1] + best3[1] < best1[0] + best3[0]) || arg1[0] < 0 ? 1 : 0;
  p.k1 = arg1[fam];
  p.k3 = arg3[fam];
  if (p.k1 < 0 || p.k3 < 0) throw std::runtime_error("BF16 autotuner found no usable launch shape");
  return t.bf16_plan[ng] = p;
}

void dfk_bf16_plan(const DevBT& t, int ng, int* plan, float* ms, int* nms)
{
  auto it = t.bf16_plan.find(ng);
  for (int i = 0; i < 8; ++i) plan[i] = 0;
  if (nms) *nms = 0;
  if (it == t.bf16_plan.end()) return;
  const ExchangePlan& p = it->second;
  const Bf16Shape& sh = kBf16Shapes[p.k1];
  plan[0] = p.k1;
  plan[1] = sh.bm();
  plan[2] = sh.bn();
  plan[3] = sh.threads();
  plan[4] = sh.DL;
  plan[5] = sh.NB;
  plan[6] = p.k3;
  plan[7] = 1;
  if (ms && nms) {
    int k = 0;
    for (float v : p.k1_ms) ms[k++] = v;
    for (float v : p.k3_ms) ms[k++] = v;
    *nms = k;
  }
}

void dfk_bf16_force(DevBT& t, int ng, int k1, int k3)
{
  if (k1 < 0 || k1 >= kNumBf16Shapes || k3 < 0 || k3 >= kNumSyrkVariants)
    throw std::runtime_error("BF16 plan: shape or SYRK variant out of range");
  if ((kBf16Shapes[k1].DL != 0) != (k3 != 0))
    throw std::runtime_error("BF16 plan: direct-to-LDS GEMM shapes pair with SYRK variants 1-4, the others with 0");
  ExchangePlan p;
  p.k1 = k1;
  p.k3 = k3;
  p.k1_ms.assign(kNumBf16Shapes, -1.f);
  p.k3_ms.assign(kNumSyrkVariants, -1.f);
  t.bf16_plan[ng] = p;
}

int dfk_bf16_num_shapes(int* nsyrk)
{
  if (nsyrk) *nsyrk = kNumSyrkVariants;
  return kNumBf16Shapes;
}

int dfk_bf16_shape_dl(int k1) { return kBf16Shapes[k1].DL; }

int dfk_bf16_tuned_bn(const DevBT& t, int ng)
{
  auto it = t.bf16_plan.find(ng);
  return it == t.bf16_plan.end() ? 0
                                 : kBf16Shapes[it->second.k1].bn() * 10000 + kBf16Shapes[it->second.k1].threads();
}

// K = -sum_X W_X^T W_X with W = B~ C, both contractions BF16x3 on MFMA, FP32
// accumulation inside a chunk of aux functions, FP64 across chunks.
// ---------------------------------------------------------------------------
// Grouped BF16x3 exchange: the builds of several tensors in one slab-GEMM
// launch and one SYRK launch (the request loop's cross-request batching).
// Every output element sums the same products in the same order as in
// dfk_mo_bf16 (the same k steps of the slab GEMM, the same 32-aux chunks of
// the SYRK), so a grouped build equals the build alone bit for bit.
// ---------------------------------------------------------------------------

// SYRK chunks of 32 aux functions at most, as in dfk_mo_bf16.
static void bf16_chunks(int naux, int nbf, int* nch_out, int* spc_out)
{
  const int nt = (nbf + SG_BM - 1) / SG_BM;
  const int ntile = nt * (nt + 1) / 2;
  int nch = std::max(1, std::min(naux, 2048 / std::max(1, ntile)));
  int spc = (naux + nch - 1) / nch;
  if (spc > 32) spc = 32;
  *nch_out = (naux + spc - 1) / spc;
  *spc_out = spc;
}

void dfk_bf16_group_prepare(Ctx& c, DevBT& t, const double* C, int ldc, int nocc, ExGroupReq& r)
{
  const BLayout& L = t.L;
  if (!t.lowp_tilde || !t.data_hi.p || t.h_ldp.size() != (size_t)L.nbb)
    throw std::runtime_error("BF16 copy of L^-1 B missing (mchip_dbt_make_lowp or mchip_dbt_make_bf16)");
  if (nocc <= 0 || nocc > ldc) throw std::runtime_error("grouped BF16 exchange: nocc out of range");
  const int naux = L.naux, nbf = L.nbf;
  const int ldw = (nocc + 31) / 32 * 32;
  const size_t nW = (size_t)naux * nbf * ldw;
  const size_t nG = (size_t)t.h_ldp_pre[L.nbb] * nocc;
  int nch = 1, spc = 1;
  bf16_chunks(naux, nbf, &nch, &spc);
  t.gx_w.ensure(nW * 2 * sizeof(uint16_t) + 256);
  t.gx_g.ensure(2 * ((nG + 127) / 128 * 128) * sizeof(uint16_t) + 256);
  t.gx_part.ensure((size_t)nch * nbf * nbf * sizeof(float));
  r.bhi = t.data_hi.as<uint16_t>();
  r.blo = t.data_lo.as<uint16_t>();
  r.off16 = t.off16.as<int64_t>();
  r.ldp = t.ldp.as<int>();
  r.ldp_pre = t.ldp_pre.as<int64_t>();
  r.bsz = t.bsz.as<int>();
  r.boff = t.boff.as<int>();
  uint16_t* ghi = t.gx_g.as<uint16_t>();
  uint16_t* glo = ghi + (nG + 127) / 128 * 128;
  r.ghi = ghi;
  r.glo = glo;
  r.whi = t.gx_w.as<uint16_t>();
  r.wlo = r.whi + nW;
  r.partial = t.gx_part.as<float>();
  r.naux = naux;
  r.nbf = nbf;
  r.ng = nocc;
  r.ldw = ldw;
  r.spc = spc;
  r.pad = nch;
  int gmax = 0;
  for (int n = 0; n < L.nbb; ++n) gmax = std::max(gmax, t.h_ldp[n]);
  if (gmax > 0 && L.nbb > 0)
    hipLaunchKernelGGL(k_gather_ct, dim3(nblk((int64_t)nocc * gmax, 256), L.nbb), dim3(256), 0, c.stream, C, ldc,
                       nocc, t.ncol.as<int>(), t.ldp.as<int>(), t.ldp_pre.as<int64_t>(), t.colmap.as<int>(),
                       t.colmap_off.as<int>(), ghi, glo);
  HIP_CHECK(hipGetLastError());
}

// Shapes of the grouped launches: slab GEMM 128 x 320 (8 waves) for more than
// 160 occupied columns, 128 x 160 (4 waves) otherwise; SYRK 256-row tiles
// above 512 basis functions, 128-row tiles otherwise.
static int group_k1_kind(const ExGroupReq& r) { return r.ng > 160 ? 0 : 1; }
static int group_k3_kind(const ExGroupReq& r) { return r.nbf > 512 ? 0 : 1; }

void dfk_bf16_group_launch(hipStream_t s, const std::vector<DevBT*>& ts, const std::vector<ExGroupReq>& reqs,
                           const std::vector<double*>& K, GroupBufs& gb)
{
  const int nj = (int)reqs.size();
  if (nj == 0) return;
  if ((int)ts.size() != nj || (int)K.size() != nj) throw std::runtime_error("grouped BF16 exchange: job lists differ");
  gb.reqs.ensure(nj * sizeof(ExGroupReq));
  HIP_CHECK(hipMemcpyAsync(gb.reqs.p, reqs.data(), nj * sizeof(ExGroupReq), hipMemcpyHostToDevice, s));
  // slab-GEMM tiles of every build, per shape: whole slabs dealt to 8 XCD
  // segments heaviest first (as bf16_tiles does for one tensor)
  for (int kind = 0; kind < 2; ++kind) {
    const int BM = 128, BN = kind == 0 ? 320 : 160;
    struct SlabWork {
      int job, slab;
      double work;
    };
    std::vector<SlabWork> sw;
    for (int j = 0; j < nj; ++j) {
      if (group_k1_kind(reqs[j]) != kind) continue;
      const BLayout& L = ts[j]->L;
      const int ncb = (reqs[j].ng + BN - 1) / BN;
      for (int n = 0; n < L.nbb; ++n) {
        const int64_t rows = (int64_t)L.naux * L.bsz[n];
        sw.push_back({j, n, (double)((rows + BM - 1) / BM * ncb) * (ts[j]->h_ldp[n] / 32 + 1)});
      }
    }
    if (sw.empty()) continue;
    std::stable_sort(sw.begin(), sw.end(), [](const SlabWork& a, const SlabWork& b) { return a.work > b.work; });
    std::vector<std::vector<Tile>> seg(8);
    double load[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (const SlabWork& w : sw) {
      int best = 0;
      for (int x = 1; x < 8; ++x)
        if (load[x] < load[best]) best = x;
      load[best] += w.work;
      const BLayout& L = ts[w.job]->L;
      const int64_t rows = (int64_t)L.naux * L.bsz[w.slab];
      for (int64_t m0 = 0; m0 < rows; m0 += BM)
        for (int n0 = 0; n0 < reqs[w.job].ng; n0 += BN) seg[best].push_back(Tile{w.slab, (int)m0, n0, w.job});
    }
    size_t len = 0;
    for (auto& v : seg) len = std::max(len, v.size());
    std::vector<Tile> v;
    v.reserve(8 * len);
    for (auto& sv : seg) {
      v.insert(v.end(), sv.begin(), sv.end());
      v.resize(v.size() + len - sv.size(), Tile{-1, 0, 0, 0});
    }
    if (v.size() >= ((size_t)1 << 31)) throw std::runtime_error("grouped slab GEMM: too many tiles");
    DevBuf& tb = gb.tiles[kind];
    tb.ensure(v.size() * sizeof(Tile));
    HIP_CHECK(hipMemcpyAsync(tb.p, v.data(), v.size() * sizeof(Tile), hipMemcpyHostToDevice, s));
    if (kind == 0)
      hipLaunchKernelGGL((k_slab_gemm_bf16x3_dla_group<2, 4, 4, 5, 3>), dim3((unsigned)v.size()), dim3(512), 0, s,
                         gb.reqs.as<ExGroupReq>(), tb.as<Tile>());
    else
      hipLaunchKernelGGL((k_slab_gemm_bf16x3_dla_group<2, 2, 4, 5, 2>), dim3((unsigned)v.size()), dim3(256), 0, s,
                         gb.reqs.as<ExGroupReq>(), tb.as<Tile>());
    HIP_CHECK(hipGetLastError());
    // the host tables must outlive the asynchronous copies above
    HIP_CHECK(hipStreamSynchronize(s));
  }
  // SYRK items (build, chunk, tile row, tile column): the tiles of one chunk
  // consecutive in one XCD segment, chunks dealt to the least loaded segment
  for (int kind = 0; kind < 2; ++kind) {
    const int BM = kind == 0 ? 256 : 128;
    std::vector<std::vector<int4>> seg(8);
    double load[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    bool any = false;
    for (int j = 0; j < nj; ++j) {
      const ExGroupReq& r = reqs[j];
      if (group_k3_kind(r) != kind) continue;
      any = true;
      const int nt = (r.nbf + BM - 1) / BM;
      const double cost = (double)(nt * (nt + 1) / 2) * r.spc * (r.ldw / 32);
      for (int ch = 0; ch < r.pad; ++ch) {
        int best = 0;
        for (int x = 1; x < 8; ++x)
          if (load[x] < load[best]) best = x;
        load[best] += cost;
        for (int a = 0; a < nt; ++a)
          for (int b = a; b < nt; ++b) seg[best].push_back(int4{j, ch, a, b});
      }
    }
    if (!any) continue;
    size_t len = 0;
    for (auto& v : seg) len = std::max(len, v.size());
    std::vector<int4> v;
    v.reserve(8 * len);
    for (auto& sv : seg) {
      v.insert(v.end(), sv.begin(), sv.end());
      v.resize(v.size() + len - sv.size(), int4{-1, 0, 0, 0});
    }
    if (v.size() >= ((size_t)1 << 31)) throw std::runtime_error("grouped SYRK: too many items");
    DevBuf& ib = gb.items[kind];
    ib.ensure(v.size() * sizeof(int4));
    HIP_CHECK(hipMemcpyAsync(ib.p, v.data(), v.size() * sizeof(int4), hipMemcpyHostToDevice, s));
    if (kind == 0)
      hipLaunchKernelGGL((k_seg_syrk_bf16x3_dl_group<2, 256, 8>), dim3((unsigned)v.size()), dim3(512), 0, s,
                         gb.reqs.as<ExGroupReq>(), ib.as<int4>());
    else
      hipLaunchKernelGGL((k_seg_syrk_bf16x3_dl_group<2, 128, 4>), dim3((unsigned)v.size()), dim3(256), 0, s,
                         gb.reqs.as<ExGroupReq>(), ib.as<int4>());
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(s));
  }
  for (int j = 0; j < nj; ++j) {
    const int nbf = reqs[j].nbf;
    hipLaunchKernelGGL(k_reduce_sym<float>, dim3(nblk((int64_t)nbf * nbf, 256)), dim3(256), 0, s, reqs[j].partial,
                       reqs[j].pad, nbf, 0, -1.0, K[j]);
    HIP_CHECK(hipGetLastError());
  }
}

void dfk_mo_bf16(Ctx& c, DevBT& t, const double* C, int ldc, int nocc, double* K, int occ_nbatches)
{
  const BLayout& L = t.L;
  if (!t.lowp_tilde || !t.data_hi.p || t.h_ldp.size() != (size_t)L.nbb)
    throw std::runtime_error("BF16 copy of L^-1 B missing (mchip_dbt_make_lowp or mchip_dbt_make_bf16)");
  if (c.bf16_exchange && nocc > 0 && occ_nbatches <= 1) {
    c.bf16_exchange(c, t, C, ldc, nocc, K);
    return;
  }
  const int naux = L.naux, nbf = L.nbf;
  if (nocc <= 0) {
    hipLaunchKernelGGL(k_fill, dim3(gs_blocks((int64_t)nbf * nbf)), dim3(256), 0, c.stream, K,
                       (int64_t)nbf * nbf, 0.0);
    return;
  }
  occ_nbatches = std::max(1, std::min(occ_nbatches, nocc));
  const int per = (nocc + occ_nbatches - 1) / occ_nbatches;
  const int nb = (nocc + per - 1) / per;
  const int ldw = (per + 31) / 32 * 32;
  const size_t nW = (size_t)naux * nbf * ldw;
  const size_t nG = (size_t)t.h_ldp_pre[L.nbb] * per;
  c.scratch_a.ensure(nW * sizeof(float) + 2 * nG * sizeof(uint16_t) + 512);
  // W as FP32 or, in the same bytes, as bf16 hi and lo halves
  float* W = c.scratch_a.as<float>();
  uint16_t* Whi = c.scratch_a.as<uint16_t>();
  uint16_t* Wlo = Whi + nW;
  uint16_t* ghi = reinterpret_cast<uint16_t*>(W + nW);
  uint16_t* glo = ghi + (nG + 127) / 128 * 128;
  // upper 128 x 128 tiles of K (table cached with the tensor)
  const int nt = (nbf + SG_BM - 1) / SG_BM;
  const int ntile = nt * (nt + 1) / 2;
  const auto tkey = std::make_pair(0, -nbf);
  if (t.tiles.find(tkey) == t.tiles.end()) {
    std::vector<int2> tl;
    for (int i = 0; i < nt; ++i)
      for (int j = i; j < nt; ++j) tl.push_back(int2{i, j});
    t.tiles[tkey] = upload_vec(tl, c.stream);
    HIP_CHECK(hipStreamSynchronize(c.stream));
  }
  const int2* d_tl = t.tiles[tkey].as<int2>();
  // chunks of at most 32 aux functions: the FP32 sum inside a chunk stays short
  // (3 MFMA accumulations per 32 occupied columns per aux function)
  int nch = 1, spc = 1;
  bf16_chunks(naux, nbf, &nch, &spc);
  // chunk partials are the FP32 chunk sums, kept exactly in FP32
  c.scratch_partial.ensure((size_t)nb * nch * nbf * nbf * sizeof(float));
  int gmax = 0;
  for (int n = 0; n < L.nbb; ++n) gmax = std::max(gmax, t.h_ldp[n]);
  auto chain = [&] {
    for (int b = 0; b < nb; ++b) {
      const int i0 = b * per;
      const int w = std::min(per, nocc - i0);
      {
        PhaseTimer tm(c, PH_SPLIT);
        if (gmax > 0 && L.nbb > 0)
          hipLaunchKernelGGL(k_gather_ct, dim3(nblk((int64_t)w * gmax, 256), L.nbb), dim3(256), 0, c.stream,
                             C + i0, ldc, w, t.ncol.as<int>(), t.ldp.as<int>(), t.ldp_pre.as<int64_t>(),
                             t.colmap.as<int>(), t.colmap_off.as<int>(), ghi, glo);
        HIP_CHECK(hipGetLastError());
      }
      float* partial = c.scratch_partial.as<float>() + (size_t)b * nch * nbf * nbf;
      const ExchangePlan& plan = tune_bf16(c, t, ghi, glo, w, W, Whi, Wlo, ldw, d_tl, ntile, nch, spc, partial);
      {
        PhaseTimer tm(c, PH_K1);
        // a narrower last batch: its tiles stop at the first column-tile edge
        // past w, so the columns up to ldw still hold the previous batch
        if (w < per) zero_w(c, W, nW * sizeof(float));
        slab_gemm_bf16(c, t, plan.k1, ghi, glo, w, W, Whi, Wlo, ldw);
      }
      PhaseTimer tm(c, PH_K3);
      seg_syrk_bf16(c, t, plan.k3, W, Whi, Wlo, naux, nbf, ldw, d_tl, ntile, nch, spc, partial);
      if (b == nb - 1)
        hipLaunchKernelGGL(k_reduce_sym<float>, dim3(nblk((int64_t)nbf * nbf, 256)), dim3(256), 0, c.stream,
                           c.scratch_partial.as<float>(), nb * nch, nbf, 0, -1.0, K);
      HIP_CHECK(hipGetLastError());
    }
  };
  const int wl = nocc - (nb - 1) * per;   // width of the last batch
  const ChainKey key{2, pkey(C), ldc, nocc, pkey(K), occ_nbatches, pkey(c.scratch_a.p), pkey(c.scratch_partial.p),
                     pkey(t.data_hi.p), pkey(t.data_lo.p), plan_key(t.bf16_plan, per), plan_key(t.bf16_plan, wl),
                     0, 0};
  run_chain(c, t, key, chain);
}

// ---------------------------------------------------------------------------
// FP32 exchange on the padded FP32 copy of B~
// ---------------------------------------------------------------------------

void dbt_make_f32(Ctx& c, DevBT& t)
{
  if (!t.transformed) throw std::runtime_error("the FP32 copy is made from the transformed B");
  const BLayout& L = t.L;
  pad_layout(c, t);
  t.data32.alloc((size_t)std::max<int64_t>(padded_elems(t), 1) * sizeof(float));
  for (int n = 0; n < L.nbb; ++n) {
    const int64_t ne = (int64_t)L.naux * L.bsz[n] * t.h_ldp[n];
    if (ne == 0) continue;
    hipLaunchKernelGGL(k_pad_f32_slab, dim3(gs_blocks(ne)), dim3(256), 0, c.stream,
                       t.data64.as<double>() + L.slab_off[n], n, t.ncol.as<int>(), t.bsz.as<int>(), L.naux,
                       t.ldp.as<int>(), t.off16.as<int64_t>(), t.data32.as<float>());
    HIP_CHECK(hipGetLastError());
  }
  HIP_CHECK(hipStreamSynchronize(c.stream));
  t.f32_plan.clear();
  t.lowp_tilde = true;
}

// Padded FP32 copy (data32) of B~ from its FP32 values in the slab layout
// (tilde_to_f32's output): the operand of the direct-to-LDS FP32 slab GEMMs.
void dbt_make_f32_from(Ctx& c, DevBT& t, const float* b32)
{
  const BLayout& L = t.L;
  pad_layout(c, t);
  t.data32.alloc((size_t)std::max<int64_t>(padded_elems(t), 1) * sizeof(float));
  for (int n = 0; n < L.nbb; ++n) {
    const int64_t ne = (int64_t)L.naux * L.bsz[n] * t.h_ldp[n];
    if (ne == 0) continue;
    hipLaunchKernelGGL(k_pad_f32_rows, dim3(gs_blocks(ne)), dim3(256), 0, c.stream, b32 + L.slab_off[n], n,
                       t.ncol.as<int>(), t.bsz.as<int>(), L.naux, t.ldp.as<int>(), t.off16.as<int64_t>(),
                       t.data32.as<float>());
    HIP_CHECK(hipGetLastError());
  }
  t.f32_plan.clear();
  t.lowp_tilde = true;
}

int64_t dbt_padded_cols(const DevBT& t) { return t.h_ldp_pre.empty() ? 0 : t.h_ldp_pre.back(); }

// W = B~ C (W[X][nu][i], rows of ldo) on the padded FP32 copy with FP64 sums
// of every 16-product step; C is FP32 n x ldc (ng columns), gathered per slab
// into G (dbt_padded_cols(t) * ng floats).  Asynchronous; the tile table is
// made on the first call (synchronising) and cached.
void slab_gemm_f32_f64sum(Ctx& c, DevBT& t, const float* C, int ldc, int ng, float* G, float* out, int ldo,
                          int depth)
{
  const BLayout& L = t.L;
  if (!t.data32.p || t.h_ldp_pre.size() != (size_t)L.nbb + 1)
    throw std::runtime_error("slab_gemm_f32_f64sum: no padded FP32 copy");
  if (ng <= 0) return;
  int gmax = 0;
  for (int n = 0; n < L.nbb; ++n) gmax = std::max(gmax, t.h_ldp[n]);
  if (gmax > 0 && L.nbb > 0)
    hipLaunchKernelGGL(k_gather_ct_f32f, dim3(nblk((int64_t)ng * gmax, 256), L.nbb), dim3(256), 0, c.stream, C, ldc,
                       ng, t.ncol.as<int>(), t.ldp.as<int>(), t.ldp_pre.as<int64_t>(), t.colmap.as<int>(),
                       t.colmap_off.as<int>(), G);
  int nt = 0;
  const Tile* tiles = bf16_tiles(c, t, 128, 128, ng, &nt);
  if (nt == 0) return;
  if (depth == 2)
    hipLaunchKernelGGL(k_slab_gemm_f32_dl_f64<2>, dim3(nt), dim3(512), 0, c.stream, t.data32.as<float>(),
                       t.off16.as<int64_t>(), t.ldp.as<int>(), t.ldp_pre.as<int64_t>(), t.bsz.as<int>(),
                       t.boff.as<int>(), tiles, L.naux, L.nbf, G, ng, out, ldo);
  else
    hipLaunchKernelGGL(k_slab_gemm_f32_dl_f64<4>, dim3(nt), dim3(512), 0, c.stream, t.data32.as<float>(),
                       t.off16.as<int64_t>(), t.ldp.as<int>(), t.ldp_pre.as<int64_t>(), t.bsz.as<int>(),
                       t.boff.as<int>(), tiles, L.naux, L.nbf, G, ng, out, ldo);
  HIP_CHECK(hipGetLastError());
}

// Launch shapes of the FP32 and FP64 slab GEMMs (direct-to-LDS): waves
// WM x WN, MFMA tiles per wave TM x TN, S pipeline stages.
struct DlShape {
  int WM, WN, TM, TN, S;
  int bm() const { return 16 * WM * TM; }
  int bn() const { return 16 * WN * TN; }
  int threads() const { return 64 * WM * WN; }
};
static const DlShape kDlShapes[] = {{2, 2, 4, 2, 3}, {2, 4, 4, 2, 3}, {2, 2, 4, 5, 3},
                                      {2, 4, 4, 5, 3}, {2, 4, 6, 5, 3}, {2, 4, 4, 5, 2}};
static constexpr int kNumDlShapes = sizeof(kDlShapes) / sizeof(kDlShapes[0]);
static constexpr int kNumF32Syrk = 3;   // direct-to-LDS SYRK with 2, 3, 4 stages

static void slab_gemm_f32dl(Ctx& c, DevBT& t, int shape, const float* G, int ng, float* W, int ldw)
{
  const DlShape sh = kDlShapes[shape];
  int nt = 0;
  const Tile* tiles = bf16_tiles(c, t, sh.bm(), sh.bn(), ng, &nt);
  if (nt == 0) return;
#define MCHIP_F32_SLAB(WM, WN, TM, TN, S)                                                                     \
  hipLaunchKernelGGL((k_slab_gemm_f32_dl<WM, WN, TM, TN, S>), dim3(nt), dim3(64 * WM * WN), 0, c.stream,      \
                     t.data32.as<float>(), t.off16.as<int64_t>(), t.ldp.as<int>(), t.ldp_pre.as<int64_t>(),   \
                     t.bsz.as<int>(), t.boff.as<int>(), tiles, t.L.naux, t.L.nbf, G, ng, W, ldw)
  switch (shape) {
    case 0: MCHIP_F32_SLAB(2, 2, 4, 2, 3); break;
    case 1: MCHIP_F32_SLAB(2, 4, 4, 2, 3); break;
    case 2: MCHIP_F32_SLAB(2, 2, 4, 5, 3); break;
    case 3: MCHIP_F32_SLAB(2, 4, 4, 5, 3); break;
    case 4: MCHIP_F32_SLAB(2, 4, 6, 5, 3); break;
    case 5: MCHIP_F32_SLAB(2, 4, 4, 5, 2); break;
    default: throw std::runtime_error("FP32 slab GEMM: unknown launch shape");
  }
#undef MCHIP_F32_SLAB
  HIP_CHECK(hipGetLastError());
}

static void seg_syrk_f32dl(Ctx& c, int variant, const float* W, int naux, int nbf, int ldw, const int2* tl,
                           int ntile, int nch, int spc, double* partial)
{
  const dim3 grid8((unsigned)((ntile * nch + 7) / 8 * 8));
  const int64_t ss = (int64_t)nbf * ldw;
  switch (variant) {
    case 0:
      hipLaunchKernelGGL(k_seg_syrk_f32_dl<2>, grid8, dim3(512), 0, c.stream, W, naux, ss, ldw, nbf, tl, ntile, spc,
                         partial);
      break;
    case 1:
      hipLaunchKernelGGL(k_seg_syrk_f32_dl<3>, grid8, dim3(512), 0, c.stream, W, naux, ss, ldw, nbf, tl, ntile, spc,
                         partial);
      break;
    case 2:
      hipLaunchKernelGGL(k_seg_syrk_f32_dl<4>, grid8, dim3(512), 0, c.stream, W, naux, ss, ldw, nbf, tl, ntile, spc,
                         partial);
      break;
    default: throw std::runtime_error("FP32 SYRK: unknown variant");
  }
  HIP_CHECK(hipGetLastError());
}

// Autotuner of the FP32 exchange for ng occupied columns: every slab-GEMM shape
// whose column tile is not mostly empty and every SYRK variant is timed once
// on the live operands (after one warm launch); the fastest of each is kept
// per ng, with the timings.
static const ExchangePlan& tune_f32(Ctx& c, DevBT& t, const float* G, int ng, float* W, int ldw, const int2* tl,
                                int ntile, int nch, int spc, double* partial)
{
  auto it = t.f32_plan.find(ng);
  if (it != t.f32_plan.end()) return it->second;
  ExchangePlan p;
  p.k1_ms.assign(kNumDlShapes, -1.f);
  p.k3_ms.assign(kNumF32Syrk, -1.f);
  float best = 1e30f;
  for (int v = 0; v < kNumDlShapes; ++v) {
    const int bn = kDlShapes[v].bn();
    if (bn > 64 && (ng + bn - 1) / bn * bn - ng >= bn / 2) continue;   // mostly empty tiles
    const float ms = time_launch(c, [&] { slab_gemm_f32dl(c, t, v, G, ng, W, ldw); });
    p.k1_ms[v] = ms;
    if (ms < best) {
      best = ms;
      p.k1 = v;
    }
  }
  best = 1e30f;
  for (int v = 0; v < kNumF32Syrk; ++v) {
    const float ms = time_launch(c, [&] { seg_syrk_f32dl(c, v, W, t.L.naux, t.L.nbf, ldw, tl, ntile, nch, spc, partial); });
    p.k3_ms[v] = ms;
    if (ms < best) {
      best = ms;
      p.k3 = v;
    }
  }
  if (p.k1 < 0 || p.k3 < 0) throw std::runtime_error("FP32 autotuner found no usable launch shape");
  return t.f32_plan[ng] = p;
}

void dfk_f32_plan(const DevBT& t, int ng, int* plan, float* ms, int* nms)
{
  auto it = t.f32_plan.find(ng);
  for (int i = 0; i < 8; ++i) plan[i] = 0;
  if (nms) *nms = 0;
  if (it == t.f32_plan.end()) return;
  const ExchangePlan& p = it->second;
  const DlShape& sh = kDlShapes[p.k1];
  plan[0] = p.k1;
  plan[1] = sh.bm();
  plan[2] = sh.bn();
  plan[3] = sh.threads();
  plan[4] = 1;
  plan[5] = sh.S;
  plan[6] = p.k3;
  plan[7] = 1;
  if (ms && nms) {
    int k = 0;
    for (float v : p.k1_ms) ms[k++] = v;
    for (float v : p.k3_ms) ms[k++] = v;
    *nms = k;
  }
}

void dfk_f32_force(DevBT& t, int ng, int k1, int k3)
{
  if (k1 < 0 || k1 >= kNumDlShapes || k3 < 0 || k3 >= kNumF32Syrk)
    throw std::runtime_error("FP32 plan: shape or SYRK variant out of range");
  ExchangePlan p;
  p.k1 = k1;
  p.k3 = k3;
  p.k1_ms.assign(kNumDlShapes, -1.f);
  p.k3_ms.assign(kNumF32Syrk, -1.f);
  t.f32_plan[ng] = p;
}

int dfk_f32_num_shapes(int* nsyrk)
{
  if (nsyrk) *nsyrk = kNumF32Syrk;
  return kNumDlShapes;
}

// K = -sum_X W_X^T W_X with W = B~ C (DFMO_K::compute_K, jkdf.cpp:221-462, on L^-1 B): both
// contractions in FP32 on MFMA; the SYRK's FP32 sums are flushed into FP64
// every 256 products and its aux chunks are summed in FP64.
void dfk_mo_f32(Ctx& c, DevBT& t, const double* C, int ldc, int nocc, double* K, int occ_nbatches)
{
  const BLayout& L = t.L;
  if (!t.lowp_tilde || (!t.data32.p && L.total) || t.h_ldp_pre.size() != (size_t)L.nbb + 1)
    throw std::runtime_error("FP32 copy of L^-1 B missing (mchip_dbt_make_lowp or mchip_dbt_make_f32)");
  const int naux = L.naux, nbf = L.nbf;
  if (nocc <= 0) {
    hipLaunchKernelGGL(k_fill, dim3(gs_blocks((int64_t)nbf * nbf)), dim3(256), 0, c.stream, K,
                       (int64_t)nbf * nbf, 0.0);
    return;
  }
  occ_nbatches = std::max(1, std::min(occ_nbatches, nocc));
  const int per = (nocc + occ_nbatches - 1) / occ_nbatches;
  const int nb = (nocc + per - 1) / per;
  const int ldw = (per + 15) / 16 * 16;   // W rows of whole 64-byte steps, zero past the batch width
  const size_t nW = (size_t)naux * nbf * ldw;
  const size_t nG = (size_t)t.h_ldp_pre[L.nbb] * per;
  c.scratch_a.ensure((nW + nG) * sizeof(float) + 256);
  float* W = c.scratch_a.as<float>();
  float* G = W + nW;   // 64-byte aligned: nW is a multiple of 16
  int ntile = 0;
  const int2* tl = syrk_tiles(c, t, nbf, 128, &ntile);
  int nch = std::max(1, std::min(naux, 2048 / std::max(1, ntile)));
  const int spc = (naux + nch - 1) / nch;
  nch = (naux + spc - 1) / spc;
  c.scratch_partial.ensure((size_t)nb * nch * nbf * nbf * sizeof(double));
  int gmax = 0;
  for (int n = 0; n < L.nbb; ++n) gmax = std::max(gmax, t.h_ldp[n]);
  auto chain = [&] {
    for (int b = 0; b < nb; ++b) {
      const int i0 = b * per;
      const int w = std::min(per, nocc - i0);
      {
        PhaseTimer tm(c, PH_SPLIT);
        if (gmax > 0 && L.nbb > 0)
          hipLaunchKernelGGL(k_gather_ct_f32, dim3(nblk((int64_t)w * gmax, 256), L.nbb), dim3(256), 0, c.stream,
                             C + i0, ldc, w, t.ncol.as<int>(), t.ldp.as<int>(), t.ldp_pre.as<int64_t>(),
                             t.colmap.as<int>(), t.colmap_off.as<int>(), G);
        HIP_CHECK(hipGetLastError());
      }
      double* partial = c.scratch_partial.as<double>() + (size_t)b * nch * nbf * nbf;
      const ExchangePlan& plan = tune_f32(c, t, G, w, W, ldw, tl, ntile, nch, spc, partial);
      {
        PhaseTimer tm(c, PH_K1);
        if (w < per) zero_w(c, W, nW * sizeof(float));   // see dfk_mo_bf16
        slab_gemm_f32dl(c, t, plan.k1, G, w, W, ldw);
      }
      PhaseTimer tm(c, PH_K3);
      seg_syrk_f32dl(c, plan.k3, W, naux, nbf, ldw, tl, ntile, nch, spc, partial);
      if (b == nb - 1)
        hipLaunchKernelGGL(k_reduce_sym<double>, dim3(nblk((int64_t)nbf * nbf, 256)), dim3(256), 0, c.stream,
                           c.scratch_partial.as<double>(), nb * nch, nbf, 0, -1.0, K);
      HIP_CHECK(hipGetLastError());
    }
  };
  const int wl = nocc - (nb - 1) * per;
  const ChainKey key{3, pkey(C), ldc, nocc, pkey(K), occ_nbatches, pkey(c.scratch_a.p), pkey(c.scratch_partial.p),
                     pkey(t.data32.p), 0, plan_key(t.f32_plan, per), plan_key(t.f32_plan, wl), 0, 0};
  run_chain(c, t, key, chain);
}

// ---------------------------------------------------------------------------
// FP64 exchange pieces on padded operands (direct-to-LDS)
// ---------------------------------------------------------------------------

// Launch shapes of the FP64 slab GEMM: waves WM x WN, MFMA tiles per wave
// TM x TN, S pipeline stages.
static const DlShape kF64Shapes[] = {{2, 2, 4, 2, 3}, {2, 4, 4, 2, 3}, {2, 2, 4, 5, 3}, {2, 4, 4, 5, 3},
                                      {2, 4, 4, 5, 2}};
static constexpr int kNumF64Shapes = sizeof(kF64Shapes) / sizeof(kF64Shapes[0]);
static constexpr int kNumF64Syrk = 3;   // direct-to-LDS SYRK with 2, 3, 4 stages

static void slab_gemm_f64dl(Ctx& c, DevBT& t, int shape, const double* G, int ng, double* W, int ldw)
{
  int nt = 0;
  const Tile* tiles = bf16_tiles(c, t, kF64Shapes[shape].bm(), kF64Shapes[shape].bn(), ng, &nt);
  if (nt == 0) return;
#define MCHIP_F64_SLAB(WM, WN, TM, TN, S)                                                                     \
  hipLaunchKernelGGL((k_slab_gemm_f64_dl<WM, WN, TM, TN, S>), dim3(nt), dim3(64 * WM * WN), 0, c.stream,      \
                     t.data64t.as<double>(), t.off16.as<int64_t>(), t.ldp.as<int>(), t.ldp_pre.as<int64_t>(), \
                     t.bsz.as<int>(), t.boff.as<int>(), tiles, t.L.naux, t.L.nbf, G, ng, W, ldw)
  switch (shape) {
    case 0: MCHIP_F64_SLAB(2, 2, 4, 2, 3); break;
    case 1: MCHIP_F64_SLAB(2, 4, 4, 2, 3); break;
    case 2: MCHIP_F64_SLAB(2, 2, 4, 5, 3); break;
    case 3: MCHIP_F64_SLAB(2, 4, 4, 5, 3); break;
    case 4: MCHIP_F64_SLAB(2, 4, 4, 5, 2); break;
    default: throw std::runtime_error("FP64 slab GEMM: unknown launch shape");
  }
#undef MCHIP_F64_SLAB
  HIP_CHECK(hipGetLastError());
}

static void seg_syrk_f64dl(Ctx& c, int variant, const double* W, int naux, int nbf, int ldw, const int2* tl,
                           int ntile, int nch, int spc, double* partial, const double* W2 = nullptr)
{
  const dim3 grid8((unsigned)((ntile * nch + 7) / 8 * 8));
  const int64_t ss = (int64_t)nbf * ldw;
  if (!W2) W2 = W;
  switch (variant) {
    case 0:
      hipLaunchKernelGGL(k_seg_syrk_f64_dl<2>, grid8, dim3(512), 0, c.stream, W, W2, naux, ss, ldw, nbf, tl, ntile,
                         spc, partial);
      break;
    case 1:
      hipLaunchKernelGGL(k_seg_syrk_f64_dl<3>, grid8, dim3(512), 0, c.stream, W, W2, naux, ss, ldw, nbf, tl, ntile,
                         spc, partial);
      break;
    case 2:
      hipLaunchKernelGGL(k_seg_syrk_f64_dl<4>, grid8, dim3(512), 0, c.stream, W, W2, naux, ss, ldw, nbf, tl, ntile,
                         spc, partial);
      break;
    default: throw std::runtime_error("FP64 SYRK: unknown variant");
  }
  HIP_CHECK(hipGetLastError());
}

// FP64 plan entry for ng columns (k1 / k3 = -1 until tuned or forced).
static ExchangePlan& f64_plan_entry(DevBT& t, int ng)
{
  auto it = t.f64_plan.find(ng);
  if (it != t.f64_plan.end()) return it->second;
  ExchangePlan p;
  p.k1_ms.assign(kNumF64Shapes, -1.f);
  p.k3_ms.assign(kNumF64Syrk, -1.f);
  return t.f64_plan[ng] = p;
}

// K3 of an FP64 exchange batch: partial sums of -W_X W_X^T over aux chunks on
// the upper 128 x 128 tiles (FP64 MFMA), reduced into K after the last batch.
// The SYRK variant is timed once per ng on the live operand (autotuner).
// Aux chunks of the FP64 exchange SYRK: nch chunks of spc aux functions.
static void f64_syrk_chunks(int naux, int ntile, int* nch_out, int* spc_out)
{
  int nch = std::max(1, std::min(naux, 2048 / std::max(1, ntile)));
  const int spc = (naux + nch - 1) / nch;
  *nch_out = (naux + spc - 1) / spc;
  *spc_out = spc;
}

// Bytes of the FP64 SYRK's chunk partials for nb occupied batches (128-row tiles).
static size_t f64_partial_bytes(const DevBT& t, int nb)
{
  const int nt = (t.L.nbf + 127) / 128;
  int nch = 1, spc = 1;
  f64_syrk_chunks(t.L.naux, nt * (nt + 1) / 2, &nch, &spc);
  return (size_t)nb * nch * t.L.nbf * t.L.nbf * sizeof(double);
}

static void syrk_f64_exchange(Ctx& c, DevBT& t, const double* W, int ng, int ldw, int nb, int b, double* K)
{
  const int naux = t.L.naux, nbf = t.L.nbf;
  int ntile = 0;
  const int2* tl = syrk_tiles(c, t, nbf, 128, &ntile);
  int nch = 1, spc = 1;
  f64_syrk_chunks(naux, ntile, &nch, &spc);
  if (b == 0) c.scratch_partial.ensure((size_t)nb * nch * nbf * nbf * sizeof(double));
  double* partial = c.scratch_partial.as<double>() + (size_t)b * nch * nbf * nbf;
  ExchangePlan& p = f64_plan_entry(t, ng);
  if (p.k3 < 0) {
    float best = 1e30f;
    for (int v = 0; v < kNumF64Syrk; ++v) {
      const float ms = time_launch(c, [&] { seg_syrk_f64dl(c, v, W, naux, nbf, ldw, tl, ntile, nch, spc, partial); });
      p.k3_ms[v] = ms;
      if (ms < best) {
        best = ms;
        p.k3 = v;
      }
    }
  }
  seg_syrk_f64dl(c, p.k3, W, naux, nbf, ldw, tl, ntile, nch, spc, partial);
  if (b == nb - 1)
    hipLaunchKernelGGL(k_reduce_sym<double>, dim3(nblk((int64_t)nbf * nbf, 256)), dim3(256), 0, c.stream,
                       c.scratch_partial.as<double>(), nb * nch, nbf, 0, -1.0, K);
  HIP_CHECK(hipGetLastError());
}

// DFMO-K (DFMO_K::compute_K, jkdf.cpp:221-462) on the padded FP64 copy of
// L^-1 B: K1 = direct-to-LDS FP64 slab GEMM with the gathered occupied
// coefficients, no K2, K3 = FP64 SYRK.  The slab-GEMM shape is timed once per
// ng on the live operands (autotuner).
static void dfk_mo_tilde64_dl(Ctx& c, DevBT& t, const double* C, int ldc, int nocc, double* K, int occ_nbatches)
{
  const BLayout& L = t.L;
  if (!t.data64t.p || t.h_ldp_pre.size() != (size_t)L.nbb + 1)
    throw std::runtime_error("FP64 copy of L^-1 B missing (mchip_dbt_make_lowp with f64)");
  const int naux = L.naux, nbf = L.nbf;
  if (nocc <= 0) {
    hipLaunchKernelGGL(k_fill, dim3(gs_blocks((int64_t)nbf * nbf)), dim3(256), 0, c.stream, K,
                       (int64_t)nbf * nbf, 0.0);
    return;
  }
  occ_nbatches = std::max(1, std::min(occ_nbatches, nocc));
  const int per = (nocc + occ_nbatches - 1) / occ_nbatches;
  const int nb = (nocc + per - 1) / per;
  const int ldw = (per + 7) / 8 * 8;
  const size_t nW = (size_t)naux * nbf * ldw;
  const size_t nG = (size_t)t.h_ldp_pre[L.nbb] * per;
  c.scratch_a.ensure((nW + nG) * sizeof(double) + 256);
  double* W = c.scratch_a.as<double>();
  double* G = W + nW;   // 64-byte aligned: nW is a multiple of 8
  int gmax = 0;
  for (int n = 0; n < L.nbb; ++n) gmax = std::max(gmax, t.h_ldp[n]);
  c.scratch_partial.ensure(f64_partial_bytes(t, nb));   // the SYRK's chunk partials (before the key)
  auto chain = [&] {
    for (int b = 0; b < nb; ++b) {
      const int i0 = b * per;
      const int w = std::min(per, nocc - i0);
      {
        PhaseTimer tm(c, PH_SPLIT);
        if (gmax > 0 && L.nbb > 0)
          hipLaunchKernelGGL(k_gather_ct_f64, dim3(nblk((int64_t)w * gmax, 256), L.nbb), dim3(256), 0, c.stream,
                             C + i0, ldc, w, t.ncol.as<int>(), t.ldp.as<int>(), t.ldp_pre.as<int64_t>(),
                             t.colmap.as<int>(), t.colmap_off.as<int>(), G);
        HIP_CHECK(hipGetLastError());
      }
      ExchangePlan& p = f64_plan_entry(t, w);
      if (p.k1 < 0) {
        float best = 1e30f;
        for (int v = 0; v < kNumF64Shapes; ++v) {
          const int bn = kF64Shapes[v].bn();
          if (bn > 64 && (w + bn - 1) / bn * bn - w >= bn / 2) continue;   // mostly empty tiles
          const float ms = time_launch(c, [&] { slab_gemm_f64dl(c, t, v, G, w, W, ldw); });
          p.k1_ms[v] = ms;
          if (ms < best) {
            best = ms;
            p.k1 = v;
          }
        }
      }
      {
        PhaseTimer tm(c, PH_K1);
        if (w < per) zero_w(c, W, nW * sizeof(double));   // see dfk_mo_bf16
        slab_gemm_f64dl(c, t, p.k1, G, w, W, ldw);
      }
      PhaseTimer tm(c, PH_K3);
      syrk_f64_exchange(c, t, W, w, ldw, nb, b, K);
    }
  };
  const int wl = nocc - (nb - 1) * per;
  const ChainKey key{5, pkey(C), ldc, nocc, pkey(K), occ_nbatches, pkey(c.scratch_a.p), pkey(c.scratch_partial.p),
                     pkey(t.data64t.p), 0, plan_key(t.f64_plan, per), plan_key(t.f64_plan, wl), 0, 0};
  run_chain(c, t, key, chain);
}

// tr(K P) of the FP64 exchange on the FP64 copy of L^-1 B at the occupied C
// (P = C C^T), without K: W = B~ C by the copy's slab GEMM (the final
// exchange's K1, same plan), then -sum_X ||W_X^T C||^2 over the upper 128 x
// 128 tiles of each nocc x nocc W_X^T C (k_exchange_energy), summed
// compensated in a fixed order.  Host result; synchronises the stream.
double dfk_energy_tilde64(Ctx& c, DevBT& t, const double* C, int ldc, int nocc)
{
  const BLayout& L = t.L;
  if (!t.data64t.p || t.h_ldp_pre.size() != (size_t)L.nbb + 1)
    throw std::runtime_error("FP64 copy of L^-1 B missing (mchip_dbt_make_lowp with f64)");
  if (nocc <= 0) return 0.0;
  if (nocc > ldc) throw std::runtime_error("exchange energy: nocc exceeds the coefficients' columns");
  const int naux = L.naux, nbf = L.nbf;
  const int ldw = (nocc + 7) / 8 * 8;
  const size_t nW = (size_t)naux * nbf * ldw;
  const size_t nG = (size_t)t.h_ldp_pre[L.nbb] * nocc;
  c.scratch_a.ensure((nW + nG) * sizeof(double) + 256);
  double* W = c.scratch_a.as<double>();
  double* G = W + nW;
  int gmax = 0;
  for (int n = 0; n < L.nbb; ++n) gmax = std::max(gmax, t.h_ldp[n]);
  if (gmax > 0 && L.nbb > 0)
    hipLaunchKernelGGL(k_gather_ct_f64, dim3(nblk((int64_t)nocc * gmax, 256), L.nbb), dim3(256), 0, c.stream, C,
                       ldc, nocc, t.ncol.as<int>(), t.ldp.as<int>(), t.ldp_pre.as<int64_t>(), t.colmap.as<int>(),
                       t.colmap_off.as<int>(), G);
  HIP_CHECK(hipGetLastError());
  ExchangePlan& p = f64_plan_entry(t, nocc);
  if (p.k1 < 0) {
    float best = 1e30f;
    for (int v = 0; v < kNumF64Shapes; ++v) {
      const int bn = kF64Shapes[v].bn();
      if (bn > 64 && (nocc + bn - 1) / bn * bn - nocc >= bn / 2) continue;   // mostly empty tiles
      const float ms = time_launch(c, [&] { slab_gemm_f64dl(c, t, v, G, nocc, W, ldw); });
      p.k1_ms[v] = ms;
      if (ms < best) {
        best = ms;
        p.k1 = v;
      }
    }
  }
  slab_gemm_f64dl(c, t, p.k1, G, nocc, W, ldw);
  // upper tiles of the nocc x nocc products (cached in t)
  const int nt = (nocc + 127) / 128;
  const int ntile = nt * (nt + 1) / 2;
  const auto key = std::make_pair(-3000000, nocc);
  if (!t.tiles.count(key)) {
    std::vector<int2> tl;
    for (int i = 0; i < nt; ++i)
      for (int j = i; j < nt; ++j) tl.push_back(int2{i, j});
    t.tiles[key] = upload_vec(tl, c.stream);
    HIP_CHECK(hipStreamSynchronize(c.stream));
  }
  const int64_t nitem = (int64_t)naux * ntile;
  if (nitem >= ((int64_t)1 << 31)) throw std::runtime_error("exchange energy: too many work items");
  c.scratch_b.ensure((size_t)(2 * nitem + 8) * sizeof(double));
  double* part = c.scratch_b.as<double>();
  double* out = part + 2 * nitem;
  hipLaunchKernelGGL(k_exchange_energy, dim3((unsigned)((nitem + 7) / 8 * 8)), dim3(256), 0, c.stream, W, ldw, nbf,
                     nocc, C, ldc, naux, t.tiles[key].as<int2>(), ntile, part);
  HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, c.stream, part, nitem, -1.0, out);
  HIP_CHECK(hipGetLastError());
  double e = 0.0;
  d2h(&e, out, sizeof(double), c.stream);
  return e;
}

// tr(K P) of the plain FP64 exchange (untransformed B, metric step), without
// K.  Per occupied batch i0 <= i < i0 + w:
//   W[x][nu][i]  = sum_mu B(x|mu nu) C(mu, i0 + i)      (K1 of the plain path)
//   M[x][j][i]   = sum_nu C(nu, j) W[x][nu][i]          (one strided-batched DGEMM)
//   Mt           = L^-1 M over x                        (the K2 triangular multiply
//                                                        on nocc instead of nbf rows)
//   E           -= sum Mt^2                             (compensated, fixed order)
// The metric step then costs naux^2 nocc^2 instead of naux^2 nbf nocc flops.
// Work: W (naux nbf ldw) and M (naux nocc ldw) in scratch_a, Mt in scratch_b,
// within the plain exchange's HT1 + HT2 of the same batch width.
double dfk_energy_metric(Ctx& c, DevBT& t, const Metric& m, const double* C, int ldc, int nocc, int occ_nbatches)
{
  const BLayout& L = t.L;
  if (t.transformed) throw std::runtime_error("exchange energy: B is transformed (use the FP64 copy form)");
  if (m.naux != L.naux) throw std::runtime_error("exchange energy: metric and B have different aux dimensions");
  if (nocc <= 0) return 0.0;
  if (nocc > ldc) throw std::runtime_error("exchange energy: nocc exceeds the coefficients' columns");
  const int naux = L.naux, nbf = L.nbf;
  occ_nbatches = std::max(1, std::min(occ_nbatches, nocc));
  const int per = (nocc + occ_nbatches - 1) / occ_nbatches;
  const int nb = (nocc + per - 1) / per;
  const int ldw = (per + 7) / 8 * 8;
  const size_t nW = (size_t)naux * nbf * ldw, nM = (size_t)naux * nocc * ldw;
  c.scratch_a.ensure((nW + nM) * sizeof(double));
  c.scratch_b.ensure(nM * sizeof(double));
  constexpr int NPART = 1024;
  c.scratch_partial.ensure((size_t)(2 * NPART * nb + 8) * sizeof(double));
  double* W = c.scratch_a.as<double>();
  double* M = W + nW;
  double* Mt = c.scratch_b.as<double>();
  double* part = c.scratch_partial.as<double>();
  double* out = part + 2 * NPART * nb;
  const double one = 1.0, zero = 0.0;
  for (int b = 0; b < nb; ++b) {
    const int i0 = b * per;
    const int w = std::min(per, nocc - i0);
    if (ldw != w)
      hipLaunchKernelGGL(k_fill, dim3(gs_blocks((int64_t)nW)), dim3(256), 0, c.stream, W, (int64_t)nW, 0.0);
    slab_gemm_t<double>(c, t, t.data64.as<double>(), C + i0, ldc, w, W, ldw);
    // column-major per x: M_x (ldw x nocc) = W_x (ldw x nbf) * C^T (nbf x nocc),
    // C^T being the column-major view of row-major C (ldc x nbf)
    RB_CHECK(rocblas_dgemm_strided_batched(c.blas, rocblas_operation_none, rocblas_operation_transpose, ldw, nocc,
                                           nbf, &one, W, ldw, (rocblas_stride)nbf * ldw, C, ldc, 0, &zero, M, ldw,
                                           (rocblas_stride)nocc * ldw, naux));
    // Mt = L^-1 M over x: column-major (nocc ldw x naux) view times (L^-1)^T
    RB_CHECK(rocblas_dtrmm_64(c.blas, rocblas_side_right, rocblas_fill_lower, rocblas_operation_transpose,
                              rocblas_diagonal_non_unit, (int64_t)nocc * ldw, naux, &one, m.Linv.as<double>(), naux,
                              M, (int64_t)nocc * ldw, Mt, (int64_t)nocc * ldw));
    hipLaunchKernelGGL(k_sumsq_pairs, dim3(NPART), dim3(256), 0, c.stream, Mt, (int64_t)nM,
                       part + (size_t)2 * NPART * b);
    HIP_CHECK(hipGetLastError());
  }
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, c.stream, part, (int64_t)NPART * nb, -1.0, out);
  HIP_CHECK(hipGetLastError());
  double e = 0.0;
  d2h(&e, out, sizeof(double), c.stream);
  return e;
}

// Full 128 x 128 tile table of an n x n product (cached in t).
static const int2* full_tiles(Ctx& c, DevBT& t, int n, int* ntile)
{
  const auto key = std::make_pair(-2000000, n);
  const int nt = (n + 127) / 128;
  *ntile = nt * nt;
  if (!t.tiles.count(key)) {
    std::vector<int2> tl;
    for (int i = 0; i < nt; ++i)
      for (int j = 0; j < nt; ++j) tl.push_back(int2{i, j});
    t.tiles[key] = upload_vec(tl, c.stream);
    HIP_CHECK(hipStreamSynchronize(c.stream));
  }
  return t.tiles[key].as<int2>();
}

// The two-sided exchange of DFLMO_K (jkdf.cpp:1093-1331, set_sym(false)):
// K(m, n) = -sum_{X,Y,i} (B u)(X|m i) V^-1(X, Y) (B v^T)(Y|n i), with
// V^-1 = L^-T L^-1 it is -sum_{Z,i} Wu(Z, m, i) Wv(Z, n, i) for Wu = L^-1 B u and
// Wv = L^-1 B v^T.  U and Vt are device arrays nbf x ldu / nbf x ldv (row-major,
// first nocc columns; Vt is v transposed).  FP64 throughout: on the FP64 copy
// of L^-1 B when fp64_copy (no metric step), else on B with L^-1 applied to
// each half-transformed tensor (as K2 of the plain DFMO path).  K is nbf x nbf
// row-major and not symmetric.
void dfk_uv(Ctx& c, DevBT& t, const Metric* m, const double* U, int ldu, const double* Vt, int ldv, int nocc,
            double* K, int occ_nbatches, bool fp64_copy)
{
  const BLayout& L = t.L;
  const int naux = L.naux, nbf = L.nbf;
  if (fp64_copy && (!t.data64t.p || t.h_ldp_pre.size() != (size_t)L.nbb + 1))
    throw std::runtime_error("two-sided exchange: FP64 copy of L^-1 B missing (mchip_dbt_make_lowp with f64)");
  if (!fp64_copy && !t.transformed && !m)
    throw std::runtime_error("two-sided exchange: the metric is needed for an untransformed B");
  if (nocc <= 0) {
    hipLaunchKernelGGL(k_fill, dim3(gs_blocks((int64_t)nbf * nbf)), dim3(256), 0, c.stream, K,
                       (int64_t)nbf * nbf, 0.0);
    return;
  }
  occ_nbatches = std::max(1, std::min(occ_nbatches, nocc));
  const int per = (nocc + occ_nbatches - 1) / occ_nbatches;
  const int nb = (nocc + per - 1) / per;
  const int ldw = (per + 7) / 8 * 8;
  const size_t nW = (size_t)naux * nbf * ldw;
  const bool metric_step = !fp64_copy && !t.transformed;
  // Wu | Wv (| half-transformed staging for the metric step)
  c.scratch_a.ensure(nW * sizeof(double) * (metric_step ? 3 : 2) + 256);
  double* Wu = c.scratch_a.as<double>();
  double* Wv = Wu + nW;
  double* HT1 = Wv + nW;
  size_t nG = 0;
  if (fp64_copy) nG = (size_t)t.h_ldp_pre[L.nbb] * per;
  c.scratch_b.ensure(std::max<size_t>(nG, 1) * sizeof(double));
  double* G = c.scratch_b.as<double>();
  int gmax = 0;
  for (int n = 0; n < (int)t.h_ldp.size(); ++n) gmax = std::max(gmax, t.h_ldp[n]);
  int ntile = 0;
  const int2* tl = full_tiles(c, t, nbf, &ntile);
  int nch = std::max(1, std::min(naux, 2048 / std::max(1, ntile)));
  const int spc = (naux + nch - 1) / nch;
  nch = (naux + spc - 1) / spc;
  c.scratch_partial.ensure((size_t)nb * nch * nbf * nbf * sizeof(double));
  // W = (L^-1) B C for one batch of columns of C into out (zero past w)
  auto half = [&](const double* C, int ldc, int i0, int w, double* out) {
    zero_w(c, out, nW * sizeof(double));
    if (fp64_copy) {
      if (gmax > 0 && L.nbb > 0)
        hipLaunchKernelGGL(k_gather_ct_f64, dim3(nblk((int64_t)w * gmax, 256), L.nbb), dim3(256), 0, c.stream,
                           C + i0, ldc, w, t.ncol.as<int>(), t.ldp.as<int>(), t.ldp_pre.as<int64_t>(),
                           t.colmap.as<int>(), t.colmap_off.as<int>(), G);
      ExchangePlan& p = f64_plan_entry(t, w);
      if (p.k1 < 0) p.k1 = 0;
      slab_gemm_f64dl(c, t, p.k1, G, w, out, ldw);
    }
    else if (metric_step) {
      zero_w(c, HT1, nW * sizeof(double));
      slab_gemm_t<double>(c, t, t.data64.as<double>(), C + i0, ldc, w, HT1, ldw);
      // L^-1 over x, as the plain DFMO path's K2
      const double one = 1.0;
      RB_CHECK(rocblas_dtrmm_64(c.blas, rocblas_side_right, rocblas_fill_lower, rocblas_operation_transpose,
                                rocblas_diagonal_non_unit, (int64_t)nbf * ldw, naux, &one, m->Linv.as<double>(), naux,
                                HT1, (int64_t)nbf * ldw, out, (int64_t)nbf * ldw));
    }
    else {
      slab_gemm_t<double>(c, t, t.data64.as<double>(), C + i0, ldc, w, out, ldw);
    }
    HIP_CHECK(hipGetLastError());
  };
  for (int b = 0; b < nb; ++b) {
    const int i0 = b * per;
    const int w = std::min(per, nocc - i0);
    half(U, ldu, i0, w, Wu);
    half(Vt, ldv, i0, w, Wv);
    double* partial = c.scratch_partial.as<double>() + (size_t)b * nch * nbf * nbf;
    seg_syrk_f64dl(c, 0, Wu, naux, nbf, ldw, tl, ntile, nch, spc, partial, Wv);
  }
  hipLaunchKernelGGL(k_reduce_full, dim3(gs_blocks((int64_t)nbf * nbf)), dim3(256), 0, c.stream,
                     c.scratch_partial.as<double>(), nb * nch, (int64_t)nbf * nbf, -1.0, K);
  HIP_CHECK(hipGetLastError());
}

void dfk_f64_plan(const DevBT& t, int ng, int* plan, float* ms, int* nms)
{
  auto it = t.f64_plan.find(ng);
  for (int i = 0; i < 8; ++i) plan[i] = 0;
  if (nms) *nms = 0;
  if (it == t.f64_plan.end()) return;
  const ExchangePlan& p = it->second;
  plan[0] = p.k1;
  if (p.k1 >= 0) {
    const DlShape& sh = kF64Shapes[p.k1];
    plan[1] = sh.bm();
    plan[2] = sh.bn();
    plan[3] = sh.threads();
    plan[5] = sh.S;
  }
  plan[4] = 1;
  plan[6] = p.k3;
  plan[7] = 1;
  if (ms && nms) {
    int k = 0;
    for (float v : p.k1_ms) ms[k++] = v;
    for (float v : p.k3_ms) ms[k++] = v;
    *nms = k;
  }
}

void dfk_f64_force(DevBT& t, int ng, int k1, int k3)
{
  if (k1 < 0 || k1 >= kNumF64Shapes || k3 < 0 || k3 >= kNumF64Syrk)
    throw std::runtime_error("FP64 plan: shape or SYRK variant out of range");
  ExchangePlan& p = f64_plan_entry(t, ng);
  p.k1 = k1;
  p.k3 = k3;
}

int dfk_f64_num_shapes(int* nsyrk)
{
  if (nsyrk) *nsyrk = kNumF64Syrk;
  return kNumF64Shapes;
}

}  // namespace mchip

namespace mchip {

// Exchange from a density (DFAO_K, jkdf.cpp:510-648, and DFMEM_K,
// jkdf.cpp:719-859): K = -sum_{X,Y} B(X|m l) V^-1(X,Y) P(l,s) B(Y|s n).  K is
// linear in P, so any factorisation P = C C^T turns it into the
// occupied-factor exchange with C, term by term.  A positive semidefinite P
// is factorised on the GPU by pivoted Cholesky; an indefinite one (a
// difference density, say: the pivoted Cholesky stops at a negative pivot)
// by its eigenvectors, P = C+ C+^T - C- C-^T with C+- = V+- |w+-|^1/2, and
// K = K(C+) - K(C-); eigenvalues below n eps max|w| are dropped.  form: 0
// FP64, 1 FP32, 2 BF16x3 (the low-precision forms need the transformed B and
// its copy).
void dfk_ao(Ctx& c, DevBT& t, const Metric* m, const double* P_dev, double* K, int occ_nbatches, int form)
{
  const int n = t.L.nbf;
  auto exchange = [&](const double* C, int r, double* Kout) {
    switch (form) {
      case 0: dfk_mo(c, t, t.transformed ? nullptr : m, C, r, r, Kout, occ_nbatches); break;
      case 1: dfk_mo_f32(c, t, C, r, r, Kout, occ_nbatches); break;
      case 2: dfk_mo_bf16(c, t, C, r, r, Kout, occ_nbatches); break;
      default: throw std::runtime_error("dfk_ao: unknown precision form");
    }
  };
  if (form < 0 || form > 2) throw std::runtime_error("dfk_ao: unknown precision form");
  const size_t nn = (size_t)n * n;
  DevBuf U(nn * sizeof(double));
  HIP_CHECK(hipMemcpyAsync(U.p, P_dev, U.bytes, hipMemcpyDeviceToDevice, c.stream));
  int rank = 0;
  bool neg = false;
  std::vector<double> hL = pivoted_cholesky_dev(c, U, n, 0.0, &rank, &neg);
  if (!neg) {
    DevBuf L((size_t)n * std::max(rank, 1) * sizeof(double));
    if (rank > 0) h2d(L.p, hL.data(), hL.size() * sizeof(double), c.stream);
    exchange(L.as<double>(), rank, K);
    HIP_CHECK(hipStreamSynchronize(c.stream));   // L and the host factor die here
    return;
  }
  // indefinite: eigenvectors of P (column-major view of the symmetric
  // row-major P is P itself)
  DevBuf W(nn * sizeof(double)), w((size_t)n * sizeof(double)), ework((size_t)n * sizeof(double));
  HIP_CHECK(hipMemcpyAsync(W.p, P_dev, W.bytes, hipMemcpyDeviceToDevice, c.stream));
  RB_CHECK(rocsolver_dsyevd(c.blas, rocblas_evect_original, rocblas_fill_upper, n, W.as<double>(), n,
                            w.as<double>(), ework.as<double>(), c.info.as<int>()));
  std::vector<double> hw(n), hV(nn);
  int info = 0;
  d2h(&info, c.info.p, sizeof(int), c.stream);
  if (info != 0) throw std::runtime_error("dfk_ao: eigendecomposition of the density did not converge");
  d2h(hw.data(), w.p, hw.size() * sizeof(double), c.stream);
  d2h(hV.data(), W.p, hV.size() * sizeof(double), c.stream);
  double wmax = 0.0;
  for (double v : hw) wmax = std::max(wmax, std::fabs(v));
  const double tol = n * 2.220446049250313e-16 * wmax;
  std::vector<int> ip, in;
  for (int k = 0; k < n; ++k) {
    if (hw[k] > tol) ip.push_back(k);
    else if (hw[k] < -tol) in.push_back(k);
  }
  auto factor = [&](const std::vector<int>& cols) {   // row-major n x r, column j = V_k |w_k|^1/2
    const int r = (int)cols.size();
    std::vector<double> C((size_t)n * std::max(r, 1), 0.0);
    for (int j = 0; j < r; ++j) {
      const double f = std::sqrt(std::fabs(hw[cols[j]]));
      for (int i = 0; i < n; ++i) C[(size_t)i * r + j] = hV[(size_t)cols[j] * n + i] * f;
    }
    return C;
  };
  const std::vector<double> Cp = factor(ip), Cn = factor(in);
  const int rp = (int)ip.size(), rn = (int)in.size();
  DevBuf dCp(Cp.size() * sizeof(double)), dCn(Cn.size() * sizeof(double)), Kn(nn * sizeof(double));
  h2d(dCp.p, Cp.data(), Cp.size() * sizeof(double), c.stream);
  h2d(dCn.p, Cn.data(), Cn.size() * sizeof(double), c.stream);
  exchange(dCp.as<double>(), rp, K);
  exchange(dCn.as<double>(), rn, Kn.as<double>());
  hipLaunchKernelGGL(k_axpby, dim3(gs_blocks((int64_t)nn)), dim3(256), 0, c.stream, K, Kn.as<double>(),
                     (int64_t)nn, -1.0, 1.0);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(c.stream));
}

}  // namespace mchip
