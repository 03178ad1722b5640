// Request loop serving DF-HF (SCF) requests on one GPU.
//
// A request is a molecule (AO and aux basis, point charges, atom blocks).
// nslots worker threads each own a device context (stream, rocBLAS handle)
// and take requests in arrival order.  A worker prepares the host side of its
// request (one-electron integrals, Coulomb metric, Schwarz block mask and an
// HBM estimate), waits until the estimate fits the HBM budget next to the
// requests already resident (admission is in arrival order: a request is
// admitted only when every earlier one has been), then builds B on the GPU
// and runs the SCF with B, the metric factors and the densities resident in
// HBM until it converges.  Requests on different workers overlap on the
// device through their streams.  Besides the budget, admission checks the
// device's free memory, so an estimate that falls short cannot oversubscribe
// HBM.
//
// The path per request is the one of hf::hfmod::compute with the DF
// builders (src/hf/hfscf.cpp:326-552, src/fock/jkdf.cpp).  The loop itself has
// no counterpart in the reference (which runs one molecule per process).
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../ints/btensor.hpp"
#include "../ints/ints_host.hpp"
#include "dfock.hpp"
#include "ints12_dev.hpp"
#include "ints_dev.hpp"
#include "scf.hpp"
#include "serve.hpp"

namespace mchip {

static double now_s()
{
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// Frobenius norms of a shell-level quantity over blocks of consecutive shells.
static std::vector<double> block_norms_2d(const std::vector<double>& Q, int nsh, const std::vector<int>& blk)
{
  const int nb = (int)blk.size();
  std::vector<int> off(nb + 1, 0);
  for (int i = 0; i < nb; ++i) off[i + 1] = off[i] + blk[i];
  std::vector<double> out((size_t)nb * nb);
  for (int a = 0; a < nb; ++a)
    for (int b = 0; b < nb; ++b) {
      double s = 0.0;
      for (int i = off[a]; i < off[a + 1]; ++i)
        for (int j = off[b]; j < off[b + 1]; ++j) s += Q[(size_t)i * nsh + j] * Q[(size_t)i * nsh + j];
      out[(size_t)a * nb + b] = std::sqrt(s);
    }
  return out;
}

static std::vector<double> block_norms_1d(const std::vector<double>& Q, const std::vector<int>& blk)
{
  std::vector<double> out(blk.size());
  int o = 0;
  for (size_t a = 0; a < blk.size(); ++a) {
    double s = 0.0;
    for (int i = o; i < o + blk[a]; ++i) s += Q[i] * Q[i];
    o += blk[a];
    out[a] = std::sqrt(s);
  }
  return out;
}

// Functions per block of consecutive shells.
static std::vector<int> block_functions(const Basis& b, const std::vector<int>& blk)
{
  std::vector<int> out(blk.size());
  int s = 0;
  for (size_t a = 0; a < blk.size(); ++a) {
    int f = 0;
    for (int i = s; i < s + blk[a]; ++i) f += 2 * b.shells[i].l + 1;
    s += blk[a];
    out[a] = f;
  }
  return out;
}

ExchangeBatcher::ExchangeBatcher(int device) : device_(device)
{
  HIP_CHECK(hipSetDevice(device_));
  HIP_CHECK(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking));
  th_ = std::thread([this] { loop(); });
}

ExchangeBatcher::~ExchangeBatcher()
{
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
  }
  cv_.notify_all();
  th_.join();
  (void)hipStreamDestroy(s_);
}

void ExchangeBatcher::stats(double* s) const
{
  std::lock_guard<std::mutex> lk(mu_);
  s[0] = (double)nbatch_;
  s[1] = (double)njob_;
  s[2] = (double)maxbatch_;
}

void ExchangeBatcher::exchange(Ctx& c, DevBT& t, const double* C, int ldc, int nocc, double* K)
{
  // one pair of events per worker thread (each worker owns one context)
  static thread_local hipEvent_t ready = nullptr, done = nullptr;
  if (!ready) {
    HIP_CHECK(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
    HIP_CHECK(hipEventCreateWithFlags(&done, hipEventDisableTiming));
  }
  Job j;
  j.t = &t;
  j.K = K;
  j.ready = ready;
  j.done = done;
  dfk_bf16_group_prepare(c, t, C, ldc, nocc, j.r);
  HIP_CHECK(hipEventRecord(ready, c.stream));
  {
    std::unique_lock<std::mutex> lk(mu_);
    q_.push_back(&j);
    cv_.notify_all();
    cv_.wait(lk, [&] { return j.finished; });
  }
  if (!j.error.empty()) throw std::runtime_error("batched exchange: " + j.error);
  HIP_CHECK(hipStreamWaitEvent(c.stream, done, 0));
}

void ExchangeBatcher::loop()
{
  HIP_CHECK(hipSetDevice(device_));
  while (true) {
    std::vector<Job*> batch;
    {
      std::unique_lock<std::mutex> lk(mu_);
      cv_.wait(lk, [&] { return stop_ || !q_.empty(); });
      if (q_.empty()) return;   // stopping
      batch.assign(q_.begin(), q_.end());
      q_.clear();
    }
    std::string err;
    try {
      std::vector<DevBT*> ts;
      std::vector<ExGroupReq> reqs;
      std::vector<double*> K;
      for (Job* j : batch) {
        HIP_CHECK(hipStreamWaitEvent(s_, j->ready, 0));
        ts.push_back(j->t);
        reqs.push_back(j->r);
        K.push_back(j->K);
      }
      dfk_bf16_group_launch(s_, ts, reqs, K, gb_);
      for (Job* j : batch) HIP_CHECK(hipEventRecord(j->done, s_));
    }
    catch (const std::exception& e) {
      err = e.what();
    }
    {
      std::lock_guard<std::mutex> lk(mu_);
      for (Job* j : batch) {
        j->error = err;
        j->finished = true;
      }
      ++nbatch_;
      njob_ += (long)batch.size();
      maxbatch_ = std::max(maxbatch_, (int)batch.size());
    }
    cv_.notify_all();
    // one batch in flight: builds handed in meanwhile form the next batch
    if (err.empty()) (void)hipStreamSynchronize(s_);
  }
}

Server::Server(int device, size_t budget_bytes, int nslots, int cpu_threads)
    : device_(device), budget_(budget_bytes), nslots_(std::max(1, nslots)), t0_(now_s())
{
  batcher_ = std::make_unique<ExchangeBatcher>(device_);
  cpu_per_worker_ = std::max(1, cpu_threads / nslots_);
  ctx_.resize(nslots_);
  for (int i = 0; i < nslots_; ++i) {
    ctx_[i] = std::make_unique<Ctx>();
    ctx_init(*ctx_[i], device_);
    // Requests carry no FP32 copy, so the escalation mode's RMS trigger would
    // hand the last iterations to the plain FP64 exchange (L^-1 applied per
    // build, about 15x a BF16x3 build).  BF16x3 runs until the SCF converges
    // or stalls instead; the energy comes from the final FP64 exchange.
    ctx_[i]->escalate_rms = 0.0;
  }
  for (int i = 0; i < nslots_; ++i) workers_.emplace_back([this, i] { worker(i); });
}

Server::~Server()
{
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
  }
  cv_.notify_all();
  for (auto& w : workers_) w.join();
  batcher_.reset();
  for (auto& c : ctx_) ctx_free(*c);
}

double Server::clock() const { return now_s() - t0_; }

int Server::submit(std::unique_ptr<SrvJob> job)
{
  std::lock_guard<std::mutex> lk(mu_);
  const int id = (int)jobs_.size();
  job->id = id;
  job->t_submit = now_s() - t0_;
  jobs_.push_back(std::move(job));
  pending_.push_back(id);
  cv_.notify_all();
  return id;
}

bool Server::result(int id, bool wait, SrvJob& out)
{
  std::unique_lock<std::mutex> lk(mu_);
  if (id < 0 || id >= (int)jobs_.size()) throw std::runtime_error("server: unknown request id");
  if (wait) cv_.wait(lk, [&] { return jobs_[id]->done; });
  if (!jobs_[id]->done) return false;
  const SrvJob& j = *jobs_[id];
  out.id = j.id;
  out.t_submit = j.t_submit;
  out.t_admit = j.t_admit;
  out.t_done = j.t_done;
  out.energy = j.energy;
  out.energy_iter = j.energy_iter;
  out.iterations = j.iterations;
  out.converged = j.converged;
  out.failed = j.failed;
  out.error = j.error;
  out.est_bytes = j.est_bytes;
  out.nbf = j.nbf;
  out.naux = j.naux;
  out.slot = j.slot;
  out.t_host = j.t_host;
  out.t_build = j.t_build;
  out.t_setup = j.t_setup;
  out.t_scf = j.t_scf;
  out.done = true;
  return true;
}

void Server::stats(double* s) const
{
  std::lock_guard<std::mutex> lk(mu_);
  s[0] = (double)in_use_;
  s[1] = (double)peak_;
  s[2] = (double)completed_;
  s[3] = (double)failed_;
  s[4] = (double)max_live_;
}

// Next request in arrival order (blocks); -1 when stopping with nothing left.
int Server::take()
{
  std::unique_lock<std::mutex> lk(mu_);
  cv_.wait(lk, [&] { return stop_ || !pending_.empty(); });
  if (pending_.empty()) return -1;
  const int id = pending_.front();
  pending_.pop_front();
  waiting_.push_back(id);
  return id;
}

// Admission: the request waits until every earlier request has been admitted
// and its estimate fits the free budget.
void Server::admit(SrvJob& j)
{
  std::unique_lock<std::mutex> lk(mu_);
  if (j.est_bytes > budget_) throw std::runtime_error("request does not fit the HBM budget");
  // the estimate must fit the budget next to the resident requests, and the
  // device must really have it free (estimates are not exact); the free-memory
  // check is re-polled while waiting
  const size_t margin = (size_t)4 << 30;
  auto device_free = [&]() {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return (size_t)0;
    return fr;
  };
  while (true) {
    cv_.wait(lk, [&] { return waiting_.front() == j.id && in_use_ + j.est_bytes <= budget_; });
    if (live_ == 0 || device_free() >= j.est_bytes + margin) break;
    cv_.wait_for(lk, std::chrono::milliseconds(20));
  }
  waiting_.pop_front();
  in_use_ += j.est_bytes;
  peak_ = std::max(peak_, in_use_);
  ++live_;
  max_live_ = std::max(max_live_, live_);
  j.t_admit = now_s() - t0_;
  cv_.notify_all();
}

void Server::finish(SrvJob& j, bool admitted)
{
  std::lock_guard<std::mutex> lk(mu_);
  if (admitted) {
    in_use_ -= j.est_bytes;
    --live_;
  }
  else {
    auto it = std::find(waiting_.begin(), waiting_.end(), j.id);
    if (it != waiting_.end()) waiting_.erase(it);
  }
  j.t_done = now_s() - t0_;
  j.done = true;
  if (j.failed) ++failed_;
  else ++completed_;
  cv_.notify_all();
}

// Page-locked host memory that only grows; HostArray views its first n
// doubles.
struct PinnedArena {
  double* p = nullptr;
  size_t n = 0;
  double* get(size_t m)
  {
    if (m > n) {
      if (p) (void)hipHostFree(p);
      p = nullptr;
      n = 0;
      HIP_CHECK(hipHostMalloc((void**)&p, m * sizeof(double), hipHostMallocDefault));
      n = m;
    }
    return p;
  }
  ~PinnedArena()
  {
    if (p) (void)hipHostFree(p);
  }
};

struct HostArray {
  double* p;
  size_t n;
  HostArray(PinnedArena& a, size_t m) : p(a.get(m)), n(m) {}
  double* data() { return p; }
  size_t size() const { return n; }
  double& operator[](size_t i) { return p[i]; }
};

void Server::worker(int slot)
{
  HIP_CHECK(hipSetDevice(device_));
  Ctx& c = *ctx_[slot];
  alloc_stream() = c.stream;   // stream-ordered device allocations for this worker
  omp_set_num_threads(cpu_per_worker_);
  while (true) {
    const int id = take();
    if (id < 0) return;
    SrvJob* j;
    {
      std::lock_guard<std::mutex> lk(mu_);
      j = jobs_[id].get();
    }
    j->slot = slot;
    bool admitted = false;
    try {
      run(c, *j, admitted);
    }
    catch (const std::exception& e) {
      j->failed = true;
      j->error = e.what();
      // the request's buffers are freed in stream order: let the frees land
      // before its memory leaves the budget
      c.scratch_a.release();
      c.scratch_b.release();
      c.scratch_c.release();
      c.scratch_partial.release();
      c.syrk_tiles.clear();
      (void)hipStreamSynchronize(c.stream);
    }
    finish(*j, admitted);
  }
}

void Server::run(Ctx& c, SrvJob& j, bool& admitted)
{
  // ---- before admission: one-electron integrals and the metric (GPU, a few
  // hundred MB at most, inside the admission margin), Schwarz mask, estimate
  const double th0 = now_s();
  const Basis& ao = j.ao;
  const Basis& aux = j.aux;
  const int n = ao.nbf, naux = aux.nbf, nocc = j.nocc;
  j.nbf = n;
  j.naux = naux;
  // host arrays that go to the device live in page-locked memory (DMA copies)
  // (kept per worker across requests: hipHostFree synchronises the whole device)
  static thread_local PinnedArena pS, pT, pM;
  HostArray S(pS, (size_t)n * n), T(pT, (size_t)n * n), M(pM, (size_t)naux * naux);
  std::vector<double> V((size_t)n * n);
  ints_1e_device(c, ao, j.charges, S.data(), T.data(), V.data());
  for (size_t i = 0; i < S.size(); ++i) T[i] += V[i];   // H = T + V
  ints_2c2e_device(c, aux, M.data());
  const int nsa = (int)ao.shells.size(), nsx = (int)aux.shells.size();
  std::vector<double> Qao((size_t)nsa * nsa), Qx(nsx);
  schwarz_ao_device(c, ao, Qao.data());
  schwarz_aux(aux, Qx.data());
  const std::vector<double> nmn = block_norms_2d(Qao, nsa, j.ao_blk);
  const std::vector<double> nx = block_norms_1d(Qx, j.aux_blk);
  const int nbb = (int)j.ao_blk.size(), nxb = (int)j.aux_blk.size();
  std::vector<uint8_t> mask((size_t)nxb * nbb * nbb);
  for (int x = 0; x < nxb; ++x)
    for (int p = 0; p < nbb * nbb; ++p)   // screening rule (float)(|Q_MN| |Q_X|) > eps
      mask[(size_t)x * nbb * nbb + p] = (double)(float)(nx[x] * nmn[p]) > j.eps ? 1 : 0;
  BLayout L;
  L.build(block_functions(aux, j.aux_blk), block_functions(ao, j.ao_blk), mask);
  const double tot = (double)L.total;
  const int ldw = (nocc + 31) / 32 * 32;
  double est = 8.0 * tot                                   // B (FP64)
               + 2.0 * 8.0 * naux * (double)naux           // L^-1, V^-1
               + 60.0 * 8.0 * n * (double)n                // SCF matrices, DIIS, split-K partials
               + 64.0 * 1048576.0;                         // tables and small buffers
  // FP32 copy and BF16 hi/lo copy (both with rows padded to 32 elements), the
  // BF16 split-K partials (one n x n FP32 block per 32 aux functions)
  // (the escalation mode runs without the FP32 copy here: BF16x3, then FP64)
  double ptot = 0.0;
  for (int b = 0; b < L.nbb; ++b) ptot += (double)naux * L.bsz[b] * ((L.ncol[b] + 31) / 32 * 32);
  const double copies = (j.kmode == 2 ? 4.0 * ptot : 0.0) +
                        (j.kmode == 3 || j.kmode == 4
                             ? 4.0 * ptot + 2.0 * 32.0 * naux * nbb * 8.0 + 4.0 * n * (double)n * ((naux + 31) / 32)
                             : 0.0);
  // FP64 exchange work space (HT1 and HT2 unless B is transformed; the final
  // exchange of the low-precision modes is the plain FP64 one, run in occupied
  // batches of at most 8 GB per work array) and the per-slab transform buffers
  // of the low-precision copies (packed canonical columns and their transform,
  // at most two slabs' worth)
  const double ht_full = 8.0 * naux * (double)n * nocc;
  j.final_occ_nbatches = j.kmode >= 2 ? std::max(1, (int)std::ceil(ht_full / 8e9)) : 1;
  const double ht = (j.kmode == 1 ? 1.0 : 2.0) * ht_full / j.final_occ_nbatches +
                    (j.kmode >= 2 ? 4.0 * naux * (double)n * ldw : 0.0);
  double maxslab = 0.0;
  for (int b = 0; b < L.nbb; ++b) maxslab = std::max(maxslab, (double)(L.slab_off[b + 1] - L.slab_off[b]));
  est += std::max(4.0 * tot, copies + ht + 16.0 * maxslab);   // build scratch vs. resident copies + work
  j.est_bytes = (size_t)est;
  j.t_host = now_s() - th0;

  // ---- admission, then the device part
  admit(j);
  admitted = true;
  {
    double tp = now_s();
    DevBT t;
    build_B_device(c, ao, aux, j.ao_blk, j.aux_blk, mask, j.eps, t);
    j.t_build = now_s() - tp;
    tp = now_s();
    Metric m;
    metric_create(c, naux, M.data(), m);
    if (j.kmode == 1) transform_B(c, t, m);
    if (j.kmode >= 2) dbt_make_lowp(c, t, m, j.kmode == 2, j.kmode == 3 || j.kmode == 4);
    j.t_setup = now_s() - tp;
    tp = now_s();
    ScfInput in;
    in.nbf = n;
    in.nocc = nocc;
    in.bsz = block_functions(ao, j.ao_blk);
    in.S = S.data();
    in.H = T.data();
    in.enuc = j.enuc;
    in.threshold = j.threshold;
    in.max_iter = j.max_iter;
    in.filter_eps = j.eps;
    KBuildFn kb, kfinal;
    EnergyFn kenergy;
    select_k_builders(c, t, &m, j.kmode, nocc, 1, j.threshold, kb, kfinal, j.final_occ_nbatches, &kenergy);
    ScfOutput out;
    // the BF16x3 builds of this request join the cross-request batches
    if (batching_)
      c.bf16_exchange = [this](Ctx& cc, DevBT& tt, const double* C, int ldc, int no, double* K) {
        batcher_->exchange(cc, tt, C, ldc, no, K);
      };
    try {
      scf_run(c, t, t.transformed ? nullptr : &m, in, out, kb, kfinal, kenergy);
    }
    catch (...) {
      c.bf16_exchange = nullptr;
      throw;
    }
    c.bf16_exchange = nullptr;
    j.t_scf = now_s() - tp;
    j.energy = out.energy;
    j.energy_iter = out.energy_iter;
    j.iterations = out.iterations;
    j.converged = out.converged;
    // the request's device memory goes back before it leaves the budget
    c.scratch_a.release();
    c.scratch_b.release();
    c.scratch_c.release();
    c.scratch_partial.release();
    c.syrk_tiles.clear();
  }
  HIP_CHECK(hipStreamSynchronize(c.stream));
}

}  // namespace mchip
