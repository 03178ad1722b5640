// Request loop serving DF-HF requests on one GPU (see serve.hip).
#pragma once

#include <atomic>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../ints/basis.hpp"
#include "../ints/ints_host.hpp"
#include "dev.hpp"
#include "dfock.hpp"

namespace mchip {

struct SrvJob {
  // request
  Basis ao, aux;
  std::vector<Charge> charges;
  std::vector<int> ao_blk, aux_blk;   // shells per block
  int nocc = 0;
  double enuc = 0.0, threshold = 1e-6, eps = 1e-10;
  int kmode = 4, max_iter = 100;
  int final_occ_nbatches = 1;          // occupied batches of the final FP64 exchange
  // bookkeeping (seconds on the server clock)
  int id = -1, slot = -1;
  double t_submit = 0.0, t_admit = 0.0, t_done = 0.0;
  double t_host = 0.0, t_build = 0.0, t_setup = 0.0, t_scf = 0.0;   // phase seconds
  size_t est_bytes = 0;
  // result
  bool done = false, failed = false, converged = false;
  std::string error;
  double energy = 0.0, energy_iter = 0.0;
  int iterations = 0, nbf = 0, naux = 0;
};

// Cross-request batching of the BF16x3 exchange: workers hand their builds
// in, one thread launches everything queued as one grouped slab-GEMM and one
// grouped SYRK (dfk_bf16_group_launch) on its own stream, one batch in flight
// at a time, so that builds arriving while a batch runs form the next one.
class ExchangeBatcher {
 public:
  explicit ExchangeBatcher(int device);
  ~ExchangeBatcher();
  // From a worker with context c: gathers the coefficients on c.stream, joins
  // the next batch and returns once it is launched; c.stream then waits for
  // the batch before K is read.  Results equal the build alone bit for bit.
  void exchange(Ctx& c, DevBT& t, const double* C, int ldc, int nocc, double* K);
  void stats(double* s) const;   // batches, builds, largest batch

 private:
  struct Job {
    DevBT* t = nullptr;
    ExGroupReq r{};
    double* K = nullptr;
    hipEvent_t ready = nullptr, done = nullptr;
    bool finished = false;
    std::string error;
  };
  void loop();
  int device_;
  hipStream_t s_ = nullptr;
  mutable std::mutex mu_;
  std::condition_variable cv_;
  std::deque<Job*> q_;
  bool stop_ = false;
  GroupBufs gb_;
  long nbatch_ = 0, njob_ = 0;
  int maxbatch_ = 0;
  std::thread th_;
};

class Server {
 public:
  Server(int device, size_t budget_bytes, int nslots, int cpu_threads);
  ~Server();   // finishes every submitted request, then stops the workers
  int submit(std::unique_ptr<SrvJob> job);
  bool result(int id, bool wait, SrvJob& out);   // copy of the bookkeeping and result
  void stats(double* s) const;   // in_use, peak, completed, failed, max live requests
  double clock() const;          // seconds since the server started
  // Cross-request batching of the BF16x3 exchange builds (on by default);
  // applies to requests that start after the call.
  void set_batching(bool on) { batching_ = on; }
  void batch_stats(double* s) const { batcher_->stats(s); }

 private:
  int take();
  void admit(SrvJob& j);
  void finish(SrvJob& j, bool admitted);
  void worker(int slot);
  void run(Ctx& c, SrvJob& j, bool& admitted);

  int device_;
  size_t budget_;
  int nslots_, cpu_per_worker_ = 1;
  double t0_;
  mutable std::mutex mu_;
  std::condition_variable cv_;
  std::vector<std::unique_ptr<SrvJob>> jobs_;
  std::deque<int> pending_, waiting_;
  size_t in_use_ = 0, peak_ = 0;
  int live_ = 0, max_live_ = 0, completed_ = 0, failed_ = 0;
  bool stop_ = false;
  std::atomic<bool> batching_{true};
  std::unique_ptr<ExchangeBatcher> batcher_;
  std::vector<std::unique_ptr<Ctx>> ctx_;
  std::vector<std::thread> workers_;
};

}  // namespace mchip
