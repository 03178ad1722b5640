// Device context and device-resident B tensor.
#pragma once

#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <cstdint>
#include <algorithm>
#include <array>
#include <cstring>
#include <functional>
#include <map>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "ints/btensor.hpp"

namespace mchip {

#define HIP_CHECK(x)                                                              \
  do {                                                                            \
    hipError_t e_ = (x);                                                          \
    if (e_ != hipSuccess)                                                         \
      throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e_) + \
                               " at " + __FILE__ + ":" + std::to_string(__LINE__)); \
  } while (0)

#define RB_CHECK(x)                                                               \
  do {                                                                            \
    rocblas_status s_ = (x);                                                      \
    if (s_ != rocblas_status_success)                                             \
      throw std::runtime_error(std::string("rocBLAS/rocSOLVER error ") +          \
                               rocblas_status_to_string(s_) + " at " + __FILE__ + \
                               ":" + std::to_string(__LINE__));                   \
  } while (0)

// Owning device buffer.
// Stream on which this thread's DevBuf allocations are made and freed in
// stream order (hipMallocAsync / hipFreeAsync).  Null (the default): hipMalloc
// / hipFree, and every hipFree synchronises the whole device.  The request
// workers set their own stream so that one request's frees do not wait for
// the kernels of the others.  A buffer is freed on the stream it was made on.
inline hipStream_t& alloc_stream()
{
  static thread_local hipStream_t s = nullptr;
  return s;
}

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  hipStream_t s = nullptr;   // allocation stream (null: hipMalloc)
  DevBuf() = default;
  explicit DevBuf(size_t n) { alloc(n); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes), s(o.s)
  {
    o.p = nullptr;
    o.bytes = 0;
    o.s = nullptr;
  }
  DevBuf& operator=(DevBuf&& o) noexcept
  {
    release();
    p = o.p;
    bytes = o.bytes;
    s = o.s;
    o.p = nullptr;
    o.bytes = 0;
    o.s = nullptr;
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t n)
  {
    release();
    if (n) {
      s = alloc_stream();
      if (s)
        HIP_CHECK(hipMallocAsync(&p, n, s));
      else
        HIP_CHECK(hipMalloc(&p, n));
    }
    bytes = n;
  }
  // grow-only
  void ensure(size_t n)
  {
    if (n > bytes) alloc(n);
  }
  void release()
  {
    if (p) {
      if (s)
        (void)hipFreeAsync(p, s);
      else
        (void)hipFree(p);
    }
    p = nullptr;
    bytes = 0;
    s = nullptr;
  }
  template <class T> T* as() const { return reinterpret_cast<T*>(p); }
};

// Page-locked host buffer (hipHostMalloc): host data for large transfers,
// copied by DMA without staging through the runtime's pageable path.
template <class T>
struct PinnedVec {
  T* p = nullptr;
  size_t n = 0;
  explicit PinnedVec(size_t count) : n(count)
  {
    if (n) HIP_CHECK(hipHostMalloc((void**)&p, n * sizeof(T), hipHostMallocDefault));
  }
  PinnedVec(const PinnedVec&) = delete;
  PinnedVec& operator=(const PinnedVec&) = delete;
  ~PinnedVec()
  {
    if (p) (void)hipHostFree(p);
  }
  T* data() { return p; }
  size_t size() const { return n; }
  T& operator[](size_t i) { return p[i]; }
};

// Host<->device copies through a per-thread page-locked staging buffer, in
// order on stream s and complete on return: the runtime's pageable-copy path
// (staging of ordinary host memory) is never used, which keeps host transfers
// of concurrent request workers independent of each other.
struct HostStage {
  static constexpr size_t kBytes = (size_t)4 << 20;
  uint8_t* p = nullptr;
  HostStage() { HIP_CHECK(hipHostMalloc((void**)&p, kBytes, hipHostMallocDefault)); }
  ~HostStage()
  {
    if (p) (void)hipHostFree(p);
  }
};

inline HostStage& host_stage()
{
  static thread_local HostStage st;
  return st;
}

inline void h2d(void* dev, const void* host, size_t bytes, hipStream_t s)
{
  HostStage& st = host_stage();
  for (size_t off = 0; off < bytes; off += HostStage::kBytes) {
    const size_t len = std::min(HostStage::kBytes, bytes - off);
    std::memcpy(st.p, (const uint8_t*)host + off, len);
    HIP_CHECK(hipMemcpyAsync((uint8_t*)dev + off, st.p, len, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipStreamSynchronize(s));
  }
}

inline void d2h(void* host, const void* dev, size_t bytes, hipStream_t s)
{
  HostStage& st = host_stage();
  for (size_t off = 0; off < bytes; off += HostStage::kBytes) {
    const size_t len = std::min(HostStage::kBytes, bytes - off);
    HIP_CHECK(hipMemcpyAsync(st.p, (const uint8_t*)dev + off, len, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    std::memcpy((uint8_t*)host + off, st.p, len);
  }
  if (bytes == 0) HIP_CHECK(hipStreamSynchronize(s));
}

template <class T>
DevBuf upload_vec(const std::vector<T>& v, hipStream_t s)
{
  DevBuf b(v.size() * sizeof(T));
  if (!v.empty()) h2d(b.p, v.data(), b.bytes, s);
  return b;
}

struct DevBT;

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  rocblas_handle blas = nullptr;
  DevBuf scratch_a, scratch_b, scratch_c;   // reusable workspaces
  DevBuf scratch_partial;                   // split-K partial sums
  DevBuf info;                              // rocsolver info
  std::map<int, DevBuf> syrk_tiles;         // n -> upper-tile table of the FP32 SYRK
  // BF16 -> FP32 switch of the escalation mode: at an error RMS <= escalate_rms
  // (0: off) or after escalate_stall iterations without progress (0: off).
  // The FP32 tail keeps |dE| at the convergence noise of the plain SCF (with
  // BF16x3 alone it reached 1.5 uEh for one 53-water cluster at 1e-6).
  double escalate_rms = 1e-5;
  int escalate_stall = 3;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // When set, the BF16x3 exchange of one occupied batch is handed to this
  // function (the request loop's cross-request batcher) instead of being
  // launched on this context's stream.
  std::function<void(Ctx&, DevBT&, const double*, int, int, double*)> bf16_exchange;
  // Replay the launch chains of repeated exchange and DF-J builds from
  // captured hipGraphs (the chain of one build is captured on its second
  // call with the same operands and replayed from the third on).
  bool graphs = true;
  // optional per-phase kernel timing (HIP events on this stream), seconds
  bool timing = false;
  hipEvent_t tev[2] = {nullptr, nullptr};
  double phase[8] = {0, 0, 0, 0, 0, 0, 0, 0};
};

// Accumulates the duration of the enclosed stream work into c.phase[slot].
struct PhaseTimer {
  Ctx& c;
  int slot;
  PhaseTimer(Ctx& cc, int s) : c(cc), slot(s)
  {
    if (c.timing) HIP_CHECK(hipEventRecord(c.tev[0], c.stream));
  }
  ~PhaseTimer() noexcept(false)
  {
    if (!c.timing) return;
    HIP_CHECK(hipEventRecord(c.tev[1], c.stream));
    HIP_CHECK(hipEventSynchronize(c.tev[1]));
    float ms = 0.f;
    HIP_CHECK(hipEventElapsedTime(&ms, c.tev[0], c.tev[1]));
    c.phase[slot] += ms * 1e-3;
  }
};

// phase slots
enum { PH_K1 = 0, PH_K2 = 1, PH_K3 = 2, PH_J = 3, PH_KINC = 4, PH_SPLIT = 5 };

enum class Prec : int { F64 = 0, F32 = 1, BF16X3 = 2 };

// Tile descriptor for grouped launches over slabs.
struct Tile {
  int slab;
  int m0;
  int n0;
  int pad;
};

// One build of a grouped BF16x3 exchange (several tensors in one launch):
// the tensor's B~ hi/lo copy and layout, its gathered coefficients, W and
// the SYRK chunk partials.
struct ExGroupReq {
  const uint16_t* bhi;
  const uint16_t* blo;
  const int64_t* off16;
  const int* ldp;
  const int64_t* ldp_pre;
  const int* bsz;
  const int* boff;
  const uint16_t* ghi;
  const uint16_t* glo;
  uint16_t* whi;
  uint16_t* wlo;
  float* partial;
  int naux, nbf, ng, ldw, spc, pad;
};

// Autotuned exchange pipeline of one precision (BF16x3, FP32 or FP64): slab-GEMM
// shape, SYRK variant (-1 = not chosen yet) and the timings (ms, -1 = not
// tried) they were chosen from.
struct ExchangePlan {
  int k1 = -1, k3 = -1;
  std::vector<float> k1_ms, k3_ms;
};

// Captured launch chains of one tensor's builds: hipGraph executables keyed
// by everything the chain's launch arguments derive from (operation, operand
// and work-array pointers, widths, tuned shapes), and the keys run once
// uncaptured (their first run builds the tile tables and tunes the shapes).
using ChainKey = std::array<int64_t, 14>;
struct GraphCache {
  std::map<ChainKey, hipGraphExec_t> exec;
  std::set<ChainKey> seen;
  int64_t replays = 0;
  GraphCache() = default;
  GraphCache(const GraphCache&) = delete;
  GraphCache& operator=(const GraphCache&) = delete;
  GraphCache(GraphCache&& o) noexcept : exec(std::move(o.exec)), seen(std::move(o.seen)), replays(o.replays)
  {
    o.exec.clear();
  }
  GraphCache& operator=(GraphCache&& o) noexcept
  {
    clear();
    exec = std::move(o.exec);
    seen = std::move(o.seen);
    replays = o.replays;
    o.exec.clear();
    return *this;
  }
  ~GraphCache() { clear(); }
  void clear()
  {
    for (auto& kv : exec) (void)hipGraphExecDestroy(kv.second);
    exec.clear();
    seen.clear();
  }
};

struct DevBT {
  BLayout L;                 // host copy of the layout
  DevBuf data64;             // FP64 slabs (B or B~)
  DevBuf data32;             // FP32 slabs (optional), rows padded to ldp[N] at off16[N]
  DevBuf data64t;            // FP64 L^-1 B slabs beside the untransformed B (optional), rows padded like data32
  DevBuf data_hi, data_lo;   // BF16 split slabs (optional), rows padded to ldp[N]
  DevBuf ldp, off16, ldp_pre;               // padded row length, slab offsets (elements), prefix of ldp
  std::vector<int> h_ldp;
  std::vector<int64_t> h_off16, h_ldp_pre;
  std::map<int, ExchangePlan> bf16_plan;        // ng -> tuned BF16 exchange pipeline
  std::map<int, ExchangePlan> f32_plan;         // ng -> tuned FP32 exchange pipeline
  std::map<int, ExchangePlan> f64_plan;         // ng -> tuned FP64 slab GEMM / SYRK (-1: not yet)
  DevBuf slab_off, ncol, bsz, boff, colmap, colmap_off, pg_off;
  std::vector<int> h_colmap, h_colmap_off;
  std::vector<int64_t> h_pg_off;            // prefix of bsz[N]*ncol[N]
  int64_t colmap_total = 0;
  int64_t pg_total = 0;                     // sum of bsz[N]*ncol[N]
  int64_t maxR = 0;                         // max bsz[N]*ncol[N]
  DevBuf cdn;                               // int2 per slab: first column of block N, first past it
  int64_t maxRc = 0;                        // max bsz[N]*cdn[N].y (canonical half of a slab)
  bool sym_mask = false;                    // block mask symmetric in (M, N): DF-J reads the canonical half
  bool transformed = false;  // data64 holds L^-1 B
  bool lowp_tilde = false;   // data32 / data_hi,lo hold L^-1 B (from either route)
  DevBuf gx_w, gx_g, gx_part;   // work arrays of this tensor's builds in a grouped BF16x3 exchange
  std::map<std::pair<int, int>, DevBuf> tiles;   // (BN, ng) -> tile table
  std::map<std::pair<int, int>, int> ntiles;
  GraphCache graphs;                             // captured build chains on this tensor
};

struct Metric {
  int naux = 0;
  DevBuf Linv;   // column-major lower-triangular L^-1 (V = L L^T)
  DevBuf Vinv;   // V^-1 = L^-T L^-1 (symmetric)
};

}  // namespace mchip
