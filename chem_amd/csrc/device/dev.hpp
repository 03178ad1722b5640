// Device context and device-resident B tensor.
#pragma once

#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <cstdint>
#include <map>
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

// Allocation stream of the calling thread.  When set (the request workers of
// the server set it to their context's stream), DevBuf allocates and frees
// stream-ordered on it, so freeing a buffer does not wait for the work other
// threads have queued on the device.
inline hipStream_t& alloc_stream()
{
  static thread_local hipStream_t s = nullptr;
  return s;
}

// Owning device buffer.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  hipStream_t astream = nullptr;   // stream of a stream-ordered allocation
  DevBuf() = default;
  explicit DevBuf(size_t n) { alloc(n); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes), astream(o.astream)
  {
    o.p = nullptr;
    o.bytes = 0;
    o.astream = nullptr;
  }
  DevBuf& operator=(DevBuf&& o) noexcept
  {
    release();
    p = o.p;
    bytes = o.bytes;
    astream = o.astream;
    o.p = nullptr;
    o.bytes = 0;
    o.astream = nullptr;
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t n)
  {
    release();
    if (n) {
      hipStream_t s = alloc_stream();
      if (s) {
        HIP_CHECK(hipMallocAsync(&p, n, s));
        astream = s;
      }
      else {
        HIP_CHECK(hipMalloc(&p, n));
      }
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
      if (astream) (void)hipFreeAsync(p, astream);
      else (void)hipFree(p);
    }
    p = nullptr;
    bytes = 0;
    astream = nullptr;
  }
  template <class T> T* as() const { return reinterpret_cast<T*>(p); }
};

template <class T>
DevBuf upload_vec(const std::vector<T>& v, hipStream_t s)
{
  DevBuf b(v.size() * sizeof(T));
  if (!v.empty()) HIP_CHECK(hipMemcpyAsync(b.p, v.data(), b.bytes, hipMemcpyHostToDevice, s));
  return b;
}

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  rocblas_handle blas = nullptr;
  DevBuf scratch_a, scratch_b, scratch_c;   // reusable workspaces
  DevBuf scratch_partial;                   // split-K partial sums
  DevBuf info;                              // rocsolver info
  std::map<int, DevBuf> syrk_tiles;         // n -> upper-tile table of the FP32 SYRK
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
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

struct DevBT {
  BLayout L;                 // host copy of the layout
  DevBuf data64;             // FP64 slabs (B or B~)
  DevBuf data32;             // FP32 slabs (optional)
  DevBuf data_hi, data_lo;   // BF16 split slabs (optional), rows padded to ldp[N]
  DevBuf ldp, off16, ldp_pre;               // padded row length, slab offsets, prefix of ldp
  std::vector<int> h_ldp;
  std::vector<int64_t> h_off16, h_ldp_pre;
  std::map<int, int> bf16_bn;               // ng -> tuned launch shape of the BF16 slab GEMM
  DevBuf slab_off, ncol, bsz, boff, colmap, colmap_off, pg_off;
  std::vector<int> h_colmap, h_colmap_off;
  std::vector<int64_t> h_pg_off;            // prefix of bsz[N]*ncol[N]
  int64_t colmap_total = 0;
  int64_t pg_total = 0;                     // sum of bsz[N]*ncol[N]
  int64_t maxR = 0;                         // max bsz[N]*ncol[N]
  bool transformed = false;  // data64 holds L^-1 B
  bool lowp_tilde = false;   // data32 / data_hi,lo hold L^-1 B (from either route)
  std::map<std::pair<int, int>, DevBuf> tiles;   // (BN, ng) -> tile table
  std::map<std::pair<int, int>, int> ntiles;
};

struct Metric {
  int naux = 0;
  DevBuf Linv;   // column-major lower-triangular L^-1 (V = L L^T)
  DevBuf Vinv;   // V^-1 = L^-T L^-1 (symmetric)
};

}  // namespace mchip
