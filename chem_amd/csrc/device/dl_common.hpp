// Helpers shared by the direct-to-LDS pipelined kernels (gfx950):
// XCD-aware tile mapping, global_load_lds_dwordx4 stage loads and their
// vmcnt waits, and the LDS chunk swizzle.
#pragma once

#include <hip/hip_runtime.h>

#include <utility>

namespace mchip {

// Workgroups are dispatched round-robin over the 8 XCDs; consecutive tile
// indices are mapped onto one XCD so that tiles sharing operand rows share an
// L2.  grid must be a multiple of 8.
__device__ __forceinline__ int xcd_tile(int bid, int grid)
{
  return (bid & 7) * (grid >> 3) + (bid >> 3);
}

// Position of 16-byte chunk c of row r (64-byte rows) in an LDS stage tile.
// A fragment read takes one chunk from each of 16 consecutive rows (from a
// multiple of 16); with f(r) = ((r >> 1) ^ (r >> 2)) & 3 those 16 chunks fill
// 16 distinct 16-byte bank groups of a 64-bank LDS, and every 8 of them
// (lanes 0-7, 8-15) 8 distinct groups of a 32-bank LDS: conflict-free either way.
__device__ __forceinline__ int swz(int r, int c) { return c ^ (((r >> 1) ^ (r >> 2)) & 3); }

typedef __attribute__((address_space(3))) void* lds_ptr_t;

__device__ __forceinline__ void lds_load16(const uint4* g, uint4* l)
{
  __builtin_amdgcn_global_load_lds((const void*)g, (lds_ptr_t)l, 16, 0, 0);
}

// s_waitcnt vmcnt(n) with the other counters left at their maximum (gfx9 encoding)
template <int N>
__device__ __forceinline__ void wait_vm()
{
  static_assert(N >= 0 && N < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((N & 0xF) | ((N >> 4) << 14) | (0x7 << 4) | (0xF << 8));
}

// Wait until at most `after` stages' loads (per wave: LPW each) are in flight.
template <int LPW, int S>
__device__ __forceinline__ void wait_stages(int after)
{
  if constexpr (S >= 4) {
    if (after >= 3) { wait_vm<3 * LPW>(); return; }
  }
  if constexpr (S >= 3) {
    if (after >= 2) { wait_vm<2 * LPW>(); return; }
  }
  if (after >= 1) { wait_vm<LPW>(); return; }
  wait_vm<0>();
}

// Workgroup barrier without the memory fence of __syncthreads(): that fence
// makes every wave wait for all of its outstanding global loads (vmcnt(0)),
// the prefetched pipeline stages included, at every step.  The pipelined
// kernels order their direct-to-LDS loads themselves (wait_vm / wait_stages
// before the barrier that publishes a stage; the barrier before a stage is
// overwritten comes after the MFMAs that consumed its LDS reads); the empty
// asm statements keep the compiler from moving memory accesses across it.
__device__ __forceinline__ void block_sync()
{
  asm volatile("" ::: "memory");
  // lgkmcnt(0), vmcnt and expcnt at their maximum: this wave's LDS reads are done
  __builtin_amdgcn_s_waitcnt((0xF) | (0x3 << 14) | (0x7 << 4));
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

// f(integral_constant<int, 0>), ..., f(integral_constant<int, N - 1>): a loop
// unrolled at compile time, so that each pipeline stage can live in an LDS
// array of its own.  The compiler's wait-count insertion tells LDS arrays
// apart and then waits, before a stage's LDS reads, only for the loads into
// that stage; with one array indexed by a run-time stage number it waits for
// every outstanding load.
template <class F, int... I>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, I...>)
{
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void static_for(F&& f)
{
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// Runs an S-stage direct-to-LDS pipeline over nsteps steps.  issue(s, k)
// starts the loads of step s into stage k (an integral_constant; the stage
// is k mod S, each stage an LDS array of its own), LPW loads per wave;
// compute(s, k) consumes step s from stage k.  Whole stage cycles in which
// every step issues run first, without branches (the compiler's wait-count
// model then stays exact across the loop's back edge and adds no waits of
// its own); the last steps follow with the issues and waits guarded.
template <int S, int LPW, class Issue, class Compute>
__device__ __forceinline__ void dl_pipeline(int nsteps, Issue&& issue, Compute&& compute)
{
  static_for<S - 1>([&](auto k) {
    if (k.value < nsteps) issue(k.value, k);
  });
  int s0 = 0;
  for (; s0 + 2 * S - 1 <= nsteps; s0 += S) {
    static_for<S>([&](auto k) {
      const int s = s0 + k.value;
      block_sync();   // every wave is done with the stage the next issue overwrites
      issue(s + S - 1, std::integral_constant<int, k.value + S - 1>{});
      wait_vm<(S - 1) * LPW>();
      block_sync();   // step s has landed for every wave
      compute(s, k);
    });
  }
  for (; s0 < nsteps; s0 += S) {
    static_for<S>([&](auto k) {
      const int s = s0 + k.value;
      if (s < nsteps) {
        block_sync();
        if (s + S - 1 < nsteps) issue(s + S - 1, std::integral_constant<int, k.value + S - 1>{});
        wait_stages<LPW, S>(min(S - 1, nsteps - 1 - s));
        block_sync();
        compute(s, k);
      }
    });
  }
}

// Loads of one stage of two single-plane operands (64-byte rows: 16 FP32 or
// 8 FP64 of each row): wave-loads j = wave + NW*u over [A rows/16][B rows/16];
// a wave-load brings 16 rows x 64 bytes.  Rows past alast / blast are
// clamped.  When NW does not divide the count, the first waves issue one load
// more than the others.
template <int BM, int BN, int NW>
__device__ __forceinline__ void issue_rows(uint4* st, int wave, int lane, const uint4* A, const uint4* B,
                                                int64_t a0, int64_t alast, int b0, int blast, int lda4, int ldb4,
                                                int kc0)
{
  constexpr int QA = BM / 16, QB = BN / 16, NJ = QA + QB;
  constexpr int UMAX = (NJ + NW - 1) / NW;
  const int rr = lane >> 2, pos = lane & 3;
#pragma unroll
  for (int u = 0; u < UMAX; ++u) {
    const int j = wave + NW * u;
    if (j >= NJ) break;
    if (j < QA) {
      const int r = 16 * j + rr;
      const int64_t gr = a0 + r < alast ? a0 + r : alast;
      lds_load16(A + gr * lda4 + kc0 + swz(r, pos), st + j * 64);
    }
    else {
      const int q = j - QA;
      const int r = 16 * q + rr;
      const int64_t gr = b0 + r < blast ? b0 + r : blast;
      lds_load16(B + gr * ldb4 + kc0 + swz(r, pos), st + BM * 4 + q * 64);
    }
  }
}

// Stage k mod S of up to four LDS arrays.
template <int S, class K, class T>
__device__ __forceinline__ T* pick_stage(K, T* a0, T* a1, T* a2, T* a3)
{
  constexpr int i = K::value % S;
  if constexpr (i == 0) return a0;
  else if constexpr (i == 1) return a1;
  else if constexpr (i == 2) return a2;
  else return a3;
}

}  // namespace mchip
