// Helpers shared by the direct-to-LDS pipelined kernels (gfx950):
// XCD-aware tile mapping, global_load_lds_dwordx4 stage loads and their
// vmcnt waits, and the LDS chunk swizzle.
#pragma once

#include <hip/hip_runtime.h>

namespace mchip {

// Workgroups are dispatched round-robin over the 8 XCDs; consecutive tile
// indices are mapped onto one XCD so that tiles sharing operand rows share an
// L2.  grid must be a multiple of 8.
__device__ __forceinline__ int xcd_tile(int bid, int grid)
{
  return (bid & 7) * (grid >> 3) + (bid >> 3);
}

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

__device__ __forceinline__ int swz(int r, int c) { return c ^ ((r >> 2) & 3); }

}  // namespace mchip
