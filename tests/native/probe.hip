// Test-only probes of the gfx950 builtins the kernels of libmchip.so rely on:
// the lane layouts of the FP64, FP32 and BF16 MFMA instructions and the LDS
// placement of global_load_lds_dwordx4.  Built by the package build into
// tests/native/libmchip_probe.so, which only the tests load; nothing of it is
// part of libmchip.so.  Each entry point returns 0, or 1 with a message for
// mchip_probe_last_error().
#include <stdexcept>
#include <string>

#include "device/dev.hpp"

static thread_local std::string g_probe_err;

extern "C" const char* mchip_probe_last_error(void) { return g_probe_err.c_str(); }

#define MCHIP_TRY try {
#define MCHIP_CATCH                  \
  }                                  \
  catch (const std::exception& e)    \
  {                                  \
    g_probe_err = e.what();          \
    return 1;                        \
  }                                  \
  return 0;

namespace mchip {

typedef double d4p __attribute__((ext_vector_type(4)));

// a_lane[l], b_lane[l]: the double lane l supplies as A and B operand;
// out[l*4+e]: accumulator element e of lane l after one MFMA from zero.
__global__ void k_probe_mfma_f64(const double* a_lane, const double* b_lane, double* out)
{
  const int l = threadIdx.x;
  d4p acc = {0.0, 0.0, 0.0, 0.0};
  acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a_lane[l], b_lane[l], acc, 0, 0, 0);
  for (int e = 0; e < 4; ++e) out[l * 4 + e] = acc[e];
}

}  // namespace mchip

using namespace mchip;

extern "C" int mchip_probe_mfma_f64(const double* a_lane, const double* b_lane, double* out)
{
  MCHIP_TRY
  DevBuf a(64 * sizeof(double)), b(64 * sizeof(double)), o(256 * sizeof(double));
  HIP_CHECK(hipMemcpy(a.p, a_lane, 64 * sizeof(double), hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(b.p, b_lane, 64 * sizeof(double), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_probe_mfma_f64, dim3(1), dim3(64), 0, 0, a.as<double>(), b.as<double>(), o.as<double>());
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipDeviceSynchronize());
  HIP_CHECK(hipMemcpy(out, o.p, 256 * sizeof(double), hipMemcpyDeviceToHost));
  MCHIP_CATCH
}

namespace mchip {

typedef float f4p __attribute__((ext_vector_type(4)));
typedef __bf16 bf8p __attribute__((ext_vector_type(8)));

// FP32 16x16x4: lane l supplies a_lane[l], b_lane[l]; out[l*4+e].
__global__ void k_probe_mfma_f32(const float* a_lane, const float* b_lane, float* out)
{
  const int l = threadIdx.x;
  f4p acc = {0.f, 0.f, 0.f, 0.f};
  acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a_lane[l], b_lane[l], acc, 0, 0, 0);
  for (int e = 0; e < 4; ++e) out[l * 4 + e] = acc[e];
}

// BF16 16x16x32: lane l supplies 8 bf16 of A (a_lane[l*8+q]) and of B; out[l*4+e].
__global__ void k_probe_mfma_bf16(const float* a_lane, const float* b_lane, float* out)
{
  const int l = threadIdx.x;
  bf8p a, b;
  for (int q = 0; q < 8; ++q) {
    a[q] = (__bf16)a_lane[l * 8 + q];
    b[q] = (__bf16)b_lane[l * 8 + q];
  }
  f4p acc = {0.f, 0.f, 0.f, 0.f};
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
  for (int e = 0; e < 4; ++e) out[l * 4 + e] = acc[e];
}

}  // namespace mchip

extern "C" int mchip_probe_mfma_lp(int kind, const float* a_lane, const float* b_lane, float* out)
{
  MCHIP_TRY
  const int per = (kind == 1) ? 8 : 1;
  DevBuf a(64 * per * sizeof(float)), b(64 * per * sizeof(float)), o(256 * sizeof(float));
  HIP_CHECK(hipMemcpy(a.p, a_lane, a.bytes, hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(b.p, b_lane, b.bytes, hipMemcpyHostToDevice));
  if (kind == 0)
    hipLaunchKernelGGL(k_probe_mfma_f32, dim3(1), dim3(64), 0, 0, a.as<float>(), b.as<float>(), o.as<float>());
  else if (kind == 1)
    hipLaunchKernelGGL(k_probe_mfma_bf16, dim3(1), dim3(64), 0, 0, a.as<float>(), b.as<float>(), o.as<float>());
  else
    throw std::runtime_error("mchip_probe_mfma_lp: kind must be 0 (f32) or 1 (bf16)");
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipDeviceSynchronize());
  HIP_CHECK(hipMemcpy(out, o.p, 256 * sizeof(float), hipMemcpyDeviceToHost));
  MCHIP_CATCH
}

namespace mchip {

// global_load_lds_dwordx4 placement: each of the 64 lanes loads the 16 bytes
// src[perm[lane]] into LDS; the LDS image (64 x 16 bytes from the destination
// base) is copied out.  Expected: lane i's 16 bytes at base + 16 i.
__global__ void k_probe_lds_load(const uint4* __restrict__ src, const int* __restrict__ perm, uint4* __restrict__ out)
{
  __shared__ uint4 lds[128];
  const int lane = threadIdx.x;
  lds[lane] = uint4{0u, 0u, 0u, 0u};
  lds[64 + lane] = uint4{0u, 0u, 0u, 0u};
  __syncthreads();
  __builtin_amdgcn_global_load_lds((const void*)(src + perm[lane]),
                                   (__attribute__((address_space(3))) void*)(lds + 64), 16, 0, 0);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  out[lane] = lds[lane];
  out[64 + lane] = lds[64 + lane];
}

}  // namespace mchip

extern "C" int mchip_probe_lds_load(const uint32_t* src, const int* perm, uint32_t* out)
{
  MCHIP_TRY
  using namespace mchip;
  DevBuf s(64 * 16), p(64 * sizeof(int)), o(128 * 16);
  HIP_CHECK(hipMemcpy(s.p, src, 64 * 16, hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(p.p, perm, 64 * sizeof(int), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_probe_lds_load, dim3(1), dim3(64), 0, 0, s.as<uint4>(), p.as<int>(), o.as<uint4>());
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipDeviceSynchronize());
  HIP_CHECK(hipMemcpy(out, o.p, 128 * 16, hipMemcpyDeviceToHost));
  MCHIP_CATCH
}
