// Lane-layout probe of the FP64 MFMA builtin (one wave, fixed 16x4 and 4x16
// operands).  Used by the tests to pin the operand/accumulator layout the
// kernels assume.
#include "../../../include/mchip.h"
#include "../capi_common.hpp"
#include "dev.hpp"

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
