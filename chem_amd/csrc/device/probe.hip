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
