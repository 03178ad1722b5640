// Three-centre integral classes with AO angular momentum 0 on the first
// shell of the pair (kernels: ints3c_cls.hpp).
#include <stdexcept>

#include "ints3c_cls.hpp"

namespace mchip {

void launch_int3c2e_la0(int lb, int lc, dim3 grid, hipStream_t st, const I3cArgs& a)
{
  launch_int3c2e_la<0>(lb, lc, grid, st, a);
}

}  // namespace mchip
