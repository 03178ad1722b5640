// GPU three-centre integrals.
#pragma once

#include <vector>

#include "../ints/basis.hpp"
#include "dev.hpp"

namespace mchip {

struct Task3 {
  int sa, sb, cp, mu0, nu0, M, N, pad;
};

// Compute, screen, filter and lay out B(X|mu nu) entirely on the device.
void build_B_device(Ctx& c, const Basis& ao, const Basis& aux, const std::vector<int>& ao_blk_shells,
                    const std::vector<int>& aux_blk_shells, const std::vector<uint8_t>& schwarz_mask,
                    double filter_eps, DevBT& t);

// Take ownership of device slabs laid out as L and build the index arrays.
void dbt_adopt(Ctx& c, const BLayout& L, DevBuf&& data, DevBT& t);

}  // namespace mchip
