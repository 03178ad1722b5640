// Block-sparse layout of the three-centre tensor B(X|mu nu).
//
// Replaces the reference's dbcsr::btensor<3> container in core mode
// (src/dbcsrx/dbcsr_btensor.hpp:101-616): blocks are (aux block, AO block,
// AO block) triples, one AO block per atom (src/desc/basis.cpp:52-80).
//
// HBM layout ("pair slabs"): for every AO block N there is one dense slab
//     slab_N[x][nu][c],  x in [0,naux), nu in block N, c in [0, ncol[N])
// whose column index c runs over the AO functions mu of the blocks M that form
// a significant pair (M,N), concatenated in ascending M.  A pair is significant
// when the screening mask keeps at least one aux block for it.  Triples
// screened out inside a significant pair are stored as zeros.  Both (M,N) and
// (N,M) are stored, so every slab is a contiguous (naux*|N|) x ncol[N]
// row-major matrix with the contracted AO index mu fastest.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <vector>

namespace mchip {

struct BLayout {
  int nxb = 0, nbb = 0;
  int naux = 0, nbf = 0;
  std::vector<int> xsz, xoff;    // aux block sizes / offsets
  std::vector<int> bsz, boff;    // AO block sizes / offsets
  std::vector<uint8_t> mask;     // [nxb][nbb][nbb], 1 = block kept
  std::vector<int> pair_ptr;     // slab N owns pairs pair_ptr[N]..pair_ptr[N+1]-1
  std::vector<int> pair_m;       // M block of each pair
  std::vector<int> pair_col;     // column offset of block M inside slab N
  std::vector<int> ncol;         // columns of slab N
  std::vector<int64_t> slab_off; // element offset of slab N
  int64_t total = 0;             // elements in all slabs

  bool kept(int x, int m, int n) const
  {
    return mask[((size_t)x * nbb + m) * nbb + n] != 0;
  }

  // Build pair structure from block sizes and mask.
  void build(const std::vector<int>& xs, const std::vector<int>& bs,
             const std::vector<uint8_t>& msk)
  {
    xsz = xs;
    bsz = bs;
    nxb = (int)xs.size();
    nbb = (int)bs.size();
    if (msk.size() != (size_t)nxb * nbb * nbb)
      throw std::runtime_error("block mask has the wrong size");
    mask = msk;
    xoff.assign(nxb + 1, 0);
    boff.assign(nbb + 1, 0);
    for (int i = 0; i < nxb; ++i) xoff[i + 1] = xoff[i] + xsz[i];
    for (int i = 0; i < nbb; ++i) boff[i + 1] = boff[i] + bsz[i];
    naux = xoff[nxb];
    nbf = boff[nbb];
    pair_ptr.assign(nbb + 1, 0);
    pair_m.clear();
    pair_col.clear();
    ncol.assign(nbb, 0);
    slab_off.assign(nbb + 1, 0);
    for (int n = 0; n < nbb; ++n) {
      int col = 0;
      for (int m = 0; m < nbb; ++m) {
        bool sig = false;
        for (int x = 0; x < nxb && !sig; ++x) sig = kept(x, m, n);
        if (!sig) continue;
        pair_m.push_back(m);
        pair_col.push_back(col);
        col += bsz[m];
      }
      pair_ptr[n + 1] = (int)pair_m.size();
      ncol[n] = col;
      slab_off[n + 1] = slab_off[n] + (int64_t)naux * bsz[n] * col;
    }
    total = slab_off[nbb];
  }

  // Column offset of block m inside slab n, or -1 if (m,n) is not significant.
  int col_of(int m, int n) const
  {
    for (int p = pair_ptr[n]; p < pair_ptr[n + 1]; ++p)
      if (pair_m[p] == m) return pair_col[p];
    return -1;
  }
};

}  // namespace mchip
