// Fine-grid cell -> covering subgrid tiles, per dimension (host construction
// and device copy), for the tile reduction that forms the fine grid from the
// stored spreading tiles.
#pragma once

#include <algorithm>
#include <memory>
#include <vector>

#include "kernels.hpp"

namespace nnf {

// The tile of bin b along one dimension starts ns/2 cells before the bin and
// spans min(bin, n - b * bin) + ns cells, periodically. Cell c is covered by
// bin b at every tile position l < extent with l = (c - start) mod n + k n.
// Only bins within ceil(ns / bin) + 1 of c's own bin can cover it, so those
// are the candidates (all bins when there are few).
struct CoverHost {
  std::vector<int> start[3];
  std::vector<int2> entry[3];
};

inline CoverHost build_cover(int dim, const BinGeom& g, int ns) {
  CoverHost h;
  const int half = ns / 2;
  for (int d = 0; d < 3; ++d) {
    if (d >= dim) {
      h.start[d] = {0, 1};
      h.entry[d] = {make_int2(0, 0)};
      continue;
    }
    const int n = g.n[d], B = g.bin[d], nb = g.nbins[d];
    const int J = (ns + B - 1) / B + 1;
    h.start[d].assign(static_cast<std::size_t>(n) + 1, 0);
    std::vector<int> cand;
    for (int c = 0; c < n; ++c) {
      h.start[d][c] = static_cast<int>(h.entry[d].size());
      cand.clear();
      if (nb <= 2 * J + 1) {
        for (int b = 0; b < nb; ++b) cand.push_back(b);
      } else {
        const int b0 = c / B;
        for (int j = -J; j <= J; ++j) cand.push_back(((b0 + j) % nb + nb) % nb);
      }
      for (int b : cand) {
        const long long org = static_cast<long long>(b) * B - half;
        const int ext = std::min(B, n - b * B) + ns;
        long long l = ((c - org) % n + n) % n;
        for (; l < ext; l += n) h.entry[d].push_back(make_int2(b, static_cast<int>(l)));
      }
    }
    h.start[d][n] = static_cast<int>(h.entry[d].size());
  }
  return h;
}

class CoverBuffers {
public:
  // Builds the tables and copies them to the device on `stream`; the host
  // copies are kept until the next build so the asynchronous copy stays valid.
  void build(int dim, const BinGeom& g, int ns, hipStream_t stream,
             const std::shared_ptr<neonufft::Allocator>& alloc) {
    host_ = build_cover(dim, g, ns);
    for (int d = 0; d < 3; ++d) {
      start_[d].reset(host_.start[d].size(), alloc);
      entry_[d].reset(host_.entry[d].size(), alloc);
      check_hip(hipMemcpyAsync(start_[d].data(), host_.start[d].data(), start_[d].bytes(),
                               hipMemcpyHostToDevice, stream),
                "hipMemcpyAsync(cover)");
      check_hip(hipMemcpyAsync(entry_[d].data(), host_.entry[d].data(), entry_[d].bytes(),
                               hipMemcpyHostToDevice, stream),
                "hipMemcpyAsync(cover)");
    }
  }

  CoverTable table() const {
    return CoverTable{{start_[0].data(), start_[1].data(), start_[2].data()},
                      {entry_[0].data(), entry_[1].data(), entry_[2].data()}};
  }

private:
  CoverHost host_;
  DeviceBuffer<int> start_[3];
  DeviceBuffer<int2> entry_[3];
};

}  // namespace nnf
