// ww3-gpu_amd/csrc/snl1_device.hpp
// Device-side helpers shared by the DIA kernels (included by .hip files only).
#pragma once

#include "snl1.hpp"

namespace ww3gpu {
namespace snl1 {

/// Workgroup index -> point index under `L` (-1: no point for this workgroup).  With
/// xcd_chunk set, consecutive workgroups of one XCD (dispatched round-robin over the 8 XCDs)
/// take consecutive slots; with a list, slot i is point list[i].
__device__ inline int layout_point(const Layout& L, int npts) {
  int p = static_cast<int>(blockIdx.x);
  if (L.xcd_chunk > 0) p = (p % 8) * L.xcd_chunk + p / 8;
  if (p >= npts) return -1;
  if (L.list != nullptr) p = L.list[p];
  if (L.skip != nullptr && L.skip[p] != 0) return -1;
  return p;
}

}  // namespace snl1
}  // namespace ww3gpu
