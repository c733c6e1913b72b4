// The tiles of 64 wavenumbers a batched or adjoint launch covers: the whole
// grid, or, for a point whose kept forward quantities exceed the adjoint's
// memory budget, one window of its grid at a time.
#pragma once

#include "device_math.h"

namespace bear_amd {
namespace dev {

// tiles [first, first + count) (in the launch's tile width); the kept buffer
// starts at the kept quantities of wavenumber `kept_first` of point 0, so
// that keptAt subtracts keptColumn's offset of that wavenumber
struct TileWindow {
  int first;
  int count;
  size_t kept_origin;
};

__host__ __device__ inline TileWindow wholeGrid(int tiles) { return TileWindow{0, tiles, 0}; }

__device__ __forceinline__ bool windowItem(size_t item, int batch, const TileWindow& w, int& point, int& tile,
                                           int group = kPointGroup) {
  int local = 0;
  const bool ok = workItem(item, batch, w.count, point, local, group);
  tile = w.first + local;
  return ok;
}

__device__ __forceinline__ size_t keptAt(const TileWindow& w, int m, size_t stride, int levels, int i) {
  return keptColumn(m, stride, levels, i) - w.kept_origin;
}

}  // namespace dev
}  // namespace bear_amd
