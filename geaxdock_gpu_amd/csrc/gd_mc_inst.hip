// gd_mc_inst.hip -- the MC kernel instances of one atoms-per-lane count,
// GD_MC_APL (set by the build: this file is compiled once per count, so the
// instances build in parallel). Each instance keeps only the registers its
// lane shape needs.
#include "gd_mc.h"

namespace gd {

// Host handle of gd_mc_kernel<GD_MC_APL, da, tiled>: da 0 plain MC and pose
// evaluation, 1 delayed acceptance with the network per atom, 2 with its
// tabulated outputs, 3 and 4 the same with the receptor split's groups
// sharing the surrogate's KDE and MCS sums; tiled: the receptor streams
// through LDS tiles.
template <>
const void *mc_kernel_handle<GD_MC_APL>(int da, bool tiled) {
  if (tiled) {
    switch (da) {
      case 1: return (const void *)gd_mc_kernel<GD_MC_APL, 1, true>;
      case 2: return (const void *)gd_mc_kernel<GD_MC_APL, 2, true>;
      case 3: return (const void *)gd_mc_kernel<GD_MC_APL, 3, true>;
      case 4: return (const void *)gd_mc_kernel<GD_MC_APL, 4, true>;
      default: return (const void *)gd_mc_kernel<GD_MC_APL, 0, true>;
    }
  }
  switch (da) {
    case 1: return (const void *)gd_mc_kernel<GD_MC_APL, 1, false>;
    case 2: return (const void *)gd_mc_kernel<GD_MC_APL, 2, false>;
    case 3: return (const void *)gd_mc_kernel<GD_MC_APL, 3, false>;
    case 4: return (const void *)gd_mc_kernel<GD_MC_APL, 4, false>;
    default: return (const void *)gd_mc_kernel<GD_MC_APL, 0, false>;
  }
}

}  // namespace gd
