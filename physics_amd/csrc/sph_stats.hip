// sph_stats.hip — the statistics forms of the density and force gathers
// (sph_gathers.h with STATS): the step's own scans, lists and exact accept
// tests, recording each particle's neighbour count, index sum and hashed-index
// xor instead of the step's outputs. Compiled apart from sph_kernels.hip,
// whose instantiations are the step's.
#include <hip/hip_runtime.h>

#include "sph_gathers.h"

namespace phy {

void launch_gather_stats(hipStream_t st, const SceneTable& T, int pass, const ForceArgs& a,
                         bool shear, const Tuning& tu, const GatherStats& gs) {
    if (pass == 0)
        density_any<true>(st, T, a.posm, a.posl, a.cell, a.start, nullptr, nullptr, nullptr, nullptr,
                          nullptr, tu, gs);
    else
        force_any<true>(st, T, 0.f, a, shear, tu, gs);
}

}  // namespace phy
