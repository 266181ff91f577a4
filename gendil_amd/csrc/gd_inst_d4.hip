// Kernel instances for 4D meshes: (nodes per dim, quadrature points per dim),
// including GenDiL's 4D range benchmarks on the device: advection with q = n
// (benchmarks/range-benchmark-advection-4d.cpp:120-143, p = 0..4) and mass with
// q = n + 1 (range-benchmark-mass-4d.cpp:116-141, p = 0..4).
#include "gd_dispatch.hpp"

namespace gd {

static const KernelSet kSets[] = {
    GD_KSET(4, 1, 1),
    GD_KSET(4, 1, 2),
    GD_KSET(4, 2, 2),
    GD_KSET(4, 2, 3),
    GD_KSET(4, 3, 3),
    GD_KSET(4, 3, 4),
    GD_KSET(4, 4, 4),
    GD_KSET(4, 4, 5),
    GD_KSET(4, 5, 5),
    GD_KSET(4, 5, 6)};

KernelList kernels_d4() { return KernelList{kSets, (int)(sizeof(kSets) / sizeof(kSets[0]))}; }

}  // namespace gd
