// Kernel instances for 6D meshes: (nodes per dim, quadrature points per dim),
// including GenDiL's 6D range benchmarks on the device: advection with q = n
// (benchmarks/range-benchmark-advection-6d.cpp:124-147, p = 0..2) and mass with
// q = n + 1 (range-benchmark-mass-6d.cpp:119-141, p = 0..2).
#include "gd_dispatch.hpp"

namespace gd {

static const KernelSet kSets[] = {
    GD_KSET(6, 1, 1),
    GD_KSET(6, 1, 2),
    GD_KSET(6, 2, 2),
    GD_KSET(6, 2, 3),
    GD_KSET(6, 3, 3),
    GD_KSET(6, 3, 4),
    GD_KSET(6, 4, 4)};

KernelList kernels_d6() { return KernelList{kSets, (int)(sizeof(kSets) / sizeof(kSets[0]))}; }

}  // namespace gd
